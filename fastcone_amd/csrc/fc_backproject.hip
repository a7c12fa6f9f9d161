// Voxel-driven backprojectors for gfx950: FDK (weight U^2), Katsevich (U^1) and unweighted
// general geometry (mode 0), coneBeamGPU.cu:178-420.
//
//   vol(x) += sum_t norm_t * w_t(x) * B_t(a/c, b/c),   (a, b, c) = M_t (x, 1)
//
// M_t is the view's FP64 3x4 projective matrix (built on the host from the source / detector
// vectors). For the reference's circular/helical geometry a/c and 1/c reproduce
// w = (y cos - x sin) U / pw and U = D / (D + x cos + y sin) of coneBeamGPU.cu:398-399, and the
// mode-0 ray/plane intersection of coneBeamGPU.cu:235-243 in closed form. B_t is the bilinear
// interpolant of window t extended by zeros; the per-view norm multiplies the data at staging,
// as the reference's fillCache does (coneBeamGPU.cu:1081).
//
// Work decomposition: one thread owns a column (x, y) and ZR = 16 consecutive z voxels, held in
// registers over every view ("wavefront-level accumulation over views"); a block owns a
// 16 x 16 x 16 voxel cube. Per view the block stages, in LDS, the detector window that the cube
// projects onto (rows of the raw projection, zero outside the detector), double-buffered: while
// the block accumulates view t from one buffer, the rows of view t+1 are already in flight (FP32
// storage: copied straight into the other buffer; 16-bit storage: into registers, written to the
// other buffer at the top of the next iteration), so one barrier per view suffices and the load
// latency hides behind the arithmetic. Per view and column the
// projective terms are evaluated once in FP64; when the detector columns do not depend on z
// (untilted detectors) each voxel then costs one FMA for its row, a clamp and a bilinear
// interpolation from LDS (two ds_read2 pairs).
//
// The windows come from a pre-pass (bp_windows, one thread per block and view): the projective
// image of a box lies in the box of its eight corners' images, so the corners' FP64 (u, v) widened
// by a pixel each side bound every tap. A view whose window does not fit the LDS tile is gathered
// from global memory with explicit bounds (rare: very short source distances); a view that misses
// the detector entirely is skipped.
//
// "Quad" staging (stage_quads: the four bilinear corners of every base pixel stored together) is
// kept for the Katsevich backprojector's filtered data.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

constexpr int kZR = 16;   // z voxels per thread
constexpr int kBXY = 16;  // 16x16 columns per 256-thread block (8x8 per wave)

template <int DT_IN>
__global__ __launch_bounds__(256) void stage_quads(const void* __restrict__ proj, int nw, int nh,
                                                   const float* __restrict__ norm,
                                                   void* __restrict__ quad, int QW, int QH) {
  int c = blockIdx.x * 256 + threadIdx.x;
  int r = blockIdx.y, t = blockIdx.z;
  if (c >= QW) return;
  int w0 = c - 1, h0 = r - 1;
  float s = norm ? norm[t] : 1.0f;
  const size_t vbase = (size_t)t * nw * nh;
  float v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int w = w0 + (q & 1), h = h0 + (q >> 1);
    bool in = w >= 0 && w < nw && h >= 0 && h < nh;
    v[q] = in ? s * load_elem<DT_IN>(proj, vbase + (size_t)h * nw + w) : 0.0f;
  }
  size_t o = ((size_t)t * QH + r) * QW + c;
  if constexpr (DT_IN == FC_F32) {
    static_cast<float4*>(quad)[o] = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    static_cast<uint2*>(quad)[o] = make_uint2(pack2<DT_IN>(v[0], v[1]), pack2<DT_IN>(v[2], v[3]));
  }
}

constexpr int kTileW = 64;                 // widest staged window (one pixel per lane)
constexpr int kPre = 12;                   // rows per wave and view: window height <= 4 kPre
// Row pitch of a staged window (floats): fixed, so a bilinear sample's two row pairs are one
// base address with immediate offsets {0, 1, kTileP, kTileP + 1}. 76 = 12 mod 32 banks: timed at
// 512^3 x 720 against pitches 65-80, it leaves the fewest bank conflicts between the lanes' rows
// (65: 51.5 ms, 68: 49.9, 70: 48.6, 72: 48.2, 74-78: 47.7, 80: 48.0).
constexpr uint32_t kTileP = 76;
constexpr int kTileCap = kTileP * 4 * kPre;  // floats per LDS buffer (14.3 KiB)
constexpr uint32_t kWinSkip = 0u;          // window word y: the cube misses the detector
constexpr uint32_t kWinGlobal = 0xffffu;   // window word y low half: gather from global memory

struct BpArgs {
  float* vol;
  const void* proj;    // [nviews][nh][nw] in the storage dtype
  const float* norm;   // per-view scale or null
  const double* views;
  const uint2* win;    // per (block, view): x = u0 | v0 << 16 (int16 each), y = Wu | Wv << 16
  double ox, oy, oz;
  int nx, ny, nz;
  int nw, nh, nviews;
  int nbx, nby;
};

struct WinArgs {
  const double* views;
  uint2* win;
  double ox, oy, oz;
  int nx, ny, nz;
  int nw, nh, nviews;
  int nbx, nby, nbz;
};

// One thread per (block, view): the detector window of the block's voxel cube. Window pixels
// [u0, u0 + Wu) x [v0, v0 + Wv) with u0 = floor(min u) - 1 and Wu = floor(max u) - u0 + 3, so every
// tap (iu, iu + 1) of a voxel in the cube lies at least one pixel inside; likewise v.
__global__ __launch_bounds__(256) void bp_windows(WinArgs a) {
  const size_t total = (size_t)a.nbx * a.nby * a.nbz * a.nviews;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  const int t = (int)(idx % (size_t)a.nviews);
  const size_t b = idx / (size_t)a.nviews;
  const int bx = (int)(b % a.nbx), by = (int)((b / a.nbx) % a.nby), bz = (int)(b / ((size_t)a.nbx * a.nby));
  const double xs[2] = {bx * kBXY + a.ox, min(bx * kBXY + kBXY - 1, a.nx - 1) + a.ox};
  const double ys[2] = {by * kBXY + a.oy, min(by * kBXY + kBXY - 1, a.ny - 1) + a.oy};
  const double zs[2] = {bz * kZR + a.oz, min(bz * kZR + kZR - 1, a.nz - 1) + a.oz};
  const double* M = a.views + (size_t)t * FC_VIEW_DOUBLES + 12;
  double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double X = xs[c & 1], Y = ys[(c >> 1) & 1], Z = zs[c >> 2];
    const double C = M[8] * X + M[9] * Y + M[10] * Z + M[11];
    ok = ok && C > 0.0;
    const double u = (M[0] * X + M[1] * Y + M[2] * Z + M[3]) / C;
    const double v = (M[4] * X + M[5] * Y + M[6] * Z + M[7]) / C;
    umin = fmin(umin, u); umax = fmax(umax, u);
    vmin = fmin(vmin, v); vmax = fmax(vmax, v);
  }
  uint2 w;
  w.x = 0u;
  w.y = kWinGlobal;
  if (ok && umin > -1e6 && umax < 1e6 && vmin > -1e6 && vmax < 1e6) {
    if (umax <= -1.0 || umin >= (double)a.nw || vmax <= -1.0 || vmin >= (double)a.nh) {
      w.y = kWinSkip;  // zero-extended data: every sample of the cube is 0
    } else {
      const int u0 = (int)floor(umin) - 1, v0 = (int)floor(vmin) - 1;
      const int Wu = (int)floor(umax) - u0 + 3, Wv = (int)floor(vmax) - v0 + 3;
      if (Wu <= kTileW && Wv <= 4 * kPre && u0 >= -32768 && u0 <= 32767 && v0 >= -32768 &&
          v0 <= 32767) {
        w.x = ((uint32_t)u0 & 0xffffu) | ((uint32_t)v0 << 16);
        w.y = (uint32_t)Wu | ((uint32_t)Wv << 16);
      }
    }
  }
  a.win[idx] = w;
}

// Zero-extended bilinear sample of one view straight from global memory (window fallback).
template <int DT>
__device__ __forceinline__ float bilinear_global(const void* proj, size_t vbase, int nw, int nh,
                                                 float u, float v) {
  if (!(u > -1.0f && u < (float)nw && v > -1.0f && v < (float)nh)) return 0.0f;
  const float fu = floorf(u), fv = floorf(v);
  const int iu = (int)fu, iv = (int)fv;
  const float au = u - fu, av = v - fv;
  float p[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int cu = iu + (q & 1), cv = iv + (q >> 1);
    const bool in = cu >= 0 && cu < nw && cv >= 0 && cv < nh;
    p[q] = in ? load_elem<DT>(proj, vbase + (size_t)cv * nw + cu) : 0.0f;
  }
  return lerp(lerp(p[0], p[1], au), lerp(p[2], p[3], au), av);
}

// Clamp to [lo, hi]; NaN maps to lo (fmaxf returns the non-NaN operand).
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
  return fminf(fmaxf(v, lo), hi);
}

// Bilinear sample at tile coordinates (u, v) >= 0 of an LDS window (row pitch kTileP).
__device__ __forceinline__ float tile_bilinear(const float* T, uint32_t iu, float au, float v) {
  const uint32_t iv = (uint32_t)(int)v;
  const float av = __builtin_amdgcn_fractf(v);
  const float* p = T + mad_u24(iv, kTileP, iu);
  return lerp(lerp(p[0], p[1], au), lerp(p[kTileP], p[kTileP + 1], au), av);
}

// One dword per active lane from global memory at row + lane_bytes (row wave-uniform, in scalar
// registers: the scalar-base form of the instruction, so a row costs no vector arithmetic) into
// LDS at lds_dst + 4 lane (lds_dst wave-uniform), by the load-to-LDS path: no registers hold the
// data and no LDS write instruction is issued. Written as assembly rather than the compiler
// builtin, with which the compiler waits for every outstanding load-to-LDS before each later LDS
// read (it cannot tell that the reads address the other window buffer). The instruction has no
// register outputs, its address registers are read at issue, and the LDS it writes is read only
// after the explicit s_waitcnt and the barrier of the next view. Loads complete in order, so the
// compiler's own vmcnt waits stay conservative with these loads outstanding.
// M0 is declared clobbered; the compiler reports that it reserves M0, and it uses M0 nowhere
// else in this kernel.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void copy_dword_to_lds(const float* row, uint32_t lane_bytes,
                                                  float* lds_dst) {
  const unsigned m0 = __builtin_amdgcn_readfirstlane((unsigned)(size_t)lds_dst);
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, %2" ::"s"(m0), "v"(lane_bytes),
               "s"(row)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// The per-view words every lane of a block needs (the view's 3 x 4 matrix, the next view's
// window word and the view's weight), by scalar loads waited for in the same statement. With
// the copies above in the loop the compiler cannot prove these words unchanged and reads them
// with vector loads; vector loads count down in issue order, so waiting for one issued after
// the next view's copies would also wait for every copy, and the view would not overlap them.
typedef unsigned int fc_u32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int fc_u32x16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ void load_view_words(const double* M, const uint2* Wn, const float* S,
                                                double m[12], uint2& wn, float& s) {
  fc_u32x16 lo;
  fc_u32x8 hi;
  unsigned long long w;
  unsigned int sv;
  asm volatile(
      "s_load_dwordx16 %0, %4, 0x0\n\t"
      "s_load_dwordx8 %1, %4, 0x40\n\t"
      "s_load_dwordx2 %2, %5, 0x0\n\t"
      "s_load_dword %3, %6, 0x0\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&s"(lo), "=&s"(hi), "=&s"(w), "=&s"(sv)
      : "s"(M), "s"(Wn), "s"(S)
      : "memory");
#pragma unroll
  for (int i = 0; i < 8; ++i) m[i] = __hiloint2double((int)lo[2 * i + 1], (int)lo[2 * i]);
#pragma unroll
  for (int i = 0; i < 4; ++i) m[8 + i] = __hiloint2double((int)hi[2 * i + 1], (int)hi[2 * i]);
  wn = make_uint2((unsigned)w, (unsigned)(w >> 32));
  s = __uint_as_float(sv);
}

template <int DT, int MODE>
__global__ __launch_bounds__(256) void bp_kernel(BpArgs a) {
  __shared__ float tile[2][kTileCap];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * kBXY + (wave & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * kBXY + (wave >> 1) * 8 + (lane >> 3);
  const int z0 = blockIdx.z * kZR;
  const double X = (double)x + a.ox, Y = (double)y + a.oy, Z0 = (double)z0 + a.oz;
  // the whole 16 x 16 x 16 cube lies in the volume (no clamping of tile coordinates needed)
  const bool inner = (int)(blockIdx.x + 1) * kBXY <= a.nx &&
                     (int)(blockIdx.y + 1) * kBXY <= a.ny && z0 + kZR <= a.nz;
  const uint2* W = a.win + ((size_t)(blockIdx.z * a.nby + blockIdx.y) * a.nbx + blockIdx.x) *
                               (size_t)a.nviews;
  float acc[kZR];
#pragma unroll
  for (int k = 0; k < kZR; ++k) acc[k] = 0.0f;

  // Rows of a view's window, staged as (row wave + 4 q, column lane); zero outside the detector.
  // 16-bit storage: registers hold the raw storage bits (converting at commit keeps the loads in
  // flight); FP32 storage copies straight into LDS (issue_lds below). The row phase is
  // wave-uniform, so row checks are scalar branches and addresses advance by a scalar stride.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  uint32_t pre[kPre];
  auto issue = [&](int t, uint2 w) {
    const uint32_t Wu = w.y & 0xffffu, Wv = w.y >> 16;
    if (Wu == kWinSkip || Wu == kWinGlobal) return;
    const int u0 = (int)(int16_t)(w.x & 0xffffu), v0 = (int)(int16_t)(w.x >> 16);
    const int pu = u0 + lane;
    const bool colok = (uint32_t)lane < Wu && pu >= 0 && pu < a.nw;
    const int pv0 = v0 + wv;
    const long long row0 = ((long long)t * a.nh + pv0) * a.nw;  // element index of (pv0, 0)
    // row wv + 4 q is in the window when 4 q < rows: compared with an immediate, so no register
    // holds a per-row constant across the view loop
    const int rows = (int)Wv - wv;
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
      pre[q] = 0u;  // +0 in every storage format
      if (4 * q < rows && (uint32_t)(pv0 + 4 * q) < (uint32_t)a.nh && colok)
        pre[q] = load_bits<DT>(a.proj, (size_t)(row0 + (long long)(4 * q) * a.nw + pu));
    }
  };
  auto commit = [&](float* T, uint2 w) {
    const uint32_t Wu = w.y & 0xffffu, Wv = w.y >> 16;
    if (Wu == kWinSkip || Wu == kWinGlobal) return;
    float* Tw = T + (uint32_t)wv * kTileP + (uint32_t)lane;
    const int rows = (int)Wv - wv;
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
      if (4 * q < rows && (uint32_t)lane < Wu)
        Tw[4u * kTileP * (uint32_t)q] = bits_to_f32<DT>(pre[q]);
    }
  };

  // FP32 storage: the window rows go straight from global memory into the LDS buffer (rows of
  // lanes < Wu; pixels outside the detector are stored as zero), with no staging registers.
  auto issue_lds = [&](int t, uint2 w, float* Tn) {
    const uint32_t Wu = w.y & 0xffffu, Wv = w.y >> 16;
    if (Wu == kWinSkip || Wu == kWinGlobal) return;
    const int u0 = (int)(int16_t)(w.x & 0xffffu), v0 = (int)(int16_t)(w.x >> 16);
    const int pu = u0 + lane;
    const bool colok = pu >= 0 && pu < a.nw;
    const int pv0 = v0 + wv;
    // the window row's first pixel (wave-uniform; it may lie before the detector row, but only
    // lanes with a pixel on the detector load) and each lane's byte offset from it
    const float* row0 = static_cast<const float*>(a.proj) + ((long long)t * a.nh + pv0) * a.nw + u0;
    const uint32_t lane_bytes = 4u * (uint32_t)lane;
    // row wv + 4 q is in the window when 4 q < rows: compared with an immediate, so no register
    // holds a per-row constant across the view loop
    const int rows = (int)Wv - wv;
    float* Tw = Tn + (uint32_t)wv * kTileP;
    if ((uint32_t)lane < Wu) {
#pragma unroll
      for (int q = 0; q < kPre; ++q) {
        if (4 * q < rows) {
          // pixels outside the detector are written as zero by the lanes themselves
          float* dst = Tw + 4u * kTileP * (uint32_t)q;
          if (colok && (uint32_t)(pv0 + 4 * q) < (uint32_t)a.nh)
            copy_dword_to_lds(row0 + (long long)(4 * q) * a.nw, lane_bytes, dst);
          else
            dst[lane] = 0.0f;
        }
      }
    }
  };

  uint2 wn = make_uint2(0u, kWinSkip);
  if (a.nviews > 0) {
    wn = W[0];
    if constexpr (DT == FC_F32) issue_lds(0, wn, tile[0]);
    else issue(0, wn);
  }
  for (int t = 0; t < a.nviews; ++t) {
    const uint2 w = wn;
    float* T = tile[t & 1];
    const bool more = t + 1 < a.nviews;
    const uint32_t Wu = w.y & 0xffffu, Wv = w.y >> 16;
    const double* Mp = a.views + (size_t)t * FC_VIEW_DOUBLES + 12;
    double M[12];
    float s;
    if constexpr (DT == FC_F32) {
      // the view's words first (they do not depend on the barrier: a wave that finishes view
      // t - 1 early waits for them while the others still compute); without norms the weight's
      // load reads the matrix's first word, and the weight is 1
      load_view_words(Mp, W + (more ? t + 1 : t),
                      a.norm ? a.norm + t : reinterpret_cast<const float*>(Mp), M, wn, s);
      if (!a.norm) s = 1.0f;
      // this wave's copies of view t have landed; the barrier then covers every wave's
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
      // the other buffer's readers (view t-1) all passed the previous barrier before this write
      commit(T, w);
    }
    __syncthreads();
    // the next view's rows are in flight while view t is accumulated; the other buffer's readers
    // (view t-1) all passed the barrier above
    if constexpr (DT == FC_F32) {
      if (more) issue_lds(t + 1, wn, tile[(t + 1) & 1]);
    } else {
      if (more) {
        wn = W[t + 1];
        issue(t + 1, wn);
      }
#pragma unroll
      for (int i = 0; i < 12; ++i) M[i] = Mp[i];
      s = a.norm ? a.norm[t] : 1.0f;
    }
    if (Wu == kWinSkip) continue;  // uniform: the cube misses the detector in this view
    const double A = M[0] * X + M[1] * Y + M[2] * Z0 + M[3];
    const double B = M[4] * X + M[5] * Y + M[6] * Z0 + M[7];
    const double C = M[8] * X + M[9] * Y + M[10] * Z0 + M[11];
    if (Wu == kWinGlobal) {
      const size_t vbase = (size_t)t * a.nh * a.nw;
      const float af = (float)A, bf = (float)B, cf = (float)C;
      const float da = (float)M[2], db = (float)M[6], dc = (float)M[10];
#pragma unroll
      for (int k = 0; k < kZR; ++k) {
        const float kf = (float)k;
        const float rc = 1.0f / fmaf(kf, dc, cf);
        float wt = s;
        if constexpr (MODE == FC_BP_FDK) wt = s * rc * rc;
        if constexpr (MODE == FC_BP_KATSEVICH) wt = s * rc;
        acc[k] = fmaf(wt, bilinear_global<DT>(a.proj, vbase, a.nw, a.nh, fmaf(kf, da, af) * rc,
                                              fmaf(kf, db, bf) * rc), acc[k]);
      }
      continue;
    }
    const float u0f = (float)(int16_t)(w.x & 0xffffu), v0f = (float)(int16_t)(w.x >> 16);
    // Tile coordinates. Valid voxels lie >= 1 pixel inside the window; the clamps only keep
    // voxels outside the volume (partial blocks) and NaN inside the buffer.
    const float umax = (float)Wu - 2.0f, vmax = (float)Wv - 2.0f;
    if (M[2] == 0.0 && M[10] == 0.0) {  // uniform per view
      // Columns independent of z: u and the weight are per column, the row is linear in z.
      // FP32 reciprocal refined by one FP64 Newton step: relative error ~2^-44
      double rc = (double)__builtin_amdgcn_rcpf((float)C);
      rc = rc * fma(-C, rc, 2.0);
      const float u = clampf((float)(A * rc) - u0f, 0.0f, umax);
      const float vb = (float)(B * rc) - v0f;
      const float dv = (float)(M[6] * rc);
      float wt = s;
      if constexpr (MODE == FC_BP_FDK) wt = s * (float)(rc * rc);
      if constexpr (MODE == FC_BP_KATSEVICH) wt = s * (float)rc;
      const uint32_t iu = (uint32_t)(int)u;
      const float au = __builtin_amdgcn_fractf(u);
      if (inner) {
        // every voxel of the cube is in the volume: its taps lie >= 1 pixel inside the window
#pragma unroll
        for (int k = 0; k < kZR; ++k)
          acc[k] = fmaf(wt, tile_bilinear(T, iu, au, fmaf((float)k, dv, vb)), acc[k]);
      } else {
#pragma unroll
        for (int k = 0; k < kZR; ++k) {
          const float v = clampf(fmaf((float)k, dv, vb), 0.0f, vmax);
          acc[k] = fmaf(wt, tile_bilinear(T, iu, au, v), acc[k]);
        }
      }
    } else {
      const float af = (float)A, bf = (float)B, cf = (float)C;
      const float da = (float)M[2], db = (float)M[6], dc = (float)M[10];
#pragma unroll
      for (int k = 0; k < kZR; ++k) {
        const float kf = (float)k;
        const float rc = __builtin_amdgcn_rcpf(fmaf(kf, dc, cf));  // 1 ulp, as the positions
        const float u = clampf(fmaf(kf, da, af) * rc - u0f, 0.0f, umax);
        const float v = clampf(fmaf(kf, db, bf) * rc - v0f, 0.0f, vmax);
        float wt = s;
        if constexpr (MODE == FC_BP_FDK) wt = s * rc * rc;
        if constexpr (MODE == FC_BP_KATSEVICH) wt = s * rc;
        acc[k] = fmaf(wt, tile_bilinear(T, (uint32_t)(int)u, __builtin_amdgcn_fractf(u), v),
                      acc[k]);
      }
    }
  }
  if (x < a.nx && y < a.ny) {
#pragma unroll
    for (int k = 0; k < kZR; ++k) {
      int z = z0 + k;
      if (z < a.nz) a.vol[((size_t)z * a.ny + y) * a.nx + x] += acc[k];
    }
  }
}

// Gather-roofline microbenchmark for the backprojector: bilinear samples from one staged 64 x 48
// LDS window per block at pseudo-random positions (same tile_bilinear as bp_kernel).
__global__ __launch_bounds__(256) void gather_bench_bp(const float* __restrict__ src, int pitch,
                                                       int nsamp, float* __restrict__ sink) {
  __shared__ float T[kTileCap];
  const float* g = src + (size_t)(blockIdx.x % 64) * 48 * pitch;
  for (int e = threadIdx.x; e < kTileCap; e += 256)
    T[e] = g[(e / kTileP) * pitch + min(e % kTileP, (uint32_t)kTileW - 1u)];
  __syncthreads();
  uint32_t s = 0x9e3779b9u * (blockIdx.x * 256 + threadIdx.x + 1);
  const float su = 62.0f / 16777216.0f, sv = 46.0f / 16777216.0f;
  float acc = 0.0f;
  for (int k = 0; k < nsamp; ++k) {
    s = s * 1664525u + 1013904223u;
    const float u = (float)(s >> 8) * su;
    s = s * 1664525u + 1013904223u;
    const float v = (float)(s >> 8) * sv;
    acc = fmaf(1.0001f, tile_bilinear(T, (uint32_t)(int)u, __builtin_amdgcn_fractf(u), v), acc);
  }
  if (acc == 12345.678f) sink[0] = acc;
}

// Ray-coherent ceiling for the backprojector: bp_kernel's fast-path inner loop for blocks inside
// the volume (per column one u and weight per view, 16 z voxels whose rows advance by dv:
// bilinear sample from the LDS window, weighted accumulate) over one resident 64 x 48 window per
// block, with the backprojector's lane layout and view-to-view drift. No per-view projective
// set-up, no window staging and no barriers, so it bounds the voxel-view rate of the LDS gather
// itself. Bounds: u <= 2 + 10.5 + 4.5 + 0.35 < 62, v <= 2 + 3 + 1.5 + 0.35 + 15 * 1.3 < 46.
__global__ __launch_bounds__(256) void gather_bench_bp_ray(const float* __restrict__ src,
                                                           int pitch, int nviews,
                                                           float* __restrict__ sink) {
  __shared__ float T[kTileCap];
  const float* g = src + (size_t)(blockIdx.x % 64) * 48 * pitch;
  for (int e = threadIdx.x; e < kTileCap; e += 256)
    T[e] = g[(e / kTileP) * pitch + min(e % kTileP, (uint32_t)kTileW - 1u)];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float xl = (float)((wave & 1) * 8 + (lane & 7)), yl = (float)((wave >> 1) * 8 + (lane >> 3));
  float acc[kZR];
#pragma unroll
  for (int k = 0; k < kZR; ++k) acc[k] = 0.0f;
#pragma unroll 1
  for (int t = 0; t < nviews; ++t) {
    const float tt = (float)(t & 7);
    const float u = clampf(2.0f + 0.7f * xl + 0.3f * yl + 0.05f * tt, 0.0f, 62.0f);
    const float vb = 2.0f + 0.2f * xl + 0.1f * yl + 0.05f * tt;
    const float wt = 1.0f + 0.001f * tt;
    const uint32_t iu = (uint32_t)(int)u;
    const float au = __builtin_amdgcn_fractf(u);
#pragma unroll
    for (int k = 0; k < kZR; ++k) {
      acc[k] = fmaf(wt, tile_bilinear(T, iu, au, fmaf((float)k, 1.3f, vb)), acc[k]);
    }
  }
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < kZR; ++k) sum += acc[k];
  if (sum == 12345.678f) sink[0] = sum;
}

template <int DT>
int dispatch_bp(const BpArgs& a, int mode, dim3 grid, hipStream_t s) {
  if (mode == FC_BP_UNWEIGHTED) bp_kernel<DT, FC_BP_UNWEIGHTED><<<grid, 256, 0, s>>>(a);
  else if (mode == FC_BP_KATSEVICH) bp_kernel<DT, FC_BP_KATSEVICH><<<grid, 256, 0, s>>>(a);
  else bp_kernel<DT, FC_BP_FDK><<<grid, 256, 0, s>>>(a);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

}  // namespace

int launch_backproject(int device, float* vol, const int vs[3], const double origin[3],
                       const void* proj, int dt, const int ps[3], const double* views,
                       const float* norm, int mode, hipStream_t stream) {
  const int nw = ps[0], nh = ps[1], nv = ps[2];
  if (nv == 0 || vs[0] == 0 || vs[1] == 0 || vs[2] == 0) return FC_OK;
  const int nbx = (vs[0] + kBXY - 1) / kBXY, nby = (vs[1] + kBXY - 1) / kBXY,
            nbz = (vs[2] + kZR - 1) / kZR;
  FC_CHECK_ARG(nby <= 65535 && nbz <= 65535, "volume too large for the grid");
  const size_t nwin = (size_t)nbx * nby * nbz * nv;
  void* ws = nullptr;
  int rc = acquire_ws(device, SLOT_QUAD, nwin * sizeof(uint2), stream, &ws);
  if (rc) return rc;
  WinArgs w;
  w.views = views;
  w.win = static_cast<uint2*>(ws);
  w.ox = origin[0]; w.oy = origin[1]; w.oz = origin[2];
  w.nx = vs[0]; w.ny = vs[1]; w.nz = vs[2];
  w.nw = nw; w.nh = nh; w.nviews = nv;
  w.nbx = nbx; w.nby = nby; w.nbz = nbz;
  BpArgs a;
  a.vol = vol;
  a.proj = proj;
  a.norm = norm;
  a.views = views;
  a.win = w.win;
  a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
  a.nx = vs[0]; a.ny = vs[1]; a.nz = vs[2];
  a.nw = nw; a.nh = nh; a.nviews = nv;
  a.nbx = nbx; a.nby = nby;
  const dim3 grid(nbx, nby, nbz);
  int tok = prof_begin("backproject", stream);
  bp_windows<<<(unsigned)((nwin + 255) / 256), 256, 0, stream>>>(w);
  int r2;
  if (dt == FC_F32) r2 = dispatch_bp<FC_F32>(a, mode, grid, stream);
  else if (dt == FC_BF16) r2 = dispatch_bp<FC_BF16>(a, mode, grid, stream);
  else r2 = dispatch_bp<FC_F16>(a, mode, grid, stream);
  prof_end(tok, stream);
  if (r2) return r2;
  return release_ws(device, SLOT_QUAD, stream);
}

int launch_stage_quads(const void* proj, int dt, int nw, int nh, int nv, const float* norm,
                       void* quad, hipStream_t stream) {
  const int QW = nw + 2, QH = nh + 2;
  FC_CHECK_ARG(QH <= 65535 && nv <= 65535, "projection stack too large for the staging grid");
  if (nv == 0) return FC_OK;
  dim3 sg((QW + 255) / 256, QH, nv);
  if (dt == FC_F32) stage_quads<FC_F32><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, quad, QW, QH);
  else if (dt == FC_BF16) stage_quads<FC_BF16><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, quad, QW, QH);
  else stage_quads<FC_F16><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, quad, QW, QH);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

int launch_gather_bench_bp(int device, int dt, int ray, double* samples_per_s) {
  (void)device;
  (void)dt;  // LDS windows hold FP32 whatever the storage type; only the staging converts
  const int pitch = 1024, rows = 64 * 48;
  std::vector<float> host((size_t)pitch * rows);
  for (size_t q = 0; q < host.size(); ++q) host[q] = (float)((q * 2654435761u) % 1000) * 1e-3f;
  float* src = nullptr;
  float* sink = nullptr;
  FC_HIP(hipMalloc(&src, host.size() * sizeof(float)));
  FC_HIP(hipMalloc(&sink, sizeof(float)));
  FC_HIP(hipMemcpy(src, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  FC_HIP(hipEventCreate(&e0));
  FC_HIP(hipEventCreate(&e1));
  const int blocks = 256 * 16, nsamp = 4096;
  double best = 0.0;
  for (int rep = 0; rep < 3; ++rep) {
    FC_HIP(hipEventRecord(e0, nullptr));
    if (ray) gather_bench_bp_ray<<<blocks, 256>>>(src, pitch, nsamp / kZR, sink);
    else gather_bench_bp<<<blocks, 256>>>(src, pitch, nsamp, sink);
    FC_HIP(hipEventRecord(e1, nullptr));
    FC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FC_HIP(hipEventElapsedTime(&ms, e0, e1));
    double rate = (double)blocks * 256.0 * nsamp / (ms * 1e-3);
    if (rep > 0 && rate > best) best = rate;
  }
  FC_HIP(hipEventDestroy(e0));
  FC_HIP(hipEventDestroy(e1));
  FC_HIP(hipFree(src));
  FC_HIP(hipFree(sink));
  *samples_per_s = best;
  return FC_OK;
}

}  // namespace fc
