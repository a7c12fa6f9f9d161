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
// Work decomposition: one thread owns a column (x, y) and ZR consecutive z voxels, held in
// registers over every view ("wavefront-level accumulation over views"): per view and column the
// projective terms are evaluated once in FP64; when the detector columns do not depend on z (all
// untilted detectors) each voxel then costs one FMA for its row, a clamp and one bilinear gather.
//
// Projection staging ("quad" layout): for every base pixel (r, c) in [-1, nh] x [-1, nw] the four
// bilinear corners p[r..r+1][c..c+1] * norm_t are stored together (16 B FP32, 8 B BF16/FP16), so a
// gather is one load; rows and columns are clamped to the padded range, which keeps every load in
// bounds and returns exact zeros outside the window.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

constexpr int kZR = 16;   // z voxels per thread
constexpr int kBXY = 16;  // 16x16 columns per 256-thread block (8x8 per wave)

template <int DT>
struct Quad;
template <>
struct Quad<FC_F32> {
  using T = float4;
};
template <>
struct Quad<FC_BF16> {
  using T = uint2;
};
template <>
struct Quad<FC_F16> {
  using T = uint2;
};

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

struct BpArgs {
  float* vol;
  const void* quad;
  const double* views;
  double ox, oy, oz;
  int nx, ny, nz;
  int nw, nh, nviews;
  int QW, QH;
};

template <int DT>
__device__ __forceinline__ float bilinear_quad(const void* base, uint32_t idx, float au,
                                               float av) {
  float p00, p01, p10, p11;
  if constexpr (DT == FC_F32) {
    float4 q = static_cast<const float4*>(base)[idx];
    p00 = q.x; p01 = q.y; p10 = q.z; p11 = q.w;
  } else {
    uint2 q = static_cast<const uint2*>(base)[idx];
    p00 = lo16<DT>(q.x); p01 = hi16<DT>(q.x); p10 = lo16<DT>(q.y); p11 = hi16<DT>(q.y);
  }
  float top = lerp(p00, p01, au), bot = lerp(p10, p11, au);
  return lerp(top, bot, av);
}

// Clamp to [lo, hi]; NaN maps to lo (fmaxf returns the non-NaN operand).
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
  return fminf(fmaxf(v, lo), hi);
}

template <int DT, int MODE>
__global__ __launch_bounds__(256) void bp_kernel(BpArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * kBXY + (wave & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * kBXY + (wave >> 1) * 8 + (lane >> 3);
  const int z0 = blockIdx.z * kZR;
  const double X = (double)x + a.ox, Y = (double)y + a.oy, Z0 = (double)z0 + a.oz;
  const float wmax = (float)a.nw, hmax = (float)a.nh;
  const size_t vstride = (size_t)a.QW * a.QH;
  float acc[kZR];
#pragma unroll
  for (int k = 0; k < kZR; ++k) acc[k] = 0.0f;

  for (int t = 0; t < a.nviews; ++t) {
    const double* M = a.views + (size_t)t * FC_VIEW_DOUBLES + 12;
    const double A = M[0] * X + M[1] * Y + M[2] * Z0 + M[3];
    const double B = M[4] * X + M[5] * Y + M[6] * Z0 + M[7];
    const double C = M[8] * X + M[9] * Y + M[10] * Z0 + M[11];
    const void* base;
    if constexpr (DT == FC_F32) {
      base = static_cast<const float4*>(a.quad) + (size_t)t * vstride;
    } else {
      base = static_cast<const uint2*>(a.quad) + (size_t)t * vstride;
    }
    if (M[2] == 0.0 && M[10] == 0.0) {  // uniform per view
      // Columns independent of z: u and the weight are per column, the row is linear in z.
      const double rc = 1.0 / C;
      float u = clampf((float)(A * rc), -1.0f, wmax);
      const float v0 = (float)(B * rc);
      const float dv = (float)(M[6] * rc);
      float w = 1.0f;
      if constexpr (MODE == FC_BP_FDK) w = (float)(rc * rc);
      if constexpr (MODE == FC_BP_KATSEVICH) w = (float)rc;
      const float fu = floorf(u);
      const float au = u - fu;
      const uint32_t iu = (uint32_t)((int)fu + 1);
#pragma unroll
      for (int k = 0; k < kZR; ++k) {
        float v = clampf(fmaf((float)k, dv, v0), -1.0f, hmax);
        float fv = floorf(v);
        float av = v - fv;
        uint32_t idx = (uint32_t)((int)fv + 1) * (uint32_t)a.QW + iu;
        acc[k] = fmaf(w, bilinear_quad<DT>(base, idx, au, av), acc[k]);
      }
    } else {
      const float af = (float)A, bf = (float)B, cf = (float)C;
      const float da = (float)M[2], db = (float)M[6], dc = (float)M[10];
#pragma unroll
      for (int k = 0; k < kZR; ++k) {
        float kf = (float)k;
        float c = fmaf(kf, dc, cf);
        float rc = 1.0f / c;
        float u = clampf(fmaf(kf, da, af) * rc, -1.0f, wmax);
        float v = clampf(fmaf(kf, db, bf) * rc, -1.0f, hmax);
        float fu = floorf(u), fv = floorf(v);
        float au = u - fu, av = v - fv;
        uint32_t idx = (uint32_t)((int)fv + 1) * (uint32_t)a.QW + (uint32_t)((int)fu + 1);
        float w = 1.0f;
        if constexpr (MODE == FC_BP_FDK) w = rc * rc;
        if constexpr (MODE == FC_BP_KATSEVICH) w = rc;
        acc[k] = fmaf(w, bilinear_quad<DT>(base, idx, au, av), acc[k]);
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

// Gather-roofline microbenchmark for the backprojector: bilinear quad gathers at pseudo-random
// positions inside one B x B pixel patch per block (same fetch + interpolation code).
template <int DT>
__global__ __launch_bounds__(256) void gather_bench_bp(const void* quad, int QW, int QH, int B,
                                                       int nsamp, float* __restrict__ sink) {
  uint32_t s = 0x9e3779b9u * (blockIdx.x * 256 + threadIdx.x + 1);
  int npx = (QW - B - 1) / B, npy = (QH - B - 1) / B;
  int b = blockIdx.x % (npx * npy);
  float bu = (float)((b % npx) * B), bv = (float)((b / npx) * B);
  const float scale = (float)(B - 1) * (1.0f / 16777216.0f);
  float acc = 0.0f;
  for (int k = 0; k < nsamp; ++k) {
    s = s * 1664525u + 1013904223u;
    float u = bu + (float)(s >> 8) * scale;
    s = s * 1664525u + 1013904223u;
    float v = bv + (float)(s >> 8) * scale;
    float fu = floorf(u), fv = floorf(v);
    uint32_t idx = (uint32_t)(int)fv * (uint32_t)QW + (uint32_t)(int)fu;
    acc = fmaf(1.0001f, bilinear_quad<DT>(quad, idx, u - fu, v - fv), acc);
  }
  if (acc == 12345.678f) sink[0] = acc;
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
  const int QW = nw + 2, QH = nh + 2;
  FC_CHECK_ARG(QH <= 65535 && nv <= 65535, "projection stack too large for the staging grid");
  FC_CHECK_ARG((size_t)QW * QH < (1ull << 31), "projection too large for 32-bit offsets");
  size_t esz = (dt == FC_F32) ? 16 : 8;
  size_t bytes = (size_t)QW * QH * nv * esz;
  void* ws = nullptr;
  int rc = acquire_ws(device, SLOT_QUAD, bytes, stream, &ws);
  if (rc) return rc;
  if (nv > 0) {
    dim3 sg((QW + 255) / 256, QH, nv);
    if (dt == FC_F32) stage_quads<FC_F32><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, ws, QW, QH);
    else if (dt == FC_BF16) stage_quads<FC_BF16><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, ws, QW, QH);
    else stage_quads<FC_F16><<<sg, 256, 0, stream>>>(proj, nw, nh, norm, ws, QW, QH);
    FC_HIP(hipGetLastError());
    BpArgs a;
    a.vol = vol;
    a.quad = ws;
    a.views = views;
    a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
    a.nx = vs[0]; a.ny = vs[1]; a.nz = vs[2];
    a.nw = nw; a.nh = nh; a.nviews = nv;
    a.QW = QW; a.QH = QH;
    dim3 grid((vs[0] + kBXY - 1) / kBXY, (vs[1] + kBXY - 1) / kBXY, (vs[2] + kZR - 1) / kZR);
    FC_CHECK_ARG(grid.y <= 65535 && grid.z <= 65535, "volume too large for the grid");
    int r2;
    int tok = prof_begin("backproject", stream);
    if (dt == FC_F32) r2 = dispatch_bp<FC_F32>(a, mode, grid, stream);
    else if (dt == FC_BF16) r2 = dispatch_bp<FC_BF16>(a, mode, grid, stream);
    else r2 = dispatch_bp<FC_F16>(a, mode, grid, stream);
    prof_end(tok, stream);
    if (r2) return r2;
  }
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

// Ray-coherent ceiling for the backprojector: its fast-path inner loop (per column: one u and
// weight per view, 16 z voxels whose rows advance by dv; clamp, bilinear quad gather, weighted
// accumulate) over one 64 x 64 pixel patch per block that stays cache resident, with the
// backprojector's lane layout and view-to-view drift. No per-view projective set-up and no fresh
// projection data per view, so it bounds the voxel-view rate of the gather itself.
// Bounds: u <= bu + 15.35, v <= bv + 4.85 + 15 * 1.3 = bv + 24.35, inside the 64 x 64 patch.
template <int DT>
__global__ __launch_bounds__(256) void gather_bench_bp_ray(const void* quad, int QW, int QH,
                                                           int nviews, float* __restrict__ sink) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float xl = (float)((wave & 1) * 8 + (lane & 7)), yl = (float)((wave >> 1) * 8 + (lane >> 3));
  const int npx = (QW - 64) / 64, npy = (QH - 64) / 64;
  const int b = blockIdx.x % (npx * npy);
  const float bu = (float)((b % npx) * 64 + 2), bv = (float)((b / npx) * 64 + 2);
  const float hmax = (float)(QH - 2);
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0f;
#pragma unroll 1
  for (int t = 0; t < nviews; ++t) {
    const float tt = (float)(t & 7);
    const float u = bu + 0.7f * xl + 0.3f * yl + 0.05f * tt;
    const float v0 = bv + 0.2f * xl + 0.1f * yl + 0.05f * tt;
    const float w = 1.0f + 0.001f * tt;
    const float fu = floorf(u);
    const float au = u - fu;
    const uint32_t iu = (uint32_t)(int)fu;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      float v = clampf(fmaf((float)k, 1.3f, v0), -1.0f, hmax);
      float fv = floorf(v);
      float av = v - fv;
      uint32_t idx = (uint32_t)(int)fv * (uint32_t)QW + iu;
      acc[k] = fmaf(w, bilinear_quad<DT>(quad, idx, au, av), acc[k]);
    }
  }
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k) sum += acc[k];
  if (sum == 12345.678f) sink[0] = sum;
}

int launch_gather_bench_bp(int device, int dt, int ray, double* samples_per_s) {
  (void)device;
  const int QW = 1024, QH = 1024;
  size_t esz = (dt == FC_F32) ? 16 : 8;
  void* quad = nullptr;
  float* sink = nullptr;
  FC_HIP(hipMalloc(&quad, (size_t)QW * QH * esz));
  FC_HIP(hipMalloc(&sink, sizeof(float)));
  FC_HIP(hipMemset(quad, 0x3c, (size_t)QW * QH * esz));
  hipEvent_t e0, e1;
  FC_HIP(hipEventCreate(&e0));
  FC_HIP(hipEventCreate(&e1));
  const int blocks = 256 * 16, nsamp = 4096;
  double best = 0.0;
  for (int B : {8, 16, 32}) {
    for (int rep = 0; rep < 2; ++rep) {
      FC_HIP(hipEventRecord(e0, nullptr));
      if (ray && dt == FC_F32) gather_bench_bp_ray<FC_F32><<<blocks, 256>>>(quad, QW, QH, nsamp / 16, sink);
      else if (ray && dt == FC_BF16) gather_bench_bp_ray<FC_BF16><<<blocks, 256>>>(quad, QW, QH, nsamp / 16, sink);
      else if (ray) gather_bench_bp_ray<FC_F16><<<blocks, 256>>>(quad, QW, QH, nsamp / 16, sink);
      else if (dt == FC_F32) gather_bench_bp<FC_F32><<<blocks, 256>>>(quad, QW, QH, B, nsamp, sink);
      else if (dt == FC_BF16) gather_bench_bp<FC_BF16><<<blocks, 256>>>(quad, QW, QH, B, nsamp, sink);
      else gather_bench_bp<FC_F16><<<blocks, 256>>>(quad, QW, QH, B, nsamp, sink);
      FC_HIP(hipEventRecord(e1, nullptr));
      FC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FC_HIP(hipEventElapsedTime(&ms, e0, e1));
      double rate = (double)blocks * 256.0 * nsamp / (ms * 1e-3);
      if (rep == 1 && rate > best) best = rate;
    }
  }
  FC_HIP(hipEventDestroy(e0));
  FC_HIP(hipEventDestroy(e1));
  FC_HIP(hipFree(quad));
  FC_HIP(hipFree(sink));
  *samples_per_s = best;
  return FC_OK;
}

}  // namespace fc
