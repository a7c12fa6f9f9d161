// 3x3x3 convolution kernels for the CGLS warm-start network (gfx950, MFMA).
//
// Layout: activations N x D x H x W x C (channels last), BF16; weights [Cout][27][Cin] BF16
// (tap = 9 dz + 3 dy + dx, zero padding 1); accumulation FP32.
//
//  * forward (and input gradient): implicit GEMM  Y[voxel][co] = sum_k P[voxel][k] Wt[k][co],
//    k = tap * Cin + ci, with v_mfma_f32_16x16x32_bf16: a wave owns 16 consecutive voxels of an
//    x row and all output channels; each lane feeds 8 consecutive k (one 16-byte load of 8
//    channels of one tap). Bias and ReLU are fused in the epilogue. The input gradient is the
//    same kernel on dY with the flipped, transposed weights  W'[ci][26 - tap][co] = W[co][tap][ci].
//  * weight gradient: dW[co][k] = sum_voxel dY[voxel][co] P[voxel][k] with
//    v_mfma_f32_16x16x32_bf16 fed the raw BF16 bits (products exact in FP32): a block per (group
//    of x rows, 64 k), one 16-column slice of K per wave; each row's output gradient is staged
//    once per block in LDS, transposed to channel-major in chunks of 256 voxels, so a lane's 8
//    voxels of one channel are one 16-byte LDS read; the input is gathered per tap. Each block
//    adds its partial dW to the FP32 master gradient with float atomics.
//  * 1-channel layers (first layer Cin = 1, last layer Cout = 1): direct FP32 VALU kernels
//    (forward, input gradient, weight gradient).
//
// Operand / result lane layouts of the MFMA form used (checked by fc_mfma_layout_probe, which also
// checks the 16x16x4 FP32 form):
//   16x16x32 bf16: A lane l holds A[l % 16][8 (l / 16) + 0..7], B lane l holds
//   B[8 (l / 16) + 0..7][l % 16]; 16x16x4 f32: A[l % 16][l / 16], B[l / 16][l % 16];
//   result D lane l, register r holds D[4 (l / 16) + r][l % 16].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---- forward / input-gradient implicit GEMM --------------------------------------------------
template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(256) void conv_fwd_mfma(const uint16_t* __restrict__ x,
                                                     const uint16_t* __restrict__ w,
                                                     const float* __restrict__ bias,
                                                     uint16_t* __restrict__ y, int D, int H,
                                                     int W) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts must be multiples of 16");
  constexpr int K = 27 * CIN;
  constexpr int KSTEPS = (K + 31) / 32;
  constexpr int NB = COUT / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int xb = blockIdx.x * 64 + wave * 16;
  const int yv = blockIdx.y;
  const int n = blockIdx.z / D, z = blockIdx.z - (blockIdx.z / D) * D;
  const int xv = xb + m;
  f32x4 acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int s = 0; s < KSTEPS; ++s) {
    const int k = s * 32 + g * 8;
    bf16x8 a = zero8;
    bf16x8 b[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) b[cb] = zero8;
    if (k < K) {
      const int tap = k / CIN, ci = k - (k / CIN) * CIN;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      const int zz = z + dz - 1, yy = yv + dy - 1, xx = xv + dx - 1;
      if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W && xv < W)
        a = *reinterpret_cast<const bf16x8*>(
            x + ((((size_t)n * D + zz) * H + yy) * W + xx) * CIN + ci);
#pragma unroll
      for (int cb = 0; cb < NB; ++cb)
        b[cb] = *reinterpret_cast<const bf16x8*>(w + (size_t)(cb * 16 + m) * K + k);
    }
#pragma unroll
    for (int cb = 0; cb < NB; ++cb)
      acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[cb], acc[cb], 0, 0, 0);
  }
  const size_t rowbase = (((size_t)n * D + z) * H + yv) * W;
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) {
    const int co = cb * 16 + m;
    const float bb = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vx = xb + g * 4 + r;
      if (vx < W) {
        float v = acc[cb][r] + bb;
        if (RELU) v = fmaxf(v, 0.0f);
        y[(rowbase + vx) * COUT + co] = f32_to_bf16_bits(v);
      }
    }
  }
}

// W'[ci][(26 - tap) * COUT + co] = W[co][tap * CIN + ci]
__global__ __launch_bounds__(256) void flip_transpose_weights(const uint16_t* __restrict__ w,
                                                              uint16_t* __restrict__ wt, int cin,
                                                              int cout) {
  int q = blockIdx.x * 256 + threadIdx.x;
  int total = cout * 27 * cin;
  if (q >= total) return;
  int co = q / (27 * cin), rem = q - co * 27 * cin;
  int tap = rem / cin, ci = rem - tap * cin;
  wt[((size_t)ci * 27 + (26 - tap)) * cout + co] = w[q];
}

// ---- weight gradient -------------------------------------------------------------------------
// grid.x: groups of kRows x-rows (n, z, y); grid.y: blocks of 64 k, one 16-column slice per wave.
// The row's output gradient is staged once per block in LDS (channel-major, 256-voxel chunks) and
// read by all four waves, so it is fetched from HBM / L2 once per 64 columns of K. dW (FP32) is
// accumulated.
constexpr int kRows = 8;
constexpr int kWgChunk = 256;            // voxels of a row staged at once
constexpr int kWgPitch = kWgChunk + 8;    // LDS row pitch (elements): 16-byte aligned rows
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_wgrad_mfma(const uint16_t* __restrict__ x,
                                                       const uint16_t* __restrict__ dy,
                                                       float* __restrict__ dw, int N, int D,
                                                       int H, int W) {
  constexpr int K = 27 * CIN;
  constexpr int NB = (COUT + 15) / 16;
  // the output-gradient chunk of the current row, transposed: T[co][voxel - chunk start]
  __shared__ __align__(16) uint16_t T[NB * 16][kWgPitch];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int kk = blockIdx.y * 64 + wave * 16 + m;  // B column of this lane
  const bool kvalid = kk < K;
  const int tap = kvalid ? kk / CIN : 0, ci = kvalid ? kk - (kk / CIN) * CIN : 0;
  const int dz = tap / 9, dyy = (tap / 3) % 3, dx = tap % 3;
  const int nrows = N * D * H;
  f32x4 acc[NB];
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int rr = 0; rr < kRows; ++rr) {
    const int row = blockIdx.x * kRows + rr;
    if (row >= nrows) break;  // uniform over the block
    const int yv = row % H, nz = row / H;
    const int z = nz % D, n = nz / D;
    const int zz = z + dz - 1, yy = yv + dyy - 1;
    const bool rvalid = kvalid && zz >= 0 && zz < D && yy >= 0 && yy < H;
    const size_t xrow = (((size_t)n * D + zz) * H + yy) * W;
    const size_t yrow = (((size_t)n * D + z) * H + yv) * W;
    for (int xc = 0; xc < W; xc += kWgChunk) {
      const int nv = min(kWgChunk, W - xc);
      const int nvp = (nv + 31) & ~31;  // the x0 loop reads whole 32-voxel steps
      __syncthreads();  // the previous chunk is no longer read
      // stage: 8 consecutive channels of one voxel per 16-byte load, zero past the row's end
      for (int e = threadIdx.x; e < nvp * (COUT / 8); e += 256) {
        const int v = e / (COUT / 8), c8 = (e - v * (COUT / 8)) * 8;
        u32x4 q = u32x4{0u, 0u, 0u, 0u};
        if (v < nv) q = *reinterpret_cast<const u32x4*>(dy + (yrow + xc + v) * COUT + c8);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          T[c8 + 2 * h][v] = (uint16_t)(q[h] & 0xffffu);
          T[c8 + 2 * h + 1][v] = (uint16_t)(q[h] >> 16);
        }
      }
      __syncthreads();
      // 32 voxels per step: lane (m, g) supplies voxels x0 + 8 g + q, q = 0..7, of its output
      // channel (A, one 16-byte LDS read) and of its K column (B, gathered) as raw BF16 bits
      for (int x0 = 0; x0 < nv; x0 += 32) {
        uint32_t bw[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          uint32_t pr = 0u;
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            const int xv = xc + x0 + g * 8 + 2 * h + e;
            const int xx = xv + dx - 1;
            uint32_t v = 0u;
            if (rvalid && xv < W && xx >= 0 && xx < W) v = x[(xrow + xx) * CIN + ci];
            pr |= v << (16 * e);
          }
          bw[h] = pr;
        }
        const bf16x8 bv = __builtin_bit_cast(bf16x8, u32x4{bw[0], bw[1], bw[2], bw[3]});
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
          const u32x4 aq = *reinterpret_cast<const u32x4*>(&T[cb * 16 + m][x0 + g * 8]);
          acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, aq), bv,
                                                            acc[cb], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int cb = 0; cb < NB; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = acc[cb][r];
      const int co = cb * 16 + g * 4 + r;                  // D row
      const int kc = blockIdx.y * 64 + wave * 16 + m;      // D column
      if (co < COUT && kc < K && v != 0.0f) unsafeAtomicAdd(dw + (size_t)co * K + kc, v);
    }
}

// ---- 1-channel layers (FP32 VALU) ------------------------------------------------------------
// Cin = 1: y[v][co] = act(b[co] + sum_tap w[co][tap] x[v + tap]); one thread per voxel.
template <int COUT, bool RELU>
__global__ __launch_bounds__(256) void conv_fwd_cin1(const uint16_t* __restrict__ x,
                                                     const uint16_t* __restrict__ w,
                                                     const float* __restrict__ bias,
                                                     uint16_t* __restrict__ y, int N, int D,
                                                     int H, int W) {
  __shared__ float ws[COUT * 27];
  for (int q = threadIdx.x; q < COUT * 27; q += 256) ws[q] = bf2f(w[q]);
  __syncthreads();
  const size_t nvox = (size_t)N * D * H * W;
  const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  const int xv = (int)(v % W), yv = (int)((v / W) % H), z = (int)((v / ((size_t)W * H)) % D);
  const size_t n = v / ((size_t)W * H * D);
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = bias ? bias[co] : 0.0f;
  for (int tap = 0; tap < 27; ++tap) {
    const int zz = z + tap / 9 - 1, yy = yv + (tap / 3) % 3 - 1, xx = xv + tap % 3 - 1;
    if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const float xvv = bf2f(x[((n * D + zz) * H + yy) * W + xx]);
#pragma unroll
    for (int co = 0; co < COUT; ++co) acc[co] = fmaf(ws[co * 27 + tap], xvv, acc[co]);
  }
#pragma unroll
  for (int co = 0; co < COUT; ++co) {
    float o = acc[co];
    if (RELU) o = fmaxf(o, 0.0f);
    y[v * COUT + co] = f32_to_bf16_bits(o);
  }
}

// Cout = 1: y[v] = b + sum_{tap, ci} w[tap][ci] x[v + tap][ci]  (output FP32, no activation).
// Also the input gradient of a Cin = 1 layer reads like this with the flipped weights.
template <int CIN>
__global__ __launch_bounds__(256) void conv_fwd_cout1(const uint16_t* __restrict__ x,
                                                      const uint16_t* __restrict__ w,
                                                      const float* __restrict__ bias,
                                                      float* __restrict__ y, int N, int D, int H,
                                                      int W) {
  __shared__ float ws[27 * CIN];
  for (int q = threadIdx.x; q < 27 * CIN; q += 256) ws[q] = bf2f(w[q]);
  __syncthreads();
  const size_t nvox = (size_t)N * D * H * W;
  const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  const int xv = (int)(v % W), yv = (int)((v / W) % H), z = (int)((v / ((size_t)W * H)) % D);
  const size_t n = v / ((size_t)W * H * D);
  float acc = bias ? bias[0] : 0.0f;
  for (int tap = 0; tap < 27; ++tap) {
    const int zz = z + tap / 9 - 1, yy = yv + (tap / 3) % 3 - 1, xx = xv + tap % 3 - 1;
    if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const uint16_t* px = x + (((n * D + zz) * H + yy) * W + xx) * CIN;
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) acc = fmaf(ws[tap * CIN + ci], bf2f(px[ci]), acc);
  }
  y[v] = acc;
}

// Input gradient of a Cout = 1 layer: dx[v][ci] = sum_tap w[tap][ci] dy[v - tap] (dy FP32).
template <int CIN>
__global__ __launch_bounds__(256) void conv_dgrad_cout1(const float* __restrict__ dy,
                                                        const uint16_t* __restrict__ w,
                                                        uint16_t* __restrict__ dx, int N, int D,
                                                        int H, int W) {
  __shared__ float ws[27 * CIN];
  for (int q = threadIdx.x; q < 27 * CIN; q += 256) ws[q] = bf2f(w[q]);
  __syncthreads();
  const size_t nvox = (size_t)N * D * H * W;
  const size_t v = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  const int xv = (int)(v % W), yv = (int)((v / W) % H), z = (int)((v / ((size_t)W * H)) % D);
  const size_t n = v / ((size_t)W * H * D);
  float acc[CIN];
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.0f;
  for (int tap = 0; tap < 27; ++tap) {
    // dy at v - (tap - 13): the output voxel whose tap `tap` reads voxel v
    const int zz = z - (tap / 9 - 1), yy = yv - ((tap / 3) % 3 - 1), xx = xv - (tap % 3 - 1);
    if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const float g = dy[((n * D + zz) * H + yy) * W + xx];
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) acc[ci] = fmaf(ws[tap * CIN + ci], g, acc[ci]);
  }
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) dx[v * CIN + ci] = f32_to_bf16_bits(acc[ci]);
}

// Weight gradient of a Cout = 1 layer with FP32 dy: dw[tap][ci] = sum_v dy[v] x[v + tap][ci].
// One block per (tap, chunk of voxels); each thread a subset of voxels for all ci.
template <int CIN>
__global__ __launch_bounds__(256) void conv_wgrad_cout1(const uint16_t* __restrict__ x,
                                                        const float* __restrict__ dy,
                                                        float* __restrict__ dw, int N, int D,
                                                        int H, int W, int chunk) {
  __shared__ float red[CIN][256];
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const size_t nvox = (size_t)N * D * H * W;
  float acc[CIN];
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.0f;
  const size_t v0 = (size_t)blockIdx.x * chunk;
  for (size_t v = v0 + threadIdx.x; v < v0 + chunk && v < nvox; v += 256) {
    const int xv = (int)(v % W), yv = (int)((v / W) % H), z = (int)((v / ((size_t)W * H)) % D);
    const size_t n = v / ((size_t)W * H * D);
    const int zz = z + dz, yy = yv + dyy, xx = xv + dx;
    if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const float g = dy[v];
    const uint16_t* px = x + (((n * D + zz) * H + yy) * W + xx) * CIN;
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) acc[ci] = fmaf(g, bf2f(px[ci]), acc[ci]);
  }
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) red[ci][threadIdx.x] = acc[ci];
  __syncthreads();
  if (threadIdx.x < CIN) {
    float s = 0.0f;
    for (int q = 0; q < 256; ++q) s += red[threadIdx.x][q];
    unsafeAtomicAdd(dw + tap * CIN + threadIdx.x, s);
  }
}

// Weight gradient of a Cin = 1 layer: dw[co][tap] = sum_v dy[v][co] x[v + tap].
template <int COUT>
__global__ __launch_bounds__(256) void conv_wgrad_cin1(const uint16_t* __restrict__ x,
                                                       const uint16_t* __restrict__ dy,
                                                       float* __restrict__ dw, int N, int D,
                                                       int H, int W, int chunk) {
  __shared__ float red[COUT][256];
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const size_t nvox = (size_t)N * D * H * W;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = 0.0f;
  const size_t v0 = (size_t)blockIdx.x * chunk;
  for (size_t v = v0 + threadIdx.x; v < v0 + chunk && v < nvox; v += 256) {
    const int xv = (int)(v % W), yv = (int)((v / W) % H), z = (int)((v / ((size_t)W * H)) % D);
    const size_t n = v / ((size_t)W * H * D);
    const int zz = z + dz, yy = yv + dyy, xx = xv + dx;
    if (zz < 0 || zz >= D || yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    const float xs = bf2f(x[((n * D + zz) * H + yy) * W + xx]);
    const uint16_t* pg = dy + v * COUT;
#pragma unroll
    for (int co = 0; co < COUT; ++co) acc[co] = fmaf(bf2f(pg[co]), xs, acc[co]);
  }
#pragma unroll
  for (int co = 0; co < COUT; ++co) red[co][threadIdx.x] = acc[co];
  __syncthreads();
  if (threadIdx.x < COUT) {
    float s = 0.0f;
    for (int q = 0; q < 256; ++q) s += red[threadIdx.x][q];
    unsafeAtomicAdd(dw + threadIdx.x * 27 + tap, s);
  }
}

// MFMA lane-layout probe: D = A B for one 16x16x32 BF16 and one 16x16x4 FP32 product, with
// A and B given row-major; a single wave, every access in bounds.
__global__ __launch_bounds__(64) void mfma_probe(const uint16_t* __restrict__ A32,
                                                 const uint16_t* __restrict__ B32,
                                                 const float* __restrict__ A4,
                                                 const float* __restrict__ B4,
                                                 float* __restrict__ D32, float* __restrict__ D4) {
  const int l = threadIdx.x, m = l & 15, g = l >> 4;
  bf16x8 a, b;
  uint16_t* pa = reinterpret_cast<uint16_t*>(&a);
  uint16_t* pb = reinterpret_cast<uint16_t*>(&b);
  for (int q = 0; q < 8; ++q) {
    pa[q] = A32[m * 32 + g * 8 + q];   // A[16][32] row-major
    pb[q] = B32[(g * 8 + q) * 16 + m]; // B[32][16] row-major
  }
  f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  f32x4 e = __builtin_amdgcn_mfma_f32_16x16x4f32(A4[m * 4 + g], B4[g * 16 + m],
                                                  f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int r = 0; r < 4; ++r) {
    D32[(g * 4 + r) * 16 + m] = d[r];
    D4[(g * 4 + r) * 16 + m] = e[r];
  }
}

template <int CIN, int COUT>
bool launch_fwd_t(const void* x, const void* w, const float* b, void* y, int N, int D, int H,
                  int W, int relu, hipStream_t s) {
  dim3 grid((W + 63) / 64, H, N * D);
  if (relu)
    conv_fwd_mfma<CIN, COUT, true><<<grid, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)w,
                                                        b, (uint16_t*)y, D, H, W);
  else
    conv_fwd_mfma<CIN, COUT, false><<<grid, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)w,
                                                         b, (uint16_t*)y, D, H, W);
  return true;
}

// Supported (Cin, Cout) pairs of the MFMA kernels: the network's layers and their transposes.
bool launch_fwd(int cin, int cout, const void* x, const void* w, const float* b, void* y, int N,
                int D, int H, int W, int relu, hipStream_t s) {
#define FC_CONV_CASE(ci, co) \
  if (cin == ci && cout == co) return launch_fwd_t<ci, co>(x, w, b, y, N, D, H, W, relu, s);
  FC_CONV_CASE(16, 16) FC_CONV_CASE(16, 32) FC_CONV_CASE(32, 32) FC_CONV_CASE(32, 64)
  FC_CONV_CASE(64, 64) FC_CONV_CASE(96, 32) FC_CONV_CASE(48, 16) FC_CONV_CASE(32, 16)
  FC_CONV_CASE(64, 32) FC_CONV_CASE(32, 96) FC_CONV_CASE(16, 48)
#undef FC_CONV_CASE
  return false;
}

bool launch_wgrad(int cin, int cout, const void* x, const void* dy, float* dw, int N, int D,
                  int H, int W, hipStream_t s) {
  const int K = 27 * cin;
  const int rows = N * D * H;
  dim3 grid((rows + kRows - 1) / kRows, (K + 63) / 64);
#define FC_WG_CASE(ci, co)                                                                   \
  if (cin == ci && cout == co) {                                                             \
    conv_wgrad_mfma<ci, co><<<grid, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)dy, dw, \
                                                 N, D, H, W);                                \
    return true;                                                                             \
  }
  FC_WG_CASE(16, 16) FC_WG_CASE(16, 32) FC_WG_CASE(32, 32) FC_WG_CASE(32, 64)
  FC_WG_CASE(64, 64) FC_WG_CASE(96, 32) FC_WG_CASE(48, 16)
#undef FC_WG_CASE
  return false;
}

}  // namespace
}  // namespace fc

using namespace fc;

extern "C" {

int fc_conv3d_fwd_dev(int device, const void* x, const void* w, const float* bias, void* y,
                      int N, int D, int H, int W, int cin, int cout, int relu, void* stream) {
  clear_error();
  FC_CHECK_ARG(x && w && y, "null pointer argument");
  FC_CHECK_ARG(N > 0 && D > 0 && H > 0 && W > 0 && H <= 65535 && (long)N * D <= 65535,
               "bad activation shape");
  int rc = select_device(device);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long nvox = (long)N * D * H * W;
  const int blocks = (int)((nvox + 255) / 256);
  int tok = prof_begin("conv3d_fwd", s);
  if (cin == 1 && cout == 16) {
    if (relu) conv_fwd_cin1<16, true><<<blocks, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)w, bias, (uint16_t*)y, N, D, H, W);
    else conv_fwd_cin1<16, false><<<blocks, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)w, bias, (uint16_t*)y, N, D, H, W);
  } else if (cout == 1 && cin == 16) {
    FC_CHECK_ARG(!relu, "the Cout = 1 layer has no activation and writes FP32");
    conv_fwd_cout1<16><<<blocks, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)w, bias, (float*)y, N, D, H, W);
  } else if (!launch_fwd(cin, cout, x, w, bias, y, N, D, H, W, relu, s)) {
    prof_end(tok, s);
    FC_CHECK_ARG(false, "unsupported channel pair %d -> %d", cin, cout);
  }
  prof_end(tok, s);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

// dx = conv^T(dy): workspace wt (cin*27*cout BF16) receives the flipped transposed weights.
int fc_conv3d_dgrad_dev(int device, const void* dy, const void* w, void* wt, void* dx, int N,
                        int D, int H, int W, int cin, int cout, void* stream) {
  clear_error();
  FC_CHECK_ARG(dy && w && dx, "null pointer argument");
  FC_CHECK_ARG(N > 0 && D > 0 && H > 0 && W > 0 && H <= 65535 && (long)N * D <= 65535,
               "bad activation shape");
  int rc = select_device(device);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long nvox = (long)N * D * H * W;
  const int blocks = (int)((nvox + 255) / 256);
  int tok = prof_begin("conv3d_dgrad", s);
  if (cout == 1 && cin == 16) {
    conv_dgrad_cout1<16><<<blocks, 256, 0, s>>>((const float*)dy, (const uint16_t*)w, (uint16_t*)dx, N, D, H, W);
  } else {
    FC_CHECK_ARG(wt, "null weight workspace");
    const int total = cout * 27 * cin;
    flip_transpose_weights<<<(total + 255) / 256, 256, 0, s>>>((const uint16_t*)w,
                                                              (uint16_t*)wt, cin, cout);
    if (!launch_fwd(cout, cin, dy, wt, nullptr, dx, N, D, H, W, 0, s)) {
      prof_end(tok, s);
      FC_CHECK_ARG(false, "unsupported channel pair %d -> %d", cin, cout);
    }
  }
  prof_end(tok, s);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

// dw (FP32 [cout][27][cin]) += sum over voxels of dy x patch.
int fc_conv3d_wgrad_dev(int device, const void* x, const void* dy, float* dw, int N, int D, int H,
                        int W, int cin, int cout, void* stream) {
  clear_error();
  FC_CHECK_ARG(x && dy && dw, "null pointer argument");
  FC_CHECK_ARG(N > 0 && D > 0 && H > 0 && W > 0, "bad activation shape");
  int rc = select_device(device);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long nvox = (long)N * D * H * W;
  int tok = prof_begin("conv3d_wgrad", s);
  if (cin == 1 && cout == 16) {
    const int chunk = 16384;
    dim3 grid((unsigned)((nvox + chunk - 1) / chunk), 27);
    conv_wgrad_cin1<16><<<grid, 256, 0, s>>>((const uint16_t*)x, (const uint16_t*)dy, dw, N, D, H, W, chunk);
  } else if (cout == 1 && cin == 16) {
    const int chunk = 16384;
    dim3 grid((unsigned)((nvox + chunk - 1) / chunk), 27);
    conv_wgrad_cout1<16><<<grid, 256, 0, s>>>((const uint16_t*)x, (const float*)dy, dw, N, D, H, W, chunk);
  } else if (!launch_wgrad(cin, cout, x, dy, dw, N, D, H, W, s)) {
    prof_end(tok, s);
    FC_CHECK_ARG(false, "unsupported channel pair %d -> %d", cin, cout);
  }
  prof_end(tok, s);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

int fc_mfma_layout_probe(int device, const void* A32, const void* B32, const float* A4,
                         const float* B4, float* D32, float* D4, void* stream) {
  clear_error();
  FC_CHECK_ARG(A32 && B32 && A4 && B4 && D32 && D4, "null pointer argument");
  int rc = select_device(device);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  mfma_probe<<<1, 64, 0, s>>>((const uint16_t*)A32, (const uint16_t*)B32, A4, B4, D32, D4);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

}  // extern "C"
