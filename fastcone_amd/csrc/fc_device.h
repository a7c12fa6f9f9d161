// Device-side helpers shared by the projector / backprojector kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fc_internal.h"

namespace fc {

// Stored element types. 16-bit storage keeps values as raw bits; arithmetic is always FP32.
template <int DT>
struct Store;
template <>
struct Store<FC_F32> {
  using T = float;
};
template <>
struct Store<FC_BF16> {
  using T = uint16_t;
};
template <>
struct Store<FC_F16> {
  using T = uint16_t;
};

__device__ __forceinline__ float bf16_bits_to_f32(uint32_t b) { return __uint_as_float(b << 16); }

// FP32 -> BF16 with round-to-nearest-even; NaN stays NaN.
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float f16_bits_to_f32(uint32_t b) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xffffu));
}
__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) {
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}

template <int DT>
__device__ __forceinline__ float load_elem(const void* p, size_t i) {
  if constexpr (DT == FC_F32) {
    return static_cast<const float*>(p)[i];
  } else if constexpr (DT == FC_BF16) {
    return bf16_bits_to_f32(static_cast<const uint16_t*>(p)[i]);
  } else {
    return f16_bits_to_f32(static_cast<const uint16_t*>(p)[i]);
  }
}

// Raw storage bits of element i (FP32 bits, or the 16-bit pattern in the low half) and their
// conversion, split so that a prefetch can keep loads in flight and convert only when it stores.
template <int DT>
__device__ __forceinline__ uint32_t load_bits(const void* p, size_t i) {
  if constexpr (DT == FC_F32) {
    return static_cast<const uint32_t*>(p)[i];
  } else {
    return static_cast<const uint16_t*>(p)[i];
  }
}
template <int DT>
__device__ __forceinline__ float bits_to_f32(uint32_t b) {
  if constexpr (DT == FC_F32) {
    return __uint_as_float(b);
  } else if constexpr (DT == FC_BF16) {
    return __uint_as_float(b << 16);
  } else {
    return f16_bits_to_f32(b);
  }
}

template <int DT>
__device__ __forceinline__ uint16_t to_bits16(float f) {
  if constexpr (DT == FC_BF16) {
    return f32_to_bf16_bits(f);
  } else {
    return f32_to_f16_bits(f);
  }
}

// Unpack the low / high 16-bit halves of a 32-bit word to FP32.
template <int DT>
__device__ __forceinline__ float lo16(uint32_t w) {
  if constexpr (DT == FC_BF16) {
    return __uint_as_float(w << 16);
  } else {
    return f16_bits_to_f32(w);
  }
}
template <int DT>
__device__ __forceinline__ float hi16(uint32_t w) {
  if constexpr (DT == FC_BF16) {
    return __uint_as_float(w & 0xffff0000u);
  } else {
    return f16_bits_to_f32(w >> 16);
  }
}

template <int DT>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)to_bits16<DT>(a) | ((uint32_t)to_bits16<DT>(b) << 16);
}

__device__ __forceinline__ float lerp(float a, float b, float t) { return fmaf(t, b - a, a); }

// a[23:0] * b[23:0] + c (low 32 bits), one full-rate VALU instruction. The compiler lowers the
// equivalent C expression with a variable multiplier to the multi-pass v_mad_u64_u32.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// mad_u24 with a wave-uniform multiplier read from a scalar register (an operand the vector
// form would copy into a VGPR at every use); b must be the same in every active lane.
__device__ __forceinline__ uint32_t mad_u24_s(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3"
      : "=v"(r)
      : "v"(a), "s"(__builtin_amdgcn_readfirstlane(b)), "v"(c));
  return r;
}
// a[23:0] * b[23:0] (low 32 bits) with a wave-uniform b read from a scalar register.
__device__ __forceinline__ uint32_t mul_u24_s(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(__builtin_amdgcn_readfirstlane(b)), "v"(a));
  return r;
}
// a * b for a, b < 2^24 (kept as v_mul_u32_u24: the compiler would otherwise fold a product
// with a constant into a quarter-rate 32-bit multiply)
__device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Logical block index for a 1-D grid, remapped so each of the 8 XCDs (which receive blocks
// round-robin) works on one contiguous range of logical blocks.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int nxcd = 8;
  int xcd = b % nxcd, idx = b / nxcd;
  int per = nb / nxcd, rem = nb % nxcd;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Ray setup for the forward projector and its adjoint. Computes, in FP64, the range of integer k
// for which s + k*h*u lies inside the closed box [-1, n] of the local block (index space), and
// returns the first sample and the step in the padded index frame (offset `pad`) as FP32.
// nk = 0 when the ray misses the block.
struct RaySeg {
  float px, py, pz;  // first sample, padded index frame
  float sx, sy, sz;  // step h*u
  int nk;
};

__device__ __forceinline__ void ray_dir(const double* V, double j, double i, double* dx,
                                        double* dy, double* dz) {
  if (V[24] == (double)FC_DET_CYLINDER) {
    double g = (j - V[25]) * V[26];
    double sg, cg;
    sincos(g, &sg, &cg);
    double r = i - V[27];
    *dx = cg * V[3] + sg * V[6] + r * V[9];
    *dy = cg * V[4] + sg * V[7] + r * V[10];
    *dz = cg * V[5] + sg * V[8] + r * V[11];
  } else {
    *dx = V[3] + j * V[6] + i * V[9] - V[0];
    *dy = V[4] + j * V[7] + i * V[10] - V[1];
    *dz = V[5] + j * V[8] + i * V[11] - V[2];
  }
}

__device__ __forceinline__ RaySeg ray_segment(const double* V, int j, int i, double ox,
                                              double oy, double oz, int nx, int ny, int nz,
                                              double h, int max_steps, float pad) {
  RaySeg r;
  r.nk = 0;
  r.px = r.py = r.pz = 0.f;
  r.sx = r.sy = r.sz = 0.f;
  double dx, dy, dz;
  ray_dir(V, (double)j, (double)i, &dx, &dy, &dz);
  double len = sqrt(dx * dx + dy * dy + dz * dz);
  if (!(len > 0.0)) return r;
  double inv = 1.0 / len;
  double ux = dx * inv, uy = dy * inv, uz = dz * inv;
  double lx = V[0] - ox, ly = V[1] - oy, lz = V[2] - oz;  // source, local index space
  double tmin = -1e300, tmax = 1e300;
  const double l[3] = {lx, ly, lz};
  const double u[3] = {ux, uy, uz};
  const double n[3] = {(double)nx, (double)ny, (double)nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (fabs(u[a]) < 1e-300) {
      if (l[a] <= -1.0 || l[a] >= n[a]) return r;
    } else {
      double ru = 1.0 / u[a];
      double t1 = (-1.0 - l[a]) * ru, t2 = (n[a] - l[a]) * ru;
      tmin = fmax(tmin, fmin(t1, t2));
      tmax = fmin(tmax, fmax(t1, t2));
    }
  }
  if (!(tmin <= tmax)) return r;
  double k0 = ceil(tmin / h), k1 = floor(tmax / h);
  if (!(k1 >= k0)) return r;
  double nkd = k1 - k0 + 1.0;
  if (!(nkd <= (double)max_steps)) return r;  // only for corrupt geometry
  double t0 = k0 * h;
  r.px = (float)(lx + t0 * ux + pad);
  r.py = (float)(ly + t0 * uy + pad);
  r.pz = (float)(lz + t0 * uz + pad);
  r.sx = (float)(h * ux);
  r.sy = (float)(h * uy);
  r.sz = (float)(h * uz);
  int nk = (int)nkd;
  // Memory-safety guard. Sample k is evaluated as fmaf((float)k, s, p0), which is monotone in k,
  // so checking the first and last sample bounds every sample. In exact arithmetic they lie in
  // [pad-1, n+pad]; the guard allows up to (pad-1) voxels of FP32 rounding slack.
  float kl = (float)(nk - 1);
  float qx = fmaf(kl, r.sx, r.px), qy = fmaf(kl, r.sy, r.py), qz = fmaf(kl, r.sz, r.pz);
  float lo = 0.0f;
  float hx = (float)nx + 2.0f * pad - 1.0f, hy = (float)ny + 2.0f * pad - 1.0f,
        hz = (float)nz + 2.0f * pad - 1.0f;
  bool ok = r.px >= lo && r.py >= lo && r.pz >= lo && qx >= lo && qy >= lo && qz >= lo &&
            r.px < hx && r.py < hy && r.pz < hz && qx < hx && qy < hy && qz < hz;
  r.nk = ok ? nk : 0;
  return r;
}

}  // namespace fc
