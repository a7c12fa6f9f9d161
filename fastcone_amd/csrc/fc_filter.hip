// Row filters for gfx950: every detector row (or kappa line) convolved with a fixed real kernel,
//
//   out[r, j] = sum_{m=0}^{nw-1} kern[j - m] * wgt(r, m) * in[r, m],     j in [0, nw),
//
// with kern supported on [-(nw-1), nw-1]. Two kernels are built in:
//   * RAMP — the FDK ramp of fastcone.py:272-292: a row zero-padded to n = 2 nw (data at [0, nw)),
//     its rfft multiplied by H[k] = pi k / n (k = 1..nw), H[0] = pi / (8 n), irfft, crop [0, nw).
//     That operator is the convolution with kern[d] = irfft(H)[d mod n] for |d| < nw, times a
//     caller scale (1/pw turns the reference's cycles-per-pixel ramp into physical units). With
//     view records, wgt is the FDK cosine pre-weight: the cosine of the angle between the pixel's
//     ray and the detector normal (physical units; fastcone.py:260-267 uses pixel units);
//   * HILBERT — the curved-detector kappa-line Hilbert kernel of the Katsevich filter:
//     kern[d] = 2 / (pi d) * (d dg / sin(d dg)) for odd d, 0 for even d.
//
// Method. The linear convolution equals a circular one of any length L >= 2 nw - 1, so each row
// is zero-padded to L = 2^p (64 <= L <= 8192), and two rows share one complex transform (row a in
// the real part, row b in the imaginary part: the kernel is real, so the two convolutions do not
// mix). The transform is a radix-8 (then 4 or 2) Stockham FFT in LDS; the spectrum is multiplied
// by DFT_L(kern) / L and conjugated, transformed forward again and conjugated back
// (ifft(X) = conj(fft(conj(X))) / L): the result's real part is row a, minus its imaginary part
// row b. DFT_L(kern) is evaluated on the device in FP64 from the closed-form taps (not from the
// length-2nw bins, which zero padding to L would not reproduce), and rounded to FP32; so are the
// twiddles exp(-2 pi i m / L). A 256-thread block owns one pair of rows at a time and reads both
// completely before it writes either, so the filter may run in place.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

constexpr int kFT = 256;  // threads per block

__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }  // a * -i

// In-register forward DFTs of 2, 4 and 8 points (kernel e^{-2 pi i n k / R}).
template <int R>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<2>(float2* v) {
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<4>(float2* v) {
  const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const float2 t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<8>(float2* v) {
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e);
  dft<4>(o);
  const float c = 0.70710678118654752f;
  o[1] = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));     // * e^{-i pi/4}
  o[2] = mul_mi(o[2]);                                                  // * e^{-i pi/2}
  o[3] = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));    // * e^{-3 i pi/4}
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

// One Stockham stage of radix R after sub-transforms of length NS, in place in LDS: every item j
// reads buf[j + r L/R], applies the twiddles of its position k = j mod NS, transforms, and writes
// buf[(j - k) R + k + r NS]. All reads finish (barrier) before any write.
template <int L, int R, int NS>
__device__ __forceinline__ void stockham(float2* buf, const float2* __restrict__ tw) {
  constexpr int NI = L / R;
  constexpr int PER = (NI + kFT - 1) / kFT;
  float2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = (int)threadIdx.x + kFT * q;
    if (NI % kFT == 0 || j < NI) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[j + r * NI];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = (int)threadIdx.x + kFT * q;
    if (NI % kFT == 0 || j < NI) {
      const int k = j & (NS - 1);
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], tw[k * r * (L / (NS * R))]);
      }
      dft<R>(v[q]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

// Forward length-L FFT: radix 8 while at least 8 points remain per sub-transform, then 4 or 2.
template <int L, int NS>
__device__ __forceinline__ void fft_pass(float2* buf, const float2* __restrict__ tw) {
  if constexpr (NS < L) {
    constexpr int R = (L / NS) >= 8 ? 8 : L / NS;
    stockham<L, R, NS>(buf, tw);
    fft_pass<L, NS * R>(buf, tw);
  }
}

struct FiltArgs {
  const void* in;
  void* out;
  const double* views;  // flat-detector view records for the cosine weight, or null
  const float2* tw;     // L twiddles
  const float2* mult;   // L multipliers DFT_L(kern) / L
  long long nrows;
  int nw, nh;
  int in_dt, out_dt;
};

__device__ __forceinline__ float load_any(const void* p, int dt, size_t i) {
  if (dt == FC_F32) return load_elem<FC_F32>(p, i);
  if (dt == FC_BF16) return load_elem<FC_BF16>(p, i);
  return load_elem<FC_F16>(p, i);
}

__device__ __forceinline__ void store_any(void* p, int dt, size_t i, float v) {
  if (dt == FC_F32) static_cast<float*>(p)[i] = v;
  else if (dt == FC_BF16) static_cast<uint16_t*>(p)[i] = f32_to_bf16_bits(v);
  else static_cast<uint16_t*>(p)[i] = f32_to_f16_bits(v);
}

// Cosine pre-weight of row r: pixel j of the row sits at b + j u (relative to the source, FP64);
// its weight is dperp / |b + j u|, dperp the source's distance from the detector plane.
struct RowWeight {
  double bx, by, bz, ux, uy, uz;
  float dperp;
  bool on;
};

__device__ __forceinline__ RowWeight row_weight(const FiltArgs& a, long long r) {
  RowWeight w;
  w.on = a.views != nullptr;
  w.bx = w.by = w.bz = w.ux = w.uy = w.uz = 0.0;
  w.dperp = 1.0f;
  if (!w.on) return w;
  const long long t = r / a.nh;
  const double i = (double)(r - t * a.nh);
  const double* V = a.views + (size_t)t * FC_VIEW_DOUBLES;
  const double r0x = V[3] - V[0], r0y = V[4] - V[1], r0z = V[5] - V[2];
  w.bx = r0x + i * V[9];
  w.by = r0y + i * V[10];
  w.bz = r0z + i * V[11];
  w.ux = V[6];
  w.uy = V[7];
  w.uz = V[8];
  const double nx = V[7] * V[11] - V[8] * V[10], ny = V[8] * V[9] - V[6] * V[11],
               nz = V[6] * V[10] - V[7] * V[9];
  const double nn = sqrt(nx * nx + ny * ny + nz * nz);
  // only flat-detector records carry the plane geometry this weight needs
  w.dperp = V[24] == (double)FC_DET_FLAT ? (float)(fabs(r0x * nx + r0y * ny + r0z * nz) / nn)
                                         : __builtin_nanf("");
  return w;
}

__device__ __forceinline__ float weight_at(const RowWeight& w, int j) {
  if (!w.on) return 1.0f;
  const double jd = (double)j;
  const double x = fma(jd, w.ux, w.bx), y = fma(jd, w.uy, w.by), z = fma(jd, w.uz, w.bz);
  return w.dperp * rsqrtf((float)(x * x + y * y + z * z));
}

template <int LOG2L>
__global__ __launch_bounds__(kFT) void row_filter_kernel(FiltArgs a) {
  constexpr int L = 1 << LOG2L;
  constexpr int PERM = (L + kFT - 1) / kFT;  // spectrum bins per thread
  __shared__ float2 buf[L];
  float2 h[PERM];
#pragma unroll
  for (int q = 0; q < PERM; ++q) {
    const int m = (int)threadIdx.x + kFT * q;
    h[q] = m < L ? a.mult[m] : make_float2(0.0f, 0.0f);
  }
  const long long npairs = (a.nrows + 1) / 2;
  for (long long p = blockIdx.x; p < npairs; p += gridDim.x) {
    const long long r0 = 2 * p, r1 = r0 + 1;
    const bool two = r1 < a.nrows;
    const RowWeight w0 = row_weight(a, r0);
    const RowWeight w1 = row_weight(a, two ? r1 : r0);
    const size_t o0 = (size_t)r0 * (size_t)a.nw, o1 = o0 + (size_t)a.nw;
#pragma unroll
    for (int q = 0; q < PERM; ++q) {
      const int m = (int)threadIdx.x + kFT * q;
      if (m < L) {
        float2 v = make_float2(0.0f, 0.0f);
        if (m < a.nw) {
          v.x = load_any(a.in, a.in_dt, o0 + m) * weight_at(w0, m);
          if (two) v.y = load_any(a.in, a.in_dt, o1 + m) * weight_at(w1, m);
        }
        buf[m] = v;
      }
    }
    __syncthreads();
    fft_pass<L, 1>(buf, a.tw);
#pragma unroll
    for (int q = 0; q < PERM; ++q) {
      const int m = (int)threadIdx.x + kFT * q;
      if (m < L) {
        const float2 z = cmul(h[q], buf[m]);
        buf[m] = make_float2(z.x, -z.y);
      }
    }
    __syncthreads();
    fft_pass<L, 1>(buf, a.tw);
#pragma unroll
    for (int q = 0; q < PERM; ++q) {
      const int j = (int)threadIdx.x + kFT * q;
      if (j < a.nw) {
        const float2 z = buf[j];
        store_any(a.out, a.out_dt, o0 + j, z.x);
        if (two) store_any(a.out, a.out_dt, o1 + j, -z.y);
      }
    }
    __syncthreads();  // the next pair overwrites buf
  }
}

// ---- tables (FP64) ---------------------------------------------------------------------------

// Taps kern[q], q = d + nw - 1, d in [-(nw-1), nw-1].
__global__ __launch_bounds__(256) void filt_taps(int kind, double p0, int nw,
                                                 double* __restrict__ kern) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 2 * nw - 1) return;
  const long long d = (long long)q - (nw - 1);
  const long long ad = d < 0 ? -d : d;
  double v = 0.0;
  if (kind == FC_FILTER_RAMP) {
    // irfft of H over n = 2 nw bins: (1/n) [H0 + 2 sum_{k=1}^{nw-1} H_k cos(2 pi k d / n)
    //                                      + H_nw cos(pi d)],  H_k = (pi / n) k, H0 = (pi / n) / 8
    const long long n = 2LL * nw;
    double s = 0.0;
    for (long long k = 1; k < nw; ++k) s += (double)k * cospi(2.0 * (double)((k * ad) % n) / (double)n);
    const double c = M_PI / (double)n;
    v = p0 * (0.125 * c + 2.0 * c * s + c * (double)nw * ((ad & 1) ? -1.0 : 1.0)) / (double)n;
  } else if (ad & 1) {
    const double x = (double)d * p0;
    v = 2.0 / (M_PI * (double)d) * (x / sin(x));
  }
  kern[q] = v;
}

// mult[k] = (1/L) sum_d kern[d] e^{-2 pi i k d / L}, with exact integer reduction of k d mod L.
__global__ __launch_bounds__(256) void filt_multiplier(const double* __restrict__ kern, int nw,
                                                       int L, float2* __restrict__ mult) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= L) return;
  double re = 0.0, im = 0.0;
  for (int q = 0; q < 2 * nw - 1; ++q) {
    const double c0 = kern[q];
    if (c0 == 0.0) continue;
    const long long d = (long long)q - (nw - 1);
    long long m = ((long long)k * d) % L;
    if (m < 0) m += L;
    double s, c;
    sincospi(2.0 * (double)m / (double)L, &s, &c);
    re = fma(c0, c, re);
    im = fma(-c0, s, im);
  }
  mult[k] = make_float2((float)(re / L), (float)(im / L));
}

__global__ __launch_bounds__(256) void filt_twiddles(int L, float2* __restrict__ tw) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= L) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)L, &s, &c);
  tw[m] = make_float2((float)c, (float)-s);
}

int filter_log2(int nw) {
  int p = 6;
  while ((1 << p) < 2 * nw - 1) ++p;
  return p;
}

inline constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Tables of the widest supported row (FC_FILTER_MAX_COLS columns, FFT length 8192): twiddles and
// multipliers (8192 float2 each) and the 8191 FP64 taps. Every row filter needs at most this.
constexpr size_t kMaxTableBytes = 2 * align256((size_t)8192 * sizeof(float2)) +
                                  align256((2 * (size_t)FC_FILTER_MAX_COLS - 1) * sizeof(double));

// Device-resident tables for fc_fdk_filter_noalloc_dev (one copy per GPU, held until the process
// ends). In host code the name is only a stand-in: the device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kMaxTableBytes];

template <int LOG2L>
void launch_rows(const FiltArgs& a, hipStream_t s) {
  const long long npairs = (a.nrows + 1) / 2;
  const int blocks = (int)(npairs < 2048 ? npairs : 2048);
  row_filter_kernel<LOG2L><<<blocks, kFT, 0, s>>>(a);
}

std::mutex g_scratch_mutex;
void* g_scratch_ptr[64] = {nullptr};

// The calling thread's current device's address of `scratch`, looked up on first use.
int scratch_address(int device, void** out) {
  FC_CHECK_ARG(device >= 0 && device < 64, "device id %d too large", device);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    FC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return FC_OK;
}

}  // namespace

size_t filter_table_bytes(int nw) {
  const size_t L = (size_t)1 << filter_log2(nw);
  return 2 * align256(L * sizeof(float2)) + align256((2 * (size_t)nw - 1) * sizeof(double));
}

int launch_row_filter(const void* in, int in_dt, void* out, int out_dt, int nw, int nh,
                      long long nrows, const double* views, int kind, double p0, void* tables,
                      hipStream_t stream) {
  FC_CHECK_ARG(nw >= 1 && nw <= FC_FILTER_MAX_COLS, "row length %d out of [1, %d]", nw,
               FC_FILTER_MAX_COLS);
  FC_CHECK_ARG(nh >= 1, "rows per view must be positive");
  if (nrows == 0) return FC_OK;
  const int p = filter_log2(nw);
  const int L = 1 << p;
  char* base = static_cast<char*>(tables);
  float2* tw = reinterpret_cast<float2*>(base);
  float2* mult = reinterpret_cast<float2*>(base + align256((size_t)L * sizeof(float2)));
  double* kern = reinterpret_cast<double*>(base + 2 * align256((size_t)L * sizeof(float2)));
  filt_taps<<<(2 * nw - 1 + 255) / 256, 256, 0, stream>>>(kind, p0, nw, kern);
  filt_twiddles<<<(L + 255) / 256, 256, 0, stream>>>(L, tw);
  filt_multiplier<<<(L + 255) / 256, 256, 0, stream>>>(kern, nw, L, mult);
  FiltArgs a;
  a.in = in;
  a.out = out;
  a.views = views;
  a.tw = tw;
  a.mult = mult;
  a.nrows = nrows;
  a.nw = nw;
  a.nh = nh;
  a.in_dt = in_dt;
  a.out_dt = out_dt;
  switch (p) {
    case 6: launch_rows<6>(a, stream); break;
    case 7: launch_rows<7>(a, stream); break;
    case 8: launch_rows<8>(a, stream); break;
    case 9: launch_rows<9>(a, stream); break;
    case 10: launch_rows<10>(a, stream); break;
    case 11: launch_rows<11>(a, stream); break;
    case 12: launch_rows<12>(a, stream); break;
    default: launch_rows<13>(a, stream); break;
  }
  FC_HIP(hipGetLastError());
  return FC_OK;
}

}  // namespace fc

using namespace fc;

extern "C" {

int fc_fdk_filter_dev(int device, const void* proj, int proj_dtype, void* out, int out_dtype,
                      const int proj_shape[3], const double* views, double ramp_scale,
                      void* stream) {
  clear_error();
  FC_CHECK_ARG(proj_shape, "null shape");
  const int nw = proj_shape[0], nh = proj_shape[1], nv = proj_shape[2];
  FC_CHECK_ARG(nw >= 1 && nw <= FC_FILTER_MAX_COLS, "detector width %d out of [1, %d]", nw,
               FC_FILTER_MAX_COLS);
  FC_CHECK_ARG(nh >= 1 && nv >= 0, "projection shape invalid (%d,%d,%d)", nw, nh, nv);
  FC_CHECK_ARG(dtype_ok(proj_dtype) && dtype_ok(out_dtype), "unknown dtype");
  FC_CHECK_ARG(std::isfinite(ramp_scale), "ramp scale not finite");
  FC_CHECK_ARG(nv == 0 || (proj && out), "null pointer argument");
  int rc = select_device(device);
  if (rc) return rc;
  if (nv == 0) return FC_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* tables = nullptr;
  if ((rc = acquire_ws(device, SLOT_FILTER, filter_table_bytes(nw), s, &tables))) return rc;
  int tok = prof_begin("fdk_filter", s);
  rc = launch_row_filter(proj, proj_dtype, out, out_dtype, nw, nh, (long long)nh * nv, views,
                         FC_FILTER_RAMP, ramp_scale, tables, s);
  prof_end(tok, s);
  if (rc) return rc;
  return release_ws(device, SLOT_FILTER, s);
}

// The same filter with no device allocation: its tables live in the file-scope `scratch` array
// (kMaxTableBytes, enough for FC_FILTER_MAX_COLS columns) instead of a library workspace. No
// counterpart in the reference.
int fc_fdk_filter_noalloc_dev(int device, const void* proj, int proj_dtype, void* out,
                              int out_dtype, const int proj_shape[3], const double* views,
                              double ramp_scale, void* stream) {
  clear_error();
  FC_CHECK_ARG(proj_shape, "null shape");
  const int nw = proj_shape[0], nh = proj_shape[1], nv = proj_shape[2];
  FC_CHECK_ARG(nw >= 1 && nw <= FC_FILTER_MAX_COLS,
               "detector width %d out of [1, %d] (the no-allocation tables hold %zu bytes)", nw,
               FC_FILTER_MAX_COLS, kMaxTableBytes);
  FC_CHECK_ARG(filter_table_bytes(nw) <= kMaxTableBytes, "filter tables exceed the scratch array");
  FC_CHECK_ARG(nh >= 1 && nv >= 0, "projection shape invalid (%d,%d,%d)", nw, nh, nv);
  FC_CHECK_ARG(dtype_ok(proj_dtype) && dtype_ok(out_dtype), "unknown dtype");
  FC_CHECK_ARG(std::isfinite(ramp_scale), "ramp scale not finite");
  FC_CHECK_ARG(nv == 0 || (proj && out), "null pointer argument");
  int rc = select_device(device);
  if (rc) return rc;
  void* tables = nullptr;
  if ((rc = scratch_address(device, &tables))) return rc;
  if (nv == 0) return FC_OK;
  // every table entry the filter reads is rewritten by this call's table kernels (same stream)
  return launch_row_filter(proj, proj_dtype, out, out_dtype, nw, nh, (long long)nh * nv, views,
                           FC_FILTER_RAMP, ramp_scale, tables, static_cast<hipStream_t>(stream));
}

}  // extern "C"
