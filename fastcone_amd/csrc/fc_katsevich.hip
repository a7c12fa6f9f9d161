// Katsevich helical reconstruction on a cylindrical (source-centred) detector, for gfx950.
//
// The reference only backprojects with a U^1 weight (coneBeamGPU.cu:258-338) and leaves the
// derivative and kappa-line filtering to the caller; it cannot represent a curved detector. This
// file implements the whole exact algorithm (Katsevich 2002, in the native curved-detector form of
// Noo, Pack & Heuscher 2003, Phys. Med. Biol. 48:3787):
//
//   g1 = [G(k+1, g + dl) - G(k-1, g - dl)] / (2 dl)        derivative at a fixed ray direction
//   g2 = g1 SDD / sqrt(SDD^2 + w^2)                         cone-angle weight
//   g3(psi_m) = g2 at w = w_k(g, psi_m)                      forward height rebinning (kappa lines)
//   g4 = sum_{n odd} 2/(pi n) (n dg / sin(n dg)) g3[j - n]   Hilbert transform, kernel 1/(pi sin)
//   g5(w) = g4 at the psi with w_k(g, psi) = w               backward height rebinning
//   f(x) += dl/(2 pi) sum_k g5(k, g*, w*) / v*   over views where x projects into the
//           Tam-Danielsson window (the PI interval of x)
//
// Helix: a(l) = (R cos l, R sin l, zs0 + hp (l - l0)), l_k = l0 + k dl. Detector: column j at fan
// angle g = (j - col_c) dg from e_v = -(cos l, sin l, 0) towards e_u = (-sin l, cos l, 0), row i
// at height w = (i - row_c) drow, SDD from the source. w_k(g, psi) = (SDD hp / R)(psi cos g +
// psi cot psi sin g), psi_m uniform on [-(pi/2 + am), pi/2 + am], am = largest |fan angle|.
//
// Layout: every stage is a dense [view][row][column] FP32 array in HBM (c2: 0.57 GB per stage,
// 1.1 GB for the kappa-line stages); the backprojector reads g5 through the quad layout of
// fc_backproject.hip. The Hilbert stage is the FFT row filter of fc_filter.hip (one convolution
// per kappa line), the other filter stages are elementwise gathers.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

struct KatGeom {
  double l0, dl, zs0, hp, R, SDD, dg, drow, col_c, row_c, pmax, dpsi;
  int nw, nh, nv, npsi;
};

__device__ __forceinline__ double kappa_w(const KatGeom& g, double gam, double psi) {
  double pc = fabs(psi) < 1e-12 ? 1.0 : psi / tan(psi);
  return g.SDD * g.hp / g.R * (psi * cos(gam) + pc * sin(gam));
}

// Linear interpolation along a row of nw values at column coordinate c, zero outside.
__device__ __forceinline__ float row_lerp(const float* row, int nw, float c) {
  float fc = floorf(c);
  int j0 = (int)fc;
  float t = c - fc, s = 0.0f;
  if (j0 >= 0 && j0 < nw) s += (1.0f - t) * row[j0];
  if (j0 + 1 >= 0 && j0 + 1 < nw) s += t * row[j0 + 1];
  return s;
}

// g2 = cone weight x fixed-direction derivative. One thread per (view, row, column).
template <int DT>
__global__ __launch_bounds__(256) void kat_derivative(const void* __restrict__ proj, KatGeom g,
                                                      float* __restrict__ g2) {
  int j = blockIdx.x * 256 + threadIdx.x;
  int i = blockIdx.y, k = blockIdx.z;
  if (j >= g.nw) return;
  size_t o = ((size_t)k * g.nh + i) * g.nw + j;
  if (k == 0 || k == g.nv - 1) {
    g2[o] = 0.0f;
    return;
  }
  const float shift = (float)(g.dl / g.dg);
  auto val = [&](int kk, float c) {
    float fc = floorf(c);
    int j0 = (int)fc;
    float t = c - fc, s = 0.0f;
    size_t rb = ((size_t)kk * g.nh + i) * g.nw;
    if (j0 >= 0 && j0 < g.nw) s += (1.0f - t) * load_elem<DT>(proj, rb + j0);
    if (j0 + 1 >= 0 && j0 + 1 < g.nw) s += t * load_elem<DT>(proj, rb + j0 + 1);
    return s;
  };
  float d = (val(k + 1, (float)j + shift) - val(k - 1, (float)j - shift)) * (float)(0.5 / g.dl);
  double w = ((double)i - g.row_c) * g.drow;
  g2[o] = d * (float)(g.SDD / sqrt(g.SDD * g.SDD + w * w));
}

// Tables: fwd[m][j] = detector row coordinate of kappa line m at column j;
// bwd[i][j] = (m + t) with w_k(g_j, psi_m..m+1) bracketing row i's height, or -1 outside.
__global__ __launch_bounds__(256) void kat_tables(KatGeom g, float* __restrict__ fwd,
                                                  float* __restrict__ bwd) {
  int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= g.nw) return;
  const double gam = ((double)j - g.col_c) * g.dg;
  for (int m = 0; m < g.npsi; ++m)
    fwd[(size_t)m * g.nw + j] = (float)(kappa_w(g, gam, -g.pmax + m * g.dpsi) / g.drow + g.row_c);
  const double wlo = kappa_w(g, gam, -g.pmax), whi = kappa_w(g, gam, g.pmax);
  for (int i = 0; i < g.nh; ++i) {
    double w = ((double)i - g.row_c) * g.drow;
    float r = -1.0f;
    if (w >= wlo && w <= whi) {
      int lo = 0, hi = g.npsi - 1;  // w_k is increasing in psi for every detector column
      while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (kappa_w(g, gam, -g.pmax + mid * g.dpsi) <= w) lo = mid; else hi = mid;
      }
      double wl = kappa_w(g, gam, -g.pmax + lo * g.dpsi);
      double wh = kappa_w(g, gam, -g.pmax + hi * g.dpsi);
      r = (float)((double)lo + (w - wl) / (wh - wl));
    }
    bwd[(size_t)i * g.nw + j] = r;
  }
}

// g3[k][m][j] = g2[k][row(m, j)][j], linear in the row, zero outside.
__global__ __launch_bounds__(256) void kat_forward_rebin(const float* __restrict__ g2, KatGeom g,
                                                         const float* __restrict__ fwd,
                                                         float* __restrict__ g3) {
  int j = blockIdx.x * 256 + threadIdx.x;
  int m = blockIdx.y, k = blockIdx.z;
  if (j >= g.nw) return;
  float r = fwd[(size_t)m * g.nw + j];
  float fr = floorf(r);
  int i0 = (int)fr;
  float t = r - fr, s = 0.0f;
  const float* col = g2 + (size_t)k * g.nh * g.nw + j;
  if (i0 >= 0 && i0 < g.nh) s += (1.0f - t) * col[(size_t)i0 * g.nw];
  if (i0 + 1 >= 0 && i0 + 1 < g.nh) s += t * col[(size_t)(i0 + 1) * g.nw];
  g3[((size_t)k * g.npsi + m) * g.nw + j] = s;
}

// The Hilbert transform of each kappa line, g4[j] = sum_{n odd} h[n] g3[j - n] with
// h[n] = 2/(pi n) * n dg / sin(n dg), is the FFT row filter's FC_FILTER_HILBERT kernel.
constexpr int kMaxCols = 2048;  // detector columns (FFT length <= 4096 in the Hilbert stage)

// g5[k][i][j] = g4 interpolated between kappa lines bracketing row i at column j.
__global__ __launch_bounds__(256) void kat_backward_rebin(const float* __restrict__ g4, KatGeom g,
                                                          const float* __restrict__ bwd,
                                                          float* __restrict__ g5) {
  int j = blockIdx.x * 256 + threadIdx.x;
  int i = blockIdx.y, k = blockIdx.z;
  if (j >= g.nw) return;
  float r = bwd[(size_t)i * g.nw + j];
  float v = 0.0f;
  if (r >= 0.0f) {
    int m0 = min((int)r, g.npsi - 2);
    float t = r - (float)m0;
    const float* col = g4 + (size_t)k * g.npsi * g.nw + j;
    v = (1.0f - t) * col[(size_t)m0 * g.nw] + t * col[(size_t)(m0 + 1) * g.nw];
  }
  g5[((size_t)k * g.nh + i) * g.nw + j] = v;
}

// PI-window backprojection. One thread per column (x, y) and kZR consecutive z voxels; per view
// and column: v* = R - x cos l - y sin l, g* = atan(u* / v*), the column coordinate and the
// Tam-Danielsson row bounds; per voxel: the row (linear in z), the window test, one quad gather.
// Each block visits only the views whose source height can see its z range through the window.
constexpr int kZR = 16;
struct KatBpArgs {
  float* vol;
  const float4* quad;    // g5 staged as bilinear quads, (nh+2) x (nw+2) per view
  const float4* vtab;    // per view: cos l, sin l, source z, 0
  KatGeom g;
  float ox, oy, oz;
  int nx, ny, nz;
  double zmax;  // largest |z - zs| a voxel can have inside the window (voxels)
};

__global__ __launch_bounds__(256) void kat_view_table(KatGeom g, float4* __restrict__ vt) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.nv) return;
  double l = g.l0 + k * g.dl;
  vt[k] = make_float4((float)cos(l), (float)sin(l), (float)(g.zs0 + k * g.dl * g.hp), 0.0f);
}

__global__ __launch_bounds__(256) void kat_bp_kernel(KatBpArgs a) {
  const KatGeom& g = a.g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
  const int y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
  const int z0 = blockIdx.z * kZR;
  const float X = (float)x + a.ox, Y = (float)y + a.oy, Z0 = (float)z0 + a.oz;
  // views whose source height lies within zmax of the block's z range (block-uniform)
  const double dz = g.hp * g.dl;
  int klo = (int)floor(((double)z0 + a.oz - a.zmax - g.zs0) / dz) - 1;
  int khi = (int)ceil(((double)z0 + a.oz + kZR - 1 + a.zmax - g.zs0) / dz) + 1;
  klo = max(klo, 0);
  khi = min(khi, g.nv - 1);
  const int QW = g.nw + 2, QH = g.nh + 2;
  const float wmax = (float)g.nw, hmax = (float)g.nh;
  const float Rf = (float)g.R, inv_dg = (float)(1.0 / g.dg), colc = (float)g.col_c;
  const float rowc = (float)g.row_c, inv_drow = (float)(1.0 / g.drow);
  const float sdd_drow = (float)(g.SDD / g.drow);
  const float c0r = (float)(g.SDD * g.hp / g.R / g.drow);
  const float hpi = 1.57079632679489662f;
  float acc[kZR];
#pragma unroll
  for (int q = 0; q < kZR; ++q) acc[q] = 0.0f;
  (void)inv_drow;
  for (int k = klo; k <= khi; ++k) {
    const float4 vt = a.vtab[k];
    const float vs = fmaf(-X, vt.x, fmaf(-Y, vt.y, Rf));  // R - x cos l - y sin l
    const float us = fmaf(-X, vt.y, Y * vt.x);            // -x sin l + y cos l
    const float t = us / vs;
    const float gam = atanf(t);
    const float sec = sqrtf(fmaf(t, t, 1.0f));              // 1 / cos g
    // w* / drow = SDD (z - zs) cos g / (v* drow): the row is linear in the voxel index
    const float dr = sdd_drow / (vs * sec);
    const float r0 = fmaf(Z0 - vt.z, dr, rowc);
    const float rtop = fmaf(c0r * (hpi - gam), sec, rowc);
    const float rbot = fmaf(-c0r * (hpi + gam), sec, rowc);
    const float wgt = 1.0f / vs;
    const float u = fminf(fmaxf(fmaf(gam, inv_dg, colc), -1.0f), wmax);
    const float fu = floorf(u), au = u - fu;
    const uint32_t iu = (uint32_t)((int)fu + 1);
    const float4* base = a.quad + (size_t)k * QW * QH;
#pragma unroll
    for (int q = 0; q < kZR; ++q) {
      const float rr = fmaf((float)q, dr, r0);
      const float w = (rr >= rbot && rr <= rtop) ? wgt : 0.0f;
      const float v = fminf(fmaxf(rr, -1.0f), hmax);
      const float fv = floorf(v), av = v - fv;
      const float4 p = base[(uint32_t)((int)fv + 1) * (uint32_t)QW + iu];
      const float top = lerp(p.x, p.y, au), bot = lerp(p.z, p.w, au);
      acc[q] = fmaf(w, lerp(top, bot, av), acc[q]);
    }
  }
  if (x < a.nx && y < a.ny) {
    const float s = (float)(g.dl / (2.0 * M_PI));
#pragma unroll
    for (int q = 0; q < kZR; ++q) {
      const int z = z0 + q;
      if (z < a.nz) a.vol[((size_t)z * a.ny + y) * a.nx + x] += s * acc[q];
    }
  }
}

}  // namespace
}  // namespace fc

using namespace fc;

extern "C" int fc_katsevich_dev(int device, const void* proj, int proj_dtype, int nw, int nh,
                                int nviews, double l0, double dl, double zs0, double hp, double R,
                                double SDD, double dgamma, double drow, double col_c,
                                double row_c, int npsi, float* vol, const int vol_shape[3],
                                const double origin[3], float* g5_out, void* stream_) {
  clear_error();
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  FC_CHECK_ARG(dtype_ok(proj_dtype), "unknown projection dtype %d", proj_dtype);
  FC_CHECK_ARG(nw > 1 && nw <= kMaxCols && nh > 1 && nh <= 65535 && nviews >= 3 &&
                   nviews <= 65535,
               "detector %d x %d / %d views out of range", nw, nh, nviews);
  FC_CHECK_ARG(npsi >= 3 && npsi <= 65535, "npsi must be in [3, 65535]");
  FC_CHECK_ARG(std::isfinite(R) && R > 0 && std::isfinite(SDD) && SDD > R, "need 0 < R < SDD");
  FC_CHECK_ARG(dl > 0 && hp > 0 && dgamma > 0 && drow > 0 && std::isfinite(l0) &&
                   std::isfinite(zs0) && std::isfinite(col_c) && std::isfinite(row_c),
               "bad helix / detector parameters");
  FC_CHECK_ARG(proj && vol && vol_shape && origin, "null pointer argument");
  FC_CHECK_ARG(vol_shape[0] > 0 && vol_shape[1] > 0 && vol_shape[2] > 0, "empty volume");
  const double am = fmax(fabs(0 - col_c), fabs(nw - 1 - col_c)) * dgamma;
  FC_CHECK_ARG(am < 1.2, "fan angle too large");
  // every voxel of the block must lie inside the helix cylinder
  double r2 = 0.0;
  for (int ax = 0; ax < 2; ++ax) {
    double lo = origin[ax], hi = origin[ax] + vol_shape[ax] - 1;
    double m = fmax(fabs(lo), fabs(hi));
    r2 += m * m;
  }
  const double rmax = sqrt(r2);
  FC_CHECK_ARG(rmax < 0.99 * R, "volume block reaches the source path");
  int rc = select_device(device);
  if (rc) return rc;

  KatGeom g;
  g.l0 = l0; g.dl = dl; g.zs0 = zs0; g.hp = hp; g.R = R; g.SDD = SDD; g.dg = dgamma;
  g.drow = drow; g.col_c = col_c; g.row_c = row_c;
  g.pmax = M_PI / 2 + am;
  g.dpsi = 2 * g.pmax / (npsi - 1);
  g.nw = nw; g.nh = nh; g.nv = nviews; g.npsi = npsi;

  const size_t plane = (size_t)nh * nw, kplane = (size_t)npsi * nw;
  void *A = nullptr, *B = nullptr, *T = nullptr;
  // A: g2, then g4 then g5 ... B: g3 ... (sizes: max of the stages)
  const size_t big = (size_t)nviews * (plane > kplane ? plane : kplane) * sizeof(float);
  if ((rc = acquire_ws(device, SLOT_KAT_A, big, stream, &A))) return rc;
  if ((rc = acquire_ws(device, SLOT_KAT_B, big, stream, &B))) return rc;
  const size_t tab = (kplane + plane + 4 * (size_t)nviews + 4) * sizeof(float);
  if ((rc = acquire_ws(device, SLOT_KAT_TAB, tab, stream, &T))) return rc;
  void* ftab = nullptr;
  if ((rc = acquire_ws(device, SLOT_FILTER, filter_table_bytes(nw), stream, &ftab))) return rc;
  float* fwd = static_cast<float*>(T);
  float* bwd = fwd + kplane;
  float4* vtab = reinterpret_cast<float4*>(fwd + ((kplane + plane + 3) & ~(size_t)3));

  int tok = prof_begin("katsevich_filter", stream);
  const int bx = (nw + 255) / 256;
  kat_tables<<<bx, 256, 0, stream>>>(g, fwd, bwd);
  float* g2 = static_cast<float*>(A);
  if (proj_dtype == FC_F32) kat_derivative<FC_F32><<<dim3(bx, nh, nviews), 256, 0, stream>>>(proj, g, g2);
  else if (proj_dtype == FC_BF16) kat_derivative<FC_BF16><<<dim3(bx, nh, nviews), 256, 0, stream>>>(proj, g, g2);
  else kat_derivative<FC_F16><<<dim3(bx, nh, nviews), 256, 0, stream>>>(proj, g, g2);
  float* g3 = static_cast<float*>(B);
  kat_forward_rebin<<<dim3(bx, npsi, nviews), 256, 0, stream>>>(g2, g, fwd, g3);
  float* g4 = static_cast<float*>(A);  // g2 is no longer needed
  // Hilbert transform along every kappa line: FFT convolution with the curved-detector kernel
  if ((rc = launch_row_filter(g3, FC_F32, g4, FC_F32, nw, npsi, (long long)npsi * nviews, nullptr,
                              FC_FILTER_HILBERT, dgamma, ftab, stream)))
    return rc;
  float* g5 = static_cast<float*>(B);  // g3 is no longer needed
  kat_backward_rebin<<<dim3(bx, nh, nviews), 256, 0, stream>>>(g4, g, bwd, g5);
  prof_end(tok, stream);
  FC_HIP(hipGetLastError());
  if (g5_out)
    FC_HIP(hipMemcpyAsync(g5_out, g5, (size_t)nviews * plane * sizeof(float),
                          hipMemcpyDeviceToDevice, stream));

  // stage g5 as quads and backproject over the PI windows
  void* Q = nullptr;
  const size_t qbytes = (size_t)(nw + 2) * (nh + 2) * nviews * 16;
  if ((rc = acquire_ws(device, SLOT_QUAD, qbytes, stream, &Q))) return rc;
  if ((rc = launch_stage_quads(g5, FC_F32, nw, nh, nviews, nullptr, Q, stream))) return rc;
  KatBpArgs a;
  a.vol = vol;
  a.quad = static_cast<const float4*>(Q);
  a.vtab = vtab;
  a.g = g;
  a.ox = (float)origin[0]; a.oy = (float)origin[1]; a.oz = (float)origin[2];
  kat_view_table<<<(nviews + 255) / 256, 256, 0, stream>>>(g, vtab);
  a.nx = vol_shape[0]; a.ny = vol_shape[1]; a.nz = vol_shape[2];
  // |z - zs| <= (hp/R) (pi/2 + am) / cos^2(am) * v*,  v* <= R + rmax
  a.zmax = hp / R * (M_PI / 2 + am) / (cos(am) * cos(am)) * (R + rmax) + 2.0 * drow;
  dim3 grid((a.nx + 15) / 16, (a.ny + 15) / 16, (a.nz + kZR - 1) / kZR);
  FC_CHECK_ARG(grid.y <= 65535 && grid.z <= 65535, "volume too large for the grid");
  tok = prof_begin("katsevich_bp", stream);
  kat_bp_kernel<<<grid, 256, 0, stream>>>(a);
  prof_end(tok, stream);
  FC_HIP(hipGetLastError());
  if ((rc = release_ws(device, SLOT_QUAD, stream))) return rc;
  if ((rc = release_ws(device, SLOT_KAT_A, stream))) return rc;
  if ((rc = release_ws(device, SLOT_KAT_B, stream))) return rc;
  if ((rc = release_ws(device, SLOT_FILTER, stream))) return rc;
  return release_ws(device, SLOT_KAT_TAB, stream);
}
