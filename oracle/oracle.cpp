// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement (C++ / OpenMP) of the reference cone-beam projector and backprojectors.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
// and only as the checker or as the timed CPU baseline, never as the product path.
//
// Conventions (all from the reference):
//   * a voxel with global index g along an axis of N voxels has coordinate g - N/2 (integer
//     division), coneBeamGPU.cu:597-599, 669-671;
//   * view geometry: source, detector centre, wStep and hStep vectors per view,
//     coneBeamGPU.cu:1672-1734;
//   * forward projection: sum of zero-extended trilinear samples at integer multiples of the step
//     from the source, coneBeamGPU.cu:96-176 (the reference step is 1 voxel with no factor; this
//     restatement takes a step h and multiplies the sum by h, which is the reference at h = 1);
//   * FDK / Katsevich voxel-driven backprojection with weight U^2 / U^1, coneBeamGPU.cu:258-420;
//   * unweighted general-geometry backprojection (kernel mode 0), coneBeamGPU.cu:178-255;
//   * per-view normalisation factor applied to the projection data for backprojection
//     (coneBeamGPU.cu:1081) and to the projector output (coneBeamGPU.cu:1165);
//   * outputs are accumulated (add-and-save), unit_coneBeamGPU.cu:81-86.
//
// Pinning: see tests/test_oracle.py (closed forms, scipy map_coordinates brute force, adjointness,
// linearity, symmetry).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return (1.0 / len(a)) * a; }
inline V3 load3(const double* p) { return {p[0], p[1], p[2]}; }
inline void store3(double* p, V3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

// Zero-extended trilinear interpolant of a block stored x fastest, at index-space point p.
inline double trilinear(const double* v, int nx, int ny, int nz, double px, double py,
                        double pz) {
  double fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
  long ix = (long)fx, iy = (long)fy, iz = (long)fz;
  double ax = px - fx, ay = py - fy, az = pz - fz;
  double s = 0.0;
  for (int dz = 0; dz < 2; ++dz) {
    long z = iz + dz;
    if (z < 0 || z >= nz) continue;
    double wz = dz ? az : 1.0 - az;
    for (int dy = 0; dy < 2; ++dy) {
      long y = iy + dy;
      if (y < 0 || y >= ny) continue;
      double wy = dy ? ay : 1.0 - ay;
      for (int dx = 0; dx < 2; ++dx) {
        long x = ix + dx;
        if (x < 0 || x >= nx) continue;
        double wx = dx ? ax : 1.0 - ax;
        s += wz * wy * wx * v[(z * ny + y) * (long)nx + x];
      }
    }
  }
  return s;
}

// Scatter counterpart of trilinear(): adds val * weight to each of the 8 neighbours.
inline void trilinear_scatter(double* v, int nx, int ny, int nz, double px, double py,
                              double pz, double val) {
  double fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
  long ix = (long)fx, iy = (long)fy, iz = (long)fz;
  double ax = px - fx, ay = py - fy, az = pz - fz;
  for (int dz = 0; dz < 2; ++dz) {
    long z = iz + dz;
    if (z < 0 || z >= nz) continue;
    double wz = dz ? az : 1.0 - az;
    for (int dy = 0; dy < 2; ++dy) {
      long y = iy + dy;
      if (y < 0 || y >= ny) continue;
      double wy = dy ? ay : 1.0 - ay;
      for (int dx = 0; dx < 2; ++dx) {
        long x = ix + dx;
        if (x < 0 || x >= nx) continue;
        double wx = dx ? ax : 1.0 - ax;
        v[(z * ny + y) * (long)nx + x] += val * wz * wy * wx;
      }
    }
  }
}

// Zero-extended bilinear interpolant of one nw x nh projection window (row = h, column = w).
inline double bilinear(const double* p, int nw, int nh, double w, double h) {
  double fw = std::floor(w), fh = std::floor(h);
  long iw = (long)fw, ih = (long)fh;
  double aw = w - fw, ah = h - fh;
  double s = 0.0;
  for (int dh = 0; dh < 2; ++dh) {
    long r = ih + dh;
    if (r < 0 || r >= nh) continue;
    double wh = dh ? ah : 1.0 - ah;
    for (int dw = 0; dw < 2; ++dw) {
      long c = iw + dw;
      if (c < 0 || c >= nw) continue;
      double ww = dw ? aw : 1.0 - aw;
      s += wh * ww * p[r * (long)nw + c];
    }
  }
  return s;
}

// Range of integer k for which src + k*h*u lies in the closed box [-1, n] (index space) on every
// axis; the zero-extended interpolant vanishes outside the open box. Returns false when empty.
// The range is widened by one step on each side: the zero extension makes the extra samples 0,
// so the sum is exact whatever rounding does at the faces.
inline bool ray_k_range(V3 s, V3 u, const int n[3], double h, long* k0, long* k1) {
  double tmin = -INFINITY, tmax = INFINITY;
  const double sv[3] = {s.x, s.y, s.z};
  const double uv[3] = {u.x, u.y, u.z};
  for (int a = 0; a < 3; ++a) {
    double lo = -1.0, hi = (double)n[a];
    if (std::fabs(uv[a]) < 1e-300) {
      if (sv[a] <= lo || sv[a] >= hi) return false;
      continue;
    }
    double t1 = (lo - sv[a]) / uv[a], t2 = (hi - sv[a]) / uv[a];
    tmin = std::max(tmin, std::min(t1, t2));
    tmax = std::min(tmax, std::max(t1, t2));
  }
  if (!(tmin <= tmax)) return false;
  *k0 = (long)std::ceil(tmin / h) - 1;
  *k1 = (long)std::floor(tmax / h) + 1;
  return true;
}

// Pixel (j, i) of a window whose pixel (0,0) has global index (woff, hoff):
// detCentre + (j + woff - Nw/2) * wStep + (i + hoff - Nh/2) * hStep   (coneBeamGPU.cu:157)
inline V3 pixel_pos(const double* det, const double* ws, const double* hs, long gw, long gh) {
  return load3(det) + (double)gw * load3(ws) + (double)gh * load3(hs);
}

}  // namespace

extern "C" {

// Per-view source position, detector centre, wStep and hStep in voxel coordinates, with the six
// misalignment parameters applied in order Dw, Dh, Dl, Dphi, Dtheta, Dpsi.
// Restates coneBeamGPU.cu:1672-1734 in FP64 (the reference computes it in FP32).
void orc_geometry(int nviews, const double* angles, const double* hmid, double D, double pw,
                  const double* mis, int nz_total, double* src, double* det, double* wstep,
                  double* hstep) {
  const double zshift = (double)(nz_total / 2);
  for (int t = 0; t < nviews; ++t) {
    double a = angles[t];
    V3 s = {-std::cos(a) * D, -std::sin(a) * D, hmid[t] - zshift};
    V3 d = {0.0, 0.0, hmid[t] - zshift};
    V3 hsv = {0.0, 0.0, pw};
    V3 wsv = {-pw * std::sin(a), pw * std::cos(a), 0.0};
    d = d + mis[0] * wsv;
    d = d + mis[1] * hsv;
    d = d + mis[2] * normalize(d - s);
    V3 whp = cross(wsv, hsv);
    hsv = len(hsv) * (std::cos(mis[3]) * normalize(hsv) + std::sin(mis[3]) * normalize(whp));
    whp = cross(wsv, hsv);
    wsv = len(wsv) * (std::cos(mis[4]) * normalize(wsv) + std::sin(mis[4]) * normalize(whp));
    whp = cross(wsv, hsv);
    V3 wnp = cross(wsv, whp);
    V3 hnp = cross(hsv, whp);
    V3 w2 = len(wsv) * (std::cos(mis[5]) * normalize(wsv) - std::sin(mis[5]) * normalize(wnp));
    V3 h2 = len(hsv) * (std::cos(mis[5]) * normalize(hsv) + std::sin(mis[5]) * normalize(hnp));
    store3(src + 3 * t, s);
    store3(det + 3 * t, d);
    store3(wstep + 3 * t, w2);
    store3(hstep + 3 * t, h2);
  }
}

// Forward projection (coneBeamGPU.cu:96-176, generalised to step h):
//   proj[t,i,j] += norm[t] * h * sum_k T(src_t + k*h*u_tij),  k over all integers,
// T = zero-extended trilinear interpolant of the local block vol[nz][ny][nx] whose voxel (0,0,0)
// has coordinate origin[]. Window pixel (j,i) of view t has global index (j+woff[t], i+hoff[t]).
void orc_project(const double* vol, int nx, int ny, int nz, const double* origin, double* proj,
                 int nw, int nh, int nviews, int Nw, int Nh, const int* woff, const int* hoff,
                 const double* src, const double* det, const double* wstep, const double* hstep,
                 double h, const double* norm) {
  const int n[3] = {nx, ny, nz};
  const V3 o = load3(origin);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int t = 0; t < nviews; ++t) {
    for (int i = 0; i < nh; ++i) {
      V3 s = load3(src + 3 * t);
      V3 sl = s - o;
      for (int j = 0; j < nw; ++j) {
        V3 P = pixel_pos(det + 3 * t, wstep + 3 * t, hstep + 3 * t, (long)j + woff[t] - Nw / 2,
                         (long)i + hoff[t] - Nh / 2);
        V3 u = normalize(P - s);
        long k0, k1;
        double sum = 0.0;
        if (ray_k_range(sl, u, n, h, &k0, &k1)) {
          for (long k = k0; k <= k1; ++k) {
            double kh = (double)k * h;
            sum += trilinear(vol, nx, ny, nz, sl.x + kh * u.x, sl.y + kh * u.y, sl.z + kh * u.z);
          }
        }
        proj[((long)t * nh + i) * nw + j] += norm[t] * h * sum;
      }
    }
  }
}

// Exact transpose of orc_project: vol += A^T proj (the matched adjoint the reference lacks).
// Each view is scattered into a private buffer, then the buffers are summed in a fixed order.
void orc_adjoint(double* vol, int nx, int ny, int nz, const double* origin, const double* proj,
                 int nw, int nh, int nviews, int Nw, int Nh, const int* woff, const int* hoff,
                 const double* src, const double* det, const double* wstep, const double* hstep,
                 double h, const double* norm) {
  const int n[3] = {nx, ny, nz};
  const V3 o = load3(origin);
  const long nvox = (long)nx * ny * nz;
  int nthreads = omp_get_max_threads();
  std::vector<std::vector<double>> part(nthreads, std::vector<double>());
#pragma omp parallel
  {
    int tid = omp_get_thread_num();
    std::vector<double>& acc = part[tid];
    acc.assign(nvox, 0.0);
#pragma omp for schedule(static)
    for (int t = 0; t < nviews; ++t) {
      V3 s = load3(src + 3 * t);
      V3 sl = s - o;
      for (int i = 0; i < nh; ++i) {
        for (int j = 0; j < nw; ++j) {
          double y = proj[((long)t * nh + i) * nw + j] * norm[t] * h;
          if (y == 0.0) continue;
          V3 P = pixel_pos(det + 3 * t, wstep + 3 * t, hstep + 3 * t,
                           (long)j + woff[t] - Nw / 2, (long)i + hoff[t] - Nh / 2);
          V3 u = normalize(P - s);
          long k0, k1;
          if (!ray_k_range(sl, u, n, h, &k0, &k1)) continue;
          for (long k = k0; k <= k1; ++k) {
            double kh = (double)k * h;
            trilinear_scatter(acc.data(), nx, ny, nz, sl.x + kh * u.x, sl.y + kh * u.y,
                              sl.z + kh * u.z, y);
          }
        }
      }
    }
  }
  for (int p = 0; p < nthreads; ++p) {
    const std::vector<double>& acc = part[p];
    if (acc.empty()) continue;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nvox; ++q) vol[q] += acc[q];
  }
}

// Number of samples the projector evaluates per view with the exact (unwidened) k range of the
// closed support box [-1, n]: the algorithmic sample count used for throughput figures.
void orc_sample_count(int nx, int ny, int nz, const double* origin, int nw, int nh, int nviews,
                      int Nw, int Nh, const int* woff, const int* hoff, const double* src,
                      const double* det, const double* wstep, const double* hstep, double h,
                      long long* counts) {
  const int n[3] = {nx, ny, nz};
  const V3 o = load3(origin);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < nviews; ++t) {
    long long c = 0;
    V3 s = load3(src + 3 * t);
    V3 sl = s - o;
    for (int i = 0; i < nh; ++i) {
      for (int j = 0; j < nw; ++j) {
        V3 P = pixel_pos(det + 3 * t, wstep + 3 * t, hstep + 3 * t, (long)j + woff[t] - Nw / 2,
                         (long)i + hoff[t] - Nh / 2);
        V3 u = normalize(P - s);
        long k0, k1;
        if (ray_k_range(sl, u, n, h, &k0, &k1)) {
          k0 += 1;  // undo the one-step widening
          k1 -= 1;
          if (k1 >= k0) c += (k1 - k0 + 1);
        }
      }
    }
    counts[t] = c;
  }
}

// FDK (power = 2) and Katsevich (power = 1) voxel-driven backprojection,
// coneBeamGPU.cu:258-420:
//   U = D / (D + x cos a + y sin a)
//   w = (y cos a - x sin a) * U / pw,   h = (z - zc) * U / pw      (pixels from index N/2)
//   vol(x,y,z) += sum_t norm[t] * U^power * B_t(w + Nw/2 - woff[t], h + Nh/2 - hoff[t])
// with B_t the zero-extended bilinear interpolant of window t and zc[t] = hMid[t] - Nz/2.
void orc_backproject_fdk(double* vol, int nx, int ny, int nz, const double* origin,
                         const double* proj, int nw, int nh, int nviews, int Nw, int Nh,
                         const int* woff, const int* hoff, const double* angles,
                         const double* zc, double D, double pw, const double* norm, int power) {
  std::vector<double> ca(nviews), sa(nviews);
  for (int t = 0; t < nviews; ++t) {
    ca[t] = std::cos(angles[t]);
    sa[t] = std::sin(angles[t]);
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < nz; ++iz) {
    for (int iy = 0; iy < ny; ++iy) {
      double z = iz + origin[2], y = iy + origin[1];
      for (int ix = 0; ix < nx; ++ix) {
        double x = ix + origin[0];
        double acc = 0.0;
        for (int t = 0; t < nviews; ++t) {
          double U = D / (D + x * ca[t] + y * sa[t]);
          double w = (y * ca[t] - x * sa[t]) * U / pw;
          double hh = (z - zc[t]) * U / pw;
          double wl = w + (double)(Nw / 2) - woff[t];
          double hl = hh + (double)(Nh / 2) - hoff[t];
          double wgt = power == 2 ? U * U : (power == 1 ? U : 1.0);
          acc += norm[t] * wgt * bilinear(proj + (long)t * nw * nh, nw, nh, wl, hl);
        }
        vol[((long)iz * ny + iy) * nx + ix] += acc;
      }
    }
  }
}

// Unweighted general-geometry backprojection (kernel mode 0), coneBeamGPU.cu:178-255: the
// source-to-voxel ray is intersected with the (possibly misaligned) detector plane and the
// detector coordinates follow from the 2x2 normal equations in wStep and hStep.
void orc_backproject_general(double* vol, int nx, int ny, int nz, const double* origin,
                             const double* proj, int nw, int nh, int nviews, int Nw, int Nh,
                             const int* woff, const int* hoff, const double* src,
                             const double* det, const double* wstep, const double* hstep,
                             const double* norm) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < nz; ++iz) {
    for (int iy = 0; iy < ny; ++iy) {
      for (int ix = 0; ix < nx; ++ix) {
        V3 X = {ix + origin[0], iy + origin[1], iz + origin[2]};
        double acc = 0.0;
        for (int t = 0; t < nviews; ++t) {
          V3 s = load3(src + 3 * t), d = load3(det + 3 * t);
          V3 ws = load3(wstep + 3 * t), hs = load3(hstep + 3 * t);
          V3 np = cross(ws, hs);
          double dist = dot(d - s, np) / dot(X - s, np);
          V3 I = dist * (X - s) + s - d;
          double wdw = dot(ws, ws), wdh = dot(ws, hs), hdh = dot(hs, hs);
          double idw = dot(I, ws), idh = dot(I, hs);
          double det2 = wdw * hdh - wdh * wdh;
          double w = (idw * hdh - idh * wdh) / det2;
          double hh = (idh * wdw - idw * wdh) / det2;
          double wl = w + (double)(Nw / 2) - woff[t];
          double hl = hh + (double)(Nh / 2) - hoff[t];
          acc += norm[t] * bilinear(proj + (long)t * nw * nh, nw, nh, wl, hl);
        }
        vol[((long)iz * ny + iy) * nx + ix] += acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Helical cone beam on a cylindrical detector centred on the source (no reference counterpart:
// the reference ABI represents only a flat detector with one pitch, fastcone.py:72, and applies
// no Katsevich filtering, coneBeamGPU.cu:258-338). Helix a(l) = (R cos l, R sin l, zs);
// e_v = -(cos l, sin l, 0), e_u = (-sin l, cos l, 0); column j looks along fan angle
// g = (j - col_c) * dg from e_v towards e_u; row i is at height w = (i - row_c) * drow on the
// detector at distance SDD from the source.
// ---------------------------------------------------------------------------------------------

// Forward projection h * sum_k T(a + k h u) along the helical curved-detector rays.
void orc_project_helical(const double* vol, int nx, int ny, int nz, const double* origin,
                         double* proj, int nw, int nh, int nviews, const double* lambdas,
                         const double* zs, double R, double SDD, double dg, double drow,
                         double col_c, double row_c, double h) {
  const int n[3] = {nx, ny, nz};
  const V3 o = load3(origin);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int t = 0; t < nviews; ++t) {
    for (int i = 0; i < nh; ++i) {
      const double l = lambdas[t];
      V3 a = {R * std::cos(l), R * std::sin(l), zs[t]};
      V3 ev = {-std::cos(l), -std::sin(l), 0.0}, eu = {-std::sin(l), std::cos(l), 0.0};
      V3 al = a - o;
      for (int j = 0; j < nw; ++j) {
        double g = (j - col_c) * dg, w = (i - row_c) * drow;
        V3 d = (SDD * std::cos(g)) * ev + (SDD * std::sin(g)) * eu + V3{0.0, 0.0, w};
        V3 u = normalize(d);
        long k0, k1;
        double sum = 0.0;
        if (ray_k_range(al, u, n, h, &k0, &k1)) {
          for (long k = k0; k <= k1; ++k) {
            double kh = (double)k * h;
            sum += trilinear(vol, nx, ny, nz, al.x + kh * u.x, al.y + kh * u.y, al.z + kh * u.z);
          }
        }
        proj[((long)t * nh + i) * nw + j] = h * sum;
      }
    }
  }
}

// kappa-line height on the detector: w_k(g, psi) = (SDD hp / R) (psi cos g + psi cot psi sin g)
static inline double kappa_w(double g, double psi, double SDD, double hp, double R) {
  double pc = std::fabs(psi) < 1e-12 ? 1.0 : psi / std::tan(psi);
  return SDD * hp / R * (psi * std::cos(g) + pc * std::sin(g));
}

// Katsevich reconstruction (Katsevich 2002; native curved-detector form of Noo, Pack & Heuscher
// 2003, Phys. Med. Biol. 48:3787), discretised as:
//   g1 = [G(k+1, g + dl) - G(k-1, g - dl)] / (2 dl)     derivative at a fixed ray direction
//        (G linear in the column, zero outside; g1 = 0 for the first and last view)
//   g2 = g1 * SDD / sqrt(SDD^2 + w^2)                    cone-angle weight
//   g3(psi_m) = g2 at w = w_k(g, psi_m), linear in the row   (forward height rebinning)
//   g4 = sum_{n odd} 2/(pi n) * (n dg / sin(n dg)) g3[j-n]     Hilbert kernel 1/(pi sin)
//   g5(w) = g4 at psi with w_k(g, psi) = w, linear between psi_m   (backward rebinning)
//   f(x) = sign/(2 pi) * sum_k dl * g5(k, g*, w*) / v*  over views whose projection of x lies in
//          the Tam-Danielsson window  -(SDD hp/R)(pi/2+g)/cos g <= w <= (SDD hp/R)(pi/2-g)/cos g
// with v* = R - x cos l - y sin l, g* = atan(u*/v*), w* = SDD (z - zs) cos g* / v*.
// psi_m uniform on [-(pi/2 + am), pi/2 + am], am = the largest |fan angle| of the detector.
void orc_katsevich(const double* proj, int nw, int nh, int nviews, const double* lambdas,
                   const double* zs, double R, double SDD, double dg, double drow, double col_c,
                   double row_c, double hp, double dl, int npsi, double sign, double* vol,
                   int nx, int ny, int nz, const double* origin, double* g5_out) {
  const long plane = (long)nh * nw;
  std::vector<double> g2((size_t)nviews * plane, 0.0);
  auto G = [&](int k, int i, double c) {  // linear in the column, zero-extended
    double fc = std::floor(c);
    long j0 = (long)fc;
    double t = c - fc, s = 0.0;
    const double* row = proj + ((long)k * nh + i) * nw;
    if (j0 >= 0 && j0 < nw) s += (1.0 - t) * row[j0];
    if (j0 + 1 >= 0 && j0 + 1 < nw) s += t * row[j0 + 1];
    return s;
  };
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < nviews - 1; ++k) {
    for (int i = 0; i < nh; ++i) {
      double w = (i - row_c) * drow;
      double cw = SDD / std::sqrt(SDD * SDD + w * w);
      for (int j = 0; j < nw; ++j) {
        double shift = dl / dg;
        double d = (G(k + 1, i, j + shift) - G(k - 1, i, j - shift)) / (2.0 * dl);
        g2[(size_t)k * plane + (long)i * nw + j] = d * cw;
      }
    }
  }
  double am = std::max(std::fabs(0 - col_c), std::fabs(nw - 1 - col_c)) * dg;
  double pmax = M_PI / 2 + am;
  double dpsi = 2 * pmax / (npsi - 1);
  // kappa-line tables per column
  std::vector<double> W((size_t)nw * npsi);
  for (int j = 0; j < nw; ++j)
    for (int m = 0; m < npsi; ++m)
      W[(size_t)j * npsi + m] = kappa_w((j - col_c) * dg, -pmax + m * dpsi, SDD, hp, R);
  // Hilbert kernel
  std::vector<double> hk(2 * nw - 1, 0.0);
  for (int n = -(nw - 1); n <= nw - 1; ++n) {
    if (n % 2 == 0) continue;
    double x = n * dg;
    hk[n + nw - 1] = 2.0 / (M_PI * n) * (x / std::sin(x));
  }
  std::vector<double> g5((size_t)nviews * plane, 0.0);
#pragma omp parallel
  {
    std::vector<double> g3((size_t)npsi * nw), g4((size_t)npsi * nw);
#pragma omp for schedule(static)
    for (int k = 1; k < nviews - 1; ++k) {
      const double* G2 = g2.data() + (size_t)k * plane;
      for (int m = 0; m < npsi; ++m)
        for (int j = 0; j < nw; ++j) {
          double r = W[(size_t)j * npsi + m] / drow + row_c;
          double fr = std::floor(r);
          long i0 = (long)fr;
          double t = r - fr, v = 0.0;
          if (i0 >= 0 && i0 < nh) v += (1 - t) * G2[i0 * nw + j];
          if (i0 + 1 >= 0 && i0 + 1 < nh) v += t * G2[(i0 + 1) * nw + j];
          g3[(size_t)m * nw + j] = v;
        }
      for (int m = 0; m < npsi; ++m)
        for (int j = 0; j < nw; ++j) {
          double acc = 0.0;
          for (int jj = 0; jj < nw; ++jj) acc += hk[j - jj + nw - 1] * g3[(size_t)m * nw + jj];
          g4[(size_t)m * nw + j] = acc;
        }
      for (int i = 0; i < nh; ++i) {
        double w = (i - row_c) * drow;
        for (int j = 0; j < nw; ++j) {
          const double* Wj = W.data() + (size_t)j * npsi;
          double v = 0.0;
          if (w >= Wj[0] && w <= Wj[npsi - 1]) {
            int lo = 0, hi = npsi - 1;
            while (hi - lo > 1) {
              int mid = (lo + hi) / 2;
              if (Wj[mid] <= w) lo = mid; else hi = mid;
            }
            double t = (w - Wj[lo]) / (Wj[hi] - Wj[lo]);
            v = (1 - t) * g4[(size_t)lo * nw + j] + t * g4[(size_t)hi * nw + j];
          }
          g5[(size_t)k * plane + (long)i * nw + j] = v;
        }
      }
    }
  }
  if (g5_out) std::memcpy(g5_out, g5.data(), sizeof(double) * g5.size());
  const double c0 = SDD * hp / R;
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < nz; ++iz) {
    for (int iy = 0; iy < ny; ++iy) {
      double z = iz + origin[2], y = iy + origin[1];
      for (int ix = 0; ix < nx; ++ix) {
        double x = ix + origin[0];
        double acc = 0.0;
        for (int k = 0; k < nviews; ++k) {
          double l = lambdas[k];
          double vs = R - x * std::cos(l) - y * std::sin(l);
          double us = -x * std::sin(l) + y * std::cos(l);
          double g = std::atan(us / vs);
          double cg = std::cos(g);
          double w = SDD * (z - zs[k]) * cg / vs;
          double wt = c0 * (M_PI / 2 - g) / cg, wb = -c0 * (M_PI / 2 + g) / cg;
          if (w < wb || w > wt) continue;
          double c = g / dg + col_c, r = w / drow + row_c;
          acc += bilinear(g5.data() + (size_t)k * plane, nw, nh, c, r) / vs;
        }
        vol[((long)iz * ny + iy) * nx + ix] += sign * dl / (2 * M_PI) * acc;
      }
    }
  }
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
