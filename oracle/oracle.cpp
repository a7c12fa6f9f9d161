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

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
