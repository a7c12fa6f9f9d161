// FP64 per-view geometry records (host side).
//
// The reference builds, per view, a source position, a detector centre and one-pixel step vectors
// in FP32 and applies six misalignment parameters (coneBeamGPU.cu:1672-1734). The same chain is
// computed here in FP64, then condensed into the record the kernels read: the source, the position
// of the window's pixel (0,0), the two pixel steps and a 3x4 projective matrix for the
// voxel-driven backprojectors (so no per-voxel ray/plane intersection is needed on the device).
#include <cmath>
#include <cstring>

#include "fc_internal.h"

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
inline double norm2(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return (1.0 / norm2(a)) * a; }
inline void put(double* p, V3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

// Projective matrix of a flat detector: a point X maps to the window pixel (u, v) where the ray
// from s through X meets the plane d0 + u*du + v*dv. Solving u*du + v*dv - lam*(X-s) = s - d0 by
// Cramer's rule gives, with n = du x dv and e = n.(d0 - s),
//   u = (X-s).((s-d0) x dv) / n.(X-s),  v = (X-s).(du x (s-d0)) / n.(X-s),  lam = e / n.(X-s),
// so the rows ((s-d0) x dv)/e, (du x (s-d0))/e, n/e applied to X - s give (u c, v c, c) with
// c = 1/lam, the inverse magnification (the FDK "U" of coneBeamGPU.cu:398 is 1/c).
void flat_matrix(V3 s, V3 d0, V3 du, V3 dv, double* M) {
  V3 n = cross(du, dv);
  double e = dot(n, d0 - s);
  V3 r0 = (1.0 / e) * cross(s - d0, dv);
  V3 r1 = (1.0 / e) * cross(du, s - d0);
  V3 r2 = (1.0 / e) * n;
  const V3 rows[3] = {r0, r1, r2};
  for (int r = 0; r < 3; ++r) {
    M[4 * r + 0] = rows[r].x;
    M[4 * r + 1] = rows[r].y;
    M[4 * r + 2] = rows[r].z;
    M[4 * r + 3] = -dot(rows[r], s);
  }
}

}  // namespace

extern "C" int fc_views_flat(int nviews, const double* angles, const double* hmid, double D,
                             double pw_w, double pw_h, const double* mis, int vol_dim_z,
                             const int proj_dim[2], const int* woff, const int* hoff,
                             double* out) {
  fc::clear_error();
  FC_CHECK_ARG(nviews >= 0, "nviews must be >= 0 (got %d)", nviews);
  FC_CHECK_ARG(nviews == 0 || (angles && hmid && out && proj_dim), "null geometry pointer");
  FC_CHECK_ARG(std::isfinite(D) && D > 0.0, "source distance must be > 0 (got %g)", D);
  FC_CHECK_ARG(std::isfinite(pw_w) && pw_w > 0.0 && std::isfinite(pw_h) && pw_h > 0.0,
               "pixel width must be > 0");
  double m[6] = {0, 0, 0, 0, 0, 0};
  if (mis) std::memcpy(m, mis, sizeof(m));
  for (int k = 0; k < 6; ++k) FC_CHECK_ARG(std::isfinite(m[k]), "misalignment %d not finite", k);
  const double zshift = (double)(vol_dim_z / 2);
  const int Nw = proj_dim ? proj_dim[0] : 0, Nh = proj_dim ? proj_dim[1] : 0;
  for (int t = 0; t < nviews; ++t) {
    const double a = angles[t];
    FC_CHECK_ARG(std::isfinite(a) && std::isfinite(hmid[t]), "view %d: non-finite angle/hMid", t);
    // coneBeamGPU.cu:1697-1700: aligned system, detector centre in the source plane on the axis.
    V3 s = {-std::cos(a) * D, -std::sin(a) * D, hmid[t] - zshift};
    V3 d = {0.0, 0.0, hmid[t] - zshift};
    V3 hs = {0.0, 0.0, pw_h};
    V3 ws = {-pw_w * std::sin(a), pw_w * std::cos(a), 0.0};
    // coneBeamGPU.cu:1701-1718: Dw, Dh, Dl shifts, then Dphi, Dtheta, Dpsi rotations.
    d = d + m[0] * ws;
    d = d + m[1] * hs;
    d = d + m[2] * unit(d - s);
    V3 whp = cross(ws, hs);
    hs = norm2(hs) * (std::cos(m[3]) * unit(hs) + std::sin(m[3]) * unit(whp));
    whp = cross(ws, hs);
    ws = norm2(ws) * (std::cos(m[4]) * unit(ws) + std::sin(m[4]) * unit(whp));
    whp = cross(ws, hs);
    V3 wnp = cross(ws, whp), hnp = cross(hs, whp);
    V3 w2 = norm2(ws) * (std::cos(m[5]) * unit(ws) - std::sin(m[5]) * unit(wnp));
    V3 h2 = norm2(hs) * (std::cos(m[5]) * unit(hs) + std::sin(m[5]) * unit(hnp));
    // Window pixel (0,0) has global index (woff, hoff); global pixel g sits at (g - N/2) steps
    // from the detector centre (coneBeamGPU.cu:157, 523-524).
    const double gw0 = (double)((woff ? woff[t] : 0) - Nw / 2);
    const double gh0 = (double)((hoff ? hoff[t] : 0) - Nh / 2);
    V3 d0 = d + gw0 * w2 + gh0 * h2;
    double* rec = out + (size_t)t * FC_VIEW_DOUBLES;
    std::memset(rec, 0, sizeof(double) * FC_VIEW_DOUBLES);
    put(rec + 0, s);
    put(rec + 3, d0);
    put(rec + 6, w2);
    put(rec + 9, h2);
    V3 n = cross(w2, h2);
    FC_CHECK_ARG(std::fabs(dot(n, d0 - s)) > 1e-12 * norm2(n) * norm2(d0 - s),
                 "view %d: source lies in the detector plane", t);
    flat_matrix(s, d0, w2, h2, rec + 12);
    rec[24] = (double)FC_DET_FLAT;
  }
  return FC_OK;
}

extern "C" int fc_views_cylinder(int nviews, const double* lambdas, const double* src_z,
                                 double R, double SDD, double dgamma, double drow,
                                 const double* col_c, const double* row_c, double* out) {
  fc::clear_error();
  FC_CHECK_ARG(nviews >= 0, "nviews must be >= 0 (got %d)", nviews);
  FC_CHECK_ARG(nviews == 0 || (lambdas && src_z && out), "null geometry pointer");
  FC_CHECK_ARG(std::isfinite(R) && R > 0 && std::isfinite(SDD) && SDD > R,
               "need 0 < R < SDD (got R=%g SDD=%g)", R, SDD);
  FC_CHECK_ARG(std::isfinite(dgamma) && dgamma > 0 && std::isfinite(drow) && drow > 0,
               "detector pitches must be > 0");
  for (int t = 0; t < nviews; ++t) {
    const double l = lambdas[t];
    FC_CHECK_ARG(std::isfinite(l) && std::isfinite(src_z[t]), "view %d: non-finite", t);
    // helix a(l) = (R cos l, R sin l, z); e_v = -(cos l, sin l, 0) points at the axis,
    // e_u = (-sin l, cos l, 0); a column at fan angle g looks along cos g e_v + sin g e_u
    V3 s = {R * std::cos(l), R * std::sin(l), src_z[t]};
    V3 ev = {-std::cos(l), -std::sin(l), 0.0};
    V3 eu = {-std::sin(l), std::cos(l), 0.0};
    double* rec = out + (size_t)t * FC_VIEW_DOUBLES;
    std::memset(rec, 0, sizeof(double) * FC_VIEW_DOUBLES);
    put(rec + 0, s);
    put(rec + 3, SDD * ev);
    put(rec + 6, SDD * eu);
    put(rec + 9, V3{0.0, 0.0, drow});
    rec[24] = (double)FC_DET_CYLINDER;
    rec[25] = col_c ? col_c[t] : 0.0;
    rec[26] = dgamma;
    rec[27] = row_c ? row_c[t] : 0.0;
    rec[28] = SDD;
    FC_CHECK_ARG(std::isfinite(rec[25]) && std::isfinite(rec[27]), "view %d: bad centre", t);
  }
  return FC_OK;
}
