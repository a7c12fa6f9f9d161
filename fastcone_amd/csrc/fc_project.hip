// Forward projector (ray marching, trilinear) and its exact transpose, for gfx950.
//
// Discretisation (coneBeamGPU.cu:96-176, generalised to step h): for each detector pixel the ray
// from the source is sampled at integer multiples k*h of the step, and
//   proj += norm * h * sum_k T(src + k h u),
// T being the trilinear interpolant of the block extended by zeros. The reference sums hardware
// texture fetches (low-precision fixed-point weights) over a fixed bounding-sphere sweep with an
// accumulated position; here each ray is clipped in FP64 to the exact support box [-1, n]^3 and
// positions are recomputed per sample with one FMA, and the interpolation weights are FP32.
//
// Volume staging ("gather layouts", chosen because HBM is plentiful and gathers are the cost):
//   OCTET: for every base voxel b in [-2, n] the 8 corner values v[b + {0,1}^3] are stored
//          contiguously (32 B in FP32, 16 B in BF16/FP16), so one trilinear sample is one
//          (FP32: two adjacent) 16-byte load. 8x the volume's bytes.
//   PLAIN: the volume with a 2-voxel zero halo; one sample is four dword-aligned 8-byte loads
//          of x-neighbour pairs. FP32 only for global gathers.
// The default projector (LDS-staged, below) copies bricks of the PLAIN staging, in any storage
// type, into LDS and gathers from there. The halo / padding makes every sample's corners
// in-bounds without per-tap checks.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

constexpr int kTile = 16;     // 16x16 pixel tiles, 8x8 pixels per wave
constexpr int kLdsCap = 6144;  // fp_lds_kernel's LDS brick budget, floats per block (24 KiB)

typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef float f32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u32x3_a4 __attribute__((ext_vector_type(3), aligned(4)));

enum FpLayout { FP_OCTET = 0, FP_PLAIN = 1, FP_LDS = 2 };

// FASTCONE_FP_LAYOUT selects the projector variant: "lds" (default: LDS-staged slab bricks over
// the plain volume), "octet" (global gathers of 8-corner records) or "plain" (global gathers).
int fp_layout_choice() {
  const char* e = std::getenv("FASTCONE_FP_LAYOUT");
  if (e && std::strcmp(e, "plain") == 0) return FP_PLAIN;
  if (e && std::strcmp(e, "octet") == 0) return FP_OCTET;
  return FP_LDS;
}

// ---- staging ---------------------------------------------------------------------------------

// Octets over base positions [-2, n] on each axis (padded index b + 2 in [0, n+2]).
template <int DT_IN>
__global__ __launch_bounds__(256) void stage_octets(const void* __restrict__ vol, int nx, int ny,
                                                    int nz, void* __restrict__ oct, int PX,
                                                    int PY) {
  int bx = blockIdx.x * 256 + threadIdx.x;
  int by = blockIdx.y, bz = blockIdx.z;
  if (bx >= PX) return;
  int x0 = bx - 2, y0 = by - 2, z0 = bz - 2;
  float v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int x = x0 + (c & 1), y = y0 + ((c >> 1) & 1), z = z0 + (c >> 2);
    bool in = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
    v[c] = in ? load_elem<DT_IN>(vol, ((size_t)z * ny + y) * nx + x) : 0.0f;
  }
  size_t o = ((size_t)bz * PY + by) * PX + bx;
  if constexpr (DT_IN == FC_F32) {
    float4* d = static_cast<float4*>(oct) + 2 * o;
    d[0] = make_float4(v[0], v[1], v[2], v[3]);
    d[1] = make_float4(v[4], v[5], v[6], v[7]);
  } else {
    uint4* d = static_cast<uint4*>(oct) + o;
    *d = make_uint4(pack2<DT_IN>(v[0], v[1]), pack2<DT_IN>(v[2], v[3]), pack2<DT_IN>(v[4], v[5]),
                    pack2<DT_IN>(v[6], v[7]));
  }
}

// Plain FP32 copy with a zero halo of 2 (padded dims n + 4).
template <int DT_IN>
__global__ __launch_bounds__(256) void stage_plain(const void* __restrict__ vol, int nx, int ny,
                                                   int nz, float* __restrict__ out, int PX,
                                                   int PY) {
  int px = blockIdx.x * 256 + threadIdx.x;
  int py = blockIdx.y, pz = blockIdx.z;
  if (px >= PX) return;
  int x = px - 2, y = py - 2, z = pz - 2;
  bool in = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
  out[((size_t)pz * PY + py) * PX + px] =
      in ? load_elem<DT_IN>(vol, ((size_t)z * ny + y) * nx + x) : 0.0f;
}

// Plain 16-bit copy with a zero halo of 2 (raw bits of the input storage type).
template <int DT_IN>
__global__ __launch_bounds__(256) void stage_plain16(const void* __restrict__ vol, int nx, int ny,
                                                     int nz, uint16_t* __restrict__ out, int PX,
                                                     int PY) {
  int px = blockIdx.x * 256 + threadIdx.x;
  int py = blockIdx.y, pz = blockIdx.z;
  if (px >= PX) return;
  int x = px - 2, y = py - 2, z = pz - 2;
  bool in = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
  out[((size_t)pz * PY + py) * PX + px] =
      in ? static_cast<const uint16_t*>(vol)[((size_t)z * ny + y) * nx + x] : (uint16_t)0;
}

// ---- projector -------------------------------------------------------------------------------

struct FpArgs {
  const void* data;  // staged volume
  float* proj;
  const double* views;
  const float* norm;
  double ox, oy, oz;
  double h;
  int nx, ny, nz;
  int PX, PY;  // staged pitch in base positions (octet) or elements (plain)
  int nw, nh, nviews;
  int tiles_w, tiles_per_view, group;
  int max_steps;
  int lds_cap;  // LDS brick budget in floats (<= kLdsCap)
  int pairs;    // 1: (value, difference) pair bricks where they fit; 0: plain values only
};

// Tile order: groups of `group` consecutive views; inside a group the same pixel tile of all
// its views is adjacent, so concurrently running blocks of one XCD gather the same voxels.
// The groups are dealt to the eight XCDs' contiguous block ranges by residue: the range at
// position x takes groups x, x + 8, x + 16, ..., so every XCD sees views from the whole scan.
// Views whose rays run along x fill their LDS bricks more slowly than views along y, and an XCD
// holding one contiguous eighth of a circular scan would finish its octant late or early.
// Returns the view t and the tile's first column / row (j0, i0); the thread's pixel is
// (j0 + (wave & 1) * 8 + (lane & 7), i0 + (wave >> 1) * 8 + (lane >> 3)).
__device__ __forceinline__ void fp_tile_origin(int tiles_w, int tiles_per_view, int group,
                                               int nviews, int* t, int* j0, int* i0) {
  int lb = xcd_remap(blockIdx.x, gridDim.x);
  int per_group = tiles_per_view * group;
  int p = lb / per_group;
  int r = lb - p * per_group;
  // Only the full groups are dealt (a partial last group keeps the last position, where the
  // grid has its smaller block range): position p among the nfull -> range x of eight
  // near-equal ranges, index i in it -> group x + 8 i.
  const int nfull = nviews / group;
  int g = p;
  if (p < nfull) {
    const int per = nfull >> 3, rem = nfull & 7;
    int x, i;
    if (p < rem * (per + 1)) {
      x = p / (per + 1);
      i = p - x * (per + 1);
    } else {
      const int pp = p - rem * (per + 1);
      x = rem + pp / per;
      i = pp - (x - rem) * per;
    }
    g = x + 8 * i;
  }
  int gsize = min(group, nviews - g * group);
  int tile = r / gsize;
  int v = r - tile * gsize;
  *t = g * group + v;
  int ty = tile / tiles_w, tx = tile - ty * tiles_w;
  *j0 = tx * kTile;
  *i0 = ty * kTile;
}

__device__ __forceinline__ void fp_tile(const FpArgs& a, int* t, int* j, int* i) {
  int j0, i0;
  fp_tile_origin(a.tiles_w, a.tiles_per_view, a.group, a.nviews, t, &j0, &i0);
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  *j = j0 + (wave & 1) * 8 + (lane & 7);
  *i = i0 + (wave >> 1) * 8 + (lane >> 3);
}

template <int DT, int LAYOUT>
__device__ __forceinline__ float trilinear_sample(const FpArgs& a, float qx, float qy, float qz) {
  float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  float ax = qx - fx, ay = qy - fy, az = qz - fz;
  uint32_t ix = (uint32_t)(int)fx, iy = (uint32_t)(int)fy, iz = (uint32_t)(int)fz;
  uint32_t o = (iz * (uint32_t)a.PY + iy) * (uint32_t)a.PX + ix;
  float v0, v1, v2, v3, v4, v5, v6, v7;
  if constexpr (LAYOUT == FP_OCTET) {
    if constexpr (DT == FC_F32) {
      const float4* p = static_cast<const float4*>(a.data) + 2 * (size_t)o;
      float4 lo = p[0], hi = p[1];
      v0 = lo.x; v1 = lo.y; v2 = lo.z; v3 = lo.w;
      v4 = hi.x; v5 = hi.y; v6 = hi.z; v7 = hi.w;
    } else {
      uint4 w = static_cast<const uint4*>(a.data)[o];
      v0 = lo16<DT>(w.x); v1 = hi16<DT>(w.x); v2 = lo16<DT>(w.y); v3 = hi16<DT>(w.y);
      v4 = lo16<DT>(w.z); v5 = hi16<DT>(w.z); v6 = lo16<DT>(w.w); v7 = hi16<DT>(w.w);
    }
  } else if constexpr (DT != FC_F32) {
    size_t sy = (size_t)a.PX, sz = (size_t)a.PX * a.PY;
    v0 = load_elem<DT>(a.data, o); v1 = load_elem<DT>(a.data, o + 1);
    v2 = load_elem<DT>(a.data, o + sy); v3 = load_elem<DT>(a.data, o + sy + 1);
    v4 = load_elem<DT>(a.data, o + sz); v5 = load_elem<DT>(a.data, o + sz + 1);
    v6 = load_elem<DT>(a.data, o + sz + sy); v7 = load_elem<DT>(a.data, o + sz + sy + 1);
  } else {
    const float* p = static_cast<const float*>(a.data) + o;
    size_t sy = (size_t)a.PX, sz = (size_t)a.PX * a.PY;
    f32x2_a4 r0 = *reinterpret_cast<const f32x2_a4*>(p);
    f32x2_a4 r1 = *reinterpret_cast<const f32x2_a4*>(p + sy);
    f32x2_a4 r2 = *reinterpret_cast<const f32x2_a4*>(p + sz);
    f32x2_a4 r3 = *reinterpret_cast<const f32x2_a4*>(p + sz + sy);
    v0 = r0.x; v1 = r0.y; v2 = r1.x; v3 = r1.y;
    v4 = r2.x; v5 = r2.y; v6 = r3.x; v7 = r3.y;
  }
  float c00 = lerp(v0, v1, ax), c10 = lerp(v2, v3, ax);
  float c01 = lerp(v4, v5, ax), c11 = lerp(v6, v7, ax);
  float c0 = lerp(c00, c10, ay), c1 = lerp(c01, c11, ay);
  return lerp(c0, c1, az);
}

template <int DT, int LAYOUT>
__global__ __launch_bounds__(256) void fp_kernel(FpArgs a) {
  int t, j, i;
  fp_tile(a, &t, &j, &i);
  if (t >= a.nviews || j >= a.nw || i >= a.nh) return;
  const double* V = a.views + (size_t)t * FC_VIEW_DOUBLES;
  RaySeg r = ray_segment(V, j, i, a.ox, a.oy, a.oz, a.nx, a.ny, a.nz, a.h, a.max_steps, 2.0f);
  float acc = 0.0f;
  for (int k = 0; k < r.nk; ++k) {
    float kf = (float)k;
    acc += trilinear_sample<DT, LAYOUT>(a, fmaf(kf, r.sx, r.px), fmaf(kf, r.sy, r.py),
                                        fmaf(kf, r.sz, r.pz));
  }
  float scale = (a.norm ? a.norm[t] : 1.0f) * (float)a.h;
  size_t idx = ((size_t)t * a.nh + i) * a.nw + j;
  a.proj[idx] += scale * acc;
}

// ---- LDS-staged projector ------------------------------------------------------------------------
//
// Rays are marched slab by slab along the volume axis most aligned with a tile's rays (the
// "dominant" axis). For each slab of S voxel planes the block stages, in LDS, the brick that
// contains every trilinear corner its rays' samples in that slab can touch, then each thread
// evaluates its ray's samples from LDS:
//   * the brick's extent in the two minor axes is the bounding box of the corner rays of the
//     block's tile(s) between the slab planes (the frustum slice), widened by one voxel each side;
//   * along the dominant axis it spans S + 3 planes, [lo - 1, lo + S + 1];
//   * LDS layout: dominant axis fastest with compile-time pitch S + 3, so the 8 corners of a
//     sample are two base addresses with immediate offsets {0, 1, S+3, S+4} (plain values) or
//     {0, S+3} (value / difference pairs);
//   * a ray's samples are split between slabs by integer k ranges computed from the shared slab
//     planes, so every sample is evaluated exactly once.
// Work is dealt in "quads": one 16x16 pixel tile in four consecutive views. Adjacent views' rays
// cross nearly the same voxels, so in "shared" mode each 512-thread block takes one 16x8 half of
// the tile in all four views (one wave per view and 8x8 quarter) and its bricks are the union of
// the four views' frustum slices over the half: one fill serves eight waves (fp_quad_kernel, two
// blocks per quad, 8 waves per SIMD, bricks up to 38 KiB). A half shares its bricks between two
// quarters, which costs 30 % fewer brick elements per ray sample than a brick per quarter. When
// the views are far apart (sparse angular sampling) the union is wide, and the quad is left to
// fp_lds_kernel, where each 256-thread block takes the whole 16x16 tile of one view (6 waves per
// SIMD, 24 KiB bricks). fp_quad_kernel decides per quad from the brick elements of both layouts
// summed over every fourth slab (a brick over its budget counting 4x; a quad whose grid rays do
// not all advance along the dominant axis always goes to fp_lds_kernel) and records the decision
// in a byte per
// quad that fp_lds_kernel, launched next on the same stream, reads.
// Bricks larger than the LDS budget (or tiles whose rays do not all advance along the dominant
// axis) fall back to gathering from the staged volume in global memory, with the same arithmetic.

constexpr int kQuad = 4;          // views per quad
// fp_quad_kernel's LDS brick budget, floats (38 KiB): with the slab table and the quad grid a
// block takes 40528 bytes, so four 512-thread blocks (8 waves per SIMD) fill the CU's 160 KiB
constexpr int kHalfCap = 9728;
constexpr int kSlabTab = 64;      // slab boxes computed per table refill
constexpr uint8_t kQuadShared = 1;

// Frustum slice boxes. Each view's 16x16 tile is described by the rays through a 4x4 grid of
// pixels, columns j0 + {0, 7, 8, 15} and rows i0 + {0, 7, 8, 15} (clamped to the detector): their
// minor / dominant direction ratios (r1, r2), plus the view's source in the padded frame and its
// interpolation margin. A sub-tile's four corner rays are the grid points (cj, ci) with
// cj in {cj0, cj0 + 1 + cwj}, ci in {ci0, ci0 + 1 + cwi}: cwj = cwi = 2 for the whole tile
// (cj0 = ci0 = 0), cwj = 2, cwi = 0 for a 16x8 half (cj0 = 0, ci0 in {0, 2}).
struct QuadGrid {
  float r1[kQuad][16], r2[kQuad][16];
  float src[kQuad][4];  // dominant, minor 1, minor 2, margin
};

// The block's dominant axis (of the tile-centre ray of view V), the two minor axes (minor axis 1
// is x whenever x is not dominant, so brick columns load coalesced), the padded extents and
// strides along them, and the slab count.
struct SlabAxes {
  int ax, a1, a2;
  float dir;
  int PD, P1, P2, nslabs;
  uint32_t gsd, gs1, gs2;
};

template <int S>
__device__ __forceinline__ SlabAxes slab_axes(const FpArgs& a, const double* V, int j0, int i0) {
  SlabAxes X;
  double cdx, cdy, cdz;
  ray_dir(V, fmin((double)j0 + 7.5, (double)(a.nw - 1)), fmin((double)i0 + 7.5, (double)(a.nh - 1)),
          &cdx, &cdy, &cdz);
  X.ax = 0;
  double best = fabs(cdx);
  if (fabs(cdy) > best) { X.ax = 1; best = fabs(cdy); }
  if (fabs(cdz) > best) X.ax = 2;
  X.dir = (X.ax == 0 ? cdx : (X.ax == 1 ? cdy : cdz)) > 0.0 ? 1.0f : -1.0f;
  X.a1 = X.ax == 0 ? 1 : 0;
  X.a2 = X.ax == 2 ? 1 : 2;
  const int PZ = a.nz + 4;
  X.PD = X.ax == 0 ? a.PX : (X.ax == 1 ? a.PY : PZ);
  X.P1 = X.a1 == 0 ? a.PX : a.PY;
  X.P2 = X.a2 == 1 ? a.PY : PZ;
  const uint32_t sy = (uint32_t)a.PX, sz = (uint32_t)a.PX * (uint32_t)a.PY;
  X.gsd = X.ax == 0 ? 1u : (X.ax == 1 ? sy : sz);
  X.gs1 = X.a1 == 0 ? 1u : sy;
  X.gs2 = X.a2 == 1 ? sy : sz;
  X.nslabs = (X.PD - 4 + 1) / S + 1;  // slabs [1 + b S, 1 + (b+1) S) cover [1, n + 2]
  return X;
}

// One grid ray (view record V, pixel (jc, ic)) into G.r1/r2[v][p]; point 0 also stores the view's
// source and margin. Returns false if the ray does not advance along the dominant axis: the slice
// of a tile's frustum between two slab planes is the hull of its corner rays' intersections only
// if every ray of the tile crosses the planes forwards; a flat tile's directions are affine in
// (j, i), so it suffices that each corner ray does (a corner ray that misses the volume is not
// checked by its thread's own test). 0.2 matches that per-ray test.
__device__ __forceinline__ bool grid_ray(QuadGrid& G, const FpArgs& a, const SlabAxes& X,
                                         const double* V, int v, int p, double jc, double ic) {
  auto pickd = [](int k, double x, double y, double z) { return k == 0 ? x : (k == 1 ? y : z); };
  double dx, dy, dz;
  ray_dir(V, jc, ic, &dx, &dy, &dz);
  const double dd = pickd(X.ax, dx, dy, dz);
  const double rd = 1.0 / dd;
  G.r1[v][p] = (float)(pickd(X.a1, dx, dy, dz) * rd);
  G.r2[v][p] = (float)(pickd(X.a2, dx, dy, dz) * rd);
  if (p == 0) {
    const double sx = V[0] - a.ox + 2.0, sy = V[1] - a.oy + 2.0, sz = V[2] - a.oz + 2.0;
    G.src[v][0] = (float)pickd(X.ax, sx, sy, sz);
    G.src[v][1] = (float)pickd(X.a1, sx, sy, sz);
    G.src[v][2] = (float)pickd(X.a2, sx, sy, sz);
    G.src[v][3] = V[24] == (double)FC_DET_CYLINDER ? 2.0f : 1.0f;
  }
  return dd * (double)X.dir / sqrt(dx * dx + dy * dy + dz * dz) > 0.2;
}

// Union over views v0 .. v0 + nv - 1 of the sub-tile's frustum slice between slab m's planes
// (traversal order), widened by the interpolation margin and clipped to the padded volume:
// (e1lo, e2lo, E1, E2).
template <int S>
__device__ __forceinline__ int4 slab_box(const QuadGrid& G, const SlabAxes& X, int v0, int nv,
                                         int cj0, int ci0, int cwj, int cwi, int m) {
  const int b = X.dir > 0.0f ? m : X.nslabs - 1 - m;
  const float lo = 1.0f + (float)(b * S), hi = lo + (float)S;
  float mn1 = 3e38f, mx1 = -3e38f, mn2 = 3e38f, mx2 = -3e38f, mg = 0.0f;
  for (int v = v0; v < v0 + nv; ++v) {
    const float tlo = lo - G.src[v][0], thi = hi - G.src[v][0];
    const float s1 = G.src[v][1], s2 = G.src[v][2];
    mg = fmaxf(mg, G.src[v][3]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int p = (ci0 + (c >> 1) * (1 + cwi)) * 4 + cj0 + (c & 1) * (1 + cwj);
      const float q1 = G.r1[v][p], q2 = G.r2[v][p];
      float u1 = fmaf(tlo, q1, s1), w1 = fmaf(thi, q1, s1);
      float u2 = fmaf(tlo, q2, s2), w2 = fmaf(thi, q2, s2);
      mn1 = fminf(mn1, fminf(u1, w1)); mx1 = fmaxf(mx1, fmaxf(u1, w1));
      mn2 = fminf(mn2, fminf(u2, w2)); mx2 = fmaxf(mx2, fmaxf(u2, w2));
    }
  }
  // clamp in float first (rays nearly parallel to a slab give huge values)
  const float P1f = (float)X.P1, P2f = (float)X.P2;
  mn1 = fminf(fmaxf(mn1, -4.0f), P1f + 4.0f); mx1 = fminf(fmaxf(mx1, -4.0f), P1f + 4.0f);
  mn2 = fminf(fmaxf(mn2, -4.0f), P2f + 4.0f); mx2 = fminf(fmaxf(mx2, -4.0f), P2f + 4.0f);
  const int e1lo = max((int)floorf(mn1 - mg), 0);
  const int e1hi = min((int)floorf(mx1 + mg) + 1, X.P1 - 1);
  const int e2lo = max((int)floorf(mn2 - mg), 0);
  const int e2hi = min((int)floorf(mx2 + mg) + 1, X.P2 - 1);
  return make_int4(e1lo, e2lo, e1hi - e1lo + 1, e2hi - e2lo + 1);
}

// LDS bricks hold (value, forward difference) pairs along the dominant axis whenever the pairs
// fit the block's LDS budget: element (d, m1, m2) is the float2 (f(d), f(d + 1) - f(d)), so the
// first interpolation step of a sample is one FMA per corner pair, f + a (f1 - f): the operations
// of lerp(f, f1, a) in the same order (the subtraction done once, at the fill), so a sample's
// value is bit-identical to interpolating plain values, for four fewer VALU instructions per
// sample. A brick whose pairs would not fit holds plain values (one float per element) instead.
// Rows hold DP elements (DP - 1 are sampled as pairs; the last pair's difference is never read),
// planes E1 rows.
//
// Plane pitch (elements) of an LDS brick of E1 rows of DP elements: DP E1 rounded up to 8 mod 16.
// Plain bricks: consecutive planes start 8 or 24 of the 32 four-byte banks apart; a replay of the
// lanes' addresses (8 x 8 rays per wave, 0.75-voxel spacing, the eight corner reads of each
// sample) on 32 banks puts this among the lowest conflict counts of all plane offsets: 1.44 bank
// passes per ideal pass against 1.97 unpadded. Pair bricks: consecutive planes start 16 banks
// apart; timed on the same lane layout (ds_read2_b64 of each plane's two corner pairs), the 16
// plane offsets mod 16 pairs differ by up to 1.6x and 8 is the fastest.
__host__ __device__ constexpr uint32_t brick_plane(uint32_t DP, uint32_t E1) {
  return DP * E1 + ((8u - DP * E1) & 15u);
}

// Elements (pairs or floats) of an E1 x E2 brick.
__host__ __device__ constexpr uint32_t brick_elems(uint32_t DP, uint32_t E1, uint32_t E2) {
  return brick_plane(DP, E1) * E2;
}

// Element e of a brick: the pair (f0, f1 - f0), or the plain value f0.
template <bool PAIR>
__device__ __forceinline__ void put_elem(float* brick, uint32_t e, float f0, float f1) {
  if constexpr (PAIR) {
    reinterpret_cast<float2*>(brick)[e] = make_float2(f0, f1 - f0);
  } else {
    brick[e] = f0;
  }
}

// Trilinear sample from a brick laid out dominant axis fastest (pitch DP elements), then minor
// axis 1 (rows), then minor axis 2 (planes of PSu elements = zbytes bytes). `base` is the brick
// address minus the brick origin's offset, so the corner of a sample at brick-frame-relative
// coordinates (qd, q1, q2) >= 0 is element qd + DP q1 + PSu q2 (PSu wave-uniform). Truncation is
// floor for q >= 0; the 24-bit multiply-adds keep the index math full rate. Pair bricks: each
// plane's two rows are one ds_read2_b64 (offsets 0 and DP pairs).
template <int DP, bool PAIR>
__device__ __forceinline__ float lds_trilinear(const char* base, uint32_t PSu, uint32_t zbytes,
                                               float qd, float q1, float q2) {
  const uint32_t id = (uint32_t)(int)qd, i1 = (uint32_t)(int)q1, i2 = (uint32_t)(int)q2;
  const float ad = __builtin_amdgcn_fractf(qd), a1f = __builtin_amdgcn_fractf(q1),
              a2f = __builtin_amdgcn_fractf(q2);
  const uint32_t idx = mad_u24_s(i2, PSu, mad_u24(i1, (uint32_t)DP, id));
  float c00, c10, c01, c11;
  if constexpr (PAIR) {
    const float2* L0 = reinterpret_cast<const float2*>(base + 8u * idx);
    const float2* L1 = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(L0) + zbytes);
    const float2 e00 = L0[0], e10 = L0[DP], e01 = L1[0], e11 = L1[DP];
    c00 = fmaf(ad, e00.y, e00.x); c10 = fmaf(ad, e10.y, e10.x);
    c01 = fmaf(ad, e01.y, e01.x); c11 = fmaf(ad, e11.y, e11.x);
  } else {
    const float* L0 = reinterpret_cast<const float*>(base + 4u * idx);
    const float* L1 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(L0) + zbytes);
    c00 = lerp(L0[0], L0[1], ad); c10 = lerp(L0[DP], L0[DP + 1], ad);
    c01 = lerp(L1[0], L1[1], ad); c11 = lerp(L1[DP], L1[DP + 1], ad);
  }
  const float c0 = lerp(c00, c10, a1f), c1 = lerp(c01, c11, a1f);
  return lerp(c0, c1, a2f);
}

// The four corner pairs of a sample's cell in a pair brick: rows i1 and i1 + 1 of planes i2 and
// i2 + 1 (element idx = id + DP i1 + PSu i2 and its neighbours), and the trilinear value at the
// cell fractions (ad, a1f, a2f) — the arithmetic of lds_trilinear<DP, true>.
struct CellPairs {
  float2 e00, e10, e01, e11;
};

template <int DP>
__device__ __forceinline__ void load_cell(const char* base, uint32_t idx, uint32_t zbytes,
                                          CellPairs& c) {
  const float2* L0 = reinterpret_cast<const float2*>(base + 8u * idx);
  const float2* L1 = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(L0) + zbytes);
  c.e00 = L0[0]; c.e10 = L0[DP]; c.e01 = L1[0]; c.e11 = L1[DP];
}

__device__ __forceinline__ float cell_trilinear(const CellPairs& c, float ad, float a1f,
                                                float a2f) {
  const float c00 = fmaf(ad, c.e00.y, c.e00.x), c10 = fmaf(ad, c.e10.y, c.e10.x);
  const float c01 = fmaf(ad, c.e01.y, c.e01.x), c11 = fmaf(ad, c.e11.y, c.e11.x);
  const float c0 = lerp(c00, c10, a1f), c1 = lerp(c01, c11, a1f);
  return lerp(c0, c1, a2f);
}

template <bool B>
struct LayoutTag {
  static constexpr bool pair = B;
};

// Sum of this thread's ray samples, slab by slab through LDS bricks whose boxes come from G
// (views v0 .. v0 + nv - 1, sub-tile (cj0, ci0, cwj, cwi)), or from global memory when
// `fallback`. NT threads per block.
template <int DT, int S, int NT>
__device__ __forceinline__ float march(const FpArgs& a, float* brick, int4* tab,
                                       const QuadGrid& G, const SlabAxes& X, int v0, int nv,
                                       int cj0, int ci0, int cwj, int cwi, const RaySeg& r,
                                       bool fallback, int lds_cap) {
  constexpr int DP = S + 3;  // dominant-axis pitch of the LDS brick
  constexpr int NW = NT / 64;  // waves per block
  float acc = 0.0f;
  if (fallback) {
    for (int k = 0; k < r.nk; ++k) {
      float kf = (float)k;
      acc += trilinear_sample<DT, FP_PLAIN>(a, fmaf(kf, r.sx, r.px), fmaf(kf, r.sy, r.py),
                                            fmaf(kf, r.sz, r.pz));
    }
    return acc;
  }
  auto pick = [](int k, float x, float y, float z) { return k == 0 ? x : (k == 1 ? y : z); };
  const float pd0 = pick(X.ax, r.px, r.py, r.pz), p10 = pick(X.a1, r.px, r.py, r.pz),
              p20 = pick(X.a2, r.px, r.py, r.pz);
  const float sd = pick(X.ax, r.sx, r.sy, r.sz), s1 = pick(X.a1, r.sx, r.sy, r.sz),
              s2 = pick(X.a2, r.sx, r.sy, r.sz);
  const float inv_sd = 1.0f / sd;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int kprev;  // first k of the next slab: k >= its entry plane
  {
    const int b0 = X.dir > 0.0f ? 0 : X.nslabs - 1;
    const float lo0 = 1.0f + (float)(b0 * S);
    const float Pin0 = X.dir > 0.0f ? lo0 : lo0 + (float)S;
    kprev = min(max((int)ceilf((Pin0 - pd0) * inv_sd), 0), r.nk);
  }
  for (int m = 0; m < X.nslabs; ++m) {
    if (m % kSlabTab == 0) {
      __syncthreads();  // the previous table is no longer read
      const int mm = m + (int)threadIdx.x;
      if (threadIdx.x < kSlabTab && mm < X.nslabs)
        tab[threadIdx.x] = slab_box<S>(G, X, v0, nv, cj0, ci0, cwj, cwi, mm);
      __syncthreads();
    }
    // the box is block-uniform: keep it (and everything derived from it) in scalar registers
    const int4 Bl = tab[m % kSlabTab];
    const int4 B =
        make_int4(__builtin_amdgcn_readfirstlane(Bl.x), __builtin_amdgcn_readfirstlane(Bl.y),
                  __builtin_amdgcn_readfirstlane(Bl.z), __builtin_amdgcn_readfirstlane(Bl.w));
    const int E1 = B.z, E2 = B.w;
    const int b = X.dir > 0.0f ? m : X.nslabs - 1 - m;
    const float lo = 1.0f + (float)(b * S), hi = lo + (float)S;
    const float Pout = X.dir > 0.0f ? hi : lo;
    // slab m's entry plane is slab m - 1's exit plane (the same float), so its first k is the
    // previous slab's end
    const int kin = kprev;
    const int kout = min(max((int)ceilf((Pout - pd0) * inv_sd), 0), r.nk);
    kprev = kout;
    if (E1 <= 0 || E2 <= 0) continue;  // the frustum misses the block here
    const int dlo = (int)lo - 1;
    const uint32_t PS = brick_plane(DP, (uint32_t)E1), pad = PS - (uint32_t)(DP * E1);
    const int NE = (int)brick_elems(DP, (uint32_t)E1, (uint32_t)E2);
    if (NE > lds_cap) {
      for (int k = kin; k < kout; ++k) {
        float kf = (float)k;
        acc += trilinear_sample<DT, FP_PLAIN>(a, fmaf(kf, r.sx, r.px), fmaf(kf, r.sy, r.py),
                                              fmaf(kf, r.sz, r.pz));
      }
      continue;
    }
    // pairs when they fit the budget, plain values otherwise; block-uniform
    auto slab = [&](auto tag) {
      constexpr bool PAIR = decltype(tag)::pair;
      __syncthreads();  // the previous brick is no longer read
      // Fill: element (d, m1, m2) from the DP values d = 0 .. DP - 1 that each row loads along the
      // dominant axis (a pair also takes value d + 1; the row's last pair, never sampled, gets
      // difference 0 or the next value's). Dominant axis x (contiguous in HBM and in the brick):
      // lanes keep one offset along x and walk rows, so consecutive lanes read one row's
      // consecutive x and an element costs one row decomposition. Dominant axis y or z: one
      // thread per (minor1, minor2) column, minor axis 1 being x so a wave's loads stay
      // coalesced, and the DP elements along the dominant axis unrolled (LDS element d + DP c,
      // pitch DP odd so the writes of consecutive columns spread over all banks). Planes past the
      // padded block (last slab) read its last plane and are never sampled.
      // block-uniform: plain arithmetic keeps it on the scalar unit
      const uint32_t gbase =
          (uint32_t)dlo * X.gsd + (uint32_t)B.x * X.gs1 + (uint32_t)B.y * X.gs2;
      const uint32_t dmax = (uint32_t)(X.PD - 1 - dlo);  // >= 2: dlo = b S <= PD - 3
      if (DT == FC_F32 && X.gsd == 1u && dmax >= (uint32_t)(DP - 1)) {
        // FP32, dominant axis x, every row whole: a row's DP elements are kSeg four-element
        // segments starting at 4 s (the last at DP - 4, overlapping its neighbour: the shared
        // elements are written twice with the same value), one 16-byte load per lane, plus for
        // pairs the value after the segment (the last segment's fourth pair is the row's last: it
        // reloads its own value). Lanes 0 .. 3 kRowsW - 1 of each wave: segment lane % kSeg of
        // row phase lane / kSeg.
        constexpr int kSeg = (DP + 3) / 4, kRowsW = 64 / kSeg, kRowsB = NW * kRowsW;
        // floor((r + 0.5) / E1) is exact for r < 2^13 with a 1-ulp reciprocal: (r + 0.5) / E1 lies
        // at least 0.5 / E1 from an integer, and the two roundings move it by less than 2^-10 / E1
        const float invE1 = __builtin_amdgcn_rcpf((float)E1);
        if (lane < kSeg * kRowsW) {
          const uint32_t rr = (uint32_t)(((float)lane + 0.5f) * (1.0f / (float)kSeg));
          const uint32_t sg = (uint32_t)lane - mul_u24(rr, (uint32_t)kSeg);
          const uint32_t d0 = min(4u * sg, (uint32_t)(DP - 4));
          const uint32_t d4 = min(4u, (uint32_t)(DP - 1) - d0);
          const float* src = static_cast<const float*>(a.data) + (gbase + d0);
          const int nrow = E1 * E2;
#pragma unroll 1
          for (int rw = kRowsW * wave + (int)rr; rw < nrow; rw += kRowsB) {
            const uint32_t m2 = (uint32_t)(((float)rw + 0.5f) * invE1);
            const uint32_t m1 = (uint32_t)rw - mul_u24_s(m2, (uint32_t)E1);
            const uint32_t o = mad_u24_s(m1, X.gs1, mul_u24_s(m2, X.gs2));
            const f32x4_a4 v = *reinterpret_cast<const f32x4_a4*>(src + o);
            const float v4 = PAIR ? src[o + d4] : 0.0f;
            const uint32_t e = mad_u24_s(m2, pad, mad_u24((uint32_t)rw, (uint32_t)DP, d0));
            put_elem<PAIR>(brick, e, v.x, v.y);
            put_elem<PAIR>(brick, e + 1u, v.y, v.z);
            put_elem<PAIR>(brick, e + 2u, v.z, v.w);
            put_elem<PAIR>(brick, e + 3u, v.w, v4);
          }
        }
      } else if (DT != FC_F32 && X.gsd == 1u && dmax >= (uint32_t)(DP + 1)) {
        // 16-bit storage, dominant axis x, two elements of slack after every row: the same
        // four-element segments, each from three aligned dwords (the segment starts on either half
        // of the first) shifted into place with alignbit (pairs: and the fifth value from the
        // third dword), converted at the LDS store
        constexpr int kSeg = (DP + 3) / 4, kRowsW = 64 / kSeg, kRowsB = NW * kRowsW;
        const float invE1 = __builtin_amdgcn_rcpf((float)E1);  // exact as above
        if (lane < kSeg * kRowsW) {
          const uint32_t rr = (uint32_t)(((float)lane + 0.5f) * (1.0f / (float)kSeg));
          const uint32_t sg = (uint32_t)lane - mul_u24(rr, (uint32_t)kSeg);
          const uint32_t d0 = min(4u * sg, (uint32_t)(DP - 4));
          const uint32_t* src = static_cast<const uint32_t*>(a.data);
          const int nrow = E1 * E2;
#pragma unroll 1
          for (int rw = kRowsW * wave + (int)rr; rw < nrow; rw += kRowsB) {
            const uint32_t m2 = (uint32_t)(((float)rw + 0.5f) * invE1);
            const uint32_t m1 = (uint32_t)rw - mul_u24_s(m2, (uint32_t)E1);
            const uint32_t e = gbase + d0 + mad_u24_s(m1, X.gs1, mul_u24_s(m2, X.gs2));
            const u32x3_a4 q = *reinterpret_cast<const u32x3_a4*>(src + (e >> 1));
            const uint32_t sh = (e & 1u) << 4;
            const uint32_t e01 = __builtin_amdgcn_alignbit(q.y, q.x, sh);
            const uint32_t e23 = __builtin_amdgcn_alignbit(q.z, q.y, sh);
            const float f0 = bits_to_f32<DT>(e01 & 0xffffu), f1 = bits_to_f32<DT>(e01 >> 16);
            const float f2 = bits_to_f32<DT>(e23 & 0xffffu), f3 = bits_to_f32<DT>(e23 >> 16);
            const float f4 = bits_to_f32<DT>((q.z >> sh) & 0xffffu);
            const uint32_t w = mad_u24_s(m2, pad, mad_u24((uint32_t)rw, (uint32_t)DP, d0));
            put_elem<PAIR>(brick, w, f0, f1);
            put_elem<PAIR>(brick, w + 1u, f1, f2);
            put_elem<PAIR>(brick, w + 2u, f2, f3);
            put_elem<PAIR>(brick, w + 3u, f3, f4);
          }
        }
      } else if (X.gsd == 1u) {
        // lanes 0..54 of each wave: d = lane % DP (dominant offset) and row phase lane / DP; the
        // block's NW waves cover 5 NW consecutive rows per step; for pairs each lane also loads
        // the next value (clamped to the padded row)
        const float invE1 = __builtin_amdgcn_rcpf((float)E1);  // exact as above
        constexpr int kLanes = (64 / DP) * DP, kRowsW = 64 / DP, kRowsB = NW * kRowsW;
        if (lane < kLanes) {
          const uint32_t rr = (uint32_t)(((float)lane + 0.5f) * (1.0f / (float)DP));
          const uint32_t d = (uint32_t)lane - mul_u24(rr, (uint32_t)DP);
          const uint32_t gd = gbase + min(d, dmax), gd1 = gbase + min(d + 1u, dmax);
          const int nrow = E1 * E2;
#pragma unroll 1
          for (int r0 = kRowsW * wave + (int)rr; r0 < nrow; r0 += 4 * kRowsB) {
            uint32_t v[4], v1[4], w[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rw = r0 + q * kRowsB;
              v[q] = 0u;
              v1[q] = 0u;
              w[q] = 0u;
              if (rw < nrow) {
                const uint32_t m2 = (uint32_t)(((float)rw + 0.5f) * invE1);
                const uint32_t m1 = (uint32_t)rw - mul_u24_s(m2, (uint32_t)E1);
                const uint32_t o = mad_u24_s(m1, X.gs1, mul_u24_s(m2, X.gs2));
                v[q] = load_bits<DT>(a.data, gd + o);
                if constexpr (PAIR) v1[q] = load_bits<DT>(a.data, gd1 + o);
                w[q] = mad_u24_s(m2, pad, mad_u24((uint32_t)rw, (uint32_t)DP, d));
              }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rw = r0 + q * kRowsB;
              if (rw < nrow)
                put_elem<PAIR>(brick, w[q], bits_to_f32<DT>(v[q]), bits_to_f32<DT>(v1[q]));
            }
          }
        }
      } else {
        const int ncol = E1 * E2;
        const float invE1 = __builtin_amdgcn_rcpf((float)E1);  // exact as above
#pragma unroll 1
        for (int c = threadIdx.x; c < ncol; c += NT) {
          const uint32_t m2 = (uint32_t)(((float)c + 0.5f) * invE1);  // exact: c < 2^10
          const uint32_t m1 = (uint32_t)c - mul_u24_s(m2, (uint32_t)E1);
          const uint32_t g = gbase + mad_u24_s(m1, X.gs1, mul_u24_s(m2, X.gs2));
          // the column's address once; each element adds a wave-uniform offset, recomputed by
          // the scalar unit per column (the empty asm redefines its inputs), so that the DP
          // offsets do not hold scalar registers across the fill loop
          const void* colp =
              static_cast<const char*>(a.data) + (size_t)g * (DT == FC_F32 ? 4u : 2u);
          uint32_t gsd = X.gsd, dm = dmax;
          asm volatile("" : "+s"(gsd), "+s"(dm));
          float v[DP];
#pragma unroll
          for (int d = 0; d < DP; ++d)
            v[d] = load_elem<DT>(colp, (size_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                                           min((uint32_t)d, dm) * gsd));
          const uint32_t col = mad_u24_s(m2, pad, mul_u24((uint32_t)c, (uint32_t)DP));
#pragma unroll
          for (int d = 0; d < DP; ++d)
            put_elem<PAIR>(brick, col + d, v[d], v[d < DP - 1 ? d + 1 : d]);
        }
      }
      __syncthreads();
      // Brick-relative LDS byte offset of a sample's corner (0,0,0). Sample coordinates are >= 0
      // in the padded frame (ray_segment guarantees it), so truncation is floor; the 24-bit
      // multiply-adds stay full rate.
      const uint32_t boff = (uint32_t)(dlo + DP * B.x) + PS * (uint32_t)B.y;
      constexpr uint32_t EB = PAIR ? 8u : 4u;  // bytes per element
      const uint32_t zbytes = EB * PS;  // one m2 plane
      const char* base = reinterpret_cast<const char*>(brick) - EB * boff;
      // two samples per iteration; the loop counts in FP32 (exact: k < 2^24), so one add and one
      // compare serve both samples
      const float kend = (float)kout;
      float kf = (float)kin, kf1 = kf + 1.0f;
      if constexpr (PAIR) {
        // A lane reads its cell's corner pairs only when the sample's base cell differs from its
        // previous sample's (a masked LDS read): samples 0.5 voxel apart along a ray share a cell
        // about half the time, so about half of the LDS reads are skipped. The first sample of a
        // slab always reads (no element index is 2^32 - 1).
        CellPairs C;
        C.e00 = C.e10 = C.e01 = C.e11 = make_float2(0.0f, 0.0f);
        uint32_t cur = 0xffffffffu;
        auto step = [&](float k) {
          const float qd = fmaf(k, sd, pd0), q1 = fmaf(k, s1, p10), q2 = fmaf(k, s2, p20);
          const uint32_t idx =
              mad_u24_s((uint32_t)(int)q2, PS, mad_u24((uint32_t)(int)q1, (uint32_t)DP,
                                                       (uint32_t)(int)qd));
          if (idx != cur) load_cell<DP>(base, idx, zbytes, C);
          cur = idx;
          return cell_trilinear(C, __builtin_amdgcn_fractf(qd), __builtin_amdgcn_fractf(q1),
                                __builtin_amdgcn_fractf(q2));
        };
        // Two steps per iteration. The empty asm (a use of the first value, ordered against
        // memory) keeps the first step's interpolation ahead of the second step's reads, so both
        // steps keep the pairs in one set of registers (moved down, the first step's values and
        // the second step's reads are live together and the compiler copies all eight).
        for (; kf1 < kend; kf += 2.0f, kf1 += 2.0f) {
          const float r0 = step(kf);
          asm volatile("" ::"v"(r0) : "memory");
          const float r1 = step(kf1);
          acc += r0;
          acc += r1;
        }
        if (kf < kend) acc += step(kf);
      } else {
        auto sample = [&](float k) {
          return lds_trilinear<DP, false>(base, PS, zbytes, fmaf(k, sd, pd0), fmaf(k, s1, p10),
                                          fmaf(k, s2, p20));
        };
        // both samples' LDS reads are in flight together
        for (; kf1 < kend; kf += 2.0f, kf1 += 2.0f) {
          const float q0 = sample(kf), q1 = sample(kf1);
          acc += q0;
          acc += q1;
        }
        if (kf < kend) acc += sample(kf);
      }
    };
    if (a.pairs && 2 * NE <= lds_cap) slab(LayoutTag<true>{});
    else slab(LayoutTag<false>{});
  }
  return acc;
}

// Shared-mode quads. Blocks (quad, half) decide the mode and record it; single-mode quads are
// left to fp_lds_kernel. `mode` = 0: cost decision, 1: always shared. Launched with half as many
// blocks as fp_lds_kernel: block b covers "virtual view" 2 q + h (quad q, half h), so the tile
// order of fp_tile_origin (groups, XCD dealing) applies with half the group size.
template <int DT, int S>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
void fp_quad_kernel(FpArgs a, uint8_t* __restrict__ quad_mode, int mode) {
  constexpr int DP = S + 3;
  __shared__ float brick[kHalfCap];
  __shared__ int4 tab[kSlabTab];
  __shared__ QuadGrid G;
  __shared__ unsigned long long cost[2];
  const int nvq = (a.nviews + kQuad - 1) / kQuad * kQuad;
  int tv, j0, i0;
  fp_tile_origin(a.tiles_w, a.tiles_per_view, a.group / 2, nvq / 2, &tv, &j0, &i0);
  const int half = tv & 1;
  const int t0 = (tv >> 1) * kQuad;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the budget: kHalfCap unless lowered below fp_lds_kernel's (tests of the global fallback)
  const int cap_shared = a.lds_cap >= kLdsCap ? kHalfCap : a.lds_cap;
  const SlabAxes X =
      slab_axes<S>(a, a.views + (size_t)min(t0, a.nviews - 1) * FC_VIEW_DOUBLES, j0, i0);
  // the quad's 4x4 ray grid in each view, one thread per (view, point)
  bool bad = false;
  if (threadIdx.x < 16 * kQuad) {
    const int v = threadIdx.x >> 4, p = threadIdx.x & 15, cj = p & 3, ci = p >> 2;
    bad = !grid_ray(G, a, X, a.views + (size_t)min(t0 + v, a.nviews - 1) * FC_VIEW_DOUBLES, v, p,
                    (double)min(j0 + (cj >> 1) * 8 + (cj & 1) * 7, a.nw - 1),
                    (double)min(i0 + (ci >> 1) * 8 + (ci & 1) * 7, a.nh - 1));
  }
  if (threadIdx.x < 2) cost[threadIdx.x] = 0ull;
  const bool quad_bad = __syncthreads_or(bad);
  // items (kind, slab), kind 0-1: the two halves over all four views, 2-5: each view's whole
  // tile, over every fourth slab (brick sizes change slowly from slab to slab); consecutive
  // lanes take consecutive slabs of one kind; a brick over its budget counts 4x. Sums stay in
  // registers and are reduced across the wave, so each wave adds once to each LDS total.
  if (!quad_bad && mode == 0) {
    unsigned long long csh = 0ull, csi = 0ull;
    const int nsub = (X.nslabs + 3) / 4;
    for (int w = threadIdx.x; w < 6 * nsub; w += 512) {
      const int kind = w / nsub, m = 4 * (w - kind * nsub);
      if (kind >= 2 && t0 + kind - 2 >= a.nviews) continue;  // a missing view of the last quad
      const int4 B = kind < 2 ? slab_box<S>(G, X, 0, kQuad, 0, kind * 2, 2, 0, m)
                              : slab_box<S>(G, X, kind - 2, 1, 0, 0, 2, 2, m);
      if (B.z > 0 && B.w > 0) {
        const unsigned long long ne = brick_elems(DP, (uint32_t)B.z, (uint32_t)B.w);
        const int cap = kind < 2 ? cap_shared : a.lds_cap;
        const unsigned long long c = ne > (unsigned long long)cap ? 4ull * ne : ne;
        if (kind < 2) csh += c; else csi += c;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      csh += __shfl_xor(csh, off);
      csi += __shfl_xor(csi, off);
    }
    if (lane == 0) {
      atomicAdd(&cost[0], csh);
      atomicAdd(&cost[1], csi);
    }
  }
  __syncthreads();
  // A quad whose grid rays do not all advance along the dominant axis goes to fp_lds_kernel,
  // which checks each view's tile on its own. The last quad of a view count that is not a
  // multiple of four is decided like the others (its missing views' waves idle in shared mode;
  // left to fp_lds_kernel, whose block order keeps one view group on one XCD, its tiles would
  // run on one eighth of the GPU after this kernel). Both halves' blocks evaluate the same
  // decision from the same inputs.
  const bool shared = mode == 1 || (!quad_bad && cost[0] <= cost[1]);
  const int tile = (i0 / kTile) * a.tiles_w + j0 / kTile;
  if (half == 0 && threadIdx.x == 0)
    quad_mode[(size_t)(t0 / kQuad) * a.tiles_per_view + tile] = shared ? kQuadShared : 0;
  if (!shared) return;
  // wave: view t0 + (wave & 3), quarter (wave >> 2) of the half (left / right 8 columns)
  const int t = t0 + (wave & 3);
  const int j = j0 + (wave >> 2) * 8 + (lane & 7);
  const int i = i0 + half * 8 + (lane >> 3);
  const bool pix = (t < a.nviews) && (j < a.nw) && (i < a.nh);
  RaySeg r;
  r.nk = 0;
  r.px = r.py = r.pz = r.sx = r.sy = r.sz = 0.f;
  if (pix)
    r = ray_segment(a.views + (size_t)t * FC_VIEW_DOUBLES, j, i, a.ox, a.oy, a.oz, a.nx, a.ny,
                    a.nz, a.h, a.max_steps, 2.0f);
  const float sdir = X.ax == 0 ? r.sx : (X.ax == 1 ? r.sy : r.sz);
  // every active ray must advance along the dominant axis in the same direction
  const bool fallback = __syncthreads_or(quad_bad || (r.nk > 0 && !(sdir * X.dir > 0.2f * (float)a.h)));
  const float acc = march<DT, S, 512>(a, brick, tab, G, X, 0, kQuad, 0, half * 2, 2, 0, r,
                                      fallback, cap_shared);
  if (pix) {
    float scale = (a.norm ? a.norm[t] : 1.0f) * (float)a.h;
    size_t idx = ((size_t)t * a.nh + i) * a.nw + j;
    a.proj[idx] += scale * acc;
  }
}

// Single-mode quads (all quads when quad_mode is null): the block takes the whole 16x16 tile of
// view tg.
template <int DT, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
void fp_lds_kernel(FpArgs a, const uint8_t* __restrict__ quad_mode) {
  __shared__ float brick[kLdsCap];
  __shared__ int4 tab[kSlabTab];
  __shared__ QuadGrid G;
  const int nvq = (a.nviews + kQuad - 1) / kQuad * kQuad;
  int t, j0, i0;
  fp_tile_origin(a.tiles_w, a.tiles_per_view, a.group, nvq, &t, &j0, &i0);
  if (quad_mode) {
    const int tile = (i0 / kTile) * a.tiles_w + j0 / kTile;
    if (quad_mode[(size_t)(t / kQuad) * a.tiles_per_view + tile] == kQuadShared) return;
  }
  if (t >= a.nviews) return;  // a padding view of the last quad (whole block)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = j0 + (wave & 1) * 8 + (lane & 7);
  const int i = i0 + (wave >> 1) * 8 + (lane >> 3);
  const bool pix = (t < a.nviews) && (j < a.nw) && (i < a.nh);
  const double* V = a.views + (size_t)min(t, a.nviews - 1) * FC_VIEW_DOUBLES;
  const SlabAxes X = slab_axes<S>(a, V, j0, i0);
  // the tile's four corner rays (grid points 0, 3, 12, 15 of view slot 0)
  bool bad = false;
  if (threadIdx.x < 4) {
    const int c = threadIdx.x, p = (c >> 1) * 12 + (c & 1) * 3;
    bad = !grid_ray(G, a, X, V, 0, p, (double)min(j0 + (c & 1) * (kTile - 1), a.nw - 1),
                    (double)min(i0 + (c >> 1) * (kTile - 1), a.nh - 1));
  }
  RaySeg r;
  r.nk = 0;
  r.px = r.py = r.pz = r.sx = r.sy = r.sz = 0.f;
  if (pix) r = ray_segment(V, j, i, a.ox, a.oy, a.oz, a.nx, a.ny, a.nz, a.h, a.max_steps, 2.0f);
  const float sdir = X.ax == 0 ? r.sx : (X.ax == 1 ? r.sy : r.sz);
  // every active ray must advance along the dominant axis in the same direction
  const bool fallback = __syncthreads_or(bad || (r.nk > 0 && !(sdir * X.dir > 0.2f * (float)a.h)));
  const float acc = march<DT, S, 256>(a, brick, tab, G, X, 0, 1, 0, 0, 2, 2, r, fallback,
                                      a.lds_cap);
  if (pix) {
    float scale = (a.norm ? a.norm[t] : 1.0f) * (float)a.h;
    size_t idx = ((size_t)t * a.nh + i) * a.nw + j;
    a.proj[idx] += scale * acc;
  }
}

// ---- exact transpose (scatter) ---------------------------------------------------------------

struct AdjArgs {
  float* acc;  // padded FP32 accumulator, halo 2, dims (n+4)
  const void* proj;
  const double* views;
  const float* norm;
  double ox, oy, oz;
  double h;
  int nx, ny, nz;
  int PX, PY;
  int nw, nh, nviews;
  int tiles_w, tiles_per_view, group;
  int max_steps;
  int lds_cap;
};

// Scatter of one sample's trilinear weights times y into the 8 corners (global atomics).
__device__ __forceinline__ void scatter_global(float* accp, uint32_t PX, uint32_t PXY, float qx,
                                               float qy, float qz, float y) {
  const uint32_t ix = (uint32_t)(int)qx, iy = (uint32_t)(int)qy, iz = (uint32_t)(int)qz;
  const float ax = __builtin_amdgcn_fractf(qx), ay = __builtin_amdgcn_fractf(qy),
              az = __builtin_amdgcn_fractf(qz);
  float* p = accp + ((size_t)iz * PXY + (size_t)iy * PX + ix);
  const float wz1 = y * az, wz0 = y - wz1;
  const float w10 = wz0 * ay, w00 = wz0 - w10, w11 = wz1 * ay, w01 = wz1 - w11;
  float v;
  v = w00 * ax; unsafeAtomicAdd(p + 1, v); unsafeAtomicAdd(p, w00 - v);
  v = w10 * ax; unsafeAtomicAdd(p + PX + 1, v); unsafeAtomicAdd(p + PX, w10 - v);
  v = w01 * ax; unsafeAtomicAdd(p + PXY + 1, v); unsafeAtomicAdd(p + PXY, w01 - v);
  v = w11 * ax; unsafeAtomicAdd(p + PXY + PX + 1, v); unsafeAtomicAdd(p + PXY + PX, w11 - v);
}

// Voxel-owned exact transpose of the projector. A block owns the 16^3 "base cells" of a brick
// (voxel indices [16 b - 1, 16 b + 15) per axis, which tile [-1, n - 1]): it walks every ray of
// every view whose detector window (adj_windows) can reach the brick, recomputes the ray's samples
// exactly as the projector does (the ray's ray_segment, stored once by adj_rays; the same FMA
// positions), keeps those whose base
// corner floor(q) lies in the brick — so every sample of every ray is scattered by exactly one
// block — and adds y h norm times its eight trilinear weights into a 17^3 LDS accumulator. The
// accumulator is 64-bit fixed point: each contribution is scaled by 2^28 / max|y h norm| (a device
// reduction over the projections, so |contribution| < 2^28), rounded to an integer and added with
// ds_add_u64, which runs about 30x faster than an LDS float atomic on this GPU; the resolution is
// 2^-28 of the largest datum and the 64-bit sums cannot overflow. The accumulator is converted
// back and added to the volume once at the end (one global float atomic per nonzero element;
// neighbouring bricks share one face). (ray, view) items are dealt to the threads round-robin
// across views, so lanes stay busy when windows are small.
constexpr int kAB = 16;             // base cells per brick edge
constexpr int kAL = kAB + 1;        // accumulator edge (corners reach one cell further)
constexpr int kAL3 = kAL * kAL * kAL;
constexpr int kAT = 512;           // threads per brick (8 waves share one accumulator)

struct AdjVArgs {
  float* vol;          // unpadded output block, accumulated
  unsigned int* ymax;  // [0]: bits of max |y h norm| over the stack; [1]: non-finite data flag
  float4* rays;        // per ray: (px, py, pz, sx), (sy, sz, nk bits, y h norm); see adj_rays
  const void* proj;
  const double* views;
  const float* norm;
  const uint2* win;    // per (brick, view): x = j0 | i0 << 16, y = Wj | Wi << 16
  double ox, oy, oz;
  double h;
  int nx, ny, nz;
  int nw, nh, nviews;
  int nbx, nby, nbz;
  int nsplit;          // view ranges per brick (grid z = nbz * nsplit)
  int max_steps;
};

// One thread per ray (view, row, column): the ray's FP32 first sample, step and sample count
// exactly as the projector computes them (ray_segment), and its datum y h norm, stored once so
// that every brick the ray crosses reads 32 bytes instead of redoing the FP64 set-up; plus
// max |y h norm| over the stack as the bits of a non-negative float (ordered like the value):
// a block-wide maximum, then one integer atomic per block.
template <int DT>
__global__ __launch_bounds__(256) void adj_rays(AdjVArgs a) {
  __shared__ float red[256];
  const size_t per_view = (size_t)a.nw * a.nh;
  const size_t total = per_view * (size_t)a.nviews;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  float m = 0.0f;
  if (idx < total) {
    const int t = (int)(idx / per_view);
    const int rem = (int)(idx - (size_t)t * per_view);
    const int i = rem / a.nw, j = rem - i * a.nw;
    const float y = load_elem<DT>(a.proj, idx) * (a.norm ? a.norm[t] : 1.0f) * (float)a.h;
    RaySeg r;
    r.nk = 0;
    r.px = r.py = r.pz = r.sx = r.sy = r.sz = 0.0f;
    if (y != 0.0f)
      r = ray_segment(a.views + (size_t)t * FC_VIEW_DOUBLES, j, i, a.ox, a.oy, a.oz, a.nx, a.ny,
                      a.nz, a.h, a.max_steps, 2.0f);
    a.rays[2 * idx] = make_float4(r.px, r.py, r.pz, r.sx);
    a.rays[2 * idx + 1] =
        make_float4(r.sy, r.sz, __int_as_float(r.nk), r.nk > 0 && fabsf(y) <= 3.402823466e38f ? y : 0.0f);
    // a non-finite datum on a ray that reaches the block cannot be held by the fixed-point
    // accumulator: it is flagged (the whole output becomes NaN) and kept out of the maximum
    const bool bad = r.nk > 0 && !(fabsf(y) <= 3.402823466e38f);
    if (bad) atomicOr(a.ymax + 1, 1u);
    m = r.nk > 0 && !bad ? fabsf(y) : 0.0f;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(a.ymax, __float_as_uint(red[0]));
}

// One thread per (brick, view): the detector pixels whose rays can carry samples with a base cell
// in the brick — the projective image of the brick's box [16 b - 1, 16 b + 15]^3 (voxel
// coordinates), widened by one pixel for FP32 sample rounding and clipped to the detector. A box
// not wholly in front of the source, or a cylindrical detector, gets the whole detector.
__global__ __launch_bounds__(256) void adj_windows(AdjVArgs a, uint2* __restrict__ out) {
  const size_t total = (size_t)a.nbx * a.nby * a.nbz * a.nviews;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= total) return;
  const int t = (int)(idx % (size_t)a.nviews);
  const size_t b = idx / (size_t)a.nviews;
  const int bx = (int)(b % a.nbx), by = (int)((b / a.nbx) % a.nby),
            bz = (int)(b / ((size_t)a.nbx * a.nby));
  const double* V = a.views + (size_t)t * FC_VIEW_DOUBLES;
  const double* M = V + 12;
  const double xs[2] = {kAB * bx - 1 + a.ox, kAB * bx + kAB - 1 + a.ox};
  const double ys[2] = {kAB * by - 1 + a.oy, kAB * by + kAB - 1 + a.oy};
  const double zs[2] = {kAB * bz - 1 + a.oz, kAB * bz + kAB - 1 + a.oz};
  double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
  bool ok = V[24] != (double)FC_DET_CYLINDER;
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
  int j0 = 0, j1 = a.nw - 1, i0 = 0, i1 = a.nh - 1;
  if (ok && umin > -1e9 && umax < 1e9 && vmin > -1e9 && vmax < 1e9) {
    j0 = max(j0, (int)ceil(umin) - 1); j1 = min(j1, (int)floor(umax) + 1);
    i0 = max(i0, (int)ceil(vmin) - 1); i1 = min(i1, (int)floor(vmax) + 1);
  }
  uint2 w = make_uint2(0u, 0u);
  if (j1 >= j0 && i1 >= i0) {
    w.x = (uint32_t)j0 | ((uint32_t)i0 << 16);
    w.y = (uint32_t)(j1 - j0 + 1) | ((uint32_t)(i1 - i0 + 1) << 16);
  }
  out[idx] = w;
}

// CAPRUN: cap a run at kRun samples (steps below 0.25 voxel; see the run merging below).
template <bool CAPRUN>
__global__ __launch_bounds__(kAT) void adjoint_vox_kernel(AdjVArgs a) {
  __shared__ unsigned long long acc[kAL3];
  for (int e = threadIdx.x; e < kAL3; e += kAT) acc[e] = 0ull;
  const int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z % a.nbz;
  // this block's share of the views (small volumes split the views over several blocks)
  const int split = blockIdx.z / a.nbz;
  if (a.ymax[1] != 0u) {
    // non-finite data reach the volume: the transpose is NaN (uniform exit). Each voxel
    // belongs to the base-cell range [16 b - 1, 16 b + 15) of exactly one brick.
    if (split == 0) {
      for (int e = threadIdx.x; e < kAB * kAB * kAB; e += kAT) {
        const int x = kAB * bx - 1 + e % kAB, yy = kAB * by - 1 + (e / kAB) % kAB,
                  z = kAB * bz - 1 + e / (kAB * kAB);
        if (x >= 0 && yy >= 0 && z >= 0 && x < a.nx && yy < a.ny && z < a.nz)
          a.vol[((size_t)z * a.ny + yy) * a.nx + x] = __builtin_nanf("");
      }
    }
    return;
  }
  const float ymax = __uint_as_float(a.ymax[0]);
  if (!(ymax > 0.0f)) return;                // all-zero data (uniform exit)
  const float fscale = 268435456.0f / ymax;  // 2^28 / max|y|
  __syncthreads();
  const int tbeg = (int)((long long)a.nviews * split / a.nsplit);
  const int tend = (int)((long long)a.nviews * (split + 1) / a.nsplit);
  // base-cell range in the padded sample frame (voxel index + 2)
  const int lox = kAB * bx + 1, loy = kAB * by + 1, loz = kAB * bz + 1;
  const float flx = (float)lox, fly = (float)loy, flz = (float)loz;
  const uint2* W = a.win + ((size_t)(bz * a.nby + by) * a.nbx + bx) * (size_t)a.nviews;
  int t = tbeg, start = 0;
  uint2 w = tbeg < tend ? W[tbeg] : make_uint2(0u, 0u);
  // Items are dealt in chunks of kAT: wave pair (2 q, 2 q + 1) takes items 128 q .. 128 q + 127
  // of each chunk, lane l of wave 2 q + b item 128 q + 2 l + b, so the 64 lanes of a wave hold
  // rays 2 pixels apart: far enough apart that they rarely add to the same LDS word at once
  // (same-address atomics serialise), close enough that their in-brick segments have similar
  // lengths (the wave runs for its longest). Timed at 256^3 x 360 views x 384^2: 1 pixel apart
  // 92.7 ms, 2 apart 78.7, 4 apart 86.3, 8 apart 89.0.
  const int perm =
      ((threadIdx.x >> 7) << 7) | ((threadIdx.x & 63) << 1) | ((threadIdx.x >> 6) & 1);
  for (int g = perm;; g += kAT) {
    int Wj = (int)(w.y & 0xffffu), npix = Wj * (int)(w.y >> 16);
    while (t < tend && g >= start + npix) {
      start += npix;
      if (++t < tend) {
        w = W[t];
        Wj = (int)(w.y & 0xffffu);
        npix = Wj * (int)(w.y >> 16);
      }
    }
    if (t >= tend) break;
    const int p = g - start;
    // p / Wj from a 1-ulp reciprocal, exact for p < 2^21: (p + 0.5) / Wj lies at least 0.5 / Wj
    // from an integer and the roundings move it by less than 2^-2 / Wj (integer division for
    // larger windows)
    const int pi = p < (1 << 21) ? (int)(((float)p + 0.5f) * __builtin_amdgcn_rcpf((float)Wj))
                                 : p / Wj;
    const int i = (int)(w.x >> 16) + pi, j = (int)(w.x & 0xffffu) + (p - pi * Wj);
    const size_t ray = ((size_t)t * a.nh + i) * a.nw + j;
    const float4 r1 = a.rays[2 * ray + 1];
    if (r1.w == 0.0f) continue;  // zero datum or a ray that misses the block
    const float4 r0 = a.rays[2 * ray];
    RaySeg r;
    r.px = r0.x; r.py = r0.y; r.pz = r0.z; r.sx = r0.w;
    r.sy = r1.x; r.sz = r1.y; r.nk = __float_as_int(r1.z);
    const float y = r1.w * fscale;  // |y| <= 2^28
    // k range of samples whose base cell can lie in the brick: per axis, q in [lo, lo + 16);
    // widened by one step for FP32 rounding, then every sample is checked exactly
    float klo = 0.0f, khi = (float)(r.nk - 1);
    auto clip = [&](float p0, float st, float lo) {
      if (st != 0.0f) {
        // a 1-ulp reciprocal: the range is widened by one step and every sample checked, and
        // |ka| > 2^22 lies far outside [0, nk)
        const float ist = __builtin_amdgcn_rcpf(st);
        const float ka = (lo - p0) * ist, kb = (lo + (float)kAB - p0) * ist;
        klo = fmaxf(klo, fminf(ka, kb) - 1.0f);
        khi = fminf(khi, fmaxf(ka, kb) + 1.0f);
      } else if (!(p0 >= lo - 1.0f && p0 < lo + (float)kAB + 1.0f)) {
        khi = -1.0f;
      }
    };
    clip(r.px, r.sx, flx);
    clip(r.py, r.sy, fly);
    clip(r.pz, r.sz, flz);
    if (!(khi >= klo)) continue;
    const int k0 = max(0, (int)ceilf(klo)), k1 = min(r.nk - 1, (int)floorf(khi));
    // Consecutive samples (0.5 voxel apart) often share a base cell: their weights are summed in
    // registers and added to LDS when the cell changes. A sum must stay below 2^31 in the 2^28
    // fixed-point scale, so at most 7 samples. The step h along a unit direction advances the
    // dominant coordinate by at least h / sqrt(3), so a cell holds at most floor(sqrt(3) / h) + 1
    // samples: 7 for h >= 0.25, and the run needs no count. Shorter steps flush after kRun.
    constexpr int kRun = 4;
    float A[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) A[c] = 0.0f;
    uint32_t cur = 0xffffffffu;
    int run = 0;
    auto flush = [&]() {
      unsigned long long* L = acc + cur;
      // round to the nearest integer, sign-extend, add in 64-bit two's complement
      auto add = [](unsigned long long* p, float v) {
        atomicAdd(p, (unsigned long long)(long long)__float2int_rn(v));
      };
      add(L, A[0]); add(L + 1, A[1]); add(L + kAL, A[2]); add(L + kAL + 1, A[3]);
      add(L + kAL * kAL, A[4]); add(L + kAL * kAL + 1, A[5]);
      add(L + kAL * kAL + kAL, A[6]); add(L + kAL * kAL + kAL + 1, A[7]);
    };
    // the sample index counts in FP32 (exact: k < 2^24), the same values as (float)k, so one add
    // and one compare per sample control the loop
    const float kf1 = (float)k1;
    for (float kf = (float)k0; kf <= kf1; kf += 1.0f) {
      const float qx = fmaf(kf, r.sx, r.px), qy = fmaf(kf, r.sy, r.py), qz = fmaf(kf, r.sz, r.pz);
      const uint32_t cx = (uint32_t)((int)qx - lox), cy = (uint32_t)((int)qy - loy),
                     cz = (uint32_t)((int)qz - loz);
      if (cx >= (uint32_t)kAB || cy >= (uint32_t)kAB || cz >= (uint32_t)kAB) continue;
      const float ax = __builtin_amdgcn_fractf(qx), ay = __builtin_amdgcn_fractf(qy),
                  az = __builtin_amdgcn_fractf(qz);
      const uint32_t cell = (cz * kAL + cy) * kAL + cx;
      const float wz1 = y * az, wz0 = y - wz1;
      const float w10 = wz0 * ay, w00 = wz0 - w10, w11 = wz1 * ay, w01 = wz1 - w11;
      const float w[8] = {w00 - w00 * ax, w00 * ax, w10 - w10 * ax, w10 * ax,
                          w01 - w01 * ax, w01 * ax, w11 - w11 * ax, w11 * ax};
      if (cell != cur || (CAPRUN && run == kRun)) {
        if (cur != 0xffffffffu) flush();
        cur = cell;
        if constexpr (CAPRUN) run = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) A[c] = 0.0f;
      }
      if constexpr (CAPRUN) ++run;
#pragma unroll
      for (int c = 0; c < 8; ++c) A[c] += w[c];
    }
    if (cur != 0xffffffffu) flush();
  }
  __syncthreads();
  // LDS element (ex, ey, ez) is voxel (16 b - 1 + e) of the block; voxels outside [0, n) are the
  // projector's zero halo
  const double inv = (double)ymax / 268435456.0;
  for (int e = threadIdx.x; e < kAL3; e += kAT) {
    const long long q = (long long)acc[e];
    if (q == 0) continue;
    const float v = (float)((double)q * inv);
    const int ex = e % kAL, ey = (e / kAL) % kAL, ez = e / (kAL * kAL);
    const int x = kAB * bx - 1 + ex, yy = kAB * by - 1 + ey, z = kAB * bz - 1 + ez;
    if (x < 0 || yy < 0 || z < 0 || x >= a.nx || yy >= a.ny || z >= a.nz) continue;
    unsafeAtomicAdd(a.vol + ((size_t)z * a.ny + yy) * a.nx + x, v);
  }
}

template <int DT>
__global__ __launch_bounds__(256) void adjoint_kernel(AdjArgs a) {
  FpArgs tmp;
  tmp.tiles_w = a.tiles_w;
  tmp.tiles_per_view = a.tiles_per_view;
  tmp.group = a.group;
  tmp.nviews = a.nviews;
  int t, j, i;
  fp_tile(tmp, &t, &j, &i);
  if (t >= a.nviews || j >= a.nw || i >= a.nh) return;
  size_t pidx = ((size_t)t * a.nh + i) * a.nw + j;
  float y = load_elem<DT>(a.proj, pidx) * (a.norm ? a.norm[t] : 1.0f) * (float)a.h;
  if (y == 0.0f) return;
  const double* V = a.views + (size_t)t * FC_VIEW_DOUBLES;
  RaySeg r = ray_segment(V, j, i, a.ox, a.oy, a.oz, a.nx, a.ny, a.nz, a.h, a.max_steps, 2.0f);
  const size_t sy = (size_t)a.PX, sz = (size_t)a.PX * a.PY;
  for (int k = 0; k < r.nk; ++k) {
    float kf = (float)k;
    float qx = fmaf(kf, r.sx, r.px), qy = fmaf(kf, r.sy, r.py), qz = fmaf(kf, r.sz, r.pz);
    float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    float ax = qx - fx, ay = qy - fy, az = qz - fz;
    uint32_t o = ((uint32_t)(int)fz * (uint32_t)a.PY + (uint32_t)(int)fy) * (uint32_t)a.PX +
                 (uint32_t)(int)fx;
    float* p = a.acc + o;
    float w0z = y * (1.0f - az), w1z = y * az;
    float w00 = w0z * (1.0f - ay), w10 = w0z * ay, w01 = w1z * (1.0f - ay), w11 = w1z * ay;
    unsafeAtomicAdd(p, w00 * (1.0f - ax));
    unsafeAtomicAdd(p + 1, w00 * ax);
    unsafeAtomicAdd(p + sy, w10 * (1.0f - ax));
    unsafeAtomicAdd(p + sy + 1, w10 * ax);
    unsafeAtomicAdd(p + sz, w01 * (1.0f - ax));
    unsafeAtomicAdd(p + sz + 1, w01 * ax);
    unsafeAtomicAdd(p + sz + sy, w11 * (1.0f - ax));
    unsafeAtomicAdd(p + sz + sy + 1, w11 * ax);
  }
}

// vol[z][y][x] += acc[z+2][y+2][x+2]
__global__ __launch_bounds__(256) void unpad_add(const float* __restrict__ acc, int PX, int PY,
                                                 float* __restrict__ vol, int nx, int ny, int nz) {
  int x = blockIdx.x * 256 + threadIdx.x;
  int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  vol[((size_t)z * ny + y) * nx + x] += acc[((size_t)(z + 2) * PY + (y + 2)) * PX + (x + 2)];
}

template <typename T>
__global__ __launch_bounds__(256) void zero_fill(T* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t stride = (size_t)gridDim.x * 256;
  for (; i < n; i += stride) p[i] = T(0);
}

// ---- sample count ----------------------------------------------------------------------------

struct CountArgs {
  const double* views;
  long long* counts;
  double ox, oy, oz;
  double h;
  int nx, ny, nz;
  int nw, nh, nviews;
  int tiles_w, tiles_per_view;
  int max_steps;
};

__global__ __launch_bounds__(256) void count_kernel(CountArgs a) {
  __shared__ int part[4];
  int tile = blockIdx.x, t = blockIdx.y;
  int ty = tile / a.tiles_w, tx = tile - ty * a.tiles_w;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int j = tx * kTile + (wave & 1) * 8 + (lane & 7);
  int i = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
  int nk = 0;
  if (j < a.nw && i < a.nh) {
    const double* V = a.views + (size_t)t * FC_VIEW_DOUBLES;
    nk = ray_segment(V, j, i, a.ox, a.oy, a.oz, a.nx, a.ny, a.nz, a.h, a.max_steps, 2.0f).nk;
  }
  // wave reduction
  for (int off = 32; off > 0; off >>= 1) nk += __shfl_down(nk, off, 64);
  if (lane == 0) part[wave] = nk;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = (long long)part[0] + part[1] + part[2] + part[3];
    atomicAdd((unsigned long long*)&a.counts[t], (unsigned long long)s);
  }
}

// ---- gather-roofline microbenchmark -------------------------------------------------------------
// Each block draws trilinear samples at pseudo-random positions inside one B^3 brick of a staged
// volume (the brick moves with the block index), with the projector's own fetch + interpolation
// code, so the sustained rate is the ceiling for that layout's gathers served from L1/L2.
template <int DT, int LAYOUT>
__global__ __launch_bounds__(256) void gather_bench_fp(FpArgs a, int B, int nsamp,
                                                       float* __restrict__ sink) {
  uint32_t s = 0x9e3779b9u * (blockIdx.x * 256 + threadIdx.x + 1);
  int nbx = (a.nx - B) / B, nby = (a.ny - B) / B, nbz = (a.nz - B) / B;
  int b = blockIdx.x % (nbx * nby * nbz);
  float bx = 2.0f + (float)((b % nbx) * B), by = 2.0f + (float)(((b / nbx) % nby) * B),
        bz = 2.0f + (float)((b / (nbx * nby)) * B);
  const float scale = (float)(B - 1) * (1.0f / 16777216.0f);
  float acc = 0.0f;
  for (int k = 0; k < nsamp; ++k) {
    s = s * 1664525u + 1013904223u;
    float qx = bx + (float)(s >> 8) * scale;
    s = s * 1664525u + 1013904223u;
    float qy = by + (float)(s >> 8) * scale;
    s = s * 1664525u + 1013904223u;
    float qz = bz + (float)(s >> 8) * scale;
    acc += trilinear_sample<DT, LAYOUT>(a, qx, qy, qz);
  }
  if (acc == 12345.678f) sink[0] = acc;  // keep the work alive
}

// LDS variant: each block stages one DP x E x E pair brick of the plain staged volume in LDS
// (same layout as fp_lds_kernel), then draws trilinear samples at pseudo-random positions inside it
// with the projector's own LDS fetch and interpolation (two samples per iteration).
template <int S>
__global__ __launch_bounds__(256) void gather_bench_lds(FpArgs a, int E, int nsamp,
                                                        float* __restrict__ sink) {
  constexpr int DP = S + 3;
  __shared__ float brick[2 * kLdsCap];  // pair bricks up to E = 23
  const uint32_t PS = brick_plane(DP, (uint32_t)E);
  const int NE = DP * E * E;
  const int nb = (a.nx - E) / E;
  const int b = blockIdx.x % (nb * nb);
  const uint32_t g0 = (uint32_t)(2 + (b % nb) * E) * (uint32_t)a.PX +
                      (uint32_t)(2 + (b / nb) * E) * (uint32_t)a.PX * (uint32_t)a.PY + 2u;
  for (int e = threadIdx.x; e < NE; e += 256) {
    const int d = e % DP, rest = e / DP, m1 = rest % E, m2 = rest / E;
    const float* src = static_cast<const float*>(a.data) + g0 + d + (uint32_t)m1 * a.PX +
                       (uint32_t)m2 * a.PX * a.PY;
    put_elem<true>(brick, d + DP * m1 + PS * m2, src[0], src[1]);
  }
  __syncthreads();
  uint32_t s = 0x9e3779b9u * (blockIdx.x * 256 + threadIdx.x + 1);
  const float sd = (float)(DP - 1) * (1.0f / 16777216.0f), se = (float)(E - 1) * (1.0f / 16777216.0f);
  const char* base = reinterpret_cast<const char*>(brick);
  const uint32_t zbytes = 8u * PS;
  float acc = 0.0f;
  for (int k = 0; k < nsamp; k += 2) {
    float q[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      s = s * 1664525u + 1013904223u;
      q[c] = (float)(s >> 8) * ((c % 3) == 0 ? sd : se);
    }
    const float r0 = lds_trilinear<DP, true>(base, PS, zbytes, q[0], q[1], q[2]);
    const float r1 = lds_trilinear<DP, true>(base, PS, zbytes, q[3], q[4], q[5]);
    acc += r0;
    acc += r1;
  }
  if (acc == 12345.678f) sink[0] = acc;
}

// Ray-coherent LDS gather ceiling: the projector's trilinear LDS sample (same lds_trilinear,
// positions p0 + k s by FMA) over one resident 11 x 16 x 16 pair brick (padded planes, as in the
// projector), with the projector's lane layout (8 x 8 rays per wave, 16 x 16 per block) and
// ray-like positions (0.75-voxel pixel spacing, diverging directions, 16 samples per slab pass),
// U samples per iteration. The LDS
// allocation is exactly the pair brick (23 KiB). No refill,
// no ray set-up and no divergence: it bounds the tap rate of LDS trilinear sampling along rays.
// Bounds: qd <= 1 + 15 * 0.47 < 9, q1, q2 in [1.25 - 1.125, 1.25 + 11.25 + 1.125 + 0.25] =
// [0.125, 13.875], so every corner index stays in [0, 9] x [0, 14] x [0, 14].
template <int S, int U>
__global__ __launch_bounds__(256) void gather_bench_lds_ray(FpArgs a, int reps,
                                                            float* __restrict__ sink) {
  constexpr int DP = S + 3, E = 16, NE = DP * E * E;
  constexpr uint32_t PS = brick_plane(DP, E);
  __shared__ float brick[2 * brick_elems(DP, E, E)];
  const int nb = (a.nx - E) / E;
  const int b = blockIdx.x % (nb * nb);
  const uint32_t g0 = (uint32_t)(2 + (b % nb) * E) * (uint32_t)a.PX +
                      (uint32_t)(2 + (b / nb) * E) * (uint32_t)a.PX * (uint32_t)a.PY + 2u;
  for (int e = threadIdx.x; e < NE; e += 256) {
    const int d = e % DP, rest = e / DP, m1 = rest % E, m2 = rest / E;
    const float* src = static_cast<const float*>(a.data) + g0 + d + (uint32_t)m1 * a.PX +
                       (uint32_t)m2 * a.PX * a.PY;
    put_elem<true>(brick, d + DP * m1 + PS * m2, src[0], src[1]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float jj = (float)((wave & 1) * 8 + (lane & 7)), ii = (float)((wave >> 1) * 8 + (lane >> 3));
  const float sd = 0.47f, s1 = 0.01f * (jj - 7.5f), s2 = 0.01f * (ii - 7.5f);
  const float pd0 = 1.0f;
  const char* base = reinterpret_cast<const char*>(brick);
  const uint32_t zbytes = 8u * PS;
  float acc = 0.0f;
#pragma unroll 1
  for (int rep = 0; rep < reps; ++rep) {
    const float off = (float)(rep & 1) * 0.25f;  // keeps passes distinct for the compiler
    const float p10 = 1.25f + 0.75f * jj + off, p20 = 1.25f + 0.75f * ii;
#pragma unroll 1
    for (int k = 0; k < 16; k += U) {
      float r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float kf = (float)(k + u);
        r[u] = lds_trilinear<DP, true>(base, PS, zbytes, fmaf(kf, sd, pd0),
                                 fmaf(kf, s1, p10), fmaf(kf, s2, p20));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc += r[u];
    }
  }
  if (acc == 12345.678f) sink[0] = acc;
}

int max_steps_for(const int vs[3], double h) {
  double diag = sqrt((double)(vs[0] + 2) * (vs[0] + 2) + (double)(vs[1] + 2) * (vs[1] + 2) +
                     (double)(vs[2] + 2) * (vs[2] + 2));
  double m = ceil(diag / h) + 2.0;
  return m > 2e9 ? 2000000000 : (int)m;
}

}  // namespace

int launch_project(int device, const void* vol, int dt, const int vs[3], const double origin[3],
                   float* proj, const int ps[3], const double* views, double h,
                   const float* norm, hipStream_t stream) {
  const int nx = vs[0], ny = vs[1], nz = vs[2];
  int layout = fp_layout_choice();
  if (layout == FP_PLAIN && dt != FC_F32) layout = FP_LDS;
  // the LDS projector's brick addressing is 24-bit: it needs (nx + 4)(ny + 4) < 2^24
  const bool big_slice = (size_t)(nx + 4) * (size_t)(ny + 4) >= (1u << 24);
  if (layout == FP_LDS && big_slice && dt == FC_F32) layout = FP_PLAIN;
  FC_CHECK_ARG(!(layout == FP_LDS && big_slice),
               "16-bit storage needs (nx + 4)(ny + 4) < 2^24 (tile the volume in x or y)");
  int PX, PY, PZ;
  size_t bytes;
  if (layout == FP_OCTET) {
    PX = nx + 3; PY = ny + 3; PZ = nz + 3;
    bytes = (size_t)PX * PY * PZ * 8 * dtype_size(dt);
  } else {
    PX = nx + 4; PY = ny + 4; PZ = nz + 4;
    bytes = (size_t)PX * PY * PZ * (layout == FP_LDS ? dtype_size(dt) : 4);
  }
  FC_CHECK_ARG(PY <= 65535 && PZ <= 65535, "volume dimension too large for staging grid");
  FC_CHECK_ARG((size_t)PX * PY * PZ < (1ull << 32), "volume too large for 32-bit voxel offsets");
  FC_CHECK_ARG(max_steps_for(vs, h) <= (1 << 24), "step %g too small for the volume: over 2^24 "
               "samples per ray (the sample loops count in FP32)", h);
  // one mode byte per quad of the LDS projector after the staged volume
  const size_t flags_off = (bytes + 255) & ~(size_t)255;
  const size_t nquads = (size_t)((ps[0] + kTile - 1) / kTile) * ((ps[1] + kTile - 1) / kTile) *
                        (size_t)((ps[2] + kQuad - 1) / kQuad);
  if (layout == FP_LDS) bytes = flags_off + nquads;
  void* ws = nullptr;
  int rc = acquire_ws(device, SLOT_OCTET, bytes, stream, &ws);
  if (rc) return rc;
  dim3 sg((PX + 255) / 256, PY, PZ);
  if (layout == FP_OCTET) {
    if (dt == FC_F32) stage_octets<FC_F32><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, ws, PX, PY);
    else if (dt == FC_BF16) stage_octets<FC_BF16><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, ws, PX, PY);
    else stage_octets<FC_F16><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, ws, PX, PY);
  } else if (layout == FP_PLAIN || dt == FC_F32) {
    stage_plain<FC_F32><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, (float*)ws, PX, PY);
  } else if (dt == FC_BF16) {
    stage_plain16<FC_BF16><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, (uint16_t*)ws, PX, PY);
  } else {
    stage_plain16<FC_F16><<<sg, 256, 0, stream>>>(vol, nx, ny, nz, (uint16_t*)ws, PX, PY);
  }
  FC_HIP(hipGetLastError());
  FpArgs a;
  a.data = ws;
  a.proj = proj;
  a.views = views;
  a.norm = norm;
  a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
  a.h = h;
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.PX = PX; a.PY = PY;
  a.nw = ps[0]; a.nh = ps[1]; a.nviews = ps[2];
  a.tiles_w = (a.nw + kTile - 1) / kTile;
  a.tiles_per_view = a.tiles_w * ((a.nh + kTile - 1) / kTile);
  const char* g = std::getenv("FASTCONE_FP_GROUP");
  a.group = g ? std::max(1, std::atoi(g)) : 8;
  a.group = std::min(a.group, std::max(1, a.nviews));
  a.max_steps = max_steps_for(vs, h);
  // FASTCONE_FP_LDS_CAP lowers the LDS brick budget (testing the global-gather fallback)
  const char* cap = std::getenv("FASTCONE_FP_LDS_CAP");
  a.lds_cap = cap ? std::max(0, std::min(kLdsCap, std::atoi(cap))) : kLdsCap;
  // FASTCONE_FP_PAIRS=0 keeps every LDS brick in plain values (testing that the pair bricks and
  // their cell reuse give the same bits)
  const char* pe = std::getenv("FASTCONE_FP_PAIRS");
  a.pairs = pe && std::strcmp(pe, "0") == 0 ? 0 : 1;
  // The LDS projector deals quads of four views; `group` stays a multiple of four.
  const int nvq = (a.nviews + kQuad - 1) / kQuad * kQuad;
  const long long nblocks = (long long)a.tiles_per_view * (layout == FP_LDS ? nvq : a.nviews);
  FC_CHECK_ARG(nblocks < (1ll << 31), "too many projection tiles");
  if (layout == FP_LDS) a.group = std::min((a.group + kQuad - 1) / kQuad * kQuad, nvq);
  // FASTCONE_FP_QUAD = "never" / "always" forces single / shared mode for every quad (tests)
  const char* qe = std::getenv("FASTCONE_FP_QUAD");
  const int qmode = qe && std::strcmp(qe, "never") == 0 ? -1
                    : (qe && std::strcmp(qe, "always") == 0 ? 1 : 0);
  uint8_t* quad_mode = qmode < 0 ? nullptr : static_cast<uint8_t*>(ws) + flags_off;
  if (nblocks > 0) {
    int tok = prof_begin("project", stream);
    const int nb = (int)nblocks;
    if (layout == FP_OCTET) {
      if (dt == FC_F32) fp_kernel<FC_F32, FP_OCTET><<<nb, 256, 0, stream>>>(a);
      else if (dt == FC_BF16) fp_kernel<FC_BF16, FP_OCTET><<<nb, 256, 0, stream>>>(a);
      else fp_kernel<FC_F16, FP_OCTET><<<nb, 256, 0, stream>>>(a);
    } else if (layout == FP_PLAIN) {
      fp_kernel<FC_F32, FP_PLAIN><<<nb, 256, 0, stream>>>(a);
    } else {
      if (qmode >= 0) {
        // two 512-thread blocks per quad and tile (nb = tiles x nvq is a multiple of four)
        const int nh = nb / 2;
        if (dt == FC_F32) fp_quad_kernel<FC_F32, 8><<<nh, 512, 0, stream>>>(a, quad_mode, qmode);
        else if (dt == FC_BF16) fp_quad_kernel<FC_BF16, 8><<<nh, 512, 0, stream>>>(a, quad_mode, qmode);
        else fp_quad_kernel<FC_F16, 8><<<nh, 512, 0, stream>>>(a, quad_mode, qmode);
      }
      if (qmode <= 0) {
        if (dt == FC_F32) fp_lds_kernel<FC_F32, 8><<<nb, 256, 0, stream>>>(a, quad_mode);
        else if (dt == FC_BF16) fp_lds_kernel<FC_BF16, 8><<<nb, 256, 0, stream>>>(a, quad_mode);
        else fp_lds_kernel<FC_F16, 8><<<nb, 256, 0, stream>>>(a, quad_mode);
      }
    }
    prof_end(tok, stream);
    FC_HIP(hipGetLastError());
  }
  return release_ws(device, SLOT_OCTET, stream);
}

int launch_adjoint(int device, float* vol, const int vs[3], const double origin[3],
                   const void* proj, int dt, const int ps[3], const double* views, double h,
                   const float* norm, hipStream_t stream) {
  const int nx = vs[0], ny = vs[1], nz = vs[2];
  const int PX = nx + 4, PY = ny + 4, PZ = nz + 4;
  FC_CHECK_ARG(PY <= 65535 && PZ <= 65535, "volume dimension too large");
  FC_CHECK_ARG((size_t)PX * PY * PZ < (1ull << 32), "volume too large for 32-bit voxel offsets");
  FC_CHECK_ARG(ps[0] <= 65535 && ps[1] <= 65535, "detector too large for the adjoint windows");
  if (ps[2] == 0 || nx == 0 || ny == 0 || nz == 0) return FC_OK;
  FC_CHECK_ARG(max_steps_for(vs, h) <= (1 << 24), "step %g too small for the volume: over 2^24 "
               "samples per ray (the sample loops count in FP32)", h);
  if (fp_layout_choice() != FP_LDS) {
    // global-atomic scatter into a padded accumulator (reference path for tests)
    size_t n = (size_t)PX * PY * PZ;
    void* ws = nullptr;
    int rc = acquire_ws(device, SLOT_OCTET, n * sizeof(float), stream, &ws);
    if (rc) return rc;
    float* acc = static_cast<float*>(ws);
    zero_fill<<<(unsigned)std::min<size_t>((n + 255) / 256, 65536), 256, 0, stream>>>(acc, n);
    FC_HIP(hipGetLastError());
    AdjArgs a;
    a.acc = acc;
    a.proj = proj;
    a.views = views;
    a.norm = norm;
    a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
    a.h = h;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.PX = PX; a.PY = PY;
    a.nw = ps[0]; a.nh = ps[1]; a.nviews = ps[2];
    a.tiles_w = (a.nw + kTile - 1) / kTile;
    a.tiles_per_view = a.tiles_w * ((a.nh + kTile - 1) / kTile);
    a.group = std::min(8, std::max(1, a.nviews));
    a.max_steps = max_steps_for(vs, h);
    long long nblocks = (long long)a.tiles_per_view * a.nviews;
    FC_CHECK_ARG(nblocks < (1ll << 31), "too many projection tiles");
    int tok = prof_begin("adjoint", stream);
    if (dt == FC_F32) adjoint_kernel<FC_F32><<<(int)nblocks, 256, 0, stream>>>(a);
    else if (dt == FC_BF16) adjoint_kernel<FC_BF16><<<(int)nblocks, 256, 0, stream>>>(a);
    else adjoint_kernel<FC_F16><<<(int)nblocks, 256, 0, stream>>>(a);
    prof_end(tok, stream);
    FC_HIP(hipGetLastError());
    dim3 ug((nx + 255) / 256, ny, nz);
    unpad_add<<<ug, 256, 0, stream>>>(acc, PX, PY, vol, nx, ny, nz);
    FC_HIP(hipGetLastError());
    return release_ws(device, SLOT_OCTET, stream);
  }
  AdjVArgs a;
  a.vol = vol;
  a.proj = proj;
  a.views = views;
  a.norm = norm;
  a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
  a.h = h;
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.nw = ps[0]; a.nh = ps[1]; a.nviews = ps[2];
  // bricks of base cells tile [-1, n - 1]
  a.nbx = (nx + 1 + kAB - 1) / kAB;
  a.nby = (ny + 1 + kAB - 1) / kAB;
  a.nbz = (nz + 1 + kAB - 1) / kAB;
  a.max_steps = max_steps_for(vs, h);
  // at least ~8 blocks per CU: split the views of small volumes over several blocks per brick
  const long long nbricks = (long long)a.nbx * a.nby * a.nbz;
  a.nsplit = (int)std::max(1LL, std::min<long long>((2048 + nbricks - 1) / nbricks, a.nviews));
  FC_CHECK_ARG((long long)a.nbz * a.nsplit <= 65535, "volume too large for the adjoint grid");
  const size_t nwin = (size_t)a.nbx * a.nby * a.nbz * a.nviews;
  const size_t nrays = (size_t)a.nw * a.nh * a.nviews;
  void* ws = nullptr;
  // ray records (32 B each), window words, then one 256-byte-aligned word for max |y|
  const size_t woff = nrays * 2 * sizeof(float4);
  const size_t moff = woff + (nwin * sizeof(uint2) + 255) / 256 * 256;
  int rc = acquire_ws(device, SLOT_OCTET, moff + 256, stream, &ws);
  if (rc) return rc;
  a.rays = static_cast<float4*>(ws);
  uint2* win = reinterpret_cast<uint2*>(static_cast<char*>(ws) + woff);
  a.win = win;
  a.ymax = reinterpret_cast<unsigned int*>(static_cast<char*>(ws) + moff);
  int tok = prof_begin("adjoint", stream);
  zero_fill<unsigned int><<<1, 256, 0, stream>>>(a.ymax, 2);  // max |y| bits, non-finite flag
  const unsigned rblocks = (unsigned)((nrays + 255) / 256);
  if (dt == FC_F32) adj_rays<FC_F32><<<rblocks, 256, 0, stream>>>(a);
  else if (dt == FC_BF16) adj_rays<FC_BF16><<<rblocks, 256, 0, stream>>>(a);
  else adj_rays<FC_F16><<<rblocks, 256, 0, stream>>>(a);
  adj_windows<<<(unsigned)((nwin + 255) / 256), 256, 0, stream>>>(a, win);
  const dim3 grid(a.nbx, a.nby, a.nbz * a.nsplit);
  if (a.h >= 0.25) adjoint_vox_kernel<false><<<grid, kAT, 0, stream>>>(a);
  else adjoint_vox_kernel<true><<<grid, kAT, 0, stream>>>(a);
  prof_end(tok, stream);
  FC_HIP(hipGetLastError());
  return release_ws(device, SLOT_OCTET, stream);
}

// Samples per second of the FP gather (octet layout for `dt`; plain FP32 when plain != 0) on a
// synthetic n^3 volume, best over brick sizes 4, 8 and 16.
int launch_gather_bench_fp(int device, int dt, int plain, double* samples_per_s) {
  const int n = 256;
  const bool lds = plain >= 2, ray = plain == 3;
  if (lds) plain = 1;
  const int vs[3] = {n, n, n};
  std::vector<float> host((size_t)n * n * n);
  for (size_t q = 0; q < host.size(); ++q) host[q] = (float)((q * 2654435761u) % 1000) * 1e-3f;
  float* dvol = nullptr;
  float* sink = nullptr;
  FC_HIP(hipMalloc(&dvol, host.size() * sizeof(float)));
  FC_HIP(hipMalloc(&sink, sizeof(float)));
  FC_HIP(hipMemcpy(dvol, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice));
  void* src = dvol;
  void* conv = nullptr;
  if (dt != FC_F32) {
    FC_HIP(hipMalloc(&conv, host.size() * 2));
    int rc = launch_convert(dvol, FC_F32, conv, dt, (long long)host.size(), nullptr);
    if (rc) return rc;
    src = conv;
  }
  const int layout = plain ? FP_PLAIN : FP_OCTET;
  int PX = n + (plain ? 4 : 3), PY = PX, PZ = PX;
  size_t bytes = plain ? (size_t)PX * PY * PZ * 4 : (size_t)PX * PY * PZ * 8 * dtype_size(dt);
  void* ws = nullptr;
  FC_HIP(hipMalloc(&ws, bytes));
  dim3 sg((PX + 255) / 256, PY, PZ);
  if (plain) stage_plain<FC_F32><<<sg, 256>>>(src, n, n, n, (float*)ws, PX, PY);
  else if (dt == FC_F32) stage_octets<FC_F32><<<sg, 256>>>(src, n, n, n, ws, PX, PY);
  else if (dt == FC_BF16) stage_octets<FC_BF16><<<sg, 256>>>(src, n, n, n, ws, PX, PY);
  else stage_octets<FC_F16><<<sg, 256>>>(src, n, n, n, ws, PX, PY);
  FC_HIP(hipGetLastError());
  FpArgs a;
  std::memset(&a, 0, sizeof(a));
  a.data = ws;
  a.nx = a.ny = a.nz = n;
  a.PX = PX;
  a.PY = PY;
  (void)vs;
  hipEvent_t e0, e1;
  FC_HIP(hipEventCreate(&e0));
  FC_HIP(hipEventCreate(&e1));
  const int blocks = 256 * 16, nsamp = 2048;
  double best = 0.0;
  for (int B : {4, 8, 16}) {
    for (int rep = 0; rep < 2; ++rep) {
      FC_HIP(hipEventRecord(e0, nullptr));
      // LDS bricks of 11 x E x E with E = 12, 18, 23 (the projector's typical slab bricks;
      // 11 * 23^2 floats fit the LDS budget)
      if (ray && B == 4) gather_bench_lds_ray<8, 2><<<blocks, 256>>>(a, nsamp / 16, sink);
      else if (ray) gather_bench_lds_ray<8, 4><<<blocks, 256>>>(a, nsamp / 16, sink);
      else if (lds) gather_bench_lds<8><<<blocks, 256>>>(a, B == 4 ? 12 : (B == 8 ? 18 : 23), nsamp, sink);
      else if (layout == FP_PLAIN) gather_bench_fp<FC_F32, FP_PLAIN><<<blocks, 256>>>(a, B, nsamp, sink);
      else if (dt == FC_F32) gather_bench_fp<FC_F32, FP_OCTET><<<blocks, 256>>>(a, B, nsamp, sink);
      else if (dt == FC_BF16) gather_bench_fp<FC_BF16, FP_OCTET><<<blocks, 256>>>(a, B, nsamp, sink);
      else gather_bench_fp<FC_F16, FP_OCTET><<<blocks, 256>>>(a, B, nsamp, sink);
      FC_HIP(hipEventRecord(e1, nullptr));
      FC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FC_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double per_thread = ray ? (double)(nsamp / 16) * 16.0 : (double)nsamp;
      double rate = (double)blocks * 256.0 * per_thread / (ms * 1e-3);
      if (rep == 1 && rate > best) best = rate;
    }
  }
  FC_HIP(hipEventDestroy(e0));
  FC_HIP(hipEventDestroy(e1));
  FC_HIP(hipFree(ws));
  FC_HIP(hipFree(dvol));
  FC_HIP(hipFree(sink));
  if (conv) FC_HIP(hipFree(conv));
  *samples_per_s = best;
  return FC_OK;
}

int launch_sample_count(int device, const int vs[3], const double origin[3], const int ps[3],
                        const double* views, double h, long long* counts, hipStream_t stream) {
  (void)device;
  CountArgs a;
  a.views = views;
  a.counts = counts;
  a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
  a.h = h;
  a.nx = vs[0]; a.ny = vs[1]; a.nz = vs[2];
  a.nw = ps[0]; a.nh = ps[1]; a.nviews = ps[2];
  a.tiles_w = (a.nw + kTile - 1) / kTile;
  a.tiles_per_view = a.tiles_w * ((a.nh + kTile - 1) / kTile);
  a.max_steps = max_steps_for(vs, h);
  FC_CHECK_ARG(a.max_steps <= (1 << 24), "step %g too small for the volume: over 2^24 "
               "samples per ray (the sample loops count in FP32)", h);
  FC_CHECK_ARG(a.nviews <= 65535, "too many views for the count grid");
  if (a.nviews > 0) {
    zero_fill<long long><<<(a.nviews + 255) / 256, 256, 0, stream>>>(counts, (size_t)a.nviews);
    FC_HIP(hipGetLastError());
  }
  if (a.nviews > 0 && a.tiles_per_view > 0) {
    count_kernel<<<dim3(a.tiles_per_view, a.nviews), 256, 0, stream>>>(a);
    FC_HIP(hipGetLastError());
  }
  return FC_OK;
}

}  // namespace fc
