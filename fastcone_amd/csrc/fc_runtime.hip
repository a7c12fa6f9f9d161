// Runtime of libFastcone: thread-local errors, device selection, cached device workspaces, the
// additive fc_* device-pointer entry points and the host-memory problem solvers that back the
// reference ABI (wrapper.c).
//
// The reference builds a geometry object, allocates device arrays, uploads a texture per tile,
// launches, downloads and adds on the host, then destroys the CUDA context at the end of every
// call (unit_coneBeamGPU.cu:88-305, wrapper.c:144). Here a call keeps a persistent per-device
// context and grow-only workspaces, the "add-and-save" accumulation happens on the device
// (existing caller data is uploaded, the kernels accumulate into it, the result is downloaded),
// and the volume block is tiled into z slabs / view batches only when a tile limit is set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fc_internal.h"

namespace fc {

// ---- errors ------------------------------------------------------------------------------------
namespace {
thread_local int g_err = FC_OK;
thread_local char g_msg[512] = {0};
}  // namespace

void set_error(int code, const char* fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
}
void clear_error() {
  g_err = FC_OK;
  g_msg[0] = 0;
}
int last_error() { return g_err; }

// ---- devices -----------------------------------------------------------------------------------
int select_device(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    set_error(FC_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    return FC_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    set_error(FC_ERR_DEVICE, "device %d out of range (%d devices)", device, n);
    return FC_ERR_DEVICE;
  }
  FC_HIP(hipSetDevice(device));
  return FC_OK;
}

// ---- workspaces --------------------------------------------------------------------------------
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_ws_mutex;
Workspace g_ws[kMaxDevices][SLOT_COUNT];

long long g_tile_voxels = 0;  // 0 = unlimited
int g_tile_views = 0;
}  // namespace

int acquire_ws(int device, Slot slot, size_t bytes, hipStream_t stream, void** out) {
  FC_CHECK_ARG(device >= 0 && device < kMaxDevices, "device id %d too large", device);
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Workspace& w = g_ws[device][slot];
  if (w.used && w.last_stream != stream) FC_HIP(hipStreamWaitEvent(stream, w.last_use, 0));
  if (w.bytes < bytes) {
    if (w.ptr) {
      // the old buffer may still be read by queued work on w.last_stream
      if (w.used) FC_HIP(hipEventSynchronize(w.last_use));
      FC_HIP(hipFree(w.ptr));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    size_t rounded = (bytes + ((size_t)1 << 21) - 1) & ~(((size_t)1 << 21) - 1);
    hipError_t e = hipMalloc(&w.ptr, rounded);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      w.ptr = nullptr;
      set_error(FC_ERR_MEMORY, "device allocation of %zu bytes failed: %s", rounded,
                hipGetErrorString(e));
      return FC_ERR_MEMORY;
    }
    w.bytes = rounded;
  }
  if (!w.last_use) FC_HIP(hipEventCreateWithFlags(&w.last_use, hipEventDisableTiming));
  *out = w.ptr;
  return FC_OK;
}

int release_ws(int device, Slot slot, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Workspace& w = g_ws[device][slot];
  FC_HIP(hipEventRecord(w.last_use, stream));
  w.last_stream = stream;
  w.used = true;
  return FC_OK;
}

int release_all_ws(int device) {
  FC_CHECK_ARG(device >= 0 && device < kMaxDevices, "device id %d too large", device);
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (int s = 0; s < SLOT_COUNT; ++s) {
    Workspace& w = g_ws[device][s];
    if (w.used) FC_HIP(hipEventSynchronize(w.last_use));
    if (w.ptr) FC_HIP(hipFree(w.ptr));
    if (w.last_use) FC_HIP(hipEventDestroy(w.last_use));
    w = Workspace();
  }
  return FC_OK;
}

// ---- kernel timing -------------------------------------------------------------------------------
namespace {
struct ProfEntry {
  std::string name;
  hipEvent_t start = nullptr, stop = nullptr;
};
std::mutex g_prof_mutex;
bool g_prof_on = false;
std::vector<ProfEntry> g_prof_pending;
std::vector<std::pair<std::string, std::pair<long long, double>>> g_prof_totals;

void prof_accumulate(const std::string& name, double ms) {
  for (auto& kv : g_prof_totals) {
    if (kv.first == name) {
      kv.second.first += 1;
      kv.second.second += ms;
      return;
    }
  }
  g_prof_totals.push_back({name, {1, ms}});
}

// Resolve every pending entry (waits for its stop event). Caller holds g_prof_mutex.
int prof_drain() {
  for (auto& e : g_prof_pending) {
    FC_HIP(hipEventSynchronize(e.stop));
    float ms = 0.f;
    FC_HIP(hipEventElapsedTime(&ms, e.start, e.stop));
    prof_accumulate(e.name, (double)ms);
    FC_HIP(hipEventDestroy(e.start));
    FC_HIP(hipEventDestroy(e.stop));
  }
  g_prof_pending.clear();
  return FC_OK;
}
}  // namespace

int prof_begin(const char* kernel, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  if (!g_prof_on) return -1;
  ProfEntry e;
  e.name = kernel;
  if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  if (hipEventRecord(e.start, stream) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  g_prof_pending.push_back(e);
  return (int)g_prof_pending.size() - 1;
}

void prof_end(int token, hipStream_t stream) {
  if (token < 0) return;
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  if (token < (int)g_prof_pending.size()) (void)hipEventRecord(g_prof_pending[token].stop, stream);
}

namespace {

int check_shapes(const int vs[3], const int ps[3]) {
  FC_CHECK_ARG(vs && ps, "null shape");
  FC_CHECK_ARG(vs[0] > 0 && vs[1] > 0 && vs[2] > 0, "volume shape must be positive (%d,%d,%d)",
               vs[0], vs[1], vs[2]);
  FC_CHECK_ARG(ps[0] > 0 && ps[1] > 0 && ps[2] >= 0, "projection shape invalid (%d,%d,%d)", ps[0],
               ps[1], ps[2]);
  FC_CHECK_ARG(vs[0] <= 16384 && vs[1] <= 16384 && vs[2] <= 16384, "volume dimension > 16384");
  FC_CHECK_ARG(ps[0] <= 65000 && ps[1] <= 65000, "detector dimension > 65000");
  return FC_OK;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace
}  // namespace fc

using namespace fc;

// ---- additive C API ------------------------------------------------------------------------------
extern "C" {

int fc_last_error(void) { return fc::g_err; }
const char* fc_last_error_message(void) { return fc::g_msg; }
void fc_clear_error(void) { fc::clear_error(); }
const char* fc_version(void) { return "fastcone-mi355x 0.1 (gfx950)"; }

int fc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int fc_gather_bench(int device, int which, int dtype, double* samples_per_s) {
  clear_error();
  FC_CHECK_ARG(samples_per_s, "null output");
  FC_CHECK_ARG(which >= 0 && which <= 5, "unknown gather benchmark %d", which);
  FC_CHECK_ARG(dtype_ok(dtype), "unknown dtype %d", dtype);
  FC_CHECK_ARG(which == 0 || which == 2 || which == 5 || dtype == FC_F32,
               "the plain and LDS benchmarks are FP32 only");
  int rc = select_device(device);
  if (rc) return rc;
  if (which == 2 || which == 5)
    return launch_gather_bench_bp(device, dtype, which == 5, samples_per_s);
  return launch_gather_bench_fp(device, dtype, which == 0 ? 0 : (which == 1 ? 1 : which - 1),
                                samples_per_s);
}

int fc_profile_enable(int on) {
  clear_error();
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  g_prof_on = on != 0;
  return FC_OK;
}

int fc_profile_reset(void) {
  clear_error();
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  int rc = prof_drain();
  g_prof_totals.clear();
  return rc;
}

int fc_profile_read(const char* kernel, long long* count, double* total_ms) {
  clear_error();
  FC_CHECK_ARG(kernel && count && total_ms, "null argument");
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  int rc = prof_drain();
  if (rc) return rc;
  *count = 0;
  *total_ms = 0.0;
  for (auto& kv : g_prof_totals) {
    if (kv.first == kernel) {
      *count = kv.second.first;
      *total_ms = kv.second.second;
    }
  }
  return FC_OK;
}

int fc_set_tile_limits(long long max_voxels, int max_views) {
  clear_error();
  FC_CHECK_ARG(max_voxels >= 0 && max_views >= 0, "tile limits must be >= 0");
  g_tile_voxels = max_voxels;
  g_tile_views = max_views;
  return FC_OK;
}

int fc_release_workspace(int device) {
  clear_error();
  int rc = select_device(device);
  if (rc) return rc;
  return release_all_ws(device);
}

// Pointer checks for the device entry points: an array may be null only when it is empty (a zero
// extent); then the call has nothing to do once the device has been validated.
static bool is_empty(const int shape[3]) {
  return (long long)shape[0] * shape[1] * shape[2] == 0;
}

int fc_project_dev(int device, const void* vol, int vol_dtype, const int vol_shape[3],
                   const double origin[3], float* proj, const int proj_shape[3],
                   const double* views, double step, const float* norm, void* stream) {
  clear_error();
  int rc = check_shapes(vol_shape, proj_shape);
  if (rc) return rc;
  FC_CHECK_ARG(dtype_ok(vol_dtype), "unknown volume dtype %d", vol_dtype);
  FC_CHECK_ARG(origin && (vol || is_empty(vol_shape)) && (proj || is_empty(proj_shape)) &&
                   (views || proj_shape[2] == 0),
               "null pointer argument");
  FC_CHECK_ARG(std::isfinite(step) && step >= 1e-3 && step <= 1e3, "step %g out of [1e-3, 1e3]",
               step);
  for (int a = 0; a < 3; ++a) FC_CHECK_ARG(std::isfinite(origin[a]), "origin not finite");
  rc = select_device(device);
  if (rc) return rc;
  if (is_empty(vol_shape) || is_empty(proj_shape)) return FC_OK;
  return launch_project(device, vol, vol_dtype, vol_shape, origin, proj, proj_shape, views, step,
                        norm, as_stream(stream));
}

int fc_adjoint_dev(int device, float* vol, const int vol_shape[3], const double origin[3],
                   const void* proj, int proj_dtype, const int proj_shape[3], const double* views,
                   double step, const float* norm, void* stream) {
  clear_error();
  int rc = check_shapes(vol_shape, proj_shape);
  if (rc) return rc;
  FC_CHECK_ARG(dtype_ok(proj_dtype), "unknown projection dtype %d", proj_dtype);
  FC_CHECK_ARG(origin && (vol || is_empty(vol_shape)) && (proj || is_empty(proj_shape)) &&
                   (views || proj_shape[2] == 0),
               "null pointer argument");
  FC_CHECK_ARG(std::isfinite(step) && step >= 1e-3 && step <= 1e3, "step %g out of [1e-3, 1e3]",
               step);
  rc = select_device(device);
  if (rc) return rc;
  if (is_empty(vol_shape) || is_empty(proj_shape)) return FC_OK;
  return launch_adjoint(device, vol, vol_shape, origin, proj, proj_dtype, proj_shape, views, step,
                        norm, as_stream(stream));
}

int fc_backproject_dev(int device, float* vol, const int vol_shape[3], const double origin[3],
                       const void* proj, int proj_dtype, const int proj_shape[3],
                       const double* views, const float* norm, int mode, void* stream) {
  clear_error();
  int rc = check_shapes(vol_shape, proj_shape);
  if (rc) return rc;
  FC_CHECK_ARG(dtype_ok(proj_dtype), "unknown projection dtype %d", proj_dtype);
  FC_CHECK_ARG(mode >= 0 && mode <= 2, "unknown backprojection mode %d", mode);
  FC_CHECK_ARG(origin && (vol || is_empty(vol_shape)) && (proj || is_empty(proj_shape)) &&
                   (views || proj_shape[2] == 0),
               "null pointer argument");
  rc = select_device(device);
  if (rc) return rc;
  if (is_empty(vol_shape) || is_empty(proj_shape)) return FC_OK;
  return launch_backproject(device, vol, vol_shape, origin, proj, proj_dtype, proj_shape, views,
                            norm, mode, as_stream(stream));
}

int fc_sample_count_dev(int device, const int vol_shape[3], const double origin[3],
                        const int proj_shape[3], const double* views, double step,
                        long long* counts, void* stream) {
  clear_error();
  int rc = check_shapes(vol_shape, proj_shape);
  if (rc) return rc;
  FC_CHECK_ARG(views && counts && origin, "null pointer argument");
  FC_CHECK_ARG(std::isfinite(step) && step >= 1e-3 && step <= 1e3, "step %g out of [1e-3, 1e3]",
               step);
  rc = select_device(device);
  if (rc) return rc;
  return launch_sample_count(device, vol_shape, origin, proj_shape, views, step, counts,
                             as_stream(stream));
}

int fc_convert_dev(int device, const void* src, int src_dtype, void* dst, int dst_dtype,
                   long long n, void* stream) {
  clear_error();
  FC_CHECK_ARG(dtype_ok(src_dtype) && dtype_ok(dst_dtype), "unknown dtype");
  FC_CHECK_ARG(n >= 0, "negative element count");
  FC_CHECK_ARG(n == 0 || (src && dst), "null pointer argument");
  int rc = select_device(device);
  if (rc) return rc;
  return launch_convert(src, src_dtype, dst, dst_dtype, n, as_stream(stream));
}

// ---- host-memory problem solvers behind the reference ABI ---------------------------------------

// Validate the reference ABI arguments shared by every GPU_CB_* call.
static int cb_validate(float* volume, float* projections, float* angles, float* hMid,
                       float D, float pw, int volDim[3], int volStart[3], int volEnd[3],
                       int projDim[3], int projWHSize[2], int* woff, int* hoff, float* norm) {
  FC_CHECK_ARG(volDim && volStart && volEnd && projDim && projWHSize, "null dimension array");
  for (int a = 0; a < 3; ++a)
    FC_CHECK_ARG(volEnd[a] > volStart[a], "empty volume block on axis %d (%d..%d)", a,
                 volStart[a], volEnd[a]);
  FC_CHECK_ARG(projDim[2] >= 0, "negative number of views");
  FC_CHECK_ARG(projWHSize[0] > 0 && projWHSize[1] > 0, "projection window must be positive");
  FC_CHECK_ARG(projDim[0] > 0 && projDim[1] > 0, "detector size must be positive");
  FC_CHECK_ARG(std::isfinite(D) && D > 0.0f, "sourceSampleDistance must be > 0");
  FC_CHECK_ARG(std::isfinite(pw) && pw > 0.0f, "pixelWidth must be > 0");
  FC_CHECK_ARG(volume && projections, "null data pointer");
  if (projDim[2] > 0) FC_CHECK_ARG(angles && hMid && woff && hoff && norm, "null per-view array");
  // A block may reach past the source radius, as in the reference (its sweep start may be
  // negative, coneBeamGPU.cu:1781): the projector clips every ray to the block, and the
  // backprojector evaluates voxels behind the source plane with the same projective formula as
  // the reference (a voxel exactly on that plane gives an infinite weight there too).
  return FC_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      set_error(FC_ERR_MEMORY, "device allocation of %zu bytes failed", bytes);
      return FC_ERR_MEMORY;
    }
    return FC_OK;
  }
};

// Forward projection of a host block into host windows (accumulating), with optional tiling.
int fc_cb_project(int GPUID, float* volume, float* projections, float* angles, float* hMid,
                  float D, float pw, int volDim[3], int volStart[3], int volEnd[3],
                  int projDim[3], int projWHSize[2], int* woff, int* hoff, float* norm,
                  float* mis) {
  clear_error();
  int rc = cb_validate(volume, projections, angles, hMid, D, pw, volDim, volStart, volEnd,
                       projDim, projWHSize, woff, hoff, norm);
  if (rc) return rc;
  rc = select_device(GPUID);
  if (rc) return rc;
  const int nv = projDim[2];
  if (nv == 0) return FC_OK;
  const int nx = volEnd[0] - volStart[0], ny = volEnd[1] - volStart[1],
            nz = volEnd[2] - volStart[2];
  const int nw = projWHSize[0], nh = projWHSize[1];
  std::vector<double> ang(nv), hm(nv), views((size_t)nv * FC_VIEW_DOUBLES);
  for (int t = 0; t < nv; ++t) {
    ang[t] = angles[t];
    hm[t] = hMid[t];
  }
  double m6[6] = {0, 0, 0, 0, 0, 0};
  if (mis)
    for (int k = 0; k < 6; ++k) m6[k] = mis[k];
  const int pdim[2] = {projDim[0], projDim[1]};
  rc = fc_views_flat(nv, ang.data(), hm.data(), D, pw, pw, m6, volDim[2], pdim, woff, hoff,
                     views.data());
  if (rc) return rc;
  const size_t proj_n = (size_t)nv * nw * nh;
  DevBuf dviews, dnorm, dproj, dvol;
  if ((rc = dviews.alloc(views.size() * sizeof(double)))) return rc;
  if ((rc = dnorm.alloc((size_t)nv * sizeof(float)))) return rc;
  if ((rc = dproj.alloc(proj_n * sizeof(float)))) return rc;
  FC_HIP(hipMemcpy(dviews.p, views.data(), views.size() * sizeof(double), hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(dnorm.p, norm, (size_t)nv * sizeof(float), hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(dproj.p, projections, proj_n * sizeof(float), hipMemcpyHostToDevice));
  // z slabs of at most g_tile_voxels voxels; view batches of at most g_tile_views views
  int slab = nz;
  if (g_tile_voxels > 0) slab = (int)std::max<long long>(1, g_tile_voxels / ((long long)nx * ny));
  slab = std::min(slab, nz);
  int vb = g_tile_views > 0 ? std::min(g_tile_views, nv) : nv;
  if ((rc = dvol.alloc((size_t)nx * ny * slab * sizeof(float)))) return rc;
  for (int z0 = 0; z0 < nz; z0 += slab) {
    int sz = std::min(slab, nz - z0);
    FC_HIP(hipMemcpy(dvol.p, volume + (size_t)z0 * nx * ny, (size_t)nx * ny * sz * sizeof(float),
                     hipMemcpyHostToDevice));
    const double origin[3] = {(double)(volStart[0] - volDim[0] / 2),
                              (double)(volStart[1] - volDim[1] / 2),
                              (double)(volStart[2] + z0 - volDim[2] / 2)};
    const int vs[3] = {nx, ny, sz};
    for (int t0 = 0; t0 < nv; t0 += vb) {
      int nb = std::min(vb, nv - t0);
      const int ps[3] = {nw, nh, nb};
      rc = launch_project(GPUID, dvol.p, FC_F32, vs, origin,
                          static_cast<float*>(dproj.p) + (size_t)t0 * nw * nh, ps,
                          static_cast<double*>(dviews.p) + (size_t)t0 * FC_VIEW_DOUBLES, 1.0,
                          static_cast<float*>(dnorm.p) + t0, nullptr);
      if (rc) return rc;
    }
    FC_HIP(hipDeviceSynchronize());
  }
  FC_HIP(hipMemcpy(projections, dproj.p, proj_n * sizeof(float), hipMemcpyDeviceToHost));
  return FC_OK;
}

// Backprojection of host windows into a host block (accumulating), with optional tiling.
int fc_cb_backproject(int GPUID, float* volume, float* projections, float* angles, float* hMid,
                      float D, float pw, int volDim[3], int volStart[3], int volEnd[3],
                      int projDim[3], int projWHSize[2], int* woff, int* hoff, float* norm,
                      float* mis, int mode) {
  clear_error();
  int rc = cb_validate(volume, projections, angles, hMid, D, pw, volDim, volStart, volEnd,
                       projDim, projWHSize, woff, hoff, norm);
  if (rc) return rc;
  FC_CHECK_ARG(mode >= 0 && mode <= 2, "unknown kernel mode %d", mode);
  rc = select_device(GPUID);
  if (rc) return rc;
  const int nv = projDim[2];
  if (nv == 0) return FC_OK;
  const int nx = volEnd[0] - volStart[0], ny = volEnd[1] - volStart[1],
            nz = volEnd[2] - volStart[2];
  const int nw = projWHSize[0], nh = projWHSize[1];
  std::vector<double> ang(nv), hm(nv), views((size_t)nv * FC_VIEW_DOUBLES);
  for (int t = 0; t < nv; ++t) {
    ang[t] = angles[t];
    hm[t] = hMid[t];
  }
  double m6[6] = {0, 0, 0, 0, 0, 0};
  if (mis)
    for (int k = 0; k < 6; ++k) m6[k] = mis[k];
  const int pdim[2] = {projDim[0], projDim[1]};
  rc = fc_views_flat(nv, ang.data(), hm.data(), D, pw, pw, m6, volDim[2], pdim, woff, hoff,
                     views.data());
  if (rc) return rc;
  const size_t proj_n = (size_t)nv * nw * nh;
  DevBuf dviews, dnorm, dproj, dvol;
  if ((rc = dviews.alloc(views.size() * sizeof(double)))) return rc;
  if ((rc = dnorm.alloc((size_t)nv * sizeof(float)))) return rc;
  FC_HIP(hipMemcpy(dviews.p, views.data(), views.size() * sizeof(double), hipMemcpyHostToDevice));
  FC_HIP(hipMemcpy(dnorm.p, norm, (size_t)nv * sizeof(float), hipMemcpyHostToDevice));
  int vb = g_tile_views > 0 ? std::min(g_tile_views, nv) : nv;
  if ((rc = dproj.alloc((size_t)vb * nw * nh * sizeof(float)))) return rc;
  int slab = nz;
  if (g_tile_voxels > 0) slab = (int)std::max<long long>(1, g_tile_voxels / ((long long)nx * ny));
  slab = std::min(slab, nz);
  if ((rc = dvol.alloc((size_t)nx * ny * slab * sizeof(float)))) return rc;
  for (int z0 = 0; z0 < nz; z0 += slab) {
    int sz = std::min(slab, nz - z0);
    float* hv = volume + (size_t)z0 * nx * ny;
    FC_HIP(hipMemcpy(dvol.p, hv, (size_t)nx * ny * sz * sizeof(float), hipMemcpyHostToDevice));
    const double origin[3] = {(double)(volStart[0] - volDim[0] / 2),
                              (double)(volStart[1] - volDim[1] / 2),
                              (double)(volStart[2] + z0 - volDim[2] / 2)};
    const int vs[3] = {nx, ny, sz};
    for (int t0 = 0; t0 < nv; t0 += vb) {
      int nb = std::min(vb, nv - t0);
      FC_HIP(hipMemcpy(dproj.p, projections + (size_t)t0 * nw * nh,
                       (size_t)nb * nw * nh * sizeof(float), hipMemcpyHostToDevice));
      const int ps[3] = {nw, nh, nb};
      rc = launch_backproject(GPUID, static_cast<float*>(dvol.p), vs, origin, dproj.p, FC_F32,
                              ps, static_cast<double*>(dviews.p) + (size_t)t0 * FC_VIEW_DOUBLES,
                              static_cast<float*>(dnorm.p) + t0, mode, nullptr);
      if (rc) return rc;
      FC_HIP(hipDeviceSynchronize());
    }
    FC_HIP(hipMemcpy(hv, dvol.p, (size_t)nx * ny * sz * sizeof(float), hipMemcpyDeviceToHost));
  }
  return FC_OK;
}

}  // extern "C"
