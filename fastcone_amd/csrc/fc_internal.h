// Internal declarations shared by the libFastcone translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/fastcone.h"

namespace fc {

// ---- thread-local error state -------------------------------------------------------------
void set_error(int code, const char* fmt, ...);
void clear_error();
int last_error();

#define FC_CHECK_ARG(cond, ...)                       \
  do {                                                \
    if (!(cond)) {                                    \
      ::fc::set_error(FC_ERR_ARGUMENT, __VA_ARGS__);  \
      return FC_ERR_ARGUMENT;                         \
    }                                                 \
  } while (0)

#define FC_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t fc_e_ = (call);                                                         \
    if (fc_e_ != hipSuccess) {                                                         \
      ::fc::set_error(FC_ERR_HIP, "%s failed: %s (%s:%d)", #call,                      \
                      hipGetErrorString(fc_e_), __FILE__, __LINE__);                   \
      return FC_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)

// Select a device; returns FC_OK or FC_ERR_DEVICE.
int select_device(int device);

// ---- per-device workspaces ----------------------------------------------------------------
// A grow-only device buffer owned by the library for staging layouts. One per (device, slot).
// A call that uses a workspace on a stream other than the one its previous user ran on first
// makes the new stream wait on an event recorded after that previous use, so reuse is ordered.
enum Slot { SLOT_OCTET = 0, SLOT_QUAD = 1, SLOT_H2D_VOL = 2, SLOT_H2D_PROJ = 3,
            SLOT_H2D_VIEWS = 4, SLOT_H2D_NORM = 5, SLOT_KAT_A = 6, SLOT_KAT_B = 7,
            SLOT_KAT_TAB = 8, SLOT_FILTER = 9, SLOT_COUNT = 10 };

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t last_use = nullptr;
  hipStream_t last_stream = nullptr;
  bool used = false;
};

// Returns a device pointer of at least `bytes` for `slot` on the current device, ordered after
// any previous use on another stream. Must be followed by release_ws() after the last enqueue.
int acquire_ws(int device, Slot slot, size_t bytes, hipStream_t stream, void** out);
int release_ws(int device, Slot slot, hipStream_t stream);
int release_all_ws(int device);

// ---- kernel timing (fc_profile_*) ---------------------------------------------------------
// prof_begin returns a token (>= 0) when profiling is enabled, else -1; prof_end records the stop
// event. Both are no-ops for token -1.
int prof_begin(const char* kernel, hipStream_t stream);
void prof_end(int token, hipStream_t stream);

// ---- dtypes ---------------------------------------------------------------------------------
inline size_t dtype_size(int dt) { return dt == FC_F32 ? 4 : 2; }
inline bool dtype_ok(int dt) { return dt == FC_F32 || dt == FC_BF16 || dt == FC_F16; }

// ---- launchers (defined in the .hip files) --------------------------------------------------
int launch_project(int device, const void* vol, int vol_dtype, const int vol_shape[3],
                   const double origin[3], float* proj, const int proj_shape[3],
                   const double* views, double step, const float* norm, hipStream_t stream);
int launch_adjoint(int device, float* vol, const int vol_shape[3], const double origin[3],
                   const void* proj, int proj_dtype, const int proj_shape[3],
                   const double* views, double step, const float* norm, hipStream_t stream);
int launch_backproject(int device, float* vol, const int vol_shape[3], const double origin[3],
                       const void* proj, int proj_dtype, const int proj_shape[3],
                       const double* views, const float* norm, int mode, hipStream_t stream);
int launch_sample_count(int device, const int vol_shape[3], const double origin[3],
                        const int proj_shape[3], const double* views, double step,
                        long long* counts, hipStream_t stream);
// Stage projections into the backprojector's quad layout (norm may be NULL).
int launch_stage_quads(const void* proj, int dt, int nw, int nh, int nv, const float* norm,
                       void* quad, hipStream_t stream);
int launch_gather_bench_fp(int device, int dt, int plain, double* samples_per_s);
int launch_gather_bench_bp(int device, int dt, int ray, double* samples_per_s);
// Row filter (fc_filter.hip): out[r, j] = sum_m kern[j - m] wgt(r, m) in[r, m] over nrows rows
// of nw values (nh rows per view for the cosine weight, views may be null = no weight);
// kind fc_filter_kind with parameter p0 (ramp scale, or the Hilbert kernel's column pitch).
// `tables` is a device buffer of filter_table_bytes(nw) bytes. in and out may alias.
size_t filter_table_bytes(int nw);
int launch_row_filter(const void* in, int in_dt, void* out, int out_dt, int nw, int nh,
                      long long nrows, const double* views, int kind, double p0, void* tables,
                      hipStream_t stream);
int launch_convert(const void* src, int src_dtype, void* dst, int dst_dtype, long long n,
                   hipStream_t stream);

}  // namespace fc
