/* fastcone.h — C-ABI of libFastcone.so, the MI355X (gfx950) cone-beam projector/backprojector.
 *
 * Two groups of entry points:
 *
 *  1. The reference ABI, unchanged (same names, argument order and meaning, void return,
 *     outputs accumulated into caller memory). Each one replaces the function of the same name
 *     in the reference's wrapper.c. Host pointers; the library copies to and from the GPU.
 *     Dimension arrays are in C order [x,y,z] / [w,h,t] (wrapper.c:129).
 *
 *  2. Additive fc_* entry points (no reference counterpart): device pointers, an explicit HIP
 *     stream, FP64 per-view geometry records, a step size, storage dtypes and a thread-local error
 *     code. The Python package (fastcone_amd) and the torch autograd operators call these.
 *
 * Errors: the reference checks nothing and exits the process on a bad device
 * (coneBeamGPU.cu:472-475). Here every entry point validates its arguments and records failures
 * in a thread-local error code (fc_last_error); the void reference entry points then return
 * without touching caller memory.
 */
#ifndef FASTCONE_H
#define FASTCONE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ */
/* 1. Reference ABI                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* Backprojection with kernel mode: 0 unweighted general geometry, 1 Katsevich weight U,
 * otherwise FDK weight U^2. Replaces wrapper.c:112-146 (GPU_CB_K_backproject). */
void GPU_CB_K_backproject(int GPUID, float* volume, float* projections, float* angles,
                          float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                          int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                          int projWHSize[2], int* projWOffset, int* projHOffset,
                          float* normalizationFactor, int kernelMode);

/* Unweighted backprojection onto a misaligned detector (kernel mode 0).
 * Replaces wrapper.c:149-174 (GPU_CB_misalign_backproject). */
void GPU_CB_misalign_backproject(int GPUID, float* volume, float* projections, float* angles,
                                 float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                                 int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                                 int projWHSize[2], int* projWOffset, int* projHOffset,
                                 float* normalizationFactor, float* misAlignments);

/* FDK backprojection (kernel mode 2). Replaces wrapper.c:177-194 (GPU_CB_backproject). */
void GPU_CB_backproject(int GPUID, float* volume, float* projections, float* angles,
                        float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                        int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                        int projWHSize[2], int* projWOffset, int* projHOffset,
                        float* normalizationFactor);

/* Forward projection with misalignment [Dw, Dh, Dl (pixels), Dphi, Dtheta, Dpsi (radians)],
 * unit ray step. Replaces wrapper.c:203-227 (GPU_CB_misalign_project). */
void GPU_CB_misalign_project(int GPUID, float* volume, float* projections, float* angles,
                             float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                             int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                             int projWHSize[2], int* projWOffset, int* projHOffset,
                             float* normalizationFactor, float* misAlignments);

/* Forward projection, aligned detector. Replaces wrapper.c:229-249 (GPU_CB_project). */
void GPU_CB_project(int GPUID, float* volume, float* projections, float* angles,
                    float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                    int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                    int projWHSize[2], int* projWOffset, int* projHOffset,
                    float* normalizationFactor);

/* ------------------------------------------------------------------------------------------ */
/* 2. Additive API                                                                             */
/* ------------------------------------------------------------------------------------------ */

enum fc_status {
  FC_OK = 0,
  FC_ERR_ARGUMENT = 1, /* invalid argument (shape, window, dtype, mode, geometry) */
  FC_ERR_DEVICE = 2,   /* bad device id or no device */
  FC_ERR_HIP = 3,      /* a HIP runtime call failed */
  FC_ERR_MEMORY = 4    /* device allocation failed */
};

enum fc_dtype { FC_F32 = 0, FC_BF16 = 1, FC_F16 = 2 };

enum fc_bp_mode { FC_BP_UNWEIGHTED = 0, FC_BP_KATSEVICH = 1, FC_BP_FDK = 2 };

enum fc_detector { FC_DET_FLAT = 0, FC_DET_CYLINDER = 1 };

/* Row-filter kernels (internal selector of the FFT row filter). */
enum fc_filter_kind { FC_FILTER_RAMP = 0, FC_FILTER_HILBERT = 1 };

/* Per-view geometry record: FC_VIEW_DOUBLES doubles per view, FP64, voxel coordinates
 * (a voxel with global index g along an axis of N voxels has coordinate g - N/2).
 *   [0..2]   source position
 *   [3..5]   flat: position of window pixel (0,0);  cylinder: SDD * (unit source->axis vector)
 *   [6..8]   flat: step per window column;          cylinder: SDD * (unit tangential vector)
 *   [9..11]  step per window row (flat: detector h step; cylinder: row pitch * z)
 *   [12..23] 3x4 row-major projective matrix M: (x,y,z,1) -> (a,b,c); flat detector: window
 *            column a/c, row b/c, and c = (distance along the detector normal from the source) /
 *            (distance of the detector centre), so 1/c is the FDK magnification U.
 *   [24]     detector kind (fc_detector)
 *   [25]     cylinder: column of gamma = 0 (window index);  [26] cylinder: column pitch (radians)
 *   [27]     cylinder: row of v = 0 (window index);         [28] cylinder: SDD
 *   [29..31] reserved (0)
 */
#define FC_VIEW_DOUBLES 32

/* Thread-local status of the calling thread's last fc_* or GPU_CB_* call (fc_status). */
int fc_last_error(void);
const char* fc_last_error_message(void);
void fc_clear_error(void);
const char* fc_version(void);
int fc_device_count(void);

/* Fill nviews * FC_VIEW_DOUBLES records for the reference's circular/helical flat-detector
 * geometry (host memory). Follows coneBeamGPU.cu:1672-1734 (FP64 here, FP32 there):
 *   angles (radians), hmid (source/detector-centre z index from volume index 0), D source-to-axis
 *   distance (voxels), pw_w / pw_h detector pitch at the axis (voxels), mis[6] (may be NULL),
 *   vol_dim_z the whole-volume z size, proj_dim [Nw, Nh] whole-detector size, woff/hoff global
 *   index of each window's pixel (0,0) (may be NULL = 0). */
int fc_views_flat(int nviews, const double* angles, const double* hmid, double D, double pw_w,
                  double pw_h, const double* mis, int vol_dim_z, const int proj_dim[2],
                  const int* woff, const int* hoff, double* views_out);

/* Cylindrical (curved) detector centred on the source, helical or circular source path
 * a(l) = (R cos l, R sin l, src_z) (voxels). Window column j looks along the fan angle
 * g = (j - col_c[t]) * dgamma, measured from e_v = -(cos l, sin l, 0) towards
 * e_u = (-sin l, cos l, 0); row i sits at height (i - row_c[t]) * drow on the detector (voxels),
 * SDD from the source. col_c / row_c may be NULL (= 0). No reference counterpart (the reference
 * represents only a flat detector with one pitch, fastcone.py:72). */
int fc_views_cylinder(int nviews, const double* lambdas, const double* src_z, double R,
                      double SDD, double dgamma, double drow, const double* col_c,
                      const double* row_c, double* views_out);

/* Forward projection on the device:
 *   proj[t][i][j] += norm[t] * step * sum_k T(src_t + k*step*u_tij)
 * vol: device, dtype fc_dtype, [nz][ny][nx] (x fastest), voxel (0,0,0) at coordinate origin[];
 * proj: device float32 [nviews][nh][nw]; views: DEVICE pointer to nviews records; norm: device
 * float32 [nviews] or NULL (= 1); stream: hipStream_t or NULL. */
int fc_project_dev(int device, const void* vol, int vol_dtype, const int vol_shape[3],
                   const double origin[3], float* proj, const int proj_shape[3],
                   const double* views, double step, const float* norm, void* stream);

/* Exact transpose of fc_project_dev (matched adjoint): vol += A^T proj. vol float32, proj dtype
 * fc_dtype. Accumulated in 64-bit fixed point (resolution 2^-28 of max |proj * norm * step|) per
 * 16^3 brick, then one float atomic per voxel: summation order is not deterministic. A NaN or
 * infinite datum on a ray that reaches the block makes the whole output block NaN. */
int fc_adjoint_dev(int device, float* vol, const int vol_shape[3], const double origin[3],
                   const void* proj, int proj_dtype, const int proj_shape[3], const double* views,
                   double step, const float* norm, void* stream);

/* Voxel-driven backprojection on the device (mode fc_bp_mode; flat detector):
 *   vol(x) += sum_t norm[t] * w_t(x) * B_t(a/c, b/c),  w = 1 (mode 0), 1/c (mode 1), 1/c^2 (2)
 * B_t = zero-extended bilinear interpolant of window t. vol float32 [nz][ny][nx]; proj dtype
 * fc_dtype [nviews][nh][nw]; norm device float32 or NULL (= 1). */
int fc_backproject_dev(int device, float* vol, const int vol_shape[3], const double origin[3],
                       const void* proj, int proj_dtype, const int proj_shape[3],
                       const double* views, const float* norm, int mode, void* stream);

/* FDK pre-processing on the device (replaces the host-side cosine weight and rampfilter of the
 * reference's fbp, fastcone.py:260-292): every detector row is multiplied by the cosine pre-weight
 * and convolved with the reference's ramp, i.e. zero-padded to 2 Nw with the data at [0, Nw), its
 * DFT multiplied by H[k] = ramp_scale * pi * k / (2 Nw), H[0] = ramp_scale * pi / (16 Nw),
 * transformed back and cropped to [0, Nw). The weight is the cosine of the angle between the
 * pixel's ray and the detector normal, from the view records (physical units; flat-detector
 * records, kind FC_DET_FLAT; other kinds give NaN); views == NULL skips it (pure ramp).
 * proj: device [nviews][nh][nw] dtype proj_dtype; out: device, same shape, dtype out_dtype; out
 * may equal proj (in place). proj_shape = [nw, nh, nviews], 1 <= nw <= FC_FILTER_MAX_COLS.
 * FP32 arithmetic (FFT of length 2^p >= 2 Nw - 1 in LDS). */
#define FC_FILTER_MAX_COLS 4096
int fc_fdk_filter_dev(int device, const void* proj, int proj_dtype, void* out, int out_dtype,
                      const int proj_shape[3], const double* views, double ramp_scale,
                      void* stream);

/* fc_fdk_filter_dev without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops). No counterpart in the reference. The filter's tables
 * (twiddles, multipliers, taps: at most 192 KiB, enough for FC_FILTER_MAX_COLS columns) live in a
 * static device array, one per GPU, held until the process ends; widths beyond FC_FILTER_MAX_COLS
 * are rejected (FC_ERR_ARGUMENT). Every call rewrites the tables it reads, on its stream; calls
 * that overlap on one GPU (different streams or threads) share the array and race. Same results
 * as fc_fdk_filter_dev. */
int fc_fdk_filter_noalloc_dev(int device, const void* proj, int proj_dtype, void* out,
                              int out_dtype, const int proj_shape[3], const double* views,
                              double ramp_scale, void* stream);

/* Katsevich exact helical reconstruction on a cylindrical detector (no reference counterpart: the
 * reference offers only a U-weighted backprojection of pre-filtered data, GPU_CB_K_backproject
 * mode 1, coneBeamGPU.cu:258-338). proj: device [nviews][nh][nw] line integrals (dtype fc_dtype)
 * taken at helix angles l_k = l0 + k dl with source height zs_k = zs0 + k dl hp (voxels);
 * column j at fan angle (j - col_c) dgamma, row i at height (i - row_c) drow at distance SDD from
 * the source (see fc_views_cylinder). npsi kappa lines. vol (device FP32, [nz][ny][nx], voxel 0
 * at coordinate origin[]) is accumulated. g5_out (device, nviews*nh*nw FP32) may receive the
 * filtered projections (NULL to skip). */
int fc_katsevich_dev(int device, const void* proj, int proj_dtype, int nw, int nh, int nviews,
                     double l0, double dl, double zs0, double hp, double R, double SDD,
                     double dgamma, double drow, double col_c, double row_c, int npsi,
                     float* vol, const int vol_shape[3], const double origin[3], float* g5_out,
                     void* stream);

/* 3x3x3 convolution (zero padding 1) for the CGLS warm-start network; no reference counterpart.
 * Activations N x D x H x W x C (channels last) BF16 on the device, weights [cout][27][cin] BF16
 * (tap = 9 dz + 3 dy + dx), bias FP32, FP32 accumulation (MFMA). Channel pairs: the network's
 * layers (1->16, 16->16, 16->32, 32->32, 32->64, 64->64, 96->32, 48->16, 16->1) and, for the
 * input gradient, their transposes. The cout = 1 layer writes / reads FP32. */
int fc_conv3d_fwd_dev(int device, const void* x, const void* w, const float* bias, void* y,
                      int N, int D, int H, int W, int cin, int cout, int relu, void* stream);
/* dx = input gradient of the convolution for output gradient dy; wt is a device workspace of
 * cin*27*cout BF16 elements (receives the flipped, transposed weights). */
int fc_conv3d_dgrad_dev(int device, const void* dy, const void* w, void* wt, void* dx, int N,
                        int D, int H, int W, int cin, int cout, void* stream);
/* dw (FP32 [cout][27][cin]) += weight gradient for input x and output gradient dy. */
int fc_conv3d_wgrad_dev(int device, const void* x, const void* dy, float* dw, int N, int D, int H,
                        int W, int cin, int cout, void* stream);
/* One 16x16x32 BF16 and one 16x16x4 FP32 MFMA product of row-major A, B (single wave): checks the
 * operand / result lane layouts the convolution kernels assume. */
int fc_mfma_layout_probe(int device, const void* A32, const void* B32, const float* A4,
                         const float* B4, float* D32, float* D4, void* stream);

/* Exact sample count of fc_project_dev per view (int64 device array counts[nviews], overwritten). */
int fc_sample_count_dev(int device, const int vol_shape[3], const double origin[3],
                        const int proj_shape[3], const double* views, double step,
                        long long* counts, void* stream);

/* Copy/convert n elements between storage dtypes on the device (f32 <-> bf16/f16,
 * round-to-nearest-even). */
int fc_convert_dev(int device, const void* src, int src_dtype, void* dst, int dst_dtype,
                   long long n, void* stream);

/* Release the library's cached device workspaces on a device. */
int fc_release_workspace(int device);

/* Out-of-core tiling limits for the reference ABI entry points (host-memory calls): at most
 * max_voxels voxels of the volume block (split into z slabs) and max_views views per device pass.
 * 0 = no limit (default: on a 288 GB device the configurations fit whole). Process-wide. */
int fc_set_tile_limits(long long max_voxels, int max_views);

/* Gather-roofline microbenchmark: the sustained rate (samples/s) of the kernels' own fetch +
 * interpolation code inside small cache- or LDS-resident bricks.
 * which = 0: projector trilinear sample, octet layout of `dtype` (8 taps per sample), random
 *            positions; 1: the same, plain FP32 layout; 2: the backprojector's bilinear sample
 *            from a resident LDS window at random positions (FP32, 4 taps per sample); 3:
 *            projector trilinear sample from LDS bricks at random positions (FP32); 4: the LDS
 *            projector's inner loop along ray-coherent paths through one resident brick (FP32; no
 *            refill, no divergence: the ceiling of the default projector); 5: the backprojector's
 *            inner loop (per column one u per view, 16 z rows) over one resident LDS window
 *            (FP32, 4 taps per sample). Best over several runs. Synchronous. */
int fc_gather_bench(int device, int which, int dtype, double* samples_per_s);

/* Kernel timing. When enabled, the main kernel of every fc_* / GPU_CB_* call ("project",
 * "adjoint", "backproject", ...) is bracketed by HIP events recorded on the stream it is launched
 * on. fc_profile_read waits for those events and returns the launch count and the summed kernel
 * time (ms) of one kernel name since the last fc_profile_reset. */
int fc_profile_enable(int on);
int fc_profile_reset(void);
int fc_profile_read(const char* kernel, long long* count, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* FASTCONE_H */
