/* wrapper.c — the reference C-ABI of libFastcone.so (wrapper.c:112-249 of the reference), same
 * five symbols, argument order and meaning, void return. Each entry point forwards to the
 * library's host-memory problem solvers (fc_runtime.hip), which validate, record failures in the
 * thread-local fc_last_error() instead of exiting, and keep a persistent device context instead of
 * tearing it down per call (the reference's cudaCleanup, wrapper.c:144).
 *
 * Kernel modes (wrapper.c:134-143, coneBeamGPU.cu:685-688): 0 = unweighted general geometry,
 * 1 = Katsevich weighting, any other value = Feldkamp weighting. Set FASTCONE_VERBOSE=1 for the
 * reference's per-call progress line (wrapper.c:130). */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/fastcone.h"

int fc_cb_project(int GPUID, float* volume, float* projections, float* angles, float* hMid,
                  float D, float pw, int volDim[3], int volStart[3], int volEnd[3],
                  int projDim[3], int projWHSize[2], int* woff, int* hoff, float* norm,
                  float* mis);
int fc_cb_backproject(int GPUID, float* volume, float* projections, float* angles, float* hMid,
                      float D, float pw, int volDim[3], int volStart[3], int volEnd[3],
                      int projDim[3], int projWHSize[2], int* woff, int* hoff, float* norm,
                      float* mis, int mode);

static int verbose(void) {
  const char* e = getenv("FASTCONE_VERBOSE");
  return e && e[0] == '1';
}

void GPU_CB_K_backproject(int GPUID, float* volume, float* projections, float* angles,
                          float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                          int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                          int projWHSize[2], int* projWOffset, int* projHOffset,
                          float* normalizationFactor, int kernelMode) {
  int mode = (kernelMode == 0) ? FC_BP_UNWEIGHTED
                               : (kernelMode == 1 ? FC_BP_KATSEVICH : FC_BP_FDK);
  if (verbose() && volDim && projDim && volStart && volEnd && projWHSize) {
    printf("Backprojecting with volDim = %d,%d,%d, projDim = %d,%d,%d, volStart = %d,%d,%d, "
           "volEnd = %d,%d,%d, projWHSize = %d,%d, kernelMode = %d \n",
           volDim[0], volDim[1], volDim[2], projDim[0], projDim[1], projDim[2], volStart[0],
           volStart[1], volStart[2], volEnd[0], volEnd[1], volEnd[2], projWHSize[0],
           projWHSize[1], kernelMode);
    fflush(stdout);
  }
  (void)fc_cb_backproject(GPUID, volume, projections, angles, hMidPlane, sourceSampleDistance,
                          pixelWidth, volDim, volStart, volEnd, projDim, projWHSize, projWOffset,
                          projHOffset, normalizationFactor, NULL, mode);
}

void GPU_CB_misalign_backproject(int GPUID, float* volume, float* projections, float* angles,
                                 float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                                 int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                                 int projWHSize[2], int* projWOffset, int* projHOffset,
                                 float* normalizationFactor, float* misAlignments) {
  (void)fc_cb_backproject(GPUID, volume, projections, angles, hMidPlane, sourceSampleDistance,
                          pixelWidth, volDim, volStart, volEnd, projDim, projWHSize, projWOffset,
                          projHOffset, normalizationFactor, misAlignments, FC_BP_UNWEIGHTED);
}

void GPU_CB_backproject(int GPUID, float* volume, float* projections, float* angles,
                        float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                        int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                        int projWHSize[2], int* projWOffset, int* projHOffset,
                        float* normalizationFactor) {
  GPU_CB_K_backproject(GPUID, volume, projections, angles, hMidPlane, sourceSampleDistance,
                       pixelWidth, volDim, volStart, volEnd, projDim, projWHSize, projWOffset,
                       projHOffset, normalizationFactor, 2);
}

void GPU_CB_misalign_project(int GPUID, float* volume, float* projections, float* angles,
                             float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                             int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                             int projWHSize[2], int* projWOffset, int* projHOffset,
                             float* normalizationFactor, float* misAlignments) {
  (void)fc_cb_project(GPUID, volume, projections, angles, hMidPlane, sourceSampleDistance,
                      pixelWidth, volDim, volStart, volEnd, projDim, projWHSize, projWOffset,
                      projHOffset, normalizationFactor, misAlignments);
}

void GPU_CB_project(int GPUID, float* volume, float* projections, float* angles,
                    float* hMidPlane, float sourceSampleDistance, float pixelWidth,
                    int volDim[3], int volStart[3], int volEnd[3], int projDim[3],
                    int projWHSize[2], int* projWOffset, int* projHOffset,
                    float* normalizationFactor) {
  (void)fc_cb_project(GPUID, volume, projections, angles, hMidPlane, sourceSampleDistance,
                      pixelWidth, volDim, volStart, volEnd, projDim, projWHSize, projWOffset,
                      projHOffset, normalizationFactor, NULL);
}
