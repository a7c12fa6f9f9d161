// Elementwise storage conversion (FP32 <-> BF16 / FP16, round-to-nearest-even).
#include <hip/hip_runtime.h>

#include "fc_device.h"
#include "fc_internal.h"

namespace fc {
namespace {

template <int SRC, int DST>
__global__ __launch_bounds__(256) void convert_kernel(const void* __restrict__ src,
                                                      void* __restrict__ dst, long long n) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  long long stride = (long long)gridDim.x * 256;
  for (; i < n; i += stride) {
    float v = load_elem<SRC>(src, (size_t)i);
    if constexpr (DST == FC_F32) {
      static_cast<float*>(dst)[i] = v;
    } else {
      static_cast<uint16_t*>(dst)[i] = to_bits16<DST>(v);
    }
  }
}

template <int SRC>
void convert_to(const void* src, void* dst, int dst_dtype, long long n, int blocks,
                hipStream_t s) {
  if (dst_dtype == FC_F32) convert_kernel<SRC, FC_F32><<<blocks, 256, 0, s>>>(src, dst, n);
  else if (dst_dtype == FC_BF16) convert_kernel<SRC, FC_BF16><<<blocks, 256, 0, s>>>(src, dst, n);
  else convert_kernel<SRC, FC_F16><<<blocks, 256, 0, s>>>(src, dst, n);
}

}  // namespace

int launch_convert(const void* src, int src_dtype, void* dst, int dst_dtype, long long n,
                   hipStream_t stream) {
  if (n == 0) return FC_OK;
  long long b = (n + 255) / 256;
  int blocks = (int)(b < 16384 ? b : 16384);
  if (src_dtype == FC_F32) convert_to<FC_F32>(src, dst, dst_dtype, n, blocks, stream);
  else if (src_dtype == FC_BF16) convert_to<FC_BF16>(src, dst, dst_dtype, n, blocks, stream);
  else convert_to<FC_F16>(src, dst, dst_dtype, n, blocks, stream);
  FC_HIP(hipGetLastError());
  return FC_OK;
}

}  // namespace fc
