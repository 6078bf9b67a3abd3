// FP64 -> FP32 copies of the p' operator for the mixed-precision solves
// (cg1.hip, cg3.hip, fcg.hip, batch.hip).
#include "reduce.hpp"

namespace cfdhip {

namespace {

constexpr int kMaxBlocks = 2048;

template <class T>
__global__ void castKernel(int64_t n, const double* __restrict__ src, T* dst) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[k] = static_cast<T>(src[k]);
}

}  // namespace

void castToFloat(Context& ctx, int64_t n, const double* src, float* dst) {
  int64_t gb = (n + 255) / 256;
  if (gb > kMaxBlocks) gb = kMaxBlocks;
  hipLaunchKernelGGL((castKernel<float>), dim3(static_cast<unsigned>(gb)), dim3(256), 0,
                     ctx.stream, n, src, dst);
  CFDHIP_CHECK(hipGetLastError());
}

}  // namespace cfdhip
