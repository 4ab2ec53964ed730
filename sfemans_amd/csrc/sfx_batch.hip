// Persistent batched launch of the fused cross-curl kernel (body in sfx_batch_body.h).
#include "sfx_batch_body.h"

namespace sfx {

__global__ void __launch_bounds__(64 * GAUSS_WAVES) __attribute__((amdgpu_waves_per_eu(1, 1))) batch_kernel(BatchArgs b) {
  extern __shared__ double2 smem[];
  batch_body(b, threadIdx.x, blockIdx.x, smem);
}

// smem_cplx: LDS of the largest fused-kernel layout among the problems.
hipError_t launch_batch(const BatchArgs& b, size_t smem_cplx, hipStream_t st) {
  if (b.n_wg == 0) return hipSuccess;
  const size_t shmem = smem_cplx * sizeof(double2);
  if (shmem == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_kernel, dim3((unsigned)b.n_wg), dim3(64 * GAUSS_WAVES), shmem, st, b);
  return hipGetLastError();
}

}  // namespace sfx
