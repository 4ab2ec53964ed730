// Launchers of the right-hand-side assembly kernels (body in sfx_rhs_body.h).
#include "sfx_engine.h"
#include "sfx_kernels.h"
#include "sfx_rhs_body.h"

namespace sfx {

// Consecutive blocks go to the same XCD (xcd_block): the partial-line
// stores of neighbouring nodes or elements then merge in one L2.
__global__ void __launch_bounds__(256) rhs_elem_kernel(RhsArgs a) {
  rhs_elem_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) rhs_gather_kernel(RhsArgs a) {
  rhs_gather_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) qs00_elem_kernel(RhsArgs a) {
  qs00_elem_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(256) qs00_gather_kernel(RhsArgs a) {
  qs00_gather_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

// rhs_3x3: element contributions, then the node-centred gather.
hipError_t launch_rhs3x3(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me * a.M, nn = a.np * a.M;
  if (ne > 0) hipLaunchKernelGGL(rhs_elem_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nn == 0) return e;
  hipLaunchKernelGGL(rhs_gather_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_qs00(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me * a.M, nn = a.np * a.M;
  if (ne > 0) hipLaunchKernelGGL(qs00_elem_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nn == 0) return e;
  hipLaunchKernelGGL(qs00_gather_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace sfx
