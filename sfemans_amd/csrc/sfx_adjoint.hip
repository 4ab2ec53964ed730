// Launchers of the adjoint kernels (bodies in sfx_adjoint_body.h).
#include "sfx_adjoint_body.h"

namespace sfx {

template <int MP>
__global__ void __launch_bounds__(64 * GAUSS_WAVES) adjoint_kernel(AdjArgs a) {
  extern __shared__ double2 smem[];
  adjoint_body<MP>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

__global__ void __launch_bounds__(256) adjoint_gather_kernel(GatherArgs a) {
  adjoint_gather_body(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

template <int MP>
static hipError_t launch_mp(const AdjArgs& a, hipStream_t st) {
  constexpr int E = GaussLds<MP>::E;
  const int64_t nblk = (a.g.me + E - 1) / E;
  if (nblk == 0) return hipSuccess;
  const size_t shmem = (size_t)AdjLds<MP>::TOTAL * sizeof(double2);
  hipLaunchKernelGGL((adjoint_kernel<MP>), dim3((unsigned)nblk), dim3(64 * GAUSS_WAVES), shmem, st, a);
  return hipGetLastError();
}

hipError_t launch_adjoint(const AdjArgs& a, const GatherArgs& ga, hipStream_t st) {
  hipError_t e = hipErrorInvalidValue;
  switch (a.g.MP) {
    case 4: e = launch_mp<4>(a, st); break;
    case 8: e = launch_mp<8>(a, st); break;
    case 16: e = launch_mp<16>(a, st); break;
    case 32: e = launch_mp<32>(a, st); break;
    case 64: e = launch_mp<64>(a, st); break;
    case 128: e = launch_mp<128>(a, st); break;
    case 256: e = launch_mp<256>(a, st); break;
  }
  if (e != hipSuccess) return e;
  const int64_t n = ga.np * ga.M;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ga);
  return hipGetLastError();
}

}  // namespace sfx
