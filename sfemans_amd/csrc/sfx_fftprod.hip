// Launchers of the FFT_PAR_*_DCL kernels (body in sfx_fftprod_body.h).
#include "sfx_fftprod_body.h"

namespace sfx {

// Waves per work-group: 4 waves hold 16 consecutive points at MP = 64, so the
// staged output stores write whole 128-byte lines.  The kernels are held to
// 256 registers (2 waves per SIMD); left alone the compiler goes a few
// registers over and the kernel runs at 1 wave per SIMD (measured 5.9 ms vs
// 7.7 ms for the cross product at 1.4M points, m_max = 64).
constexpr int FFTPROD_WPB = 4;

template <int MP, int OP>
__global__ void __launch_bounds__(64 * FFTPROD_WPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
fftprod_kernel(FftProdArgs a) {
  extern __shared__ double2 smem[];
  fftprod_body<MP, OP>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), FFTPROD_WPB, smem);
}

template <int MP, int OP>
static hipError_t launch_mp(const FftProdArgs& a, hipStream_t st) {
  constexpr int PW = Shape<MP>::PW;
  const int64_t pts_per_block = (int64_t)PW * FFTPROD_WPB;
  const int64_t nblk = (a.np + pts_per_block - 1) / pts_per_block;
  if (nblk == 0) return hipSuccess;
  const size_t shmem = fftprod_smem_cplx<MP>(FFTPROD_WPB) * sizeof(double2);
  hipLaunchKernelGGL((fftprod_kernel<MP, OP>), dim3((unsigned)nblk), dim3(64 * FFTPROD_WPB), shmem, st, a);
  return hipGetLastError();
}

template <int OP>
static hipError_t launch_op(const FftProdArgs& a, hipStream_t st) {
  switch (a.MP) {
    case 4: return launch_mp<4, OP>(a, st);
    case 8: return launch_mp<8, OP>(a, st);
    case 16: return launch_mp<16, OP>(a, st);
    case 32: return launch_mp<32, OP>(a, st);
    case 64: return launch_mp<64, OP>(a, st);
    case 128: return launch_mp<128, OP>(a, st);
    case 256: return launch_mp<256, OP>(a, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fftprod(int op, const FftProdArgs& a, hipStream_t st) {
  switch (op) {
    case OP_CROSS: return launch_op<OP_CROSS>(a, st);
    case OP_DOT: return launch_op<OP_DOT>(a, st);
    case OP_PROD: return launch_op<OP_PROD>(a, st);
    case OP_SVECT1: return launch_op<OP_SVECT1>(a, st);
    case OP_TENSOR: return launch_op<OP_TENSOR>(a, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace sfx
