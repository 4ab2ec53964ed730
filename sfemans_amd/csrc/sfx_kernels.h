// Internal kernel interfaces shared by the HIP translation units and the C-ABI.
#pragma once
#include "sfx_portable.h"

namespace sfx {

enum { OP_CROSS = 0, OP_DOT = 1, OP_PROD = 2, OP_SVECT1 = 3, OP_TENSOR = 4 };

// Twiddles for one power-of-two MP: wN[j] = e^{2 pi i j/(3 MP)}, wMP[j] = e^{2 pi i j/MP}.
struct Twiddles {
  const double2* wN;
  const double2* wMP;
};

struct FftProdArgs {
  const double* V1;   // (np, nb1, M)
  const double* V2;   // (np, nb2, M)
  double* Vout;       // (np, 2*NO, M), or (3, np, 6, M) for the tensor product
  int64_t np;
  int M, MP, nb1, nb2;
  const double2* wN;
  const double2* wMP;
  int64_t n1, n2, nout;   // element counts of V1, V2, Vout (bounds checks on the CPU build)
};

#ifndef SFX_EMULATE
hipError_t launch_fftprod(int op, const FftProdArgs& a, hipStream_t st);
#endif

// Smallest supported power of two >= max(M, 4); 0 if M > 256.
inline int mp_for(int M) {
  int mp = 4;
  while (mp < M) mp *= 2;
  return mp <= 256 ? mp : 0;
}

}  // namespace sfx
