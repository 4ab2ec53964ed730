// Host-side twiddle tables, computed in long double and rounded once to FP64.
#pragma once
#include <cmath>
#include <vector>

#include "sfx_portable.h"

namespace sfx {

// wN[j] = e^{2 pi i j/(3 MP)}, j < 3 MP;  wMP[j] = e^{2 pi i j/MP}, j < MP.
inline void host_twiddles(int MP, std::vector<double2>& wN, std::vector<double2>& wMP) {
  const long double pi = 3.141592653589793238462643383279502884L;
  const int n = 3 * MP;
  wN.resize(n);
  wMP.resize(MP);
  for (int j = 0; j < n; ++j) {
    const long double a = 2.0L * pi * (long double)j / (long double)n;
    wN[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
  for (int j = 0; j < MP; ++j) {
    const long double a = 2.0L * pi * (long double)j / (long double)MP;
    wMP[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
}

}  // namespace sfx
