// CPU emulator of the libsfx kernel bodies (test infrastructure).
//
// Each work-group runs as 64*wpb std::threads sharing one LDS buffer, with a
// std::barrier standing in for __syncthreads.  Every global and LDS index the
// kernels compute is checked against its buffer (SFX_CHECK); a violation
// prints the site and aborts, so tests run the emulator in a child process.
#define SFX_EMULATE 1
#include <barrier>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <thread>
#include <vector>

#include "../../sfemans_amd/csrc/sfx_fftprod_body.h"
#include "../../sfemans_amd/csrc/sfx_twiddle.h"

static thread_local std::barrier<>* g_bar = nullptr;
void sfx_emu_sync() { g_bar->arrive_and_wait(); }
void sfx_emu_oob(const char* what, long long idx, long long n, const char* file, int line) {
  std::fprintf(stderr, "SFX_EMU out of bounds: %s index %lld not in [0,%lld) at %s:%d\n", what, idx, n, file, line);
  std::fflush(stderr);
  std::abort();
}

using namespace sfx;

static void run_grid(int64_t nblk, int threads, size_t smem_cplx,
                     const std::function<void(int, int64_t, double2*)>& body) {
  std::vector<double2> smem(smem_cplx);
  for (int64_t b = 0; b < nblk; ++b) {
    std::barrier<> bar(threads);
    for (auto& x : smem) x = make_double2(1e300, -1e300);   // poison: stale reads show up
    std::vector<std::thread> th;
    th.reserve(threads);
    for (int t = 0; t < threads; ++t)
      th.emplace_back([&, t] { g_bar = &bar; body(t, b, smem.data()); });
    for (auto& x : th) x.join();
  }
}

template <int MP, int OP>
static void emu_fftprod_mp(const FftProdArgs& a, int wpb) {
  const int64_t pts = (int64_t)Shape<MP>::PW * wpb;
  const int64_t nblk = (a.np + pts - 1) / pts;
  run_grid(nblk, 64 * wpb, (size_t)wpb * 3 * Shape<MP>::WAVE_CPLX,
           [&](int t, int64_t b, double2* s) { fftprod_body<MP, OP>(a, t, b, wpb, s); });
}

template <int OP>
static int emu_op(const FftProdArgs& a, int wpb) {
  switch (a.MP) {
    case 4: emu_fftprod_mp<4, OP>(a, wpb); return 0;
    case 8: emu_fftprod_mp<8, OP>(a, wpb); return 0;
    case 16: emu_fftprod_mp<16, OP>(a, wpb); return 0;
    case 32: emu_fftprod_mp<32, OP>(a, wpb); return 0;
    case 64: emu_fftprod_mp<64, OP>(a, wpb); return 0;
    case 128: emu_fftprod_mp<128, OP>(a, wpb); return 0;
    case 256: emu_fftprod_mp<256, OP>(a, wpb); return 0;
  }
  return 1;
}

extern "C" int emu_fftprod(int op, const double* V1, const double* V2, double* Vout, int64_t np, int M,
                           int nb1, int nb2, int nbout, int wpb) {
  FftProdArgs a;
  a.V1 = V1; a.V2 = V2; a.Vout = Vout;
  a.np = np; a.M = M; a.MP = mp_for(M); a.nb1 = nb1; a.nb2 = nb2;
  a.n1 = np * nb1 * (int64_t)M; a.n2 = np * nb2 * (int64_t)M; a.nout = np * nbout * (int64_t)M;
  std::vector<double2> wN, wMP;
  host_twiddles(a.MP, wN, wMP);
  a.wN = wN.data(); a.wMP = wMP.data();
  switch (op) {
    case OP_CROSS: return emu_op<OP_CROSS>(a, wpb);
    case OP_DOT: return emu_op<OP_DOT>(a, wpb);
    case OP_PROD: return emu_op<OP_PROD>(a, wpb);
    case OP_SVECT1: return emu_op<OP_SVECT1>(a, wpb);
    case OP_TENSOR: return emu_op<OP_TENSOR>(a, wpb);
  }
  return 2;
}
