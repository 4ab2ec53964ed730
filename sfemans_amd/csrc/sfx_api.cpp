// C-ABI of libsfx: argument checks, twiddle tables, host<->device staging.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_exact_body.h"
#include "sfx_kernels.h"
#include "sfx_twiddle.h"

namespace sfx {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(SFX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

void clear_err() { g_err[0] = 0; }

// ---- twiddle tables, one set per device, built on the host in long double
struct DevTables {
  double2* base = nullptr;
  size_t off_wN[9] = {0};
  size_t off_wMP[9] = {0};
};
static std::mutex g_mu;
static std::map<int, DevTables> g_tables;

static int log2i(int v) { int k = 0; while ((1 << k) < v) ++k; return k; }

int get_twiddles(int MP, Twiddles* tw) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tables.find(dev);
  if (it == g_tables.end()) {
    std::vector<double2> host, wN, wMP;
    DevTables t;
    for (int k = 2; k <= 8; ++k) {
      host_twiddles(1 << k, wN, wMP);
      t.off_wN[k] = host.size();
      host.insert(host.end(), wN.begin(), wN.end());
      t.off_wMP[k] = host.size();
      host.insert(host.end(), wMP.begin(), wMP.end());
    }
    e = hipMalloc(&t.base, host.size() * sizeof(double2));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(twiddles)");
    e = hipMemcpy(t.base, host.data(), host.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(twiddles)");
    it = g_tables.emplace(dev, t).first;
  }
  const int k = log2i(MP);
  tw->wN = it->second.base + it->second.off_wN[k];
  tw->wMP = it->second.base + it->second.off_wMP[k];
  return SFX_OK;
}

// ---- grow-only device staging buffers for the host-pointer entry points
constexpr int N_STAGING = 20;   // staging slots per device
struct Staging {
  void* p[N_STAGING] = {nullptr};
  size_t n[N_STAGING] = {0};
};
static std::map<int, Staging> g_staging;

int staging(int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= N_STAGING) return fail(SFX_ERR_ARG, "staging slot %d out of range", slot);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  Staging& s = g_staging[dev];
  if (s.n[slot] < bytes) {
    if (s.p[slot]) (void)hipFree(s.p[slot]);
    s.p[slot] = nullptr;
    s.n[slot] = 0;
    e = hipMalloc(&s.p[slot], bytes ? bytes : 8);
    if (e != hipSuccess) return fail(SFX_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    s.n[slot] = bytes;
  }
  *out = s.p[slot];
  return SFX_OK;
}

// ---- grow-only device scratch of the device-pointer entry points, one set
// per (device, stream).  Work queued on one stream runs in order, so its
// calls may reuse one buffer; calls on two streams get two buffers.  A buffer
// is replaced only after its stream has drained (the kernels queued before
// may still read the old one).
static std::map<std::pair<int, uintptr_t>, Staging> g_stream_scratch;

int stream_scratch(int slot, size_t bytes, hipStream_t st, void** out) {
  if (slot < 0 || slot >= N_STAGING) return fail(SFX_ERR_ARG, "scratch slot %d out of range", slot);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  Staging& s = g_stream_scratch[std::make_pair(dev, (uintptr_t)st)];
  if (s.n[slot] < bytes) {
    if (s.p[slot]) {
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "scratch: hipStreamSynchronize");
      (void)hipFree(s.p[slot]);
    }
    s.p[slot] = nullptr;
    s.n[slot] = 0;
    e = hipMalloc(&s.p[slot], bytes ? bytes : 8);
    if (e != hipSuccess) return fail(SFX_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    s.n[slot] = bytes;
  }
  *out = s.p[slot];
  return SFX_OK;
}

int release_stream_scratch(hipStream_t st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_stream_scratch.find(std::make_pair(dev, (uintptr_t)st));
  if (it == g_stream_scratch.end()) return SFX_OK;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "scratch release: hipStreamSynchronize");
  for (int i = 0; i < N_STAGING; ++i)
    if (it->second.p[i]) (void)hipFree(it->second.p[i]);
  g_stream_scratch.erase(it);
  return SFX_OK;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Common checks of the FFT_PAR_* argument conventions (fft_parallel.F90:186-191).
int check_modes(const char* name, int64_t np, int32_t m_max_c, int32_t m_max_pad) {
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "BUG in %s: m_max_c==0", name);
  if (np < 0) return fail(SFX_ERR_ARG, "%s: np < 0", name);
  if (m_max_c > 256) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_c=%d > 256", name, m_max_c);
  if (2 * (int64_t)m_max_pad - 1 < 3 * (int64_t)m_max_c - 2)
    return fail(SFX_ERR_UNSUPPORTED,
                "%s: m_max_pad=%d gives an aliased grid (2*m_max_pad-1 < 3*m_max-2)", name, m_max_pad);
  return SFX_OK;
}

int check_procs(const char* name, int64_t np, int32_t nb_procs, int64_t bloc_size) {
  if (nb_procs != 1) return fail(SFX_ERR_UNSUPPORTED, "%s: nb_procs=%d (one GPU: nb_procs must be 1)", name, nb_procs);
  if (bloc_size * nb_procs < np) return fail(SFX_ERR_ARG, "%s: bloc_size*nb_procs < np", name);
  return SFX_OK;
}

int run_fftprod_dev(int op, const char* name, const double* V1, const double* V2, double* Vout,
                    int64_t np, int nb1, int nb2, int nbout, int32_t M, hipStream_t st) {
  FftProdArgs a;
  a.V1 = V1; a.V2 = V2; a.Vout = Vout;
  a.np = np; a.M = M; a.MP = mp_for(M); a.nb1 = nb1; a.nb2 = nb2;
  a.n1 = np * nb1 * (int64_t)M; a.n2 = np * nb2 * (int64_t)M; a.nout = np * nbout * (int64_t)M;
  Twiddles tw;
  int rc = get_twiddles(a.MP, &tw);
  if (rc) return rc;
  a.wN = tw.wN; a.wMP = tw.wMP;
  if (np == 0) return SFX_OK;
  hipError_t e = launch_fftprod(op, a, st);
  if (e != hipSuccess) return hip_fail(e, name);
  return SFX_OK;
}

// Host-pointer path: stage, run, copy back, synchronise.
int run_fftprod_host(int op, const char* name, const double* V1, const double* V2, double* Vout,
                     int64_t np, int nb1, int nb2, int nbout, int32_t M, double* temps) {
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  const size_t n1 = (size_t)np * nb1 * M, n2 = (size_t)np * nb2 * M, no = (size_t)np * nbout * M;
  void *d1, *d2, *d3;
  int rc;
  if ((rc = staging(0, n1 * 8, &d1)) || (rc = staging(1, n2 * 8, &d2)) || (rc = staging(2, no * 8, &d3))) return rc;
  double t0 = now_s();
  hipError_t e;
  if ((e = hipMemcpy(d1, V1, n1 * 8, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "H2D");
  if ((e = hipMemcpy(d2, V2, n2 * 8, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "H2D");
  double t1 = now_s();
  rc = run_fftprod_dev(op, name, (const double*)d1, (const double*)d2, (double*)d3, np, nb1, nb2, nbout, M, nullptr);
  if (rc) return rc;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, name);
  double t2 = now_s();
  if ((e = hipMemcpy(Vout, d3, no * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  double t3 = now_s();
  if (temps) { temps[0] = (t1 - t0) + (t3 - t2); temps[1] = t2 - t1; }
  return SFX_OK;
}

}  // namespace sfx

using namespace sfx;

extern "C" {

const char* sfx_last_error(void) { return g_err; }

const char* sfx_version(void) { return "sfx 0.1 (gfx950, FP64)"; }

int sfx_stream_release(void* stream) {
  clear_err();
  return release_stream_scratch((hipStream_t)stream);
}

int sfx_init(int32_t device) {
  clear_err();
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  Twiddles tw;
  return get_twiddles(64, &tw);
}

int sfx_fft_par_cross_prod_dcl(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                               int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                               int32_t m_max_pad, double* temps) {
  clear_err();
  const char* nm = "FFT_PAR_CROSS_PROD_DCL";
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_CROSS_PROD_DCL ");
  if (nb_field != 6) return fail(SFX_ERR_ARG, "%s: the cross product needs nb_field == 6", nm);
  int rc;
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad)) || (rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  return run_fftprod_host(OP_CROSS, nm, V1_in, V2_in, V_out, np, 6, 6, 6, m_max_c, temps);
}

int sfx_fft_par_dot_prod_dcl(const double* V1_in, const double* V2_in, double* c_out, int64_t np,
                             int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                             int32_t m_max_pad, double* temps) {
  clear_err();
  const char* nm = "FFT_PAR_DOT_PROD_DCL";
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_DOT_PROD_DCL ");
  if (nb_field != 6) return fail(SFX_ERR_ARG, "%s: the dot product needs nb_field == 6", nm);
  int rc;
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad)) || (rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  return run_fftprod_host(OP_DOT, nm, V1_in, V2_in, c_out, np, 6, 6, 2, m_max_c, temps);
}

int sfx_fft_par_prod_dcl(const double* c1_in, const double* c2_in, double* c_out, int64_t np,
                         int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                         double* temps) {
  clear_err();
  const char* nm = "FFT_PAR_PROD_DCL";
  if (m_max_c == 0) return fail(SFX_ERR_ARG, " BUG FFT_PAR_PROD_DCL ");
  int rc;
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad)) || (rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  return run_fftprod_host(OP_PROD, nm, c1_in, c2_in, c_out, np, 2, 2, 2, m_max_c, temps);
}

int sfx_fft_scalar_vect_dcl(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                            int32_t nb_field1, int32_t nb_field2, int32_t m_max_c, int32_t pb,
                            int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  const char* nm = "FFT_SCALAR_VECT_DCL";
  if (nb_field1 % 2 != 0 || nb_field2 % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_SCALAR_VECT_DCL ");
  if (nb_field1 != 6 || nb_field2 != 2)
    return fail(SFX_ERR_ARG, "error in dimension of INPUT variables while calling FFT_SCALAR_VECT_DCL ");
  if (pb != 1 && pb != 2) return fail(SFX_ERR_ARG, "error in problem type while calling FFT_SCALAR_VECT_DCL ");
  int rc;
  if ((rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  // v / s is not a polynomial product: it runs on the reference's own grid
  if (pb == 2) return exact_binary_host(EX_SV2, V1_in, V2_in, V_out, np, m_max_c, m_max_pad, temps);
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad))) return rc;
  return run_fftprod_host(OP_SVECT1, nm, V1_in, V2_in, V_out, np, 6, 2, 6, m_max_c, temps);
}

int sfx_fft_tensor_dcl(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                       int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                       int32_t m_max_pad, int32_t opt_tension, double* temps) {
  clear_err();
  const char* nm = "FFT_TENSOR_DCL";
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_TENSOR_DCL ");
  if (nb_field != 6) return fail(SFX_ERR_ARG, "%s: the tensor product needs nb_field == 6", nm);
  int rc;
  if ((rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  // the tension tensor divides by |V2|: it runs on the reference's own grid
  if (opt_tension) return exact_binary_host(EX_TENSION, V1_in, V2_in, V_out, np, m_max_c, m_max_pad, temps);
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad))) return rc;
  return run_fftprod_host(OP_TENSOR, nm, V1_in, V2_in, V_out, np, 6, 6, 18, m_max_c, temps);
}

// FFT_COMPUTE_DIFFU_MOM: two band-limited products s * V1 and s * V2.
int sfx_fft_compute_diffu_mom(const double* V1_in, const double* V2_in, const double* V3_in, double* V1_out,
                              double* V2_out, int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t nb_field3,
                              int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                              double* temps) {
  clear_err();
  const char* nm = "FFT_COMPUTE_DIFFU_MOM";
  if (nb_field1 != 6 || nb_field2 != 6 || nb_field3 != 2 || m_max_c == 0)
    return fail(SFX_ERR_ARG, "error on inputs while calling FFT_COMPUTE_DIFFU_MOM");
  int rc;
  if ((rc = check_modes(nm, np, m_max_c, m_max_pad)) || (rc = check_procs(nm, np, nb_procs, bloc_size))) return rc;
  double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
  if ((rc = run_fftprod_host(OP_SVECT1, nm, V1_in, V3_in, V1_out, np, 6, 2, 6, m_max_c, temps ? t1 : nullptr)))
    return rc;
  if ((rc = run_fftprod_host(OP_SVECT1, nm, V2_in, V3_in, V2_out, np, 6, 2, 6, m_max_c, temps ? t2 : nullptr)))
    return rc;
  if (temps) for (int i = 0; i < 3; ++i) temps[i] = t1[i] + t2[i];
  return SFX_OK;
}

#define DEV_CHECK(nm)                                            \
  clear_err();                                                   \
  {                                                              \
    int rc_ = check_modes(nm, np, m_max_c, m_max_pad);           \
    if (rc_) return rc_;                                         \
  }

int sfx_fft_par_cross_prod_dcl_dev(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                                   int32_t m_max_c, int32_t m_max_pad, void* stream) {
  DEV_CHECK("FFT_PAR_CROSS_PROD_DCL");
  return run_fftprod_dev(OP_CROSS, "FFT_PAR_CROSS_PROD_DCL", V1_in, V2_in, V_out, np, 6, 6, 6, m_max_c,
                         (hipStream_t)stream);
}
int sfx_fft_par_dot_prod_dcl_dev(const double* V1_in, const double* V2_in, double* c_out, int64_t np,
                                 int32_t m_max_c, int32_t m_max_pad, void* stream) {
  DEV_CHECK("FFT_PAR_DOT_PROD_DCL");
  return run_fftprod_dev(OP_DOT, "FFT_PAR_DOT_PROD_DCL", V1_in, V2_in, c_out, np, 6, 6, 2, m_max_c,
                         (hipStream_t)stream);
}
int sfx_fft_par_prod_dcl_dev(const double* c1_in, const double* c2_in, double* c_out, int64_t np,
                             int32_t m_max_c, int32_t m_max_pad, void* stream) {
  DEV_CHECK("FFT_PAR_PROD_DCL");
  return run_fftprod_dev(OP_PROD, "FFT_PAR_PROD_DCL", c1_in, c2_in, c_out, np, 2, 2, 2, m_max_c,
                         (hipStream_t)stream);
}
int sfx_fft_scalar_vect_dcl_dev(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                                int32_t m_max_c, int32_t pb, int32_t m_max_pad, void* stream) {
  clear_err();
  if (pb == 2) return exact_sv2_dev(V1_in, V2_in, V_out, np, m_max_c, m_max_pad, (hipStream_t)stream);
  if (pb != 1) return fail(SFX_ERR_ARG, "error in problem type while calling FFT_SCALAR_VECT_DCL ");
  DEV_CHECK("FFT_SCALAR_VECT_DCL");
  return run_fftprod_dev(OP_SVECT1, "FFT_SCALAR_VECT_DCL", V1_in, V2_in, V_out, np, 6, 2, 6, m_max_c,
                         (hipStream_t)stream);
}
int sfx_fft_tensor_dcl_dev(const double* V1_in, const double* V2_in, double* V_out, int64_t np,
                           int32_t m_max_c, int32_t m_max_pad, int32_t opt_tension, void* stream) {
  clear_err();
  if (opt_tension) return exact_tension_dev(V1_in, V2_in, V_out, np, m_max_c, m_max_pad, (hipStream_t)stream);
  DEV_CHECK("FFT_TENSOR_DCL");
  return run_fftprod_dev(OP_TENSOR, "FFT_TENSOR_DCL", V1_in, V2_in, V_out, np, 6, 6, 18, m_max_c,
                         (hipStream_t)stream);
}

int sfx_fft_compute_diffu_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in, double* V1_out,
                                  double* V2_out, int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream) {
  DEV_CHECK("FFT_COMPUTE_DIFFU_MOM");
  int rc = run_fftprod_dev(OP_SVECT1, "FFT_COMPUTE_DIFFU_MOM", V1_in, V3_in, V1_out, np, 6, 2, 6, m_max_c,
                           (hipStream_t)stream);
  if (rc) return rc;
  return run_fftprod_dev(OP_SVECT1, "FFT_COMPUTE_DIFFU_MOM", V2_in, V3_in, V2_out, np, 6, 2, 6, m_max_c,
                         (hipStream_t)stream);
}

}  // extern "C"
