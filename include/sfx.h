/* sfx: MI355X (gfx950) pseudo-spectral nonlinear terms for SFEMaNS - C ABI.
 *
 * Arrays use the Fortran layout of the reference (column-major, point index
 * fastest): a Fourier field is V(np, nb_field, m_max_c) with (cos, sin)
 * column pairs per scalar component; the vector order is
 * (r_c, r_s, theta_c, theta_s, z_c, z_s).  All arithmetic is FP64.
 *
 * Every entry point returns an int status: SFX_OK, or an error code with a
 * message available from sfx_last_error().  The reference stops the program
 * on the same conditions (WRITE + STOP or error_petsc); the Fortran shim maps
 * a non-zero status to error_petsc.
 *
 * Host-pointer entry points copy inputs to the GPU, run, and copy results
 * back (synchronous).  "_dev" entry points take device pointers and a HIP
 * stream (hipStream_t passed as void*, NULL = default stream) and return
 * after the launch; they do not synchronise.
 *
 * Scope: one GPU, nb_procs = 1 (the MPI_ALLTOALL transposes of the reference
 * are identities on one rank).  Products are computed on the exactly
 * de-aliased grid N = 3*MP (MP = power of two >= m_max); callers' m_max_pad
 * must describe a de-aliased grid (2*m_max_pad-1 >= 3*m_max-2), which every
 * caller in the reference satisfies (m_max_pad = 3*m_max/2).
 */
#ifndef SFX_H
#define SFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFX_OK 0
#define SFX_ERR_ARG 1          /* invalid argument (reference: WRITE+STOP / error_petsc) */
#define SFX_ERR_UNSUPPORTED 2  /* valid for the reference, outside this library's scope */
#define SFX_ERR_HIP 3          /* HIP runtime error */
#define SFX_ERR_NOMEM 4

/* Message of the last error on the calling thread ("" if none). */
const char* sfx_last_error(void);

/* Library version string. */
const char* sfx_version(void);

/* Select and initialise the device used by subsequent calls on this thread. */
int sfx_init(int32_t device);

/* ---------------------------------------------------------------------------
 * Pseudo-spectral products at arbitrary points (Gauss points or nodes).
 * ------------------------------------------------------------------------- */

/* Replaces FFT_PAR_CROSS_PROD_DCL (SOURCES/POST_PROCESSING/fft_parallel.F90:147).
 * V_out(np,6,m) = modes 0..m_max_c-1 of V1 x V2, V1/V2 (np,6,m_max_c).
 * temps (nullable, 3 doubles) is zeroed then receives
 * [host<->device copy time, kernel time, 0] in seconds. */
int sfx_fft_par_cross_prod_dcl(const double* V1_in, const double* V2_in, double* V_out,
                               int64_t np, int32_t nb_field, int32_t m_max_c,
                               int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                               double* temps);

/* Replaces FFT_PAR_DOT_PROD_DCL (fft_parallel.F90:330): c_out(np,2,m) = V1 . V2. */
int sfx_fft_par_dot_prod_dcl(const double* V1_in, const double* V2_in, double* c_out,
                             int64_t np, int32_t nb_field, int32_t m_max_c,
                             int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                             double* temps);

/* Replaces FFT_PAR_PROD_DCL (fft_parallel.F90:502): c_out(np,2,m) = c1 * c2. */
int sfx_fft_par_prod_dcl(const double* c1_in, const double* c2_in, double* c_out,
                         int64_t np, int32_t m_max_c,
                         int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                         double* temps);

/* Replaces FFT_SCALAR_VECT_DCL (fft_parallel.F90:1021) for pb = 1:
 * V_out(np,6,m) = s * v with V1 the vector (np,6,m) and V2 the scalar (np,2,m).
 * pb = 2 (v / s, not a polynomial product) returns SFX_ERR_UNSUPPORTED. */
int sfx_fft_scalar_vect_dcl(const double* V1_in, const double* V2_in, double* V_out,
                            int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t m_max_c,
                            int32_t pb, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                            double* temps);

/* Replaces FFT_TENSOR_DCL without opt_tension (fft_parallel.F90:1543):
 * V_out(3,np,6,m), V_out(a,:,2b-1:2b,:) = coefficients of V1_a * V2_b.
 * opt_tension != 0 returns SFX_ERR_UNSUPPORTED. */
int sfx_fft_tensor_dcl(const double* V1_in, const double* V2_in, double* V_out,
                       int64_t np, int32_t nb_field, int32_t m_max_c,
                       int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                       int32_t opt_tension, double* temps);

/* Device-pointer forms of the five products above (no copies, no sync). */
int sfx_fft_par_cross_prod_dcl_dev(const double* V1_in, const double* V2_in, double* V_out,
                                   int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);
int sfx_fft_par_dot_prod_dcl_dev(const double* V1_in, const double* V2_in, double* c_out,
                                 int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);
int sfx_fft_par_prod_dcl_dev(const double* c1_in, const double* c2_in, double* c_out,
                             int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);
int sfx_fft_scalar_vect_dcl_dev(const double* V1_in, const double* V2_in, double* V_out,
                                int64_t np, int32_t m_max_c, int32_t pb, int32_t m_max_pad,
                                void* stream);
int sfx_fft_tensor_dcl_dev(const double* V1_in, const double* V2_in, double* V_out,
                           int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SFX_H */
