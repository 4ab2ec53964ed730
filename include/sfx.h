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

/* Scratch of the "_dev" entry points.  The exact-grid products (the
 * non-polynomial FFT_PAR_* routines and the entropy viscosities) keep a
 * grow-only scratch per (device, stream): calls on different streams run
 * concurrently without sharing it.  sfx_stream_release waits for the stream
 * and frees its scratch; call it before destroying a stream that ran such
 * products.  No counterpart in the reference (FFTW plans and temporary
 * arrays there are per call). */
int sfx_stream_release(void* stream);

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

/* Replaces FFT_SCALAR_VECT_DCL (fft_parallel.F90:1021): V1 the vector (np,6,m),
 * V2 the scalar (np,2,m); pb = 1: V_out(np,6,m) = s * v; pb = 2: v / s,
 * evaluated on the reference's own grid N = 2*m_max_pad-1 (the quotient is
 * not band-limited, so its coefficients depend on N). */
int sfx_fft_scalar_vect_dcl(const double* V1_in, const double* V2_in, double* V_out,
                            int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t m_max_c,
                            int32_t pb, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                            double* temps);

/* Replaces FFT_TENSOR_DCL (fft_parallel.F90:1543): V_out(3,np,6,m),
 * V_out(a,:,2b-1:2b,:) = coefficients of V1_a * V2_b, or with opt_tension != 0
 * of V1_a V2_b/(|V2| + 1e-14) - delta_ab |V2| (:1688-1706), evaluated on the
 * reference's own grid N = 2*m_max_pad-1. */
int sfx_fft_tensor_dcl(const double* V1_in, const double* V2_in, double* V_out,
                       int64_t np, int32_t nb_field, int32_t m_max_c,
                       int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                       int32_t opt_tension, double* temps);

/* Device-pointer forms of the five products above (no copies, no sync);
 * scalar-vector pb = 2 and the tension tensor use the exact-grid scratch below. */
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
                           int64_t np, int32_t m_max_c, int32_t m_max_pad, int32_t opt_tension,
                           void* stream);

/* ---------------------------------------------------------------------------
 * Non-polynomial products on the reference's own azimuthal grid
 * N = 2*m_max_pad-1 (odd), as real DFTs: their coefficients depend on N.
 * m_max_pad >= m_max_c is required.  The "_dev" forms keep a grow-only
 * scratch (up to 2 * outputs * m_max_pad * np doubles) per (device, stream):
 * calls on different streams may overlap.  A call that must grow it waits
 * for its stream first, so under HIP graph capture make one call of the
 * largest size beforehand.  sfx_stream_release frees a stream's scratch.
 * The host-pointer forms stage their arrays in per-device buffers: call
 * them from one host thread at a time per device.
 * ------------------------------------------------------------------------- */

/* Replaces FFT_PAR_DIV_DCL (fft_parallel.F90:667): c_out(np,2,m) = c1 / c2. */
int sfx_fft_par_div_dcl(const double* c1_in, const double* c2_in, double* c_out, int64_t np,
                        int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                        double* temps);
int sfx_fft_par_div_dcl_dev(const double* c1_in, const double* c2_in, double* c_out, int64_t np,
                            int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_HEAVISIDE_DCL (fft_parallel.F90:832): V1_in(n_values-1, np, 2, m)
 * level sets, values(n_values) (host array), coeff_tanh the half-width of regul
 * (:2234-2246); on the grid rho = values(1), then for each interface i
 * rho = rho*(1 - regul(phi_i)) + values(i+1)*regul(phi_i); V_out(np,2,m).
 * At most 4 interfaces. */
int sfx_fft_heaviside_dcl(const double* V1_in, const double* values, double* V_out, int32_t n_values,
                          int64_t np, int32_t nb_field, int32_t m_max_c, int32_t nb_procs,
                          int64_t bloc_size, int32_t m_max_pad, double coeff_tanh, double* temps);
int sfx_fft_heaviside_dcl_dev(const double* V1_in, const double* values, double* V_out, int32_t n_values,
                              int64_t np, int32_t m_max_c, int32_t m_max_pad, double coeff_tanh,
                              void* stream);

/* Replaces FFT_SQUAREROOT_SCALAR (fft_parallel.F90:4601): c1_inout(np,nb_field,m),
 * its first scalar (columns 1:2) overwritten by SQRT(MAX(0, MIN(1, c))). */
int sfx_fft_squareroot_scalar(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                              int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps);
int sfx_fft_squareroot_scalar_dev(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                  int32_t m_max_pad, void* stream);

/* Replaces FFT_SCALAR_VECT_NO_OVERSHOOT (fft_parallel.F90:1222): as
 * sfx_fft_scalar_vect_dcl (pb = 1 product, pb = 2 quotient) with the scalar
 * clamped on the grid to [MINVAL(scalar_bounds), MAXVAL(scalar_bounds)]
 * (:1352-1361); scalar_bounds is a host array of n_bounds values. */
int sfx_fft_scalar_vect_no_overshoot(const double* scalar_bounds, int32_t n_bounds, const double* V1_in,
                                     const double* V2_in, double* V_out, int64_t np, int32_t nb_field1,
                                     int32_t nb_field2, int32_t m_max_c, int32_t pb, int32_t nb_procs,
                                     int64_t bloc_size, int32_t m_max_pad, double* temps);
int sfx_fft_scalar_vect_no_overshoot_dev(const double* scalar_bounds, int32_t n_bounds, const double* V1_in,
                                         const double* V2_in, double* V_out, int64_t np, int32_t m_max_c,
                                         int32_t pb, int32_t m_max_pad, void* stream);

/* Replaces FFT_NO_OVERSHOOT_LEVEL_SET (fft_parallel.F90:3649): c1_inout(np,nb_field,m),
 * its first scalar overwritten by MAX(0, MIN(1, c)) on the grid (:3771). */
int sfx_fft_no_overshoot_level_set(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                   int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps);
int sfx_fft_no_overshoot_level_set_dev(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                       int32_t m_max_pad, void* stream);

/* Replaces FFT_PAR_ENTRO_VISC_DCL (fft_parallel.F90:4055): V1_in(np,6,m) vector,
 * c1_in(np,2,m) scalar, c_in_real(2*m_max_pad-1, np) a field given on the grid
 * angles; V_out(np,6,m) = coefficients of (coeff1_in_level - c_in_real) V1 and
 * c_out(np,2,m) of MAX(0, c (1 - c)) (:4209-4219).  hloc_gauss is accepted for
 * signature parity and not read (nullable). */
int sfx_fft_par_entro_visc_dcl(const double* V1_in, const double* c1_in, const double* c_in_real,
                               const double* hloc_gauss, double coeff1_in_level, double* V_out, double* c_out,
                               int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t m_max_c, int32_t nb_procs,
                               int64_t bloc_size, int32_t m_max_pad, double* temps);
int sfx_fft_par_entro_visc_dcl_dev(const double* V1_in, const double* c1_in, const double* c_in_real,
                                   double coeff1_in_level, double* V_out, double* c_out, int64_t np,
                                   int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_PAR_COMPR_ENTRO_VISC_DCL (fft_parallel.F90:3160): V1_in = grad(phi),
 * V2_in = grad(phi_reg) (np,6,m), c1_in = phi (np,2,m), c_in_real(2*m_max_pad-1, np)
 * the entropy viscosity on the grid angles, level_set_comp_factor the module
 * input the reference reads (inputs%level_set_comp_factor); on the grid
 * c2 = MAX(0, phi (1 - phi)), V_out = coefficients of
 * (coeff1_in_level - c_in_real) V1 + c_in_real comp c2/(|V2| + 1e-14) V2 and
 * c_out of c2 (:3316-3330).  hloc_gauss is accepted for signature parity. */
int sfx_fft_par_compr_entro_visc_dcl(const double* V1_in, const double* V2_in, const double* c1_in,
                                     const double* c_in_real, const double* hloc_gauss, double coeff1_in_level,
                                     double level_set_comp_factor, double* V_out, double* c_out, int64_t np,
                                     int32_t nb_field1, int32_t nb_field2, int32_t nb_field3, int32_t m_max_c,
                                     int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps);
int sfx_fft_par_compr_entro_visc_dcl_dev(const double* V1_in, const double* V2_in, const double* c1_in,
                                         const double* c_in_real, double coeff1_in_level,
                                         double level_set_comp_factor, double* V_out, double* c_out, int64_t np,
                                         int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_COMPUTE_ENTROPY_VISC (fft_parallel.F90:2418): V1_in velocity,
 * V2_in residual, V3_in..V5_in the three rows of a tensor (np,6,m), hloc_gauss(np)
 * the mesh size per Gauss point, points grouped l_G per element (np a multiple
 * of l_G, l_G <= 16).  On the grid: |v| maximised over the element's points and
 * the neighbouring angles (periodic) (:2571-2591), visco = -les_coeff1 +
 * MIN(les_coeff4 |v|max, les_coeff2 h |v . res|/(residual_normalization + 1e-14))
 * (:2601-2607); Vk_out = coefficients of visco V(k+2) (:2612-2620).  The
 * reference also reduces max |v| over the ranks without using it; that is not
 * computed.  les_coeff1/2/4 are the module inputs it reads (inputs%LES_coeff*). */
int sfx_fft_compute_entropy_visc(const double* V1_in, const double* V2_in, const double* V3_in, const double* V4_in,
                                 const double* V5_in, const double* hloc_gauss, double* V1_out, double* V2_out,
                                 double* V3_out, int64_t np, int32_t nb_field, int32_t m_max_c, int32_t nb_procs,
                                 int64_t bloc_size, int32_t m_max_pad, double residual_normalization, int32_t l_G,
                                 double les_coeff1, double les_coeff2, double les_coeff4, double* temps);
int sfx_fft_compute_entropy_visc_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                    const double* V4_in, const double* V5_in, const double* hloc_gauss,
                                    double* V1_out, double* V2_out, double* V3_out, int64_t np, int32_t m_max_c,
                                    int32_t m_max_pad, double residual_normalization, int32_t l_G,
                                    double les_coeff1, double les_coeff2, double les_coeff4, void* stream);

/* Replaces FFT_COMPUTE_ENTROPY_VISC_MOM (fft_parallel.F90:2701): velocity V1_in,
 * momentum V2_in, residual V3_in (np,6,m), mass residual c1_in (np,2,m),
 * hloc_gauss(np), points grouped l_G per element (np a multiple of l_G <= 16).
 * On the grid: |u| and |m| maximised over the element's points and the
 * neighbouring angles (:2839-2870), the raw viscosity les_coeff2 h
 * MAX(|u . res|, |c |u|^2|) maximised likewise (:2890-2916), vmax the maximum
 * of |u| (one rank: over all points and angles); outputs on the grid angles,
 * c1_real_out(2*m_max_pad-1, np) and, if not null, opt_c2_real_out (:2918-2931).
 * les_coeff2/4 are the module inputs the reference reads. */
int sfx_fft_compute_entropy_visc_mom(const double* V1_in, const double* V2_in, const double* V3_in,
                                     const double* c1_in, const double* hloc_gauss, double* c1_real_out,
                                     double* opt_c2_real_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                                     int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, int32_t l_G,
                                     double les_coeff2, double les_coeff4, double* temps);
int sfx_fft_compute_entropy_visc_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                         const double* c1_in, const double* hloc_gauss, double* c1_real_out,
                                         double* opt_c2_real_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                                         int32_t l_G, double les_coeff2, double les_coeff4, void* stream);

/* Replaces FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM (fft_parallel.F90:3419): V1_in,
 * V2_in, V3_in (np,6,m) the three momentum gradients, c_in_real(2*m_max_pad-1, np)
 * the entropy viscosity on the grid angles, les_coeff1_mom the module input the
 * reference reads (inputs%LES_coeff1_mom); V_out(3,np,6,m), V_out(a,:,2f-1:2f,:)
 * = coefficients of (c_in_real - les_coeff1_mom) V_a,f (:3552-3568). */
int sfx_fft_compute_entropy_visc_grad_mom(const double* V1_in, const double* V2_in, const double* V3_in,
                                          const double* c_in_real, double les_coeff1_mom, double* V_out, int64_t np,
                                          int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                                          int32_t m_max_pad, double* temps);
int sfx_fft_compute_entropy_visc_grad_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                              const double* c_in_real, double les_coeff1_mom, double* V_out,
                                              int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_PAR_VAR_ETA_PROD_T_DCL (fft_parallel.F90:1798) and
 * FFT_PAR_VAR_ETA_PROD_GAUSS_DCL (:2015): c_out(np,nb_field,m) = coefficients
 * of eta * c_in on the grid N = 2*m_max_pad-1, nb_field 2 or 6 (:1954-1962).
 * The reference calls a user function ETA(mesh, angles, N, nb, ne, time) for
 * eta on the angles 2*pi*j/N; here the caller passes its values,
 * eta_grid(N, np) column-major (the Fortran shim calls the same function to
 * fill it: nodes for the _T_ form, Gauss points for the _GAUSS_ form). */
int sfx_fft_par_var_eta_prod_dcl(const double* eta_grid, const double* c_in, double* c_out, int64_t np,
                                 int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                                 int32_t m_max_pad, double* temps);
int sfx_fft_par_var_eta_prod_dcl_dev(const double* eta_grid, const double* c_in, double* c_out, int64_t np,
                                     int32_t nb_field, int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Analysis of values given on the grid, the last step of every FFT_PAR_*
 * product (fft_parallel.F90:277-324 for the cross product): real_in(N, np)
 * on the angles 2*pi*j/N, N = 2*m_max_pad-1 -> c_out(np,2,m_max_c), mode 0
 * the mean, mode m the (cos, sin) coefficients.  With sfx_fft_par_real it
 * composes FFT_PAR_SCAL_FUNCT (fft_parallel.F90:4294), whose user function
 * funct(x) runs where the caller's code runs: real planes, funct applied to
 * each value, then this analysis. */
int sfx_fft_par_from_real(const double* real_in, double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad);
int sfx_fft_par_from_real_dev(const double* real_in, double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                              void* stream);

/* Replaces FFT_MAX_MIN_VEL_DCL (fft_parallel.F90:4467): V_out[0] = max of |v|
 * over the grid (|v| from the first three fields if nb_field = 6, else
 * |first field|), V_out[1] = the minimum as the reference computes it
 * (:4563-4576): its start value is the first positive |v| on the last grid
 * angle, so it is the minimum of the positive |v| if |v| > 0 somewhere on
 * the last angle, else 0.  _dev: V_out is device memory (2 doubles). */
int sfx_fft_max_min_vel_dcl(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                            int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad);
int sfx_fft_max_min_vel_dcl_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                                int32_t m_max_pad, void* stream);

/* Replaces FFT_CHECK_INTERFACE (fft_parallel.F90:2267): V1_in(nb_fluids-1,
 * np, nb_field = 2, m) level sets; interface_ok = 0 if at some point and grid
 * angle a level set in [0.1, 0.9] is below the next one (:2388-2404), else 1
 * (always 1 for two fluids).  At most 5 fluids.  _dev: interface_ok is device
 * memory (one int32). */
int sfx_fft_check_interface(const double* V1_in, int32_t nb_fluids, int32_t* interface_ok, int64_t np,
                            int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                            int32_t m_max_pad, double* temps);
int sfx_fft_check_interface_dev(const double* V1_in, int32_t nb_fluids, int32_t* interface_ok, int64_t np,
                                int32_t nb_field, int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_PAR_REAL (fft_parallel.F90:28): V1_in (np,nb_field,m_max_c)
 * -> V_out(N, nb_field/2, np), column-major, the values of each (cos, sin)
 * column pair on the N grid angles 2*pi*j/N (the synthesis of
 * :98-140).  N = sfx_fft_par_real_n(m_max_c, opt_nb_plane) = 2*m_max_pad-1
 * with m_max_pad = (opt_nb_plane+1)/2 when opt_nb_plane > 2*m_max_c-1, else
 * m_max_c (:76-84; opt_nb_plane <= 0 when absent).  The reference allocates
 * V_out itself; here the caller passes N*(nb_field/2)*np doubles.  nb_procs
 * must be 1 (the MPI_ALLTOALL is not implemented). */
int32_t sfx_fft_par_real_n(int32_t m_max_c, int32_t opt_nb_plane);
int sfx_fft_par_real(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                     int32_t nb_procs, int32_t opt_nb_plane);
int sfx_fft_par_real_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                         int32_t opt_nb_plane, void* stream);

/* Replaces FFT_COMPRESSION_LEVEL_SET_DCL (fft_parallel.F90:3821): V1_in
 * (np,6,m) the level-set gradient (components 1-3 used), V2_in (np,6,m) the
 * velocity, c_in (np,2,m) the level set phi -> c_out (np,2,m), the
 * coefficients of level_set_comp_factor*4*les_coeff4*min(1, max|v|)
 * *(2 regul(phi, 0.01) - 1)*(max(1-(2phi-1)^2, 0)/(2 h_min_distance) - |grad phi|)
 * (:3997-4006).  max|v| is taken over all points and angles (one rank): the
 * element and angle maxima the reference takes first (:3960-3984) do not
 * change it, so hloc_gauss and l_G are accepted and not read.
 * level_set_comp_factor, les_coeff4 and h_min_distance are the module inputs
 * the reference reads (inputs%...). */
int sfx_fft_compression_level_set_dcl(const double* V1_in, const double* V2_in, const double* c_in, double* c_out,
                                      const double* hloc_gauss, int32_t l_G, int64_t np, int32_t nb_field,
                                      int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                                      double level_set_comp_factor, double les_coeff4, double h_min_distance,
                                      double* temps);
int sfx_fft_compression_level_set_dcl_dev(const double* V1_in, const double* V2_in, const double* c_in,
                                          double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                                          double level_set_comp_factor, double les_coeff4, double h_min_distance,
                                          void* stream);

/* Replaces FFT_COMPUTE_DIFFU_MOM (fft_parallel.F90:2939): V1_out = s V1 and
 * V2_out = s V2 with V1, V2 (np,6,m) vectors and s = V3 (np,2,m) a scalar;
 * band-limited products, computed by the product kernel. */
int sfx_fft_compute_diffu_mom(const double* V1_in, const double* V2_in, const double* V3_in, double* V1_out,
                              double* V2_out, int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t nb_field3,
                              int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                              double* temps);
int sfx_fft_compute_diffu_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in, double* V1_out,
                                  double* V2_out, int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream);

/* Replaces FFT_MAX_VEL_DCL (fft_parallel.F90:1426): *V_out = max over points and
 * grid angles of |v| (nb_field 6) or |c| (nb_field 2).  The _dev form writes
 * the maximum to device memory. */
int sfx_fft_max_vel_dcl(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                        int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad);
int sfx_fft_max_vel_dcl_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field,
                            int32_t m_max_c, int32_t m_max_pad, void* stream);

/* ---------------------------------------------------------------------------
 * Meshes.  A P2 meridian mesh is registered once (the reference builds it at
 * initialisation: INIT, SOURCES/INITIALISATION/initialization.F90:1153-1222)
 * and is immutable afterwards.  jj(6,me) is 1-based with the reference local
 * numbering (mesh%jj, SOURCES/MESH/def_type_mesh.F90:74-112), rr(2,np) the
 * node coordinates (r, z).  Elements may be curved (midpoint nodes moved onto
 * a curved or spherical boundary, as create_iso_grid_distributed does,
 * SOURCES/MESH/refine_mesh.F90:342-353): the kernels take the Jacobian of
 * the isoparametric map at every Gauss point from all six nodes
 * (SOURCES/GAUSS_POINTS/gauss_points_2d.F90:108-122).  An element whose
 * Jacobian vanishes at a Gauss point is rejected (SFX_ERR_ARG).
 * ------------------------------------------------------------------------- */
typedef struct sfx_mesh_s* sfx_mesh_t;

int sfx_mesh_create(const int32_t* jj, const double* rr, int32_t n_w, int32_t l_G, int64_t me, int64_t np,
                    sfx_mesh_t* out);
int sfx_mesh_destroy(sfx_mesh_t mesh);
/* Boundary sides of the mesh, for the boundary term of rhs_3x3: jjs(3, mes)
 * 1-based (mesh%jjs, SOURCES/MESH/def_type_mesh.F90:76; the two ends of the
 * side, then its midpoint node).  Replaces any sides registered before. */
int sfx_mesh_set_boundary(sfx_mesh_t mesh, const int32_t* jjs, int64_t mes);
int sfx_mesh_info(sfx_mesh_t mesh, int64_t* me, int64_t* np);

/* ---------------------------------------------------------------------------
 * Fused Gauss-point nonlinear terms (interpolation + curl/gradient +
 * de-aliased product), output at the l_G*me Gauss points, index
 * (m-1)*l_G + l.  list_mode(m_max_c) holds the mode numbers (the product
 * treats slots as contiguous modes, fft_parallel.F90:19-25).
 * ------------------------------------------------------------------------- */

/* Replaces smb_cross_prod_gauss_sft_par
 * (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514): V_out(G,6,m) =
 * (curl V_in + 2 Omega) x V_in with the precession vector of taux_precession,
 * angle_s_pi when precession_in != 0. */
int sfx_smb_cross_prod_gauss_sft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                     const double* V_in, double* V_out, int32_t precession_in,
                                     double taux_precession, double angle_s_pi);
int sfx_smb_cross_prod_gauss_sft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                         const double* V_in, double* V_out, int32_t precession_in,
                                         double taux_precession, double angle_s_pi, void* stream);

/* smb_cross_prod_gauss_sft_par followed by rhs_3x3 of its output, fused
 * (subroutine_ns_with_velocity.F90:514-630 then assembling_rhs.F90:5-189
 * through rhs_ns_gauss_3x3's -rotv_v, computing_rhs_gauss.F90:100):
 * vb_145(np,3,m), vb_236(np,3,m) = scale * rhs_3x3((curl V_in + 2 Omega) x V_in)
 * for every mode, with no Gauss-point array: the element sums are formed in
 * the work-group and gathered node by node in element order, giving the
 * same values, bit for bit, as the two calls with scale passed to
 * sfx_rhs_3x3_dev (scale = -1 for the -rotv_v share of the NS right-hand
 * side).  work_elem: 36*me*m doubles. */
int sfx_nl_rhs_3x3_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* V_in,
                       double* vb_145, double* vb_236, double scale, int32_t precession_in, double taux_precession,
                       double angle_s_pi, double* work_elem, void* stream);

/* Replaces smb_CurlH_cross_B_gauss_sft_par (subroutine_ns_with_velocity.F90:632):
 * V_out(G,6,m) = (curl V_in) x W_in, V_in and W_in nodal (np,6,m). */
int sfx_smb_curlh_cross_b_gauss_sft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                        const double* V_in, const double* W_in, double* V_out);
int sfx_smb_curlh_cross_b_gauss_sft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                            const double* V_in, const double* W_in, double* V_out, void* stream);

/* Replaces smb_ugradc_gauss_fft_par
 * (SOURCES/SIMULATION/subroutine_temperature_with_T.F90:247):
 * c_out(G,2,m) = (hc grad c).u + (div u)(hc c), heat_capa_in(me) per element. */
int sfx_smb_ugradc_gauss_fft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                 const double* heat_capa_in, const double* V_in, const double* c_in,
                                 double* c_out);
int sfx_smb_ugradc_gauss_fft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                     const double* heat_capa_dev, const double* V_in, const double* c_in,
                                     double* c_out, void* stream);

/* The MHD nonlinear term of the Navier-Stokes step in one launch:
 * rotv_v(G,6,m) = (curl u) x u - coeff_lorentz (curl H) x B, as formed by
 * subroutine_ns_with_velocity.F90:227 and :246-248 from two smb_* calls. */
int sfx_nl_rotv_mhd(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* u,
                    const double* H, const double* B, double coeff_lorentz, double* rotv_v);
int sfx_nl_rotv_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* u,
                        const double* H, const double* B, double coeff_lorentz, double* rotv_v, void* stream);

/* ---------------------------------------------------------------------------
 * Tangent of the nonlinear terms (no counterpart in the reference, whose
 * LightKrylov operator binds rmatvec to the forward matvec,
 * SOURCES/LIGHTKRYLOV_OBJECTS/exponential_propagator.F90:61-62).  Around a
 * base flow U, with B(a,b) = P[(curl a) x b] the bilinear map of
 * smb_cross_prod_gauss_sft_par: out(G,6,m) = B(U,dU) + B(dU,U).  The base
 * flow's Gauss values are recomputed from its nodal coefficients.
 * ------------------------------------------------------------------------- */
int sfx_nl_tangent(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* U,
                   const double* dU, double* out);
int sfx_nl_tangent_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                       const double* dU, double* out, void* stream);
/* Exact transpose of sfx_nl_tangent in the Euclidean inner product of the
 * coefficient arrays: dbar(np,6,m) = T_U^T ybar for ybar(G,6,m).  work_elem
 * is a caller-provided (6*me*6*m) device buffer of per-element nodal
 * contributions, summed node by node in element order (deterministic). */
int sfx_nl_adjoint(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* U, const double* ybar,
                   double* dbar);
int sfx_nl_adjoint_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                       const double* ybar, double* work_elem, double* dbar, void* stream);
/* Linearized convective and Lorentz terms around (U, H0, B0):
 * out = B(U,dU) + B(dU,U) - coeff_lorentz [B(H0,dB) + B(dH,B0)]. */
int sfx_nl_tangent_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                           const double* H0, const double* B0, const double* dU, const double* dH,
                           const double* dB, double coeff_lorentz, double* out, void* stream);
/* Exact transposes of the Lorentz part and of the whole of sfx_nl_tangent_mhd_dev
 * (no counterpart in the reference, whose LightKrylov operator uses the
 * forward map for rmatvec: SOURCES/LIGHTKRYLOV_OBJECTS/exponential_propagator.F90:61-62).
 * For ybar(G,6,m): dHbar, dBbar (np,6,m) of -coeff_lorentz [B(H0,dB) + B(dH,B0)];
 * the MHD form adds dUbar of the convective part.  work_elem holds
 * 2*36*me*m doubles. */
int sfx_nl_adjoint_lorentz_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* H0,
                               const double* B0, double coeff_lorentz, const double* ybar, double* work_elem,
                               double* dHbar, double* dBbar, void* stream);
int sfx_nl_adjoint_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                           const double* H0, const double* B0, double coeff_lorentz, const double* ybar,
                           double* work_elem, double* dUbar, double* dHbar, double* dBbar, void* stream);

/* ---------------------------------------------------------------------------
 * The inner product of SFEMaNS fields and the adjoint in it.
 * ------------------------------------------------------------------------- */

/* Replaces dot_product_SF (SOURCES/SOLVER/tn_axi.F90:9-37; dot_product_champ,
 * fem_tn_ns_mhd.F90:92-135; the element loop of fem_tn_axi.F90:8-48), one rank:
 * 2 pi SUM_k w_k SUM_c SUM_(e,l) SUM(v ww) SUM(w ww) rj ray over the nodal
 * arrays v, w (np, nb_field, m_max_c), w_k = 1 if list_mode(k) = 0, else 1/2.
 * LightKrylov's dot on SFEMaNS fields is this product
 * (SOURCES/TOOLS/def_type_field.F90:176-198).  The sum is formed in a fixed
 * order (deterministic).  _dev: result_dev is one device double, work holds
 * sfx_dot_product_sf_work(...) doubles; the host form copies v and w. */
int64_t sfx_dot_product_sf_work(sfx_mesh_t mesh, int32_t m_max_c, int32_t nb_field);
int sfx_dot_product_sf_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, int32_t nb_field,
                           const double* v, const double* w, double* result_dev, double* work, void* stream);
int sfx_dot_product_sf(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, int32_t nb_field, const double* v,
                       const double* w, double* result);
/* No counterpart in the reference: sfx_dot_product_sf_dev for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  It takes no work array and allocates nothing: the
 * partial sums (at most 2048 doubles, which covers every size, since the sum
 * is split over at most 2048 work-groups) live in a fixed 16 KiB device array
 * that each GPU holds until the process ends.  Every partial is written
 * before it is read, so nothing needs clearing between calls.  Calls on one
 * GPU share that array: calls that overlap on one GPU (two streams, two
 * threads) race. */
int sfx_dot_product_sf_noalloc(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, int32_t nb_field,
                               const double* v, const double* w, double* result_dev, void* stream);

/* The r-weighted P2 mass matrix M_ij = SUM_(e,l) ww(i,l) ww(j,l) rj ray, the
 * Gram matrix of dot_product_SF per column (no counterpart as a routine in
 * the reference, which solves with PETSc): y = M x for ncols nodal columns
 * (np, ncols); work_elem: 6*me*ncols doubles. */
int sfx_mass_apply_dev(sfx_mesh_t mesh, int32_t ncols, const double* x, double* y, double* work_elem, void* stream);
/* x = M^{-1} b for ncols columns: Jacobi-preconditioned conjugate gradients on
 * every column at once, until (r, D^{-1} r) <= rtol^2 (r0, D^{-1} r0) in every
 * column (rtol <= 0: 1e-14; max_iter <= 0: 1000), checked every 8 iterations
 * (the call synchronises the stream).  work: sfx_mass_solve_work(mesh, ncols)
 * doubles; iters (nullable) receives the iteration count.  SFX_ERR_ARG if
 * not converged. */
int64_t sfx_mass_solve_work(sfx_mesh_t mesh, int32_t ncols);
int sfx_mass_solve_dev(sfx_mesh_t mesh, int32_t ncols, const double* b, double* x, double rtol, int32_t max_iter,
                       double* work, int32_t* iters, void* stream);

/* The linearized convective term as an operator on nodal fields, and its
 * adjoint in dot_product_SF, the inner product LightKrylov uses (no
 * counterpart in the reference, whose LightKrylov operator binds rmatvec to
 * the forward map, SOURCES/LIGHTKRYLOV_OBJECTS/exponential_propagator.F90:61-62):
 *   A dU = M^{-1} R T_U dU   (the L2(r dr dz) projection onto P2, mode by mode
 *                             and column by column, of sfx_nl_tangent; R f =
 *                             SUM_l ww(i,l) f(l) rj ray, qs_00_gauss's right-hand side)
 *   A* y = M^{-1} w^{-1} T_U^T R^T (w y)   so that <A x, y>_SF = <x, A* y>_SF,
 * with T_U^T the Euclidean transpose (sfx_nl_adjoint_dev) and w the mode
 * weights.  All arrays (np, 6, m) on the device; rtol as sfx_mass_solve_dev;
 * work: sfx_nl_proj_work(mesh, m_max_c) doubles.  Both synchronise the stream. */
int64_t sfx_nl_proj_work(sfx_mesh_t mesh, int32_t m_max_c);
int sfx_nl_tangent_proj_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                            const double* dU, double* out, double rtol, double* work, void* stream);
int sfx_nl_adjoint_sf_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                          const double* ybar, double* out, double rtol, double* work, void* stream);

/* ---------------------------------------------------------------------------
 * Right-hand-side assembly into dense local vectors.  On one rank the PETSc
 * index of component ki of node i is (ki-1)*np + i
 * (SOURCES/INITIALISATION/st_matrix.F90:101), so vb_145/vb_236 are (np,3)
 * per mode.  Sums are formed node by node in element order (the order of the
 * reference's VecSetValues loop), without atomics.
 * ------------------------------------------------------------------------- */

/* Replaces rhs_3x3 without opt_tensor/opt_tensor_scaln_bdy
 * (SOURCES/SOLVER/assembling_rhs.F90:5): one mode, rhs_in(G,6) ->
 * vb_145(np,3), vb_236(np,3).  `mode` is accepted for signature parity. */
int sfx_rhs_3x3(sfx_mesh_t mesh, int32_t mode, const double* rhs_in, double* vb_145, double* vb_236);
/* Replaces rhs_3x3 with all its options (SOURCES/SOLVER/assembling_rhs.F90:5-189),
 * one mode, host arrays: rhs_in(G,6), opt_tensor(3,G,6) (nullable; its 1/r
 * and mode terms use `mode`, :56-68), opt_tensor_scaln_bdy(3*mes,6) at the
 * three Gauss points of every boundary side (nullable; needs
 * sfx_mesh_set_boundary, :129-185) -> vb_145(np,3), vb_236(np,3). */
int sfx_rhs_3x3_mode(sfx_mesh_t mesh, int32_t mode, const double* rhs_in, const double* opt_tensor,
                     const double* opt_tensor_scaln_bdy, double* vb_145, double* vb_236);
/* The boundary term of rhs_3x3 for every mode, ADDED to vb_145(np,3,m) and
 * vb_236(np,3,m) (device arrays) times scale: tensor_bdy(3*mes,6,m), per side
 * and Gauss point smb = T*ray*rjs with points on the axis (ray < 1e-10)
 * skipped (assembling_rhs.F90:129-185); each boundary node adds its sides in
 * side order after the volume terms, as the reference's VecSetValues calls do. */
int sfx_rhs_3x3_bdy_dev(sfx_mesh_t mesh, int32_t m_max_c, const double* tensor_bdy, double* vb_145,
                        double* vb_236, double scale, void* stream);
/* All modes at once on the device: rhs_gauss(G,6,m) -> vb_145(np,3,m),
 * vb_236(np,3,m), each multiplied by `scale`.  work_elem is a caller-provided
 * device buffer of 36*me*m doubles (the per-element contributions). */
int sfx_rhs_3x3_dev(sfx_mesh_t mesh, int32_t m_max_c, const double* rhs_gauss, double* vb_145, double* vb_236,
                    double scale, double* work_elem, void* stream);
/* Replaces rhs_ns_gauss_3x3 without the penalty options
 * (SOURCES/SOLVER/computing_rhs_gauss.F90:7-146) followed by rhs_3x3 of its
 * output for every mode, with or without opt_tensor
 * (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:313-328,
 * SOURCES/SOLVER/assembling_rhs.F90:5-126), fused: the Gauss array rhs_gauss
 * is never stored.  Per Gauss point rhs = (ft + fp_inc + fs - rotv_v) + fp,
 * with ft, fs the Gauss values of V1m and of the nodal momentum source ff,
 * and fp, fp_inc the pressure-gradient terms of P and P_inc, the pressure
 * and its increment injected onto the velocity mesh (the reference's
 * inject_generic, done by the caller).  opt_tensor(3,G,6,m) is rhs_3x3's
 * optional tensor (the LES or variable-viscosity stress the caller passes
 * negated, subroutine_ns_with_velocity.F90:321-326), assembled in its weak
 * divergence form.  Device arrays: list_mode_dev(m), rotv_v(G,6,m),
 * V1m(np,6,m), ff(np,6,m), P(np,2,m), P_inc(np,2,m), opt_tensor; a null
 * field counts as zero.  Outputs vb_145(np,3,m), vb_236(np,3,m); work_elem
 * is a caller-provided device buffer of 36*me*m doubles. */
int sfx_rhs_ns_3x3_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* rotv_v,
                       const double* V1m, const double* ff, const double* P, const double* P_inc,
                       const double* opt_tensor, double* work_elem, double* vb_145, double* vb_236,
                       void* stream);

/* Replaces rhs_ns_gauss_3x3 with its penalty options (if_ns_penalty,
 * if_impose_vel_in_solids; SOURCES/SOLVER/computing_rhs_gauss.F90:7-146)
 * followed by rhs_3x3 for every mode: rhs0 = ft + fp_inc + fs - rotv_v at the
 * Gauss points, minus imposed_vel on the slot mode0_slot (the slot i with
 * list_mode(i) = 0, or -1) (:54-121); then the product with the penalty
 * function penal_grid(2*m_max_pad-1, G), its values on the grid angles at the
 * Gauss points, m_max_pad = 3*m_max_c/2 (FFT_PAR_VAR_ETA_PROD_GAUSS_DCL,
 * :124-129); then imposed_vel added back on mode0_slot and fp added
 * (:132-145), and the result assembled with opt_tensor as
 * sfx_rhs_ns_3x3_dev does.  imposed_vel(np,6) is the nodal array
 * 3*imposed_velocity_by_penalty/(2 dt) (:107) or null; the caller evaluates
 * penal_in_real_space on the grid (the user function of the reference).
 * work_gauss: 2*G*6*m doubles; work_elem: 36*me*m doubles.  The penalty
 * product uses the exact-grid scratch described above. */
int sfx_rhs_ns_3x3_penal_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, int32_t mode0_slot,
                             const double* rotv_v, const double* V1m, const double* ff, const double* P,
                             const double* P_inc, const double* penal_grid, const double* imposed_vel,
                             const double* opt_tensor, double* work_gauss, double* work_elem, double* vb_145,
                             double* vb_236, void* stream);

/* Replaces qs_00_gauss (SOURCES/SOLVER/fem_rhs_axi.F90:940): vect(np) from
 * heat_capa(me), nodal ff(np) and Gauss ff_gauss(G). */
int sfx_qs_00_gauss(sfx_mesh_t mesh, const double* heat_capa, const double* ff, const double* ff_gauss,
                    double* vect);
/* ncols independent right-hand sides: ff(np,ncols), ff_gauss(G,ncols) -> vect(np,ncols).
 * work_elem is a caller-provided device buffer of 6*me*ncols doubles.  A null
 * heat_capa counts as 1, a null ff or ff_gauss as zero. */
int sfx_qs_00_gauss_dev(sfx_mesh_t mesh, int32_t ncols, const double* heat_capa, const double* ff,
                        const double* ff_gauss, double* work_elem, double* vect, void* stream);
/* smb_ugradc_gauss_fft_par followed by qs_00_gauss of its output, fused
 * (subroutine_temperature_with_T.F90:247-350 then fem_rhs_axi.F90:940-985):
 * vect(np,2,m) = scale * qs_00_gauss(heat_capa, ff, c_gauss) for every mode,
 * with c_gauss = (hc grad c).u + (div u)(hc c) formed in the work-group and
 * never stored.  ff (np,2,m), the nodal source, may be null (zero).  Equal,
 * bit for bit, to the two calls (the second with ncols = 2m).  work_elem:
 * 12*me*m doubles. */
int sfx_nl_qs_00_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* heat_capa_dev,
                     const double* V_in, const double* c_in, const double* ff, double* vect, double scale,
                     double* work_elem, void* stream);

/* Nonlinear share of the Navier-Stokes right-hand sides of an MHD step,
 * all modes: vb = rhs_3x3(-[(curl u) x u - coeff_lorentz (curl H) x B]).
 * work_gauss (G*6*m doubles) and work_elem (36*me*m doubles) are
 * caller-provided device buffers. */
int sfx_nl_rhs_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* u,
                       const double* H, const double* B, double coeff_lorentz, double* work_gauss,
                       double* work_elem, double* vb_145, double* vb_236, void* stream);

/* ---------------------------------------------------------------------------
 * Batched persistent launch.  No single counterpart in the reference: a batch
 * stands for many independent calls of smb_cross_prod_gauss_sft_par
 * (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514), one per problem,
 * on meshes of different sizes and different m_max_c, run as one kernel
 * launch whose work-groups walk host-built tile lists (or, with the BY_SIZE
 * schedule, one grid per grid size), replayable as one hipGraph.  All arrays
 * are device pointers and must stay valid for the life of the batch. */
typedef struct {
  sfx_mesh_t mesh;
  const int32_t* list_mode_dev;   /* (m_max_c) */
  int32_t m_max_c;                /* 1..256 */
  const double* V_in;             /* (np, 6, m_max_c) */
  double* V_out;                  /* (7 me, 6, m_max_c) */
} sfx_batch_problem_t;
typedef struct sfx_batch_s* sfx_batch_t;

#define SFX_BATCH_LPT 0          /* longest tile first to the least-loaded list */
#define SFX_BATCH_CONTIGUOUS 1   /* equal-cost contiguous ranges, grouped by grid size */
#define SFX_BATCH_ROUND_ROBIN 2  /* tile i to list i mod n_wg */
#define SFX_BATCH_BY_SIZE 3      /* one ordinary grid per grid size, launched in turn (n_wg ignored) */

int sfx_batch_create(const sfx_batch_problem_t* probs, int32_t n_prob, sfx_batch_t* out);
/* Set the schedule: policy and number of persistent work-groups. */
int sfx_batch_schedule(sfx_batch_t batch, int32_t policy, int32_t n_wg);
/* Time every (policy, n_wg) candidate `reps` times on `stream` and keep the
 * fastest; cached per batch shape (me, m_max_c of every problem). */
int sfx_batch_autotune(sfx_batch_t batch, int32_t reps, void* stream);
/* Capture the launch into a hipGraph; sfx_batch_run then replays it. */
int sfx_batch_capture(sfx_batch_t batch);
int sfx_batch_run(sfx_batch_t batch, void* stream);
int sfx_batch_info(sfx_batch_t batch, int32_t* policy, int32_t* n_wg, int64_t* n_tiles, double* tuned_ms,
                   int32_t* captured);
int sfx_batch_destroy(sfx_batch_t batch);

#ifdef __cplusplus
}
#endif

#endif /* SFX_H */
