!===
! Drop-in replacement, for one GPU, of module fft_parallele
! (SOURCES/POST_PROCESSING/fft_parallel.F90): the same 24 PUBLIC names
! (:8-17) with the reference's argument lists, calling libsfx through
! ISO_C_BINDING.
!
! Routine by routine (reference line): FFT_PAR_REAL :28, FFT_PAR_CROSS_PROD_DCL
! :147, FFT_PAR_DOT_PROD_DCL :330, FFT_PAR_PROD_DCL :502, FFT_PAR_DIV_DCL :667,
! FFT_HEAVISIDE_DCL :832, FFT_SCALAR_VECT_DCL :1021,
! FFT_SCALAR_VECT_NO_OVERSHOOT :1222, FFT_MAX_VEL_DCL :1426, FFT_TENSOR_DCL
! :1543, FFT_PAR_VAR_ETA_PROD_T_DCL :1798, FFT_PAR_VAR_ETA_PROD_GAUSS_DCL :2015,
! regul :2234, FFT_CHECK_INTERFACE :2267, FFT_COMPUTE_ENTROPY_VISC :2418,
! FFT_COMPUTE_ENTROPY_VISC_MOM :2701, FFT_COMPUTE_DIFFU_MOM :2939,
! FFT_PAR_COMPR_ENTRO_VISC_DCL :3160, FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM :3419,
! FFT_NO_OVERSHOOT_LEVEL_SET :3649, FFT_COMPRESSION_LEVEL_SET_DCL :3821,
! FFT_PAR_ENTRO_VISC_DCL :4055, FFT_PAR_SCAL_FUNCT :4294, FFT_MAX_MIN_VEL_DCL
! :4467, FFT_SQUAREROOT_SCALAR :4601.
!
! The host's modules provide what the reference reads beyond the arguments:
! def_type_mesh (mesh_type, passed on to the user's ETA function) and
! input_data (inputs%LES_coeff1/2/4, LES_coeff1_mom, level_set_comp_factor,
! h_min_distance).  User functions (ETA, funct) run here on the host: ETA
! is evaluated on the grid angles 2*pi*(n-1)/N exactly as the reference does
! (:1865-1868, :1937-1951) and its values go to the GPU; funct is applied to
! the real planes between the GPU synthesis and analysis.
!
! The communicator is the legacy integer MPI handle and nb_procs must be 1.
! A non-zero status from libsfx stops the program with the library's
! message, as the reference does with WRITE + STOP or error_petsc.  Array
! arguments are copied to contiguous temporaries (sections are allowed).
!===
MODULE fft_parallele
  USE, INTRINSIC :: ISO_C_BINDING
  USE def_type_mesh, ONLY: mesh_type
  USE input_data, ONLY: inputs
  IMPLICIT NONE
  PUBLIC :: FFT_PAR_REAL, FFT_PAR_CROSS_PROD_DCL, &
       FFT_PAR_DOT_PROD_DCL, FFT_PAR_PROD_DCL, FFT_PAR_DIV_DCL, FFT_HEAVISIDE_DCL, &
       FFT_SCALAR_VECT_DCL, FFT_MAX_VEL_DCL, FFT_TENSOR_DCL, &
       FFT_PAR_VAR_ETA_PROD_T_DCL, FFT_PAR_VAR_ETA_PROD_GAUSS_DCL, &
       FFT_CHECK_INTERFACE, FFT_COMPUTE_ENTROPY_VISC, FFT_COMPUTE_ENTROPY_VISC_MOM, &
       FFT_COMPUTE_DIFFU_MOM, FFT_PAR_COMPR_ENTRO_VISC_DCL, &
       FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM, FFT_NO_OVERSHOOT_LEVEL_SET, &
       FFT_COMPRESSION_LEVEL_SET_DCL, regul, FFT_PAR_ENTRO_VISC_DCL, &
       FFT_PAR_SCAL_FUNCT, FFT_SCALAR_VECT_NO_OVERSHOOT, FFT_MAX_MIN_VEL_DCL, &
       FFT_SQUAREROOT_SCALAR
  PUBLIC :: sfx_stop
  PRIVATE

  INTERFACE
     FUNCTION c_cross(V1, V2, Vo, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_cross_prod_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_cross
     FUNCTION c_dot(V1, V2, Vo, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_dot_prod_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_dot
     FUNCTION c_prod(V1, V2, Vo, np, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_prod_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_prod
     FUNCTION c_svect(V1, V2, Vo, np, nb1, nb2, m, pb, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_scalar_vect_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nb1, nb2, m, pb, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_svect
     FUNCTION c_svect_no(bounds, nbnd, V1, V2, Vo, np, nb1, nb2, m, pb, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_scalar_vect_no_overshoot') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: bounds, V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbnd, nb1, nb2, m, pb, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_svect_no
     FUNCTION c_tensor(V1, V2, Vo, np, nbf, m, nbp, bloc, mpad, tens, temps) &
          BIND(C, NAME='sfx_fft_tensor_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad, tens
       INTEGER(C_INT) :: rc
     END FUNCTION c_tensor
     FUNCTION c_div(V1, V2, Vo, np, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_div_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_div
     FUNCTION c_heavi(V1, vals, Vo, nv, np, nbf, m, nbp, bloc, mpad, eps, temps) &
          BIND(C, NAME='sfx_fft_heaviside_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, vals, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nv, nbf, m, nbp, mpad
       REAL(C_DOUBLE), VALUE :: eps
       INTEGER(C_INT) :: rc
     END FUNCTION c_heavi
     FUNCTION c_sqrt(c, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_squareroot_scalar') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: c, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_sqrt
     FUNCTION c_no_ls(c, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_no_overshoot_level_set') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: c, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_no_ls
     FUNCTION c_maxvel(V1, vo, np, nbf, m, nbp, bloc, mpad) &
          BIND(C, NAME='sfx_fft_max_vel_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, vo
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_maxvel
     FUNCTION c_maxminvel(V1, vo, np, nbf, m, nbp, bloc, mpad) &
          BIND(C, NAME='sfx_fft_max_min_vel_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, vo
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_maxminvel
     FUNCTION c_checkif(V1, nbfl, ok, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_check_interface') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, ok, temps
       INTEGER(C_INT32_T), VALUE :: nbfl, nbf, m, nbp, mpad
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT) :: rc
     END FUNCTION c_checkif
     FUNCTION c_eta(eta, ci, co, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_var_eta_prod_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: eta, ci, co, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_eta
     FUNCTION c_evisc(V1, V2, V3, V4, V5, h, O1, O2, O3, np, nbf, m, nbp, bloc, mpad, resn, lG, k1, k2, k4, &
          temps) BIND(C, NAME='sfx_fft_compute_entropy_visc') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, V2, V3, V4, V5, h, O1, O2, O3, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad, lG
       REAL(C_DOUBLE), VALUE :: resn, k1, k2, k4
       INTEGER(C_INT) :: rc
     END FUNCTION c_evisc
     FUNCTION c_evisc_mom(V1, V2, V3, c1, h, r1, r2, np, nbf, m, nbp, bloc, mpad, lG, k2, k4, temps) &
          BIND(C, NAME='sfx_fft_compute_entropy_visc_mom') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, V2, V3, c1, h, r1, r2, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad, lG
       REAL(C_DOUBLE), VALUE :: k2, k4
       INTEGER(C_INT) :: rc
     END FUNCTION c_evisc_mom
     FUNCTION c_evisc_grad(V1, V2, V3, cr, k1m, Vo, np, nbf, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_compute_entropy_visc_grad_mom') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, V2, V3, cr, Vo, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, mpad
       REAL(C_DOUBLE), VALUE :: k1m
       INTEGER(C_INT) :: rc
     END FUNCTION c_evisc_grad
     FUNCTION c_diffu(V1, V2, V3, O1, O2, np, nb1, nb2, nb3, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_compute_diffu_mom') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, V2, V3, O1, O2, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nb1, nb2, nb3, m, nbp, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_diffu
     FUNCTION c_entro(V1, c1, cr, h, k1, Vo, co, np, nb1, nb2, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_entro_visc_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, c1, cr, h, Vo, co, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nb1, nb2, m, nbp, mpad
       REAL(C_DOUBLE), VALUE :: k1
       INTEGER(C_INT) :: rc
     END FUNCTION c_entro
     FUNCTION c_compr(V1, V2, c1, cr, h, k1, lscf, Vo, co, np, nb1, nb2, nb3, m, nbp, bloc, mpad, temps) &
          BIND(C, NAME='sfx_fft_par_compr_entro_visc_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, V2, c1, cr, h, Vo, co, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: nb1, nb2, nb3, m, nbp, mpad
       REAL(C_DOUBLE), VALUE :: k1, lscf
       INTEGER(C_INT) :: rc
     END FUNCTION c_compr
     FUNCTION c_cls(V1, V2, ci, co, h, lG, np, nbf, m, nbp, bloc, mpad, lscf, k4, hmin, temps) &
          BIND(C, NAME='sfx_fft_compression_level_set_dcl') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: V1, V2, ci, co, h, temps
       INTEGER(C_INT64_T), VALUE :: np, bloc
       INTEGER(C_INT32_T), VALUE :: lG, nbf, m, nbp, mpad
       REAL(C_DOUBLE), VALUE :: lscf, k4, hmin
       INTEGER(C_INT) :: rc
     END FUNCTION c_cls
     FUNCTION c_from_real(R, Co, np, m, mpad) BIND(C, NAME='sfx_fft_par_from_real') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: R, Co
       INTEGER(C_INT64_T), VALUE :: np
       INTEGER(C_INT32_T), VALUE :: m, mpad
       INTEGER(C_INT) :: rc
     END FUNCTION c_from_real
     FUNCTION c_real_n(m, nplane) BIND(C, NAME='sfx_fft_par_real_n') RESULT(n)
       IMPORT :: C_INT32_T
       INTEGER(C_INT32_T), VALUE :: m, nplane
       INTEGER(C_INT32_T) :: n
     END FUNCTION c_real_n
     FUNCTION c_real(V1, Vo, np, nbf, m, nbp, nplane) BIND(C, NAME='sfx_fft_par_real') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: V1, Vo
       INTEGER(C_INT64_T), VALUE :: np
       INTEGER(C_INT32_T), VALUE :: nbf, m, nbp, nplane
       INTEGER(C_INT) :: rc
     END FUNCTION c_real
     FUNCTION c_last_error() BIND(C, NAME='sfx_last_error') RESULT(p)
       IMPORT :: C_PTR
       TYPE(C_PTR) :: p
     END FUNCTION c_last_error
  END INTERFACE

CONTAINS

  !===Stop with libsfx's message on a non-zero status
  SUBROUTINE sfx_stop(rc, name)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: name
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    CHARACTER(LEN=512) :: txt
    INTEGER :: i
    IF (rc == 0) RETURN
    CALL C_F_POINTER(c_last_error(), msg, [512])
    txt = ' '
    DO i = 1, 512
       IF (msg(i) == C_NULL_CHAR) EXIT
       txt(i:i) = msg(i)
    END DO
    WRITE(*,*) ' BUG in ', name, ': ', TRIM(txt)
    STOP 1
  END SUBROUTINE sfx_stop

  FUNCTION temps_ptr(temps, local) RESULT(p)
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(3), TARGET, INTENT(INOUT) :: local
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    IF (PRESENT(temps)) p = C_LOC(local)
  END FUNCTION temps_ptr

  SUBROUTINE temps_out(temps, local)
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(3), INTENT(IN) :: local
    IF (PRESENT(temps)) temps(1:3) = local
  END SUBROUTINE temps_out

  !===The angles of the grid, 2*pi*(n-1)/N (fft_parallel.F90:1865-1868)
  SUBROUTINE grid_angles(angles)
    REAL(KIND=8), DIMENSION(:), INTENT(OUT) :: angles
    REAL(KIND=8) :: pi
    INTEGER :: n
    pi = ACOS(-1.d0)
    DO n = 1, SIZE(angles)
       angles(n) = 2*pi*(n-1)/SIZE(angles)
    END DO
  END SUBROUTINE grid_angles

  SUBROUTINE product3(kind, name, V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps, pb)
    INTEGER, INTENT(IN) :: kind
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    INTEGER, OPTIONAL, INTENT(IN) :: pb
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, ao
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    INTEGER(C_INT64_T) :: np
    ALLOCATE(a1, SOURCE=V1_in)
    ALLOCATE(a2, SOURCE=V2_in)
    ALLOCATE(ao(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    np = SIZE(V1_in, 1)
    SELECT CASE (kind)
    CASE (1)
       rc = c_cross(C_LOC(a1), C_LOC(a2), C_LOC(ao), np, INT(SIZE(V1_in,2),C_INT32_T), &
            INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
            INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CASE (2)
       rc = c_dot(C_LOC(a1), C_LOC(a2), C_LOC(ao), np, INT(SIZE(V1_in,2),C_INT32_T), &
            INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
            INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CASE (3)
       rc = c_prod(C_LOC(a1), C_LOC(a2), C_LOC(ao), np, INT(SIZE(V1_in,3),C_INT32_T), &
            INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), &
            temps_ptr(temps, tloc))
    CASE (4)
       rc = c_svect(C_LOC(a1), C_LOC(a2), C_LOC(ao), np, INT(SIZE(V1_in,2),C_INT32_T), &
            INT(SIZE(V2_in,2),C_INT32_T), INT(SIZE(V1_in,3),C_INT32_T), INT(pb,C_INT32_T), &
            INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), &
            temps_ptr(temps, tloc))
    CASE (5)
       rc = c_div(C_LOC(a1), C_LOC(a2), C_LOC(ao), np, INT(SIZE(V1_in,3),C_INT32_T), &
            INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), &
            temps_ptr(temps, tloc))
    END SELECT
    CALL sfx_stop(rc, name)
    V_out = ao
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, ao)
  END SUBROUTINE product3

  SUBROUTINE FFT_PAR_CROSS_PROD_DCL(communicator, V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    CALL product3(1, 'FFT_PAR_CROSS_PROD_DCL', V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps)
  END SUBROUTINE FFT_PAR_CROSS_PROD_DCL

  SUBROUTINE FFT_PAR_DOT_PROD_DCL(communicator, V1_in, V2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    CALL product3(2, 'FFT_PAR_DOT_PROD_DCL', V1_in, V2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
  END SUBROUTINE FFT_PAR_DOT_PROD_DCL

  SUBROUTINE FFT_PAR_PROD_DCL(communicator, c1_in, c2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c1_in, c2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    CALL product3(3, 'FFT_PAR_PROD_DCL', c1_in, c2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
  END SUBROUTINE FFT_PAR_PROD_DCL

  SUBROUTINE FFT_SCALAR_VECT_DCL(communicator, V1_in, V2_in, V_out, pb, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: pb, nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    CALL product3(4, 'FFT_SCALAR_VECT_DCL', V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps, pb)
  END SUBROUTINE FFT_SCALAR_VECT_DCL

  SUBROUTINE FFT_SCALAR_VECT_NO_OVERSHOOT(communicator, scalar_bounds, V1_in, V2_in, V_out, pb, nb_procs, &
       bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: scalar_bounds
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: pb, nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, ao
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: bnd
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in)
    ALLOCATE(a2, SOURCE=V2_in)
    ALLOCATE(bnd, SOURCE=scalar_bounds)
    ALLOCATE(ao(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    rc = c_svect_no(C_LOC(bnd), INT(SIZE(scalar_bounds),C_INT32_T), C_LOC(a1), C_LOC(a2), C_LOC(ao), &
         INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V2_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(pb,C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_SCALAR_VECT_NO_OVERSHOOT')
    V_out = ao
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, bnd, ao)
  END SUBROUTINE FFT_SCALAR_VECT_NO_OVERSHOOT

  SUBROUTINE FFT_TENSOR_DCL(communicator, V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps, opt_tension)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    LOGICAL, OPTIONAL, INTENT(IN) :: opt_tension
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2
    REAL(KIND=8), DIMENSION(:,:,:,:), ALLOCATABLE, TARGET :: ao
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT32_T) :: tens
    INTEGER(C_INT) :: rc
    tens = 0
    IF (PRESENT(opt_tension)) THEN
       IF (opt_tension) tens = 1
    END IF
    ALLOCATE(a1, SOURCE=V1_in)
    ALLOCATE(a2, SOURCE=V2_in)
    ALLOCATE(ao(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3), SIZE(V_out,4)))
    rc = c_tensor(C_LOC(a1), C_LOC(a2), C_LOC(ao), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), tens, temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_TENSOR_DCL')
    V_out = ao
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, ao)
  END SUBROUTINE FFT_TENSOR_DCL

  SUBROUTINE FFT_PAR_DIV_DCL(communicator, c1_in, c2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c1_in, c2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    CALL product3(5, 'FFT_PAR_DIV_DCL', c1_in, c2_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
  END SUBROUTINE FFT_PAR_DIV_DCL

  SUBROUTINE FFT_HEAVISIDE_DCL(communicator, V1_in, values, V_out, nb_procs, bloc_size, m_max_pad, &
       coeff_tanh, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(IN) :: V1_in
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: values
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), INTENT(IN) :: coeff_tanh
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: av
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: ao
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in)
    ALLOCATE(av, SOURCE=values)
    ALLOCATE(ao(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    rc = c_heavi(C_LOC(a1), C_LOC(av), C_LOC(ao), INT(SIZE(values),C_INT32_T), INT(SIZE(V1_in,2),C_INT64_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(SIZE(V1_in,4),C_INT32_T), INT(nb_procs,C_INT32_T), &
         INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), REAL(coeff_tanh,C_DOUBLE), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_HEAVISIDE_DCL')
    V_out = ao
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, av, ao)
  END SUBROUTINE FFT_HEAVISIDE_DCL

  !===The smoothed Heaviside of a level set (fft_parallel.F90:2234-2246)
  FUNCTION regul(phi, eps) RESULT(r)
    REAL(KIND=8), INTENT(IN) :: phi, eps
    REAL(KIND=8) :: x, r
    x = phi - 0.5d0
    IF (x <= -eps) THEN
       r = 0.d0
    ELSE IF (x <= eps) THEN
       r = (1 + x*(x**2 - 3*eps**2)/(-2*eps**3))/2
    ELSE
       r = 1.d0
    END IF
  END FUNCTION regul

  SUBROUTINE FFT_SQUAREROOT_SCALAR(communicator, c1_inout, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(INOUT) :: c1_inout
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=c1_inout)
    rc = c_sqrt(C_LOC(a1), INT(SIZE(c1_inout,1),C_INT64_T), INT(SIZE(c1_inout,2),C_INT32_T), &
         INT(SIZE(c1_inout,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_SQUAREROOT_SCALAR')
    c1_inout = a1
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1)
  END SUBROUTINE FFT_SQUAREROOT_SCALAR

  SUBROUTINE FFT_NO_OVERSHOOT_LEVEL_SET(communicator, c1_inout, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(INOUT) :: c1_inout
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=c1_inout)
    rc = c_no_ls(C_LOC(a1), INT(SIZE(c1_inout,1),C_INT64_T), INT(SIZE(c1_inout,2),C_INT32_T), &
         INT(SIZE(c1_inout,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_NO_OVERSHOOT_LEVEL_SET')
    c1_inout = a1
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1)
  END SUBROUTINE FFT_NO_OVERSHOOT_LEVEL_SET

  SUBROUTINE FFT_MAX_VEL_DCL(communicator, V1_in, V_out, nb_procs, bloc_size, m_max_pad)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in
    REAL(KIND=8), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), TARGET :: vo
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in)
    rc = c_maxvel(C_LOC(a1), C_LOC(vo), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T))
    CALL sfx_stop(rc, 'FFT_MAX_VEL_DCL')
    V_out = vo
    DEALLOCATE(a1)
  END SUBROUTINE FFT_MAX_VEL_DCL

  SUBROUTINE FFT_MAX_MIN_VEL_DCL(communicator, V1_in, V_out, nb_procs, bloc_size, m_max_pad)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in
    REAL(KIND=8), DIMENSION(2), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), DIMENSION(2), TARGET :: vo
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in)
    rc = c_maxminvel(C_LOC(a1), C_LOC(vo), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T))
    CALL sfx_stop(rc, 'FFT_MAX_MIN_VEL_DCL')
    V_out = vo
    DEALLOCATE(a1)
  END SUBROUTINE FFT_MAX_MIN_VEL_DCL

  SUBROUTINE FFT_PAR_REAL(communicator, V1_in, V_out, opt_nb_plane)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: V_out
    INTEGER, OPTIONAL :: opt_nb_plane
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, ao
    INTEGER(C_INT32_T) :: nplane, n
    INTEGER(C_INT) :: rc
    nplane = 0
    IF (PRESENT(opt_nb_plane)) nplane = INT(opt_nb_plane, C_INT32_T)
    n = c_real_n(INT(SIZE(V1_in,3),C_INT32_T), nplane)
    IF (ALLOCATED(V_out)) DEALLOCATE(V_out)
    ALLOCATE(V_out(n, SIZE(V1_in,2)/2, SIZE(V1_in,1)))
    ALLOCATE(a1, SOURCE=V1_in)
    ALLOCATE(ao(n, MAX(SIZE(V1_in,2)/2,1), MAX(SIZE(V1_in,1),1)))
    rc = c_real(C_LOC(a1), C_LOC(ao), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), 1_C_INT32_T, nplane)
    CALL sfx_stop(rc, 'FFT_PAR_REAL')
    V_out = ao(:, 1:SIZE(V1_in,2)/2, 1:SIZE(V1_in,1))
    DEALLOCATE(a1, ao)
  END SUBROUTINE FFT_PAR_REAL

  SUBROUTINE FFT_PAR_SCAL_FUNCT(communicator, c1_inout, funct, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(INOUT) :: c1_inout
    INTERFACE
       FUNCTION funct(x) RESULT(vv)
         REAL(KIND=8) :: x
         REAL(KIND=8) :: vv
       END FUNCTION funct
    END INTERFACE
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, ao
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: rv
    INTEGER(C_INT32_T) :: n
    INTEGER(C_INT) :: rc
    INTEGER :: i, j, np, m, nb_field
    IF (PRESENT(temps)) temps = 0.d0
    np = SIZE(c1_inout,1); nb_field = SIZE(c1_inout,2); m = SIZE(c1_inout,3)
    ! the reference's checks (fft_parallel.F90:4347-4350), then the shape of its point blocks
    IF (MOD(nb_field,2) /= 0 .OR. nb_field < 2 .OR. m == 0) THEN
       WRITE(*,*) ' BUG in FFT_PAR_SCAL_FUNCT '
       STOP 1
    END IF
    IF (nb_procs /= 1) THEN
       WRITE(*,*) ' BUG in FFT_PAR_SCAL_FUNCT: one GPU, nb_procs must be 1'
       STOP 1
    END IF
    IF (bloc_size*nb_procs < np .OR. m_max_pad < m) THEN
       WRITE(*,*) ' BUG in FFT_PAR_SCAL_FUNCT: bloc_size*nb_procs < np or m_max_pad < m_max_c'
       STOP 1
    END IF
    n = 2*m_max_pad - 1
    ALLOCATE(a1(np, 2, m), ao(np, 2, m), rv(n, MAX(np,1)))
    a1 = c1_inout(:, 1:2, :)
    ! values of the first field on the n grid angles (opt_nb_plane = n)
    rc = c_real(C_LOC(a1), C_LOC(rv), INT(np,C_INT64_T), 2_C_INT32_T, INT(m,C_INT32_T), 1_C_INT32_T, n)
    CALL sfx_stop(rc, 'FFT_PAR_SCAL_FUNCT')
    DO j = 1, np
       DO i = 1, n
          rv(i,j) = funct(rv(i,j))
       END DO
    END DO
    rc = c_from_real(C_LOC(rv), C_LOC(ao), INT(np,C_INT64_T), INT(m,C_INT32_T), INT(m_max_pad,C_INT32_T))
    CALL sfx_stop(rc, 'FFT_PAR_SCAL_FUNCT')
    c1_inout(:, 1:2, :) = ao
    DEALLOCATE(a1, ao, rv)
  END SUBROUTINE FFT_PAR_SCAL_FUNCT

  SUBROUTINE FFT_CHECK_INTERFACE(communicator, V1_in, nb_fluids, interface_ok, nb_procs, bloc_size, m_max_pad, &
       temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(IN) :: V1_in
    INTEGER, INTENT(IN) :: nb_fluids
    INTEGER, INTENT(OUT) :: interface_ok
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:,:), ALLOCATABLE, TARGET :: a1
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT32_T), TARGET :: ok
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in)
    rc = c_checkif(C_LOC(a1), INT(nb_fluids,C_INT32_T), C_LOC(ok), INT(SIZE(V1_in,2),C_INT64_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(SIZE(V1_in,4),C_INT32_T), INT(nb_procs,C_INT32_T), &
         INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_CHECK_INTERFACE')
    interface_ok = ok
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1)
  END SUBROUTINE FFT_CHECK_INTERFACE

  !===eta * c_in with eta(N, np) given on the grid angles: the product shared by
  !===both FFT_PAR_VAR_ETA_PROD_* routines (fft_parallel.F90:1953-1962)
  SUBROUTINE eta_product(name, eta_n, c_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: eta_n
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: ae
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, ao
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(ae, SOURCE=eta_n)
    ALLOCATE(a1, SOURCE=c_in)
    ALLOCATE(ao(SIZE(c_out,1), SIZE(c_out,2), SIZE(c_out,3)))
    rc = c_eta(C_LOC(ae), C_LOC(a1), C_LOC(ao), INT(SIZE(c_in,1),C_INT64_T), INT(SIZE(c_in,2),C_INT32_T), &
         INT(SIZE(c_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, name)
    c_out = ao
    CALL temps_out(temps, tloc)
    DEALLOCATE(ae, a1, ao)
  END SUBROUTINE eta_product

  SUBROUTINE FFT_PAR_VAR_ETA_PROD_T_DCL(communicator, eta, H_mesh, c_in, c_out, nb_procs, bloc_size, m_max_pad, &
       time, temps)
    INTEGER :: communicator
    TYPE(mesh_type) :: H_mesh
    INTERFACE
       FUNCTION eta(H_mesh, angles, nb_angles, nb, ne, time) RESULT(vv)
         USE def_type_mesh
         IMPLICIT NONE
         TYPE(mesh_type), INTENT(IN) :: H_mesh
         REAL(KIND=8), DIMENSION(:), INTENT(IN) :: angles
         INTEGER, INTENT(IN) :: nb_angles
         INTEGER, INTENT(IN) :: nb, ne
         REAL(KIND=8), INTENT(IN) :: time
         REAL(KIND=8), DIMENSION(nb_angles,ne-nb+1) :: vv
       END FUNCTION eta
    END INTERFACE
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8) :: time
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(2*m_max_pad-1) :: angles
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: eta_n
    INTEGER :: np
    np = SIZE(c_in,1)
    CALL grid_angles(angles)
    ! one rank: every point is local, nb = 1, ne = np (fft_parallel.F90:1937-1951)
    ALLOCATE(eta_n(2*m_max_pad-1, MAX(np,1)))
    eta_n = 1.d0
    IF (np >= 1) eta_n(:, 1:np) = eta(H_mesh, angles, 2*m_max_pad-1, 1, np, time)
    CALL eta_product('FFT_PAR_VAR_ETA_PROD_T_DCL', eta_n, c_in, c_out, nb_procs, bloc_size, m_max_pad, temps)
    DEALLOCATE(eta_n)
  END SUBROUTINE FFT_PAR_VAR_ETA_PROD_T_DCL

  SUBROUTINE FFT_PAR_VAR_ETA_PROD_GAUSS_DCL(communicator, eta, H_mesh, c_in, c_out, nb_procs, bloc_size, &
       m_max_pad, rr_gauss, time, temps)
    INTEGER :: communicator
    TYPE(mesh_type) :: H_mesh
    INTERFACE
       FUNCTION eta(H_mesh, rr_gauss, angles, nb_angles, nb, ne, time) RESULT(vv)
         USE def_type_mesh
         IMPLICIT NONE
         TYPE(mesh_type) :: H_mesh
         REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: rr_gauss
         REAL(KIND=8), DIMENSION(:), INTENT(IN) :: angles
         INTEGER, INTENT(IN) :: nb_angles
         INTEGER, INTENT(IN) :: nb, ne
         REAL(KIND=8), INTENT(IN) :: time
         REAL(KIND=8), DIMENSION(nb_angles,ne-nb+1) :: vv
       END FUNCTION eta
    END INTERFACE
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: rr_gauss
    REAL(KIND=8) :: time
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(2*m_max_pad-1) :: angles
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: eta_n
    INTEGER :: np
    np = SIZE(c_in,1)
    CALL grid_angles(angles)
    ! one rank: nb = 1, ne = np and the Gauss points rr_gauss(:, 1:np) (fft_parallel.F90:2156-2170)
    ALLOCATE(eta_n(2*m_max_pad-1, MAX(np,1)))
    eta_n = 1.d0
    IF (np >= 1) eta_n(:, 1:np) = eta(H_mesh, rr_gauss(:, 1:np), angles, 2*m_max_pad-1, 1, np, time)
    CALL eta_product('FFT_PAR_VAR_ETA_PROD_GAUSS_DCL', eta_n, c_in, c_out, nb_procs, bloc_size, m_max_pad, &
         temps)
    DEALLOCATE(eta_n)
  END SUBROUTINE FFT_PAR_VAR_ETA_PROD_GAUSS_DCL

  SUBROUTINE FFT_COMPUTE_ENTROPY_VISC(communicator, communicator_S, V1_in, V2_in, V3_in, V4_in, V5_in, &
       hloc_gauss, V1_out, V2_out, V3_out, nb_procs, bloc_size, m_max_pad, residual_normalization, l_G, temps)
    INTEGER :: communicator, communicator_S
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in, V3_in, V4_in, V5_in
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: hloc_gauss
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V1_out, V2_out, V3_out
    REAL(KIND=8), INTENT(IN) :: residual_normalization
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad, l_G
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, a3, a4, a5, o1, o2, o3
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: h
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(a3, SOURCE=V3_in)
    ALLOCATE(a4, SOURCE=V4_in); ALLOCATE(a5, SOURCE=V5_in); ALLOCATE(h, SOURCE=hloc_gauss)
    ALLOCATE(o1(SIZE(V1_out,1), SIZE(V1_out,2), SIZE(V1_out,3)))
    ALLOCATE(o2(SIZE(V2_out,1), SIZE(V2_out,2), SIZE(V2_out,3)))
    ALLOCATE(o3(SIZE(V3_out,1), SIZE(V3_out,2), SIZE(V3_out,3)))
    rc = c_evisc(C_LOC(a1), C_LOC(a2), C_LOC(a3), C_LOC(a4), C_LOC(a5), C_LOC(h), C_LOC(o1), C_LOC(o2), &
         C_LOC(o3), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V1_in,3),C_INT32_T), &
         INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), &
         REAL(residual_normalization,C_DOUBLE), INT(l_G,C_INT32_T), REAL(inputs%LES_coeff1,C_DOUBLE), &
         REAL(inputs%LES_coeff2,C_DOUBLE), REAL(inputs%LES_coeff4,C_DOUBLE), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_COMPUTE_ENTROPY_VISC')
    V1_out = o1; V2_out = o2; V3_out = o3
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, a3, a4, a5, h, o1, o2, o3)
  END SUBROUTINE FFT_COMPUTE_ENTROPY_VISC

  SUBROUTINE FFT_COMPUTE_ENTROPY_VISC_MOM(communicator, communicator_S, V1_in, V2_in, V3_in, c1_in, hloc_gauss, &
       c1_real_out, nb_procs, bloc_size, m_max_pad, l_G, opt_c2_real_out, temps)
    INTEGER :: communicator, communicator_S
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in, V3_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c1_in
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: hloc_gauss
    REAL(KIND=8), DIMENSION(:,:), INTENT(OUT) :: c1_real_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad, l_G
    REAL(KIND=8), DIMENSION(:,:), OPTIONAL, INTENT(OUT) :: opt_c2_real_out
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, a3, c1
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: h
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: r1, r2
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    TYPE(C_PTR) :: p2
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(a3, SOURCE=V3_in)
    ALLOCATE(c1, SOURCE=c1_in); ALLOCATE(h, SOURCE=hloc_gauss)
    ALLOCATE(r1(SIZE(c1_real_out,1), SIZE(c1_real_out,2)))
    p2 = C_NULL_PTR
    IF (PRESENT(opt_c2_real_out)) THEN
       ALLOCATE(r2(SIZE(opt_c2_real_out,1), SIZE(opt_c2_real_out,2)))
       p2 = C_LOC(r2)
    END IF
    IF (PRESENT(temps)) temps = 0.d0
    rc = c_evisc_mom(C_LOC(a1), C_LOC(a2), C_LOC(a3), C_LOC(c1), C_LOC(h), C_LOC(r1), p2, &
         INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V1_in,3),C_INT32_T), &
         INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), INT(l_G,C_INT32_T), &
         REAL(inputs%LES_coeff2,C_DOUBLE), REAL(inputs%LES_coeff4,C_DOUBLE), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_COMPUTE_ENTROPY_VISC_MOM')
    c1_real_out = r1
    IF (PRESENT(opt_c2_real_out)) THEN
       opt_c2_real_out = r2
       DEALLOCATE(r2)
    END IF
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, a3, c1, h, r1)
  END SUBROUTINE FFT_COMPUTE_ENTROPY_VISC_MOM

  SUBROUTINE FFT_COMPUTE_DIFFU_MOM(communicator, V1_in, V2_in, V3_in, V1_out, V2_out, nb_procs, bloc_size, &
       m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V3_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V1_out
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V2_out
    INTEGER, INTENT(IN) :: bloc_size, m_max_pad, nb_procs
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, a3, o1, o2
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(a3, SOURCE=V3_in)
    ALLOCATE(o1(SIZE(V1_out,1), SIZE(V1_out,2), SIZE(V1_out,3)))
    ALLOCATE(o2(SIZE(V2_out,1), SIZE(V2_out,2), SIZE(V2_out,3)))
    rc = c_diffu(C_LOC(a1), C_LOC(a2), C_LOC(a3), C_LOC(o1), C_LOC(o2), INT(SIZE(V1_in,1),C_INT64_T), &
         INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V2_in,2),C_INT32_T), INT(SIZE(V3_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_COMPUTE_DIFFU_MOM')
    V1_out = o1; V2_out = o2
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, a3, o1, o2)
  END SUBROUTINE FFT_COMPUTE_DIFFU_MOM

  SUBROUTINE FFT_PAR_COMPR_ENTRO_VISC_DCL(communicator, V1_in, V2_in, c1_in, c_in_real, hloc_gauss, &
       coeff1_in_level, V_out, c_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c1_in
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: c_in_real
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: hloc_gauss
    REAL(KIND=8), INTENT(IN) :: coeff1_in_level
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    INTEGER, INTENT(IN) :: bloc_size, m_max_pad, nb_procs
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, c1, vo, co
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: cr
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(c1, SOURCE=c1_in)
    ALLOCATE(cr, SOURCE=c_in_real)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)), co(SIZE(c_out,1), SIZE(c_out,2), SIZE(c_out,3)))
    rc = c_compr(C_LOC(a1), C_LOC(a2), C_LOC(c1), C_LOC(cr), C_NULL_PTR, REAL(coeff1_in_level,C_DOUBLE), &
         REAL(inputs%level_set_comp_factor,C_DOUBLE), C_LOC(vo), C_LOC(co), INT(SIZE(V1_in,1),C_INT64_T), &
         INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V2_in,2),C_INT32_T), INT(SIZE(c1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_PAR_COMPR_ENTRO_VISC_DCL')
    V_out = vo; c_out = co
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, c1, cr, vo, co)
  END SUBROUTINE FFT_PAR_COMPR_ENTRO_VISC_DCL

  SUBROUTINE FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM(communicator, V1_in, V2_in, V3_in, c_in_real, V_out, nb_procs, &
       bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in, V3_in
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: c_in_real
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(OUT) :: V_out
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, a3
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: cr
    REAL(KIND=8), DIMENSION(:,:,:,:), ALLOCATABLE, TARGET :: vo
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(a3, SOURCE=V3_in)
    ALLOCATE(cr, SOURCE=c_in_real)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3), SIZE(V_out,4)))
    rc = c_evisc_grad(C_LOC(a1), C_LOC(a2), C_LOC(a3), C_LOC(cr), REAL(inputs%LES_coeff1_mom,C_DOUBLE), &
         C_LOC(vo), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V1_in,3),C_INT32_T), &
         INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM')
    V_out = vo
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, a3, cr, vo)
  END SUBROUTINE FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM

  SUBROUTINE FFT_COMPRESSION_LEVEL_SET_DCL(communicator_F, communicator_S, V1_in, V2_in, c_in, c_out, &
       hloc_gauss, l_G, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator_F, communicator_S
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in, V2_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: hloc_gauss
    INTEGER, INTENT(IN) :: l_G
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    INTEGER, INTENT(IN) :: nb_procs, bloc_size, m_max_pad
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, a2, ci, co
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(a2, SOURCE=V2_in); ALLOCATE(ci, SOURCE=c_in)
    ALLOCATE(co(SIZE(c_out,1), SIZE(c_out,2), SIZE(c_out,3)))
    rc = c_cls(C_LOC(a1), C_LOC(a2), C_LOC(ci), C_LOC(co), C_NULL_PTR, INT(l_G,C_INT32_T), &
         INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(V1_in,3),C_INT32_T), &
         INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), INT(m_max_pad,C_INT32_T), &
         REAL(inputs%level_set_comp_factor,C_DOUBLE), REAL(inputs%LES_coeff4,C_DOUBLE), &
         REAL(inputs%h_min_distance,C_DOUBLE), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_COMPRESSION_LEVEL_SET_DCL')
    c_out = co
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, a2, ci, co)
  END SUBROUTINE FFT_COMPRESSION_LEVEL_SET_DCL

  SUBROUTINE FFT_PAR_ENTRO_VISC_DCL(communicator, V1_in, c1_in, c_in_real, hloc_gauss, coeff1_in_level, V_out, &
       c_out, nb_procs, bloc_size, m_max_pad, temps)
    INTEGER :: communicator
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V1_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: c1_in
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: c_in_real
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: hloc_gauss
    REAL(KIND=8), INTENT(IN) :: coeff1_in_level
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    REAL(KIND=8), DIMENSION(:), OPTIONAL, INTENT(INOUT) :: temps
    INTEGER, INTENT(IN) :: bloc_size, m_max_pad, nb_procs
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: a1, c1, vo, co
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: cr
    REAL(KIND=8), DIMENSION(3), TARGET :: tloc
    INTEGER(C_INT) :: rc
    ALLOCATE(a1, SOURCE=V1_in); ALLOCATE(c1, SOURCE=c1_in); ALLOCATE(cr, SOURCE=c_in_real)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)), co(SIZE(c_out,1), SIZE(c_out,2), SIZE(c_out,3)))
    rc = c_entro(C_LOC(a1), C_LOC(c1), C_LOC(cr), C_NULL_PTR, REAL(coeff1_in_level,C_DOUBLE), C_LOC(vo), &
         C_LOC(co), INT(SIZE(V1_in,1),C_INT64_T), INT(SIZE(V1_in,2),C_INT32_T), INT(SIZE(c1_in,2),C_INT32_T), &
         INT(SIZE(V1_in,3),C_INT32_T), INT(nb_procs,C_INT32_T), INT(bloc_size,C_INT64_T), &
         INT(m_max_pad,C_INT32_T), temps_ptr(temps, tloc))
    CALL sfx_stop(rc, 'FFT_PAR_ENTRO_VISC_DCL')
    V_out = vo; c_out = co
    CALL temps_out(temps, tloc)
    DEALLOCATE(a1, c1, cr, vo, co)
  END SUBROUTINE FFT_PAR_ENTRO_VISC_DCL

END MODULE fft_parallele
