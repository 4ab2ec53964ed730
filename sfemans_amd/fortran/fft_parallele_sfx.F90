!===
! Drop-in replacement, for one GPU, of the polynomial products of module
! fft_parallele (SOURCES/POST_PROCESSING/fft_parallel.F90) and of the
! Gauss-point nonlinear terms, calling libsfx through ISO_C_BINDING.
!
! The routines keep the reference's names, argument order and meaning
! (FFT_PAR_CROSS_PROD_DCL :147, FFT_PAR_DOT_PROD_DCL :330, FFT_PAR_PROD_DCL :502,
! FFT_PAR_DIV_DCL :667, FFT_HEAVISIDE_DCL :832, FFT_SCALAR_VECT_DCL :1021,
! FFT_MAX_VEL_DCL :1426, FFT_TENSOR_DCL :1543, FFT_SQUAREROOT_SCALAR :4601,
! FFT_PAR_REAL :28, FFT_CHECK_INTERFACE :2267, FFT_MAX_MIN_VEL_DCL :4467,
! FFT_PAR_SCAL_FUNCT :4294: the real planes on the GPU, the caller's funct
! here, then the analysis on the GPU).  The communicator is the
! legacy integer MPI handle; nb_procs must be 1.  A non-zero status from
! libsfx stops the program with the library's message, as the reference does
! with WRITE + STOP or error_petsc.  Non-contiguous actual arguments (array
! sections) are copied to contiguous temporaries and back.
!===
MODULE fft_parallele
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: FFT_PAR_CROSS_PROD_DCL, FFT_PAR_DOT_PROD_DCL, FFT_PAR_PROD_DCL, &
       FFT_SCALAR_VECT_DCL, FFT_TENSOR_DCL, FFT_PAR_DIV_DCL, FFT_HEAVISIDE_DCL, &
       FFT_SQUAREROOT_SCALAR, FFT_MAX_VEL_DCL, FFT_MAX_MIN_VEL_DCL, FFT_PAR_REAL, FFT_CHECK_INTERFACE, &
       FFT_PAR_SCAL_FUNCT, &
       sfx_mesh_register, sfx_mesh_release, &
       sfx_smb_cross_prod_gauss, sfx_smb_curlh_cross_b_gauss

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
     FUNCTION c_mesh_create(jj, rr, n_w, l_G, me, np, handle) BIND(C, NAME='sfx_mesh_create') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: jj, rr
       INTEGER(C_INT32_T), VALUE :: n_w, l_G
       INTEGER(C_INT64_T), VALUE :: me, np
       TYPE(C_PTR) :: handle
       INTEGER(C_INT) :: rc
     END FUNCTION c_mesh_create
     FUNCTION c_mesh_destroy(handle) BIND(C, NAME='sfx_mesh_destroy') RESULT(rc)
       IMPORT :: C_PTR, C_INT
       TYPE(C_PTR), VALUE :: handle
       INTEGER(C_INT) :: rc
     END FUNCTION c_mesh_destroy
     FUNCTION c_smb_cross(mesh, lm, m, Vi, Vo, prec, taux, ang) &
          BIND(C, NAME='sfx_smb_cross_prod_gauss_sft_par') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_DOUBLE, C_INT
       TYPE(C_PTR), VALUE :: mesh, lm, Vi, Vo
       INTEGER(C_INT32_T), VALUE :: m, prec
       REAL(C_DOUBLE), VALUE :: taux, ang
       INTEGER(C_INT) :: rc
     END FUNCTION c_smb_cross
     FUNCTION c_smb_curlh(mesh, lm, m, Vi, Wi, Vo) &
          BIND(C, NAME='sfx_smb_curlh_cross_b_gauss_sft_par') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT
       TYPE(C_PTR), VALUE :: mesh, lm, Vi, Wi, Vo
       INTEGER(C_INT32_T), VALUE :: m
       INTEGER(C_INT) :: rc
     END FUNCTION c_smb_curlh
     FUNCTION c_last_error() BIND(C, NAME='sfx_last_error') RESULT(p)
       IMPORT :: C_PTR
       TYPE(C_PTR) :: p
     END FUNCTION c_last_error
  END INTERFACE

CONTAINS

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
    IF (PRESENT(temps)) temps(1:3) = tloc
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
    IF (PRESENT(temps)) temps(1:3) = tloc
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
    IF (PRESENT(temps)) temps(1:3) = tloc
    DEALLOCATE(a1, av, ao)
  END SUBROUTINE FFT_HEAVISIDE_DCL

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
    IF (PRESENT(temps)) temps(1:3) = tloc
    DEALLOCATE(a1)
  END SUBROUTINE FFT_SQUAREROOT_SCALAR

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
    INTEGER :: i, j, np, m
    IF (PRESENT(temps)) temps = 0.d0
    np = SIZE(c1_inout,1); m = SIZE(c1_inout,3)
    IF (nb_procs /= 1 .OR. m_max_pad < m) THEN
       WRITE(*,*) 'FFT_PAR_SCAL_FUNCT: nb_procs must be 1 and m_max_pad >= m_max_c'
       STOP
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
    IF (PRESENT(temps)) temps(1:3) = tloc
    DEALLOCATE(a1)
  END SUBROUTINE FFT_CHECK_INTERFACE

  ! Register the P2 mesh once (mesh%jj, mesh%rr of def_type_mesh).
  FUNCTION sfx_mesh_register(jj, rr) RESULT(handle)
    INTEGER, DIMENSION(:,:), INTENT(IN) :: jj
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: rr
    TYPE(C_PTR) :: handle
    INTEGER(C_INT32_T), DIMENSION(:,:), ALLOCATABLE, TARGET :: jc
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: rc_
    INTEGER(C_INT) :: rc
    ALLOCATE(jc(SIZE(jj,1), SIZE(jj,2)))
    jc = INT(jj, C_INT32_T)
    ALLOCATE(rc_, SOURCE=rr)
    rc = c_mesh_create(C_LOC(jc), C_LOC(rc_), INT(SIZE(jj,1),C_INT32_T), 7_C_INT32_T, &
         INT(SIZE(jj,2),C_INT64_T), INT(SIZE(rr,2),C_INT64_T), handle)
    CALL sfx_stop(rc, 'sfx_mesh_register')
    DEALLOCATE(jc, rc_)
  END FUNCTION sfx_mesh_register

  SUBROUTINE sfx_mesh_release(handle)
    TYPE(C_PTR), INTENT(INOUT) :: handle
    INTEGER(C_INT) :: rc
    rc = c_mesh_destroy(handle)
    handle = C_NULL_PTR
  END SUBROUTINE sfx_mesh_release

  ! Body of smb_cross_prod_gauss_sft_par (subroutine_ns_with_velocity.F90:514)
  ! for a mesh registered with sfx_mesh_register.
  SUBROUTINE sfx_smb_cross_prod_gauss(mesh, list_mode, V_in, V_out, precession_in, taux_precession, angle_s_pi)
    TYPE(C_PTR), INTENT(IN) :: mesh
    INTEGER, DIMENSION(:), INTENT(IN) :: list_mode
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    LOGICAL, INTENT(IN) :: precession_in
    REAL(KIND=8), INTENT(IN) :: taux_precession, angle_s_pi
    INTEGER(C_INT32_T), DIMENSION(:), ALLOCATABLE, TARGET :: lm
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: vi, vo
    INTEGER(C_INT32_T) :: prec
    INTEGER(C_INT) :: rc
    ALLOCATE(lm(SIZE(list_mode)))
    lm = INT(list_mode, C_INT32_T)
    ALLOCATE(vi, SOURCE=V_in)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    prec = 0
    IF (precession_in) prec = 1
    rc = c_smb_cross(mesh, C_LOC(lm), INT(SIZE(list_mode),C_INT32_T), C_LOC(vi), C_LOC(vo), prec, &
         taux_precession, angle_s_pi)
    CALL sfx_stop(rc, 'smb_cross_prod_gauss_sft_par')
    V_out = vo
    DEALLOCATE(lm, vi, vo)
  END SUBROUTINE sfx_smb_cross_prod_gauss

  ! Body of smb_CurlH_cross_B_gauss_sft_par (subroutine_ns_with_velocity.F90:632).
  SUBROUTINE sfx_smb_curlh_cross_b_gauss(mesh, list_mode, V_in, W_in, V_out)
    TYPE(C_PTR), INTENT(IN) :: mesh
    INTEGER, DIMENSION(:), INTENT(IN) :: list_mode
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V_in, W_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    INTEGER(C_INT32_T), DIMENSION(:), ALLOCATABLE, TARGET :: lm
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: vi, wi, vo
    INTEGER(C_INT) :: rc
    ALLOCATE(lm(SIZE(list_mode)))
    lm = INT(list_mode, C_INT32_T)
    ALLOCATE(vi, SOURCE=V_in)
    ALLOCATE(wi, SOURCE=W_in)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    rc = c_smb_curlh(mesh, C_LOC(lm), INT(SIZE(list_mode),C_INT32_T), C_LOC(vi), C_LOC(wi), C_LOC(vo))
    CALL sfx_stop(rc, 'smb_CurlH_cross_B_gauss_sft_par')
    V_out = vo
    DEALLOCATE(lm, vi, wi, vo)
  END SUBROUTINE sfx_smb_curlh_cross_b_gauss

END MODULE fft_parallele
