!===
! The Gauss-point nonlinear terms and the right-hand-side assembly of the
! reference, for one GPU, with the reference's argument lists:
!   smb_cross_prod_gauss_sft_par     SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514
!   smb_CurlH_cross_B_gauss_sft_par  SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:632
!   smb_ugradc_gauss_fft_par         SOURCES/SIMULATION/subroutine_temperature_with_T.F90:247
!   rhs_3x3                          SOURCES/SOLVER/assembling_rhs.F90:5
!   qs_00_gauss                      SOURCES/SOLVER/fem_rhs_axi.F90:940
! In the reference the three smb_* routines are private procedures of the
! physics modules (one copy per module); a module using this one drops its
! own copy and keeps its calls.  rhs_3x3 and qs_00_gauss fill PETSc vectors
! with VecSetValues; here their vector arguments are the vectors' local
! arrays (VecGetArrayF90 on one rank), filled through the layout's
! loc_to_glob numbering, every entry written (the reference starts with
! VecSet(.., 0)).
!
! The mesh is registered with libsfx on first use and kept, keyed by the
! addresses of mesh%jj and mesh%rr and the sizes me, np (SFEMaNS builds its
! meshes once, at INIT); sfx_release_meshes frees the registrations.  The
! boundary sides mesh%jjs are registered with the mesh when rhs_3x3 first
! receives opt_tensor_scaln_bdy.  Errors stop the program with libsfx's
! message.
!===
MODULE sfx_nonlinear
  USE, INTRINSIC :: ISO_C_BINDING
  USE def_type_mesh, ONLY: mesh_type, petsc_csr_LA
  USE input_data, ONLY: inputs
  USE fft_parallele, ONLY: sfx_stop
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: smb_cross_prod_gauss_sft_par, smb_CurlH_cross_B_gauss_sft_par, smb_ugradc_gauss_fft_par, &
       rhs_3x3, qs_00_gauss, sfx_release_meshes

  TYPE registered_mesh
     TYPE(C_PTR) :: jj_addr = C_NULL_PTR, rr_addr = C_NULL_PTR
     INTEGER :: me = -1, np = -1, mes = -1
     TYPE(C_PTR) :: handle = C_NULL_PTR
  END TYPE registered_mesh
  INTEGER, PARAMETER :: max_meshes = 32
  TYPE(registered_mesh), DIMENSION(max_meshes), SAVE :: reg

  INTERFACE
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
     FUNCTION c_mesh_boundary(handle, jjs, mes) BIND(C, NAME='sfx_mesh_set_boundary') RESULT(rc)
       IMPORT :: C_PTR, C_INT64_T, C_INT
       TYPE(C_PTR), VALUE :: handle, jjs
       INTEGER(C_INT64_T), VALUE :: mes
       INTEGER(C_INT) :: rc
     END FUNCTION c_mesh_boundary
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
     FUNCTION c_smb_ugradc(mesh, lm, m, hc, Vi, ci, co) &
          BIND(C, NAME='sfx_smb_ugradc_gauss_fft_par') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT
       TYPE(C_PTR), VALUE :: mesh, lm, hc, Vi, ci, co
       INTEGER(C_INT32_T), VALUE :: m
       INTEGER(C_INT) :: rc
     END FUNCTION c_smb_ugradc
     FUNCTION c_rhs_mode(mesh, mode, rhs, tens, tbdy, v145, v236) BIND(C, NAME='sfx_rhs_3x3_mode') RESULT(rc)
       IMPORT :: C_PTR, C_INT32_T, C_INT
       TYPE(C_PTR), VALUE :: mesh, rhs, tens, tbdy, v145, v236
       INTEGER(C_INT32_T), VALUE :: mode
       INTEGER(C_INT) :: rc
     END FUNCTION c_rhs_mode
     FUNCTION c_qs00(mesh, hc, ff, ffg, vect) BIND(C, NAME='sfx_qs_00_gauss') RESULT(rc)
       IMPORT :: C_PTR, C_INT
       TYPE(C_PTR), VALUE :: mesh, hc, ff, ffg, vect
       INTEGER(C_INT) :: rc
     END FUNCTION c_qs00
  END INTERFACE

CONTAINS

  !===The libsfx handle of a mesh, registered on first use
  FUNCTION mesh_handle(mesh, with_boundary) RESULT(h)
    TYPE(mesh_type), INTENT(IN) :: mesh
    LOGICAL, INTENT(IN) :: with_boundary
    TYPE(C_PTR) :: h
    TYPE(C_PTR) :: ja, ra
    INTEGER(C_INT32_T), DIMENSION(:,:), ALLOCATABLE, TARGET :: jc, jsc
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: rc_
    INTEGER(C_INT) :: rc
    INTEGER :: i, slot
    IF (mesh%gauss%n_w /= 6 .OR. mesh%gauss%l_G /= 7) THEN
       WRITE(*,*) ' BUG in sfx_nonlinear: only P2 meshes (n_w = 6, l_G = 7)'
       STOP 1
    END IF
    ja = C_NULL_PTR; ra = C_NULL_PTR
    IF (mesh%me > 0) ja = C_LOC(mesh%jj(1,1))
    IF (mesh%np > 0) ra = C_LOC(mesh%rr(1,1))
    slot = 0
    DO i = 1, max_meshes
       IF (C_ASSOCIATED(reg(i)%handle)) THEN
          IF (C_ASSOCIATED(reg(i)%jj_addr, ja) .AND. C_ASSOCIATED(reg(i)%rr_addr, ra) .AND. &
               reg(i)%me == mesh%me .AND. reg(i)%np == mesh%np) THEN
             slot = i
             EXIT
          END IF
       END IF
    END DO
    IF (slot == 0) THEN
       DO i = 1, max_meshes
          IF (.NOT. C_ASSOCIATED(reg(i)%handle)) THEN
             slot = i
             EXIT
          END IF
       END DO
       IF (slot == 0) THEN
          WRITE(*,*) ' BUG in sfx_nonlinear: more than', max_meshes, ' meshes registered'
          STOP 1
       END IF
       ALLOCATE(jc(6, MAX(mesh%me,1)), rc_(2, MAX(mesh%np,1)))
       IF (mesh%me > 0) jc(:, 1:mesh%me) = INT(mesh%jj(:, 1:mesh%me), C_INT32_T)
       IF (mesh%np > 0) rc_(:, 1:mesh%np) = mesh%rr(:, 1:mesh%np)
       rc = c_mesh_create(C_LOC(jc), C_LOC(rc_), 6_C_INT32_T, 7_C_INT32_T, INT(mesh%me,C_INT64_T), &
            INT(mesh%np,C_INT64_T), reg(slot)%handle)
       CALL sfx_stop(rc, 'sfx_mesh_create')
       DEALLOCATE(jc, rc_)
       reg(slot)%jj_addr = ja; reg(slot)%rr_addr = ra
       reg(slot)%me = mesh%me; reg(slot)%np = mesh%np; reg(slot)%mes = -1
    END IF
    IF (with_boundary .AND. reg(slot)%mes /= mesh%mes) THEN
       ALLOCATE(jsc(3, MAX(mesh%mes,1)))
       IF (mesh%mes > 0) jsc(:, 1:mesh%mes) = INT(mesh%jjs(1:3, 1:mesh%mes), C_INT32_T)
       rc = c_mesh_boundary(reg(slot)%handle, C_LOC(jsc), INT(mesh%mes,C_INT64_T))
       CALL sfx_stop(rc, 'sfx_mesh_set_boundary')
       DEALLOCATE(jsc)
       reg(slot)%mes = mesh%mes
    END IF
    h = reg(slot)%handle
  END FUNCTION mesh_handle

  SUBROUTINE sfx_release_meshes()
    INTEGER :: i
    INTEGER(C_INT) :: rc
    DO i = 1, max_meshes
       IF (C_ASSOCIATED(reg(i)%handle)) THEN
          rc = c_mesh_destroy(reg(i)%handle)
          reg(i) = registered_mesh()
       END IF
    END DO
  END SUBROUTINE sfx_release_meshes

  SUBROUTINE check_one_rank(mesh, name)
    TYPE(mesh_type), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    IF (mesh%dom_me /= mesh%me) THEN
       WRITE(*,*) ' BUG in ', name, ': one GPU, every element must be owned (dom_me = me)'
       STOP 1
    END IF
  END SUBROUTINE check_one_rank

  SUBROUTINE smb_cross_prod_gauss_sft_par(communicator, mesh, list_mode, V_in, V_out, precession_in)
    INTEGER :: communicator
    TYPE(mesh_type), INTENT(IN) :: mesh
    INTEGER, DIMENSION(:), INTENT(IN) :: list_mode
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: V_out
    LOGICAL, INTENT(IN) :: precession_in
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
    ! the precession vector of inputs%taux_precession, inputs%angle_s_pi
    ! (subroutine_ns_with_velocity.F90:536-551)
    rc = c_smb_cross(mesh_handle(mesh, .FALSE.), C_LOC(lm), INT(SIZE(list_mode),C_INT32_T), C_LOC(vi), &
         C_LOC(vo), prec, REAL(inputs%taux_precession,C_DOUBLE), REAL(inputs%angle_s_pi,C_DOUBLE))
    CALL sfx_stop(rc, 'smb_cross_prod_gauss_sft_par')
    V_out = vo
    DEALLOCATE(lm, vi, vo)
  END SUBROUTINE smb_cross_prod_gauss_sft_par

  SUBROUTINE smb_CurlH_cross_B_gauss_sft_par(communicator, mesh, list_mode, V_in, W_in, V_out)
    INTEGER :: communicator
    TYPE(mesh_type), INTENT(IN) :: mesh
    INTEGER, DIMENSION(:), INTENT(IN) :: list_mode
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V_in, W_in
    REAL(KIND=8), DIMENSION(:,:,:) :: V_out
    INTEGER(C_INT32_T), DIMENSION(:), ALLOCATABLE, TARGET :: lm
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: vi, wi, vo
    INTEGER(C_INT) :: rc
    ALLOCATE(lm(SIZE(list_mode)))
    lm = INT(list_mode, C_INT32_T)
    ALLOCATE(vi, SOURCE=V_in)
    ALLOCATE(wi, SOURCE=W_in)
    ALLOCATE(vo(SIZE(V_out,1), SIZE(V_out,2), SIZE(V_out,3)))
    rc = c_smb_curlh(mesh_handle(mesh, .FALSE.), C_LOC(lm), INT(SIZE(list_mode),C_INT32_T), C_LOC(vi), &
         C_LOC(wi), C_LOC(vo))
    CALL sfx_stop(rc, 'smb_CurlH_cross_B_gauss_sft_par')
    V_out = vo
    DEALLOCATE(lm, vi, wi, vo)
  END SUBROUTINE smb_CurlH_cross_B_gauss_sft_par

  SUBROUTINE smb_ugradc_gauss_fft_par(communicator, mesh, list_mode, heat_capa_in, V_in, c_in, c_out)
    INTEGER :: communicator
    TYPE(mesh_type), TARGET :: mesh
    INTEGER, DIMENSION(:), INTENT(IN) :: list_mode
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: heat_capa_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: V_in, c_in
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: c_out
    INTEGER(C_INT32_T), DIMENSION(:), ALLOCATABLE, TARGET :: lm
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: vi, ci, co
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: hc
    INTEGER(C_INT) :: rc
    ALLOCATE(lm(SIZE(list_mode)))
    lm = INT(list_mode, C_INT32_T)
    ALLOCATE(vi, SOURCE=V_in)
    ALLOCATE(ci, SOURCE=c_in)
    ALLOCATE(hc(MAX(mesh%me,1)))
    hc(1:mesh%me) = heat_capa_in(1:mesh%me)
    ALLOCATE(co(SIZE(c_out,1), SIZE(c_out,2), SIZE(c_out,3)))
    rc = c_smb_ugradc(mesh_handle(mesh, .FALSE.), C_LOC(lm), INT(SIZE(list_mode),C_INT32_T), C_LOC(hc), &
         C_LOC(vi), C_LOC(ci), C_LOC(co))
    CALL sfx_stop(rc, 'smb_ugradc_gauss_fft_par')
    c_out = co
    DEALLOCATE(lm, vi, ci, hc, co)
  END SUBROUTINE smb_ugradc_gauss_fft_par

  !===rhs_3x3 (assembling_rhs.F90:5-189): vb_145, vb_236 are the local arrays
  !===of the two vectors (3*np entries each), indexed by vv_3_LA%loc_to_glob
  SUBROUTINE rhs_3x3(mesh, vv_3_LA, mode, rhs_in, vb_145, vb_236, opt_tensor, opt_tensor_scaln_bdy)
    TYPE(mesh_type) :: mesh
    TYPE(petsc_csr_LA) :: vv_3_LA
    INTEGER, INTENT(IN) :: mode
    REAL(KIND=8), DIMENSION(mesh%dom_me*mesh%gauss%l_G,6), INTENT(IN) :: rhs_in
    REAL(KIND=8), DIMENSION(:), INTENT(OUT) :: vb_145, vb_236
    REAL(KIND=8), DIMENSION(3,mesh%dom_me*mesh%gauss%l_G,6), INTENT(IN), OPTIONAL :: opt_tensor
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN), OPTIONAL :: opt_tensor_scaln_bdy
    REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE, TARGET :: ri, tb, d145, d236
    REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: ti
    TYPE(C_PTR) :: pt, pb, h
    INTEGER(C_INT) :: rc
    INTEGER :: i, ki, np
    CALL check_one_rank(mesh, 'rhs_3x3')
    np = mesh%np
    h = mesh_handle(mesh, PRESENT(opt_tensor_scaln_bdy))
    ALLOCATE(ri, SOURCE=rhs_in)
    pt = C_NULL_PTR; pb = C_NULL_PTR
    IF (PRESENT(opt_tensor)) THEN
       ALLOCATE(ti, SOURCE=opt_tensor)
       pt = C_LOC(ti)
    END IF
    IF (PRESENT(opt_tensor_scaln_bdy)) THEN
       IF (SIZE(opt_tensor_scaln_bdy,1) < mesh%gauss%l_Gs*mesh%mes .OR. SIZE(opt_tensor_scaln_bdy,2) /= 6) THEN
          WRITE(*,*) ' BUG in rhs_3x3: opt_tensor_scaln_bdy must be (l_Gs*mes, 6)'
          STOP 1
       END IF
       ALLOCATE(tb(3*MAX(mesh%mes,1), 6))
       tb(1:3*mesh%mes, :) = opt_tensor_scaln_bdy(1:3*mesh%mes, :)
       pb = C_LOC(tb)
    END IF
    ALLOCATE(d145(MAX(np,1), 3), d236(MAX(np,1), 3))
    rc = c_rhs_mode(h, INT(mode,C_INT32_T), C_LOC(ri), pt, pb, C_LOC(d145), C_LOC(d236))
    CALL sfx_stop(rc, 'rhs_3x3')
    DO ki = 1, 3
       DO i = 1, np
          vb_145(vv_3_LA%loc_to_glob(ki,i)) = d145(i,ki)
          vb_236(vv_3_LA%loc_to_glob(ki,i)) = d236(i,ki)
       END DO
    END DO
    DEALLOCATE(ri, d145, d236)
    IF (ALLOCATED(ti)) DEALLOCATE(ti)
    IF (ALLOCATED(tb)) DEALLOCATE(tb)
  END SUBROUTINE rhs_3x3

  !===qs_00_gauss (fem_rhs_axi.F90:940-987): vect is the vector's local array,
  !===indexed by LA%loc_to_glob(1, :)
  SUBROUTINE qs_00_gauss(mesh, LA, heat_capa, ff, ff_gauss, vect)
    TYPE(mesh_type), TARGET :: mesh
    TYPE(petsc_csr_LA) :: LA
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: heat_capa
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: ff, ff_gauss
    REAL(KIND=8), DIMENSION(:), INTENT(OUT) :: vect
    REAL(KIND=8), DIMENSION(:), ALLOCATABLE, TARGET :: hc, f, fg, v
    INTEGER(C_INT) :: rc
    INTEGER :: i
    CALL check_one_rank(mesh, 'qs_00_gauss')
    ALLOCATE(hc(MAX(mesh%me,1)), f(MAX(mesh%np,1)), fg(MAX(7*mesh%me,1)), v(MAX(mesh%np,1)))
    hc(1:mesh%me) = heat_capa(1:mesh%me)
    f(1:mesh%np) = ff(1:mesh%np)
    fg(1:7*mesh%me) = ff_gauss(1:7*mesh%me)
    rc = c_qs00(mesh_handle(mesh, .FALSE.), C_LOC(hc), C_LOC(f), C_LOC(fg), C_LOC(v))
    CALL sfx_stop(rc, 'qs_00_gauss')
    DO i = 1, mesh%np
       vect(LA%loc_to_glob(1,i)) = v(i)
    END DO
    DEALLOCATE(hc, f, fg, v)
  END SUBROUTINE qs_00_gauss

END MODULE sfx_nonlinear
