!===
! Calls every routine of the fft_parallele and sfx_nonlinear modules the way
! the reference's callers do (their argument lists and array shapes), on
! inputs read from raw little-endian files, and writes each result to a
! raw file, so that the calling test checks them against the oracle.
!
! usage: sfx_fortran_driver <dir>
!   <dir>/dims.txt : np m_max me npm mes   (np a multiple of 7: seven Gauss points per element)
!   inputs (float64 unless noted, Fortran order):
!     v1..v5 (np,6,m), s1 (np,2,m) in [0,1], s2 (np,2,m) positive, creal (2*mp-1,np), h (np),
!     ls (2,np,2,m) level sets, rg2 (2,np) points for the Gauss form of ETA,
!     jj (6,me) int32, rr (2,npm), jjs (3,mes) int32, vn, wn (npm,6,m), cn (npm,2,m), hc (me),
!     rhs (7me,6), ten (3,7me,6), tbd (3mes,6), ffn (npm), ffg (7me)
!   mp = 3*m/2 (the callers' m_max_pad); inputs%... are set below
! writes <name>.bin for every routine (see the calling test).
!===
PROGRAM sfx_fortran_driver
  USE, INTRINSIC :: ISO_C_BINDING
  USE def_type_mesh
  USE input_data
  USE fft_parallele
  USE sfx_nonlinear
  IMPLICIT NONE
  CHARACTER(LEN=1024) :: dir
  INTEGER :: np, m, me, npm, mes, mp, nr, u, i, k, ok
  INTEGER :: comm
  REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: v1, v2, v3, v4, v5, s1, s2, o6, o6b, o6c, o2, o2b, vn, wn, cn
  REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: vreal, og
  REAL(KIND=8), DIMENSION(:,:,:,:), ALLOCATABLE :: ls, t4, ten4
  REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: creal, rg2, rr, rhs, tbd, r1, r2
  REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: ten
  REAL(KIND=8), DIMENSION(:), ALLOCATABLE :: h, hc, ffn, ffg, vb1, vb2, vq
  INTEGER, DIMENSION(:,:), ALLOCATABLE :: jj, jjs
  INTEGER, DIMENSION(:), ALLOCATABLE :: list_mode
  REAL(KIND=8), DIMENSION(3) :: temps
  REAL(KIND=8) :: vmax
  REAL(KIND=8), DIMENSION(2) :: vmm
  TYPE(mesh_type) :: mesh
  TYPE(petsc_csr_LA) :: LA3, LA1

  CALL GET_COMMAND_ARGUMENT(1, dir)
  OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/dims.txt', STATUS='old')
  READ(u,*) np, m, me, npm, mes
  CLOSE(u)
  mp = 3*m/2
  nr = 2*mp - 1
  comm = 0
  temps = 0.d0
  inputs%LES_coeff1 = 0.3d0; inputs%LES_coeff2 = 0.9d0; inputs%LES_coeff3 = 0.d0; inputs%LES_coeff4 = 0.8d0
  inputs%LES_coeff1_mom = 0.2d0; inputs%level_set_comp_factor = 1.5d0; inputs%h_min_distance = 0.01d0
  inputs%taux_precession = 0.3d0; inputs%angle_s_pi = 0.25d0

  ALLOCATE(v1(np,6,m), v2(np,6,m), v3(np,6,m), v4(np,6,m), v5(np,6,m), s1(np,2,m), s2(np,2,m))
  ALLOCATE(creal(nr,np), h(np), ls(2,np,2,m), rg2(2,np))
  CALL rd3('v1', v1); CALL rd3('v2', v2); CALL rd3('v3', v3); CALL rd3('v4', v4); CALL rd3('v5', v5)
  CALL rd3('s1', s1); CALL rd3('s2', s2); CALL rd2('creal', creal); CALL rd1('h', h); CALL rd4('ls', ls)
  CALL rd2('rg2', rg2)
  ALLOCATE(o6(np,6,m), o6b(np,6,m), o6c(np,6,m), o2(np,2,m), o2b(np,2,m), t4(3,np,6,m))

  !===polynomial products
  CALL FFT_PAR_CROSS_PROD_DCL(comm, v1, v2, o6, 1, np, mp, temps)
  CALL wr3('cross', o6)
  CALL FFT_PAR_DOT_PROD_DCL(comm, v1, v2, o2, 1, np, mp)
  CALL wr3('dot', o2)
  CALL FFT_PAR_PROD_DCL(comm, s1, s2, o2, 1, np, mp, temps)
  CALL wr3('prod', o2)
  CALL FFT_SCALAR_VECT_DCL(comm, v1, s2, o6, 1, 1, np, mp)
  CALL wr3('svect1', o6)
  CALL FFT_TENSOR_DCL(comm, v1, v2, t4, 1, np, mp)
  CALL wr4('tensor', t4)
  !===non-polynomial products on the reference grid
  CALL FFT_PAR_DIV_DCL(comm, v1(:,1:2,:), s2, o2, 1, np, mp)
  CALL wr3('div', o2)
  CALL FFT_SCALAR_VECT_DCL(comm, v1, s2, o6, 2, 1, np, mp, temps)
  CALL wr3('svect2', o6)
  CALL FFT_SCALAR_VECT_NO_OVERSHOOT(comm, [0.2d0, 0.7d0, 0.5d0], v1, s1, o6, 1, 1, np, mp)
  CALL wr3('svect_no', o6)
  CALL FFT_TENSOR_DCL(comm, v1, v2, t4, 1, np, mp, opt_tension=.TRUE.)
  CALL wr4('tension', t4)
  o2 = s1
  CALL FFT_SQUAREROOT_SCALAR(comm, o2, 1, np, mp)
  CALL wr3('sqrt', o2)
  o2 = s1
  CALL FFT_NO_OVERSHOOT_LEVEL_SET(comm, o2, 1, np, mp, temps)
  CALL wr3('no_ls', o2)
  CALL FFT_HEAVISIDE_DCL(comm, ls, [1.d0, 3.d0, 0.5d0], o2, 1, np, mp, 0.05d0, temps)
  CALL wr3('heavi', o2)
  CALL FFT_MAX_VEL_DCL(comm, v1, vmax, 1, np, mp)
  CALL wr0('maxvel', [vmax])
  CALL FFT_MAX_MIN_VEL_DCL(comm, v1, vmm, 1, np, mp)
  CALL wr0('maxmin', vmm)
  CALL FFT_CHECK_INTERFACE(comm, ls, 3, ok, 1, np, mp, temps)
  CALL wr0('checkif', [DBLE(ok)])
  CALL wr0('regul', [regul(0.47d0, 0.05d0), regul(0.6d0, 0.05d0), regul(0.4d0, 0.05d0), regul(0.5d0, 0.05d0)])
  CALL FFT_PAR_REAL(comm, v1, vreal, 2*m+1)
  CALL wr3('real', vreal)
  o2 = s2
  CALL FFT_PAR_SCAL_FUNCT(comm, o2, square, 1, np, mp)
  CALL wr3('scal', o2)
  !===user-function products (the function below, on the grid angles)
  ALLOCATE(rr(2,npm))
  CALL rd2('rr', rr)
  mesh%np = npm
  ALLOCATE(mesh%rr(2,npm))
  mesh%rr = rr
  CALL FFT_PAR_VAR_ETA_PROD_T_DCL(comm, eta_t, mesh, v1(1:MIN(np,npm),:,:), o6(1:MIN(np,npm),:,:), 1, np, mp, &
       0.5d0, temps)
  CALL wr3('eta_t', o6(1:MIN(np,npm),:,:))
  CALL FFT_PAR_VAR_ETA_PROD_GAUSS_DCL(comm, eta_g, mesh, s1, o2, 1, np, mp, rg2, 0.5d0)
  CALL wr3('eta_g', o2)
  !===entropy viscosity and level-set compression
  CALL FFT_COMPUTE_ENTROPY_VISC(comm, comm, v1, v2, v3, v4, v5, h, o6, o6b, o6c, 1, np, mp, 1.0d0, 7, temps)
  CALL wr3('ev1', o6); CALL wr3('ev2', o6b); CALL wr3('ev3', o6c)
  ALLOCATE(r1(nr,np), r2(nr,np))
  CALL FFT_COMPUTE_ENTROPY_VISC_MOM(comm, comm, v1, v2, v3, s2, h, r1, 1, np, mp, 7, opt_c2_real_out=r2)
  CALL wr2('evm1', r1); CALL wr2('evm2', r2)
  CALL FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM(comm, v1, v2, v3, creal, t4, 1, np, mp, temps)
  CALL wr4('evgrad', t4)
  CALL FFT_COMPUTE_DIFFU_MOM(comm, v1, v2, s2, o6, o6b, 1, np, mp)
  CALL wr3('diffu1', o6); CALL wr3('diffu2', o6b)
  CALL FFT_PAR_ENTRO_VISC_DCL(comm, v1, s1, creal, h, 0.3d0, o6, o2, 1, np, mp)
  CALL wr3('entro_v', o6); CALL wr3('entro_c', o2)
  CALL FFT_PAR_COMPR_ENTRO_VISC_DCL(comm, v1, v2, s1, creal, h, 0.3d0, o6, o2, 1, np, mp, temps)
  CALL wr3('compr_v', o6); CALL wr3('compr_c', o2)
  CALL FFT_COMPRESSION_LEVEL_SET_DCL(comm, comm, v1, v2, s1, o2, h, 7, 1, np, mp, temps)
  CALL wr3('cls', o2)

  !===the Gauss-point terms and the assembly on a P2 mesh
  ALLOCATE(jj(6,me), jjs(3,mes), vn(npm,6,m), wn(npm,6,m), cn(npm,2,m), hc(me))
  CALL rdi('jj', jj); CALL rdi('jjs', jjs)
  CALL rd3('vn', vn); CALL rd3('wn', wn); CALL rd3('cn', cn); CALL rd1('hc', hc)
  mesh%me = me; mesh%dom_me = me; mesh%mes = mes; mesh%dom_mes = mes; mesh%dom_np = npm
  ALLOCATE(mesh%jj(6,me), mesh%jjs(3,mes))
  mesh%jj = jj; mesh%jjs = jjs
  mesh%gauss%k_d = 2; mesh%gauss%n_w = 6; mesh%gauss%l_G = 7; mesh%gauss%n_ws = 3; mesh%gauss%l_Gs = 3
  ALLOCATE(list_mode(m))
  list_mode = [(i - 1, i = 1, m)]
  ALLOCATE(og(7*me,6,m))
  CALL smb_cross_prod_gauss_sft_par(comm, mesh, list_mode, vn, og, .TRUE.)
  CALL wr3('smb_cross', og)
  CALL smb_CurlH_cross_B_gauss_sft_par(comm, mesh, list_mode, vn, wn, og)
  CALL wr3('smb_curlh', og)
  CALL smb_ugradc_gauss_fft_par(comm, mesh, list_mode, hc, vn, cn, og(:,1:2,:))
  CALL wr3('smb_ugradc', og(:,1:2,:))
  ! one-rank vector layouts: loc_to_glob(ki, i) = (ki-1)*np + i
  ALLOCATE(LA3%loc_to_glob(3,npm), LA1%loc_to_glob(1,npm))
  DO k = 1, 3
     DO i = 1, npm
        LA3%loc_to_glob(k,i) = (k-1)*npm + i
     END DO
  END DO
  DO i = 1, npm
     LA1%loc_to_glob(1,i) = i
  END DO
  ALLOCATE(rhs(7*me,6), ten(3,7*me,6), tbd(3*mes,6), vb1(3*npm), vb2(3*npm))
  CALL rd2('rhs', rhs); CALL rd3('ten', ten); CALL rd2('tbd', tbd)
  CALL rhs_3x3(mesh, LA3, 2, rhs, vb1, vb2)
  CALL wr0('rhs_a145', vb1); CALL wr0('rhs_a236', vb2)
  CALL rhs_3x3(mesh, LA3, 2, rhs, vb1, vb2, opt_tensor=ten, opt_tensor_scaln_bdy=tbd)
  CALL wr0('rhs_b145', vb1); CALL wr0('rhs_b236', vb2)
  ALLOCATE(ffn(npm), ffg(7*me), vq(npm))
  CALL rd1('ffn', ffn); CALL rd1('ffg', ffg)
  CALL qs_00_gauss(mesh, LA1, hc, ffn, ffg, vq)
  CALL wr0('qs00', vq)
  CALL sfx_release_meshes()
  WRITE(*,'(A)') 'sfx_fortran_driver: ok'

CONTAINS

  FUNCTION square(x) RESULT(vv)
    REAL(KIND=8) :: x
    REAL(KIND=8) :: vv
    vv = x*x
  END FUNCTION square

  !===eta(r, z, theta) = 1 + r*cos(theta)/2 + z*sin(theta)/4 + time at the nodes nb..ne
  FUNCTION eta_t(H_mesh, angles, nb_angles, nb, ne, time) RESULT(vv)
    TYPE(mesh_type), INTENT(IN) :: H_mesh
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: angles
    INTEGER, INTENT(IN) :: nb_angles, nb, ne
    REAL(KIND=8), INTENT(IN) :: time
    REAL(KIND=8), DIMENSION(nb_angles,ne-nb+1) :: vv
    INTEGER :: j
    DO j = nb, ne
       vv(:, j-nb+1) = 1 + H_mesh%rr(1,j)*COS(angles)/2 + H_mesh%rr(2,j)*SIN(angles)/4 + time
    END DO
  END FUNCTION eta_t

  !===the same function at the points rr_gauss
  FUNCTION eta_g(H_mesh, rr_gauss, angles, nb_angles, nb, ne, time) RESULT(vv)
    TYPE(mesh_type) :: H_mesh
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: rr_gauss
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: angles
    INTEGER, INTENT(IN) :: nb_angles, nb, ne
    REAL(KIND=8), INTENT(IN) :: time
    REAL(KIND=8), DIMENSION(nb_angles,ne-nb+1) :: vv
    INTEGER :: j
    DO j = 1, ne-nb+1
       vv(:, j) = 1 + rr_gauss(1,j)*COS(angles)/2 + rr_gauss(2,j)*SIN(angles)/4 + time
    END DO
  END FUNCTION eta_g

  SUBROUTINE rd_open(name, v)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(OUT) :: v
    OPEN(NEWUNIT=v, FILE=TRIM(dir)//'/'//name//'.bin', ACCESS='stream', FORM='unformatted', STATUS='old')
  END SUBROUTINE rd_open

  SUBROUTINE rd1(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:), INTENT(OUT) :: a
    INTEGER :: v
    CALL rd_open(name, v); READ(v) a; CLOSE(v)
  END SUBROUTINE rd1

  SUBROUTINE rd2(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:), INTENT(OUT) :: a
    INTEGER :: v
    CALL rd_open(name, v); READ(v) a; CLOSE(v)
  END SUBROUTINE rd2

  SUBROUTINE rd3(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: a
    INTEGER :: v
    CALL rd_open(name, v); READ(v) a; CLOSE(v)
  END SUBROUTINE rd3

  SUBROUTINE rd4(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(OUT) :: a
    INTEGER :: v
    CALL rd_open(name, v); READ(v) a; CLOSE(v)
  END SUBROUTINE rd4

  SUBROUTINE rdi(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, DIMENSION(:,:), INTENT(OUT) :: a
    INTEGER(C_INT32_T), DIMENSION(SIZE(a,1), SIZE(a,2)) :: b
    INTEGER :: v
    CALL rd_open(name, v); READ(v) b; CLOSE(v)
    a = b
  END SUBROUTINE rdi

  SUBROUTINE wr_open(name, v)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(OUT) :: v
    OPEN(NEWUNIT=v, FILE=TRIM(dir)//'/'//name//'.out', ACCESS='stream', FORM='unformatted', STATUS='replace')
  END SUBROUTINE wr_open

  SUBROUTINE wr0(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:), INTENT(IN) :: a
    INTEGER :: v
    CALL wr_open(name, v); WRITE(v) a; CLOSE(v)
  END SUBROUTINE wr0

  SUBROUTINE wr2(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:), INTENT(IN) :: a
    INTEGER :: v
    CALL wr_open(name, v); WRITE(v) a; CLOSE(v)
  END SUBROUTINE wr2

  SUBROUTINE wr3(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: a
    INTEGER :: v
    CALL wr_open(name, v); WRITE(v) a; CLOSE(v)
  END SUBROUTINE wr3

  SUBROUTINE wr4(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:,:), INTENT(IN) :: a
    INTEGER :: v
    CALL wr_open(name, v); WRITE(v) a; CLOSE(v)
  END SUBROUTINE wr4

END PROGRAM sfx_fortran_driver
