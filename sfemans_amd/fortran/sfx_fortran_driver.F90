!===
! Exercises the fft_parallele shim the way the reference's callers do: a
! cross product, a dot product, a division, a square root, the velocity
! maximum and (max, min), the real planes (on the reference's own grid), and
! the smb_cross_prod_gauss_sft_par
! body on a small P2 mesh read from a file.  Inputs are read from and outputs written to
! raw little-endian float64 files named on the command line, so that the
! calling test can check them.
!
! usage: sfx_fortran_driver <dir>
!   <dir>/dims.txt : np nb_field m_max me npm
!   <dir>/v1.bin, v2.bin (np,6,m_max); jj.bin (6,me) int32; rr.bin (2,npm);
!   <dir>/vn.bin (npm,6,m_max) nodal velocity; den.bin (np,2,m_max) a positive scalar
! writes <dir>/cross.bin (np,6,m), dot.bin (np,2,m), div.bin (np,2,m) = v1(:,1:2)/den,
!   sqrt.bin (np,2,m) = SQRT(MAX(0,MIN(1,den))), maxvel.bin (1) = max |v1|, smb.bin (7*me,6,m),
!   maxmin.bin (2) = FFT_MAX_MIN_VEL_DCL of v1, real.bin (2*m_max+1,3,np) = FFT_PAR_REAL of v1
!   with opt_nb_plane = 2*m_max+1, scal.bin (np,2,m) = FFT_PAR_SCAL_FUNCT of den with funct(x) = x*x
!===
PROGRAM sfx_fortran_driver
  USE, INTRINSIC :: ISO_C_BINDING
  USE fft_parallele
  IMPLICIT NONE
  CHARACTER(LEN=1024) :: dir
  INTEGER :: np, nbf, m_max, me, npm, u, i
  REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: v1, v2, vc, vd, vn, vs, den, vq, vr
  REAL(KIND=8), DIMENSION(1,1,1) :: vmax
  REAL(KIND=8), DIMENSION(2,1,1) :: vmm
  REAL(KIND=8), DIMENSION(:,:,:), ALLOCATABLE :: vreal
  INTEGER, DIMENSION(:,:), ALLOCATABLE :: jj
  INTEGER(C_INT32_T), DIMENSION(:,:), ALLOCATABLE :: jj32
  REAL(KIND=8), DIMENSION(:,:), ALLOCATABLE :: rr
  INTEGER, DIMENSION(:), ALLOCATABLE :: list_mode
  REAL(KIND=8), DIMENSION(3) :: temps
  TYPE(C_PTR) :: mesh
  INTEGER :: comm

  CALL GET_COMMAND_ARGUMENT(1, dir)
  OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/dims.txt', STATUS='old')
  READ(u,*) np, nbf, m_max, me, npm
  CLOSE(u)
  ALLOCATE(v1(np,6,m_max), v2(np,6,m_max), vc(np,6,m_max), vd(np,2,m_max))
  CALL read_real(TRIM(dir)//'/v1.bin', v1)
  CALL read_real(TRIM(dir)//'/v2.bin', v2)
  comm = 0
  temps = 0.d0
  CALL FFT_PAR_CROSS_PROD_DCL(comm, v1, v2, vc, 1, np, 3*m_max/2, temps)
  CALL FFT_PAR_DOT_PROD_DCL(comm, v1, v2, vd, 1, np, 3*m_max/2)
  CALL write_real(TRIM(dir)//'/cross.bin', vc)
  CALL write_real(TRIM(dir)//'/dot.bin', vd)

  ALLOCATE(den(np,2,m_max), vq(np,2,m_max), vr(np,2,m_max))
  CALL read_real(TRIM(dir)//'/den.bin', den)
  CALL FFT_PAR_DIV_DCL(comm, v1(:,1:2,:), den, vq, 1, np, 3*m_max/2)
  vr = den
  CALL FFT_SQUAREROOT_SCALAR(comm, vr, 1, np, 3*m_max/2)
  CALL FFT_MAX_VEL_DCL(comm, v1, vmax(1,1,1), 1, np, 3*m_max/2)
  CALL write_real(TRIM(dir)//'/div.bin', vq)
  CALL write_real(TRIM(dir)//'/sqrt.bin', vr)
  CALL write_real(TRIM(dir)//'/maxvel.bin', vmax)
  CALL FFT_MAX_MIN_VEL_DCL(comm, v1, vmm(:,1,1), 1, np, 3*m_max/2)
  CALL write_real(TRIM(dir)//'/maxmin.bin', vmm)
  CALL FFT_PAR_REAL(comm, v1, vreal, 2*m_max+1)
  CALL write_real(TRIM(dir)//'/real.bin', vreal)
  vr = den
  CALL FFT_PAR_SCAL_FUNCT(comm, vr, square, 1, np, 3*m_max/2)
  CALL write_real(TRIM(dir)//'/scal.bin', vr)

  ALLOCATE(jj32(6,me), jj(6,me), rr(2,npm), vn(npm,6,m_max), vs(7*me,6,m_max), list_mode(m_max))
  OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/jj.bin', ACCESS='stream', FORM='unformatted', STATUS='old')
  READ(u) jj32
  CLOSE(u)
  jj = jj32
  OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/rr.bin', ACCESS='stream', FORM='unformatted', STATUS='old')
  READ(u) rr
  CLOSE(u)
  CALL read_real(TRIM(dir)//'/vn.bin', vn)
  list_mode = [(i - 1, i = 1, m_max)]
  mesh = sfx_mesh_register(jj, rr)
  CALL sfx_smb_cross_prod_gauss(mesh, list_mode, vn, vs, .FALSE., 0.d0, 0.d0)
  CALL sfx_mesh_release(mesh)
  CALL write_real(TRIM(dir)//'/smb.bin', vs)
  WRITE(*,'(A)') 'sfx_fortran_driver: ok'

CONTAINS

  FUNCTION square(x) RESULT(vv)
    REAL(KIND=8) :: x
    REAL(KIND=8) :: vv
    vv = x*x
  END FUNCTION square

  SUBROUTINE read_real(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(OUT) :: a
    INTEGER :: v
    OPEN(NEWUNIT=v, FILE=name, ACCESS='stream', FORM='unformatted', STATUS='old')
    READ(v) a
    CLOSE(v)
  END SUBROUTINE read_real

  SUBROUTINE write_real(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(KIND=8), DIMENSION(:,:,:), INTENT(IN) :: a
    INTEGER :: v
    OPEN(NEWUNIT=v, FILE=name, ACCESS='stream', FORM='unformatted', STATUS='replace')
    WRITE(v) a
    CLOSE(v)
  END SUBROUTINE write_real

END PROGRAM sfx_fortran_driver
