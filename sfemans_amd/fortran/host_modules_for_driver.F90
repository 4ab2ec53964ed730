!===
! The two host modules the Fortran shim reads, as the test driver provides
! them: a program built inside SFEMaNS compiles the shim against SFEMaNS's
! own def_type_mesh (SOURCES/MESH/def_type_mesh.F90) and input_data
! (SOURCES/INITIALISATION/my_data_module.F90:443), which carry these fields
! among many others.  Only the fields the shim and the driver's user
! functions read are declared here, with the same names, types and ranks.
!===
MODULE def_type_mesh
  IMPLICIT NONE
  PUBLIC

  !===Vector layout of a PETSc matrix/vector block (def_type_mesh.F90:18-24)
  TYPE petsc_csr_LA
     INTEGER, DIMENSION(:), POINTER :: ia, ja
     INTEGER, DIMENSION(:, :), POINTER :: loc_to_glob
     INTEGER :: kmax
     INTEGER, DIMENSION(:), POINTER :: np
     INTEGER, DIMENSION(:), POINTER :: dom_np
  END TYPE petsc_csr_LA

  !===Quadrature tables (def_type_mesh.F90:37-53)
  TYPE gauss_type
     INTEGER :: k_d, n_w, l_G, n_ws, l_Gs
     REAL(KIND = 8), DIMENSION(:, :), POINTER :: ww
     REAL(KIND = 8), DIMENSION(:, :), POINTER :: wws
     REAL(KIND = 8), DIMENSION(:, :, :, :), POINTER :: dw
     REAL(KIND = 8), DIMENSION(:, :), POINTER :: rj
     REAL(KIND = 8), DIMENSION(:, :), POINTER :: rjs
  END TYPE gauss_type

  !===Mesh (def_type_mesh.F90:74-112)
  TYPE mesh_type
     INTEGER, POINTER, DIMENSION(:, :) :: jj, jjs
     REAL(KIND = 8), POINTER, DIMENSION(:, :) :: rr
     INTEGER :: dom_me, dom_np, dom_mes
     INTEGER :: me, mes, np
     TYPE(gauss_type) :: gauss
  END TYPE mesh_type
END MODULE def_type_mesh

MODULE input_data
  IMPLICIT NONE
  PRIVATE
  !===The data-file inputs the nonlinear terms read (my_data_module.F90:20-238)
  TYPE, PUBLIC :: my_data
     REAL(KIND = 8) :: LES_coeff1, LES_coeff2, LES_coeff3, LES_coeff4
     REAL(KIND = 8) :: LES_coeff1_mom
     REAL(KIND = 8) :: taux_precession, angle_s_pi
     REAL(KIND = 8) :: level_set_comp_factor
     REAL(KIND = 8) :: h_min_distance
  END TYPE my_data
  TYPE(my_data), PUBLIC :: inputs
END MODULE input_data
