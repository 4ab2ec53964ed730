"""CPU oracle for the pseudo-spectral nonlinear-term path.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference algorithm, used as the checker by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The product path (``sfemans_amd``) never imports it.

Modules and the reference code they restate (paths relative to the
SFEMaNS source tree):

* ``fft_par``    - ``FFT_PAR_REAL``, ``FFT_PAR_CROSS_PROD_DCL``,
  ``FFT_PAR_DOT_PROD_DCL``, ``FFT_PAR_PROD_DCL``, ``FFT_SCALAR_VECT_DCL``
  (pb=1 and pb=2), ``FFT_TENSOR_DCL`` (SOURCES/POST_PROCESSING/fft_parallel.F90),
  with ``nb_procs = 1``; two DFT engines: numpy's pocketfft and a long-double
  direct sum on the reference grid ``N = 2*m_max_pad-1``.
* ``gauss_p2``   - ``element_2d_p2`` (SOURCES/GAUSS_POINTS/GP_2D_p2.F90:9-88)
  and the volume part of ``create_gauss_points_2d``
  (SOURCES/GAUSS_POINTS/gauss_points_2d.F90:102-125).
* ``smb``        - the Gauss-point preparers ``smb_cross_prod_gauss_sft_par``,
  ``smb_CurlH_cross_B_gauss_sft_par``
  (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514-727) and
  ``smb_ugradc_gauss_fft_par``
  (SOURCES/SIMULATION/subroutine_temperature_with_T.F90:247-350).
* ``assembly``   - ``rhs_3x3`` without the tensor options
  (SOURCES/SOLVER/assembling_rhs.F90:5-189) and ``qs_00_gauss``
  (SOURCES/SOLVER/fem_rhs_axi.F90:940-987) into dense local vectors.
* ``linearized`` - tangent and adjoint of the bilinear maps (no reference
  counterpart; pinned by brute-force transposition on tiny inputs).
* ``cpu``        - ctypes wrapper of ``c/sfx_cpu.c``, a compiled C restatement
  (OpenMP over element blocks) of ``smb_cross_prod_gauss_sft_par`` on the
  reference grid, checked against ``smb`` and timed as the CPU baseline.

Parity pins (see ``tests/test_oracle_*.py``): the long-double direct DFT,
single-mode closed forms, the TEST1 known answer
(REGRESSION_SUITE/TEST1/LOCAL_SOURCES/condlim.F90:140-211), quadratic fields
for the interpolation and curl, the exactness of the 7-point rule for the
scatter, and algebraic invariants.  Every function here has at least one
such check.
"""
