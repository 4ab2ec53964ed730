"""sfemans_amd - MI355X-native pseudo-spectral nonlinear terms for SFEMaNS.

The compute path is ``libsfx.so`` (hand-written HIP kernels for gfx950 behind
the C-ABI of ``include/sfx.h``).  This package is the Python host mirror of
the reference interfaces for that path:

* ``fft_parallele`` - ``FFT_PAR_*`` products (fft_parallel.F90);
* ``mesh``          - synthetic P2 meshes with the reference numbering;
* ``synth``         - seeded synthetic Fourier fields.
"""
from . import mesh, synth  # noqa: F401

__version__ = "0.1.0"
