"""ctypes wrapper of the C restatement (``oracle/c/sfx_cpu.c``).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``: a compiled restatement
of ``smb_cross_prod_gauss_sft_par`` (without precession) on the reference
grid, used to cross-check the numpy restatement and as the CPU timing of
``bench.py``.  Build it with ``make -C oracle/c``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["LIB_PATH", "available", "smb_cross_prod_gauss", "fft_grid"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "libsfx_cpu.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sfx_cpu_smb_cross_prod_gauss.argtypes = [vp, vp, C.c_int64, C.c_int64, vp, C.c_int32, vp, vp, C.c_int32]
        L.sfx_cpu_smb_cross_prod_gauss.restype = C.c_int
        L.sfx_cpu_smb_cross_prod_gauss_fft.argtypes = L.sfx_cpu_smb_cross_prod_gauss.argtypes
        L.sfx_cpu_smb_cross_prod_gauss_fft.restype = C.c_int
        _lib = L
    return _lib


def fft_grid(M: int) -> int:
    """Grid of the FFT variant: the smallest N = 3*2^k with N >= 3M - 2."""
    n = 3
    while n < 3 * M - 2:
        n *= 2
    return n


def smb_cross_prod_gauss(jj, rr, list_mode, V_in, nthreads: int = 1, out=None, fft: bool = False):
    """``P_M[(curl V) x V]`` at the Gauss points, (7 me, 6, M), on ``nthreads`` cores.

    ``fft=False``: direct sums on the reference grid N = 2*m_max_pad - 1;
    ``fft=True``: the mixed-radix FFT on N = :func:`fft_grid` (M).
    """
    jj = np.asfortranarray(jj, dtype=np.int32)
    rr = np.asfortranarray(rr, dtype=np.float64)
    V = np.asfortranarray(V_in, dtype=np.float64)
    lm = np.ascontiguousarray(list_mode, dtype=np.int32)
    me, npn, M = jj.shape[1], rr.shape[1], V.shape[2]
    if V.shape != (npn, 6, M) or lm.size != M:
        raise ValueError("V_in must be (np, 6, M) with M = len(list_mode)")
    if out is None:
        out = np.empty((7 * me, 6, M), order="F")
    L = _load()
    fn = L.sfx_cpu_smb_cross_prod_gauss_fft if fft else L.sfx_cpu_smb_cross_prod_gauss
    rc = fn(jj.ctypes.data, rr.ctypes.data, me, npn, lm.ctypes.data, M,
                                               V.ctypes.data, out.ctypes.data, nthreads)
    if rc != 0:
        raise ValueError("sfx_cpu_smb_cross_prod_gauss: bad arguments")
    return out
