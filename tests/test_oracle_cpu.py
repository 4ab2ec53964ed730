"""The oracle's C restatement (oracle/c/sfx_cpu.c) against the numpy one.

Both restate smb_cross_prod_gauss_sft_par (subroutine_ns_with_velocity.F90:514-630)
on the reference grid N = 2*m_max_pad - 1; the numpy restatement is pinned by
tests/test_oracle_*.py (long-double direct DFT, closed forms, known answers).
The C one is the CPU baseline of bench.py, so it must compute the same thing.
Its FFT variant runs on the smooth grid N = 3*2^k >= 3M - 2, on which the
bilinear product has the same coefficients up to rounding.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import cpu, smb
from sfemans_amd.mesh import rectangle_p2
from sfemans_amd.synth import fourier_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def built():
    if not cpu.available():
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    assert cpu.available()


@pytest.mark.parametrize("nx,ny,M", [(3, 2, 1), (2, 2, 2), (2, 3, 3), (4, 4, 8), (5, 3, 13), (6, 4, 64),
                                     (2, 2, 100), (3, 3, 128)])
@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("fft", [False, True])
def test_c_restatement_matches_numpy(nx, ny, M, threads, fft):
    m = rectangle_p2(nx, ny)
    u = fourier_field(m.np, 6, M, 3, 4.0)
    lm = np.arange(M)
    a = cpu.smb_cross_prod_gauss(m.jj, m.rr, lm, u, threads, fft=fft)
    b = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, lm, u)
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(a - b)) / scale <= 1e-13
    # mode-0 sine columns are zero
    assert np.all(a[:, 1::2, 0] == 0.0)


def test_c_restatement_list_mode():
    """Non-contiguous modes: the curl uses list_mode values, the transform slots."""
    m = rectangle_p2(3, 3)
    lm = np.array([0, 2, 5, 7])
    u = fourier_field(m.np, 6, lm.size, 4, 2.0)
    b = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, lm, u)
    for fft in (False, True):
        a = cpu.smb_cross_prod_gauss(m.jj, m.rr, lm, u, 2, fft=fft)
        assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= 1e-13


def test_c_restatement_rejects_bad_shapes():
    m = rectangle_p2(2, 2)
    u = fourier_field(m.np, 6, 4, 1, 2.0)
    with pytest.raises(ValueError):
        cpu.smb_cross_prod_gauss(m.jj, m.rr, np.arange(3), u)


def test_fft_grid_sizes():
    assert [cpu.fft_grid(M) for M in (1, 2, 3, 8, 13, 64, 100, 128)] == [3, 6, 12, 24, 48, 192, 384, 384]
