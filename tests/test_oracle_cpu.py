"""The oracle's C restatement (oracle/c/sfx_cpu.c) against the numpy one.

Both restate smb_cross_prod_gauss_sft_par (subroutine_ns_with_velocity.F90:514-630)
on the reference grid N = 2*m_max_pad - 1; the numpy restatement is pinned by
tests/test_oracle_*.py (long-double direct DFT, closed forms, known answers).
The C one is the CPU baseline of bench.py, so it must compute the same thing.
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


@pytest.mark.parametrize("nx,ny,M", [(3, 2, 1), (2, 2, 2), (2, 3, 3), (4, 4, 8), (5, 3, 13), (6, 4, 64)])
@pytest.mark.parametrize("threads", [1, 3])
def test_c_restatement_matches_numpy(nx, ny, M, threads):
    m = rectangle_p2(nx, ny)
    u = fourier_field(m.np, 6, M, 3, 4.0)
    lm = np.arange(M)
    a = cpu.smb_cross_prod_gauss(m.jj, m.rr, lm, u, threads)
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
    a = cpu.smb_cross_prod_gauss(m.jj, m.rr, lm, u, 2)
    b = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, lm, u)
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) <= 1e-13


def test_c_restatement_rejects_bad_shapes():
    m = rectangle_p2(2, 2)
    u = fourier_field(m.np, 6, 4, 1, 2.0)
    with pytest.raises(ValueError):
        cpu.smb_cross_prod_gauss(m.jj, m.rr, np.arange(3), u)
