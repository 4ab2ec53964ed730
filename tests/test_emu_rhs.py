"""RHS assembly kernels on the CPU emulator vs the oracle (CPU only)."""
import numpy as np
import pytest

from oracle import assembly
from sfemans_amd.mesh import rectangle_p2
from tests import emu_runner


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("nx,ny,M", [(1, 1, 1), (3, 2, 4), (5, 7, 3)])
def test_rhs_3x3(nx, ny, M):
    m = rectangle_p2(nx, ny)
    f = np.asfortranarray(np.random.default_rng(1).standard_normal((m.n_gauss, 6, M)))
    v1 = np.zeros((m.np, 3, M), order="F"); v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_rhs", 0, m.jj, m.rr, m.me, m.np, M, f, v1, v2, None, None, None, None, -1.0)
    assert rc == 0
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, f)
    assert rel(arrs[3], -e1) <= 1e-13 and rel(arrs[4], -e2) <= 1e-13


@pytest.mark.parametrize("nx,ny,nc", [(2, 2, 1), (4, 3, 5)])
def test_qs_00_gauss(nx, ny, nc):
    m = rectangle_p2(nx, ny)
    rng = np.random.default_rng(2)
    hc = rng.uniform(0.5, 2, m.me)
    ff = np.asfortranarray(rng.standard_normal((m.np, nc)))
    fg = np.asfortranarray(rng.standard_normal((m.n_gauss, nc)))
    v = np.zeros((m.np, nc), order="F")
    rc, arrs = emu_runner.call("emu_rhs", 1, m.jj, m.rr, m.me, m.np, nc, None, None, None, hc, ff, fg, v, 1.0)
    assert rc == 0
    exp = np.stack([assembly.qs_00_gauss(m.jj, m.rr, hc, ff[:, c], fg[:, c]) for c in range(nc)], axis=1)
    assert rel(arrs[-1], exp) <= 1e-13


def _ns_inputs(m, M, seed, with_all=True):
    rng = np.random.default_rng(seed)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    rotv = F(m.n_gauss, 6, M)
    V1m = F(m.np, 6, M) if with_all else None
    ff = F(m.np, 6, M)
    P = F(m.np, 2, M)
    Pinc = F(m.np, 2, M) if with_all else None
    return rotv, V1m, ff, P, Pinc


@pytest.mark.parametrize("nx,ny,modes,with_all", [(1, 1, [0], True), (3, 2, [0, 1, 2, 3], True),
                                                  (4, 3, [0, 2, 5], False), (2, 5, list(range(9)), True)])
def test_rhs_ns_3x3(nx, ny, modes, with_all):
    """Fused rhs_ns_gauss_3x3 + rhs_3x3 (every mode) vs the oracle's two steps."""
    m = rectangle_p2(nx, ny)
    M = len(modes)
    lm = np.asarray(modes, dtype=np.int32)
    rotv, V1m, ff, P, Pinc = _ns_inputs(m, M, 3, with_all)
    v1 = np.zeros((m.np, 3, M), order="F"); v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_rhs_ns", m.jj, m.rr, m.me, m.np, lm, M, rotv, V1m, ff, P, Pinc, None, v1, v2)
    assert rc == 0
    z6, z2 = np.zeros((m.np, 6, M)), np.zeros((m.np, 2, M))
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, z6 if V1m is None else V1m, ff, P,
                                  z2 if Pinc is None else Pinc, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g)
    assert rel(arrs[-2], e1) <= 1e-12 and rel(arrs[-1], e2) <= 1e-12


@pytest.mark.parametrize("nx,ny,modes,with_all", [(1, 1, [0], False), (3, 2, [0, 1, 2, 3], True),
                                                  (2, 5, [0, 2, 3, 7, 8, 9, 11, 12, 13, 14, 15, 16, 17, 18,
                                                          19, 20, 23], True)])
def test_rhs_ns_3x3_with_tensor(nx, ny, modes, with_all):
    """opt_tensor of rhs_3x3 (assembling_rhs.F90:56-121) in the fused kernel."""
    m = rectangle_p2(nx, ny)
    M = len(modes)
    lm = np.asarray(modes, dtype=np.int32)
    rotv, V1m, ff, P, Pinc = _ns_inputs(m, M, 4, with_all)
    T = np.asfortranarray(np.random.default_rng(5).standard_normal((3, m.n_gauss, 6, M)))
    v1 = np.zeros((m.np, 3, M), order="F"); v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_rhs_ns", m.jj, m.rr, m.me, m.np, lm, M, rotv, V1m, ff, P, Pinc, T, v1, v2)
    assert rc == 0
    z6, z2 = np.zeros((m.np, 6, M)), np.zeros((m.np, 2, M))
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, z6 if V1m is None else V1m, ff, P,
                                  z2 if Pinc is None else Pinc, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g, lm, T)
    assert rel(arrs[-2], e1) <= 1e-12 and rel(arrs[-1], e2) <= 1e-12
