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


@pytest.mark.parametrize("mesh_kind,M", [("rect", 1), ("rect", 5), ("sector", 3)])
def test_rhs_3x3_boundary_term(mesh_kind, M):
    """opt_tensor_scaln_bdy (assembling_rhs.F90:129-185) added after the volume
    sums, vs the oracle, on a rectangle (sides on the axis skipped) and on a
    sector with curved sides."""
    from sfemans_amd.mesh import boundary_sides, sector_p2
    m = rectangle_p2(3, 4) if mesh_kind == "rect" else sector_p2(3, 5)
    jjs = boundary_sides(m.jj)
    rng = np.random.default_rng(11)
    tb = np.asfortranarray(rng.standard_normal((3 * jjs.shape[1], 6, M)))
    base1 = np.asfortranarray(rng.standard_normal((m.np, 3, M)))
    base2 = np.asfortranarray(rng.standard_normal((m.np, 3, M)))
    rc, arrs = emu_runner.call("emu_rhs_bdy", jjs, jjs.shape[1], m.rr, m.np, M, tb, base1.copy(order="F"),
                               base2.copy(order="F"), 2.0)
    assert rc == 0
    e1, e2 = assembly.rhs_3x3_bdy(jjs, m.rr, tb)
    assert rel(arrs[-2], base1 + 2.0 * e1) <= 1e-13 and rel(arrs[-1], base2 + 2.0 * e2) <= 1e-13
    touched = np.zeros(m.np, bool)
    touched[jjs.ravel() - 1] = True
    assert np.all(arrs[-2][~touched] == base1[~touched])


@pytest.mark.parametrize("modes,slot", [([0, 1, 2], 0), ([3, 0, 5, 1], 1), ([1, 2], -1)])
def test_rhs_ns_gauss_values_with_imposed_velocity(modes, slot):
    """The first step of rhs_ns_gauss_3x3's penalty branch: the Gauss values
    ft + fp_inc + fs - rotv_v, minus the imposed velocity on the mode-0 slot
    (computing_rhs_gauss.F90:54-121), vs the oracle."""
    m = rectangle_p2(3, 2)
    M = len(modes)
    lm = np.asarray(modes, dtype=np.int32)
    rotv, V1m, ff, P, Pinc = _ns_inputs(m, M, 13)
    imp = np.asfortranarray(np.random.default_rng(14).standard_normal((m.np, 6)))
    out = np.zeros((m.n_gauss, 6, M), order="F")
    rc, arrs = emu_runner.call("emu_rhs_ns_gauss", m.jj, m.rr, m.me, m.np, lm, M, rotv, V1m, ff, Pinc, imp, slot, out)
    assert rc == 0
    z2 = np.zeros_like(P)
    exp = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, V1m, ff, z2, Pinc, rotv)
    if slot >= 0:
        from oracle.gauss_p2 import gauss_tables
        ww = gauss_tables(m.jj, m.rr)[0]
        ig = np.einsum("enk,nl->elk", imp[(m.jj - 1).T], ww).reshape(m.n_gauss, 6)
        exp[:, :, slot] -= ig
    assert rel(arrs[-1], exp) <= 1e-13


def _fan_mesh(nt=12):
    """A vertex shared by nt triangles, in the half plane r > 0."""
    ang = np.linspace(0.0, 2 * np.pi, nt, endpoint=False)
    vr = np.concatenate([[2.0], 2.0 + np.cos(ang)])
    vz = np.concatenate([[0.0], np.sin(ang)])
    tri = np.array([[0, 1 + i, 1 + (i + 1) % nt] for i in range(nt)])
    from sfemans_amd.mesh import p2_from_p1
    return p2_from_p1(tri, vr, vz)


def test_rhs_3x3_many_entries():
    """A node with 12 (element, local node) entries, summed in element order."""
    m = _fan_mesh(12)
    M = 3
    f = np.asfortranarray(np.random.default_rng(4).standard_normal((m.n_gauss, 6, M)))
    v1 = np.zeros((m.np, 3, M), order="F"); v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_rhs", 0, m.jj, m.rr, m.me, m.np, M, f, v1, v2, None, None, None, None, -1.0)
    assert rc == 0
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, f)
    assert rel(arrs[3], -e1) <= 1e-13 and rel(arrs[4], -e2) <= 1e-13
