"""Meshes read from the reference's regression-suite .FEM files (CPU only).

The fixtures under tests/golden/meshes are data files of the reference's
regression suite (REGRESSION_SUITE/MESHES): Mesh_10_form.FEM (TEST2, TEST9,
TEST45: Navier-Stokes on domain 1) and Mesh_20_form.FEM (TEST25).  Each file
holds a P1 section, the one the reference reads, followed by a P2 section
written by the mesh generator.  The P2 section is an independent statement of
the P2 grid: the mesh built here from the P1 section must give, cell by cell
and slot by slot, the same node coordinates.

The kernel bodies then run in the CPU emulator on these unstructured meshes,
whose cells come in both orientations and whose vertex triplets are not in
ascending node order, against the oracle.
"""
import os

import numpy as np
import pytest

from oracle import assembly, smb
from sfemans_amd.mesh import load_fem_p2, read_fem_p1
from sfemans_amd.synth import fourier_field
from tests import emu_runner
from tests.test_emu_adjoint import emu_adjoint, emu_tangent
from tests.test_emu_gauss import run as emu_gauss

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def fem(name):
    return os.path.join(MESHES, name)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def read_p2_section(path):
    """The second (P2) section of a .FEM file: (jj (me, 6) 0-based, i_d, rr (np, 2))."""
    with open(path) as f:
        lines = f.read().split("\n")
    npt, _, me, _, mes = (int(x) for x in lines[0].split()[:5])
    i = 1 + me + mes + npt
    np2, nw, me2, nws, mes2 = (int(x) for x in lines[i].split()[:5])
    assert (nw, me2, nws, mes2) == (6, me, 3, mes)
    cells = np.array([lines[i + 1 + k].split()[:10] for k in range(me)], dtype=np.int64)
    rr = np.array([lines[i + 1 + me + mes + k].replace("D", "E").split()[:2] for k in range(np2)],
                  dtype=np.float64)
    return cells[:, :6] - 1, cells[:, 9], rr


@pytest.mark.parametrize("name", ["Mesh_10_form.FEM", "Mesh_20_form.FEM"])
def test_p1_section_counts(name):
    jj, i_d, rr = read_fem_p1(fem(name))
    with open(fem(name)) as f:
        npt, nw, me, nws, mes = (int(x) for x in f.readline().split()[:5])
    assert jj.shape == (3, me) and rr.shape == (2, npt) and i_d.shape == (me,)
    assert jj.min() == 1 and jj.max() == npt
    assert set(np.unique(i_d)) == {1, 2}


@pytest.mark.parametrize("name", ["Mesh_10_form.FEM", "Mesh_20_form.FEM"])
def test_p2_grid_matches_the_files_p2_section(name):
    """All domains: every cell's six node coordinates, slot by slot, equal the
    generator's P2 section (printed to 16 digits); node counts agree."""
    m = load_fem_p2(fem(name))
    jj2, i_d2, rr2 = read_p2_section(fem(name))
    assert m.np == rr2.shape[0]
    jj1, i_d, _ = read_fem_p1(fem(name))
    assert np.array_equal(i_d, i_d2)
    # cells of the loaded mesh: domain 1 then domain 2, file order within each
    order = np.concatenate([np.nonzero(i_d == d)[0] for d in (1, 2)])
    file_cells = jj2[order]
    got = m.rr[:, m.jj - 1]                          # (2, 6, me)
    # the generator's vertex order may differ: match vertices by coordinates,
    # then each edge node by the vertex opposite to it
    for e in range(m.me):
        fr = rr2[file_cells[e]]                      # (6, 2)
        gv = got[:, :3, e].T
        perm = [int(np.argmin(np.sum((fr[:3] - v) ** 2, axis=1))) for v in gv]
        assert sorted(perm) == [0, 1, 2]
        assert np.allclose(gv, fr[perm], rtol=0, atol=1e-15)
        assert np.allclose(got[:, 3:, e].T, fr[[3 + p for p in perm]], rtol=0, atol=1e-15)


def test_domain_selection_and_local_orientation():
    """Domain 1 of Mesh_10 (the Navier-Stokes domain of TEST2): 130 cells,
    nodes numbered by first appearance, the vertex triplet of each cell in
    file-index order, so some cells are not in ascending node order and both
    orientations occur."""
    m = load_fem_p2(fem("Mesh_10_form.FEM"), [1])
    assert m.me == 130
    jj = m.jj - 1
    seen = np.zeros(m.np, bool)
    nxt = 0
    for e in range(m.me):                            # first-appearance numbering
        for v in jj[:3, e]:
            if not seen[v]:
                assert v == nxt
                seen[v] = True
                nxt += 1
    assert not np.all(np.diff(jj[:3], axis=0) > 0)
    r, z = m.rr[0, jj[:3]], m.rr[1, jj[:3]]
    det = (r[1] - r[0]) * (z[2] - z[0]) - (r[2] - r[0]) * (z[1] - z[0])
    assert np.all(det != 0) and np.any(det > 0) and np.any(det < 0)
    # the P1 area of domain 1 equals the sum of the P2 cells' areas
    assert abs(0.5 * np.sum(np.abs(det)) - 0.5) < 1e-12


def test_edge_midpoint_from_the_lower_local_vertex():
    """Midpoints are rr(a) + (rr(b) - rr(a))/2 with a the lower local index
    in the first cell that meets the edge (bit-exact)."""
    m = load_fem_p2(fem("Mesh_20_form.FEM"), [1, 2])
    jj = m.jj - 1
    done = set()
    for e in range(m.me):
        for k in range(3):
            node = jj[3 + k, e]
            if node in done:
                continue
            done.add(node)
            a, b = [jj[j, e] for j in range(3) if j != k]
            assert np.array_equal(m.rr[:, node], m.rr[:, a] + (m.rr[:, b] - m.rr[:, a]) / 2.0)


# ---- kernel bodies in the emulator on the unstructured meshes -------------

MESH_CASES = [("Mesh_10_form.FEM", [1], 3), ("Mesh_20_form.FEM", [1, 2], 8)]


@pytest.mark.parametrize("name,doms,M", MESH_CASES)
def test_cross_curl_on_fem_mesh(name, doms, M):
    m = load_fem_p2(fem(name), doms)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 111, 2.0)
    got = emu_gauss(0, m, u, None, None, None, modes, 6)
    assert rel(got, smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)) <= 1e-12


@pytest.mark.parametrize("name,doms,M", MESH_CASES)
def test_curlh_cross_b_on_fem_mesh(name, doms, M):
    m = load_fem_p2(fem(name), doms)
    modes = np.arange(M)
    H, B = fourier_field(m.np, 6, M, 112, 2.0), fourier_field(m.np, 6, M, 113, 2.0)
    got = emu_gauss(1, m, H, B, None, None, modes, 6)
    assert rel(got, smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B)) <= 1e-12


def test_ugradc_on_fem_mesh():
    m = load_fem_p2(fem("Mesh_10_form.FEM"), [1, 2])
    M = 5
    modes = np.arange(M)
    u, c = fourier_field(m.np, 6, M, 114, 2.0), fourier_field(m.np, 2, M, 115, 2.0)
    hc = np.random.default_rng(116).uniform(0.5, 2.0, m.me)
    got = emu_gauss(2, m, u, c, None, hc, modes, 2)
    assert rel(got, smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, u, c)) <= 1e-12


@pytest.mark.parametrize("name,doms,M", MESH_CASES)
def test_rhs_ns_3x3_on_fem_mesh(name, doms, M):
    m = load_fem_p2(fem(name), doms)
    lm = np.arange(M, dtype=np.int32)
    rng = np.random.default_rng(117)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    rotv, V1m, ff, P, Pinc = F(m.n_gauss, 6, M), F(m.np, 6, M), F(m.np, 6, M), F(m.np, 2, M), F(m.np, 2, M)
    v1 = np.zeros((m.np, 3, M), order="F"); v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_rhs_ns", m.jj, m.rr, m.me, m.np, lm, M, rotv, V1m, ff, P, Pinc, None, v1, v2)
    assert rc == 0
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, V1m, ff, P, Pinc, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g)
    assert rel(arrs[-2], e1) <= 1e-12 and rel(arrs[-1], e2) <= 1e-12


def test_adjoint_dot_product_on_fem_mesh():
    m = load_fem_p2(fem("Mesh_10_form.FEM"), [1])
    M = 3
    modes = np.arange(M)
    U, D = fourier_field(m.np, 6, M, 118, 2.0), fourier_field(m.np, 6, M, 119, 2.0)
    Y = np.asfortranarray(np.random.default_rng(120).standard_normal((m.n_gauss, 6, M)))
    TD = emu_tangent(m, U, D, modes)
    TtY = emu_adjoint(m, U, Y, modes)
    assert abs(np.sum(TD * Y) - np.sum(D * TtY)) <= 1e-12 * np.linalg.norm(TD) * np.linalg.norm(Y)


def test_tangent_matches_the_oracle_linearization_on_fem_mesh():
    m = load_fem_p2(fem("Mesh_10_form.FEM"), [2])
    M = 4
    modes = np.arange(M)
    U, D = fourier_field(m.np, 6, M, 121, 2.0), fourier_field(m.np, 6, M, 122, 2.0)
    got = emu_tangent(m, U, D, modes)
    exp = (smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, U, D)
           + smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, D, U))
    assert rel(got, exp) <= 1e-12
