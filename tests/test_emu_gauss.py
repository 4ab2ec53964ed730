"""Fused Gauss-point kernels on the CPU emulator vs the oracle (CPU only)."""
import numpy as np
import pytest

from oracle import smb
from sfemans_amd.mesh import rectangle_p2
from sfemans_amd.synth import fourier_field
from tests import emu_runner

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def run(variant, mesh, U, U2, U3, hc, modes, ncout, cl=0.0, om=(0.0, 0.0, 0.0)):
    M = len(modes)
    out = np.zeros((mesh.n_gauss, ncout, M), order="F")
    lm = np.asarray(modes, dtype=np.int32)
    nc2 = U2.shape[1] if U2 is not None else 0
    rc, arrs = emu_runner.call("emu_gauss", variant, mesh.jj, mesh.rr, mesh.me, mesh.np, U, U.shape[1],
                               U2, nc2, U3, hc, lm, M, out, ncout, cl, *om)
    assert rc == 0
    return arrs[-1]


CASES = [(2, 3, 1), (3, 2, 3), (2, 5, 8), (3, 3, 13), (2, 3, 20), (2, 2, 64), (1, 2, 100)]


@pytest.mark.parametrize("nx,ny,M", CASES)
def test_cross_curl(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 11, 2.0)
    got = run(0, m, u, None, None, None, modes, 6)
    exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)
    assert rel(got, exp) <= TOL


def test_cross_curl_precession_and_selected_modes():
    m = rectangle_p2(3, 4)
    modes = np.array([0, 1, 3, 5, 7])
    u = fourier_field(m.np, 6, 5, 12, 2.0, modes)
    taux, ang = 0.3, 0.25
    om = smb.precession_omega(modes, taux, ang)
    got = run(0, m, u, None, None, None, modes, 6, 0.0,
              (2 * om[0, 1], 2 * om[3, 1], 2 * om[4, 0]))
    exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u, True, taux, ang)
    assert rel(got, exp) <= TOL


@pytest.mark.parametrize("nx,ny,M", CASES[:4])
def test_curlh_cross_b(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    H = fourier_field(m.np, 6, M, 21, 2.0)
    B = fourier_field(m.np, 6, M, 22, 2.0)
    got = run(1, m, H, B, None, None, modes, 6)
    exp = smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B)
    assert rel(got, exp) <= TOL


@pytest.mark.parametrize("nx,ny,M", CASES[:4] + [(2, 2, 32)])
def test_ugradc(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 31, 2.0)
    c = fourier_field(m.np, 2, M, 32, 2.0)
    hc = np.random.default_rng(3).uniform(0.5, 2.0, m.me)
    got = run(2, m, u, c, None, hc, modes, 2)
    exp = smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, u, c)
    assert rel(got, exp) <= TOL


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 3), (3, 2, 8), (2, 2, 64)])
def test_mhd_rotv(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 41, 4.0)
    H = fourier_field(m.np, 6, M, 42, 4.0)
    B = fourier_field(m.np, 6, M, 43, 4.0)
    cl = 0.7
    got = run(3, m, u, H, B, None, modes, 6, cl)
    exp = (smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)
           - cl * smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B))
    assert rel(got, exp) <= TOL


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 5), (3, 2, 32)])
def test_tangent_is_symmetrised_bilinear_form(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    U = fourier_field(m.np, 6, M, 61, 2.0)
    D = fourier_field(m.np, 6, M, 62, 2.0)
    got = run(4, m, U, D, None, None, modes, 6)
    exp = (smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, U, D)
           + smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, D, U))
    assert rel(got, exp) <= TOL


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 1), (3, 2, 3), (2, 5, 8), (2, 2, 64), (1, 2, 100)])
def test_fused_nodal_to_rhs_equals_rhs_3x3_of_the_gauss_output(nx, ny, M):
    """The fused form (no Gauss array) gives rhs_3x3 of smb_cross_prod_gauss_sft_par's
    output bit for bit, with the same summation order; and the oracle to 1e-12."""
    from oracle import assembly
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 13, 2.0)
    lm = np.asarray(modes, dtype=np.int32)
    v1 = np.zeros((m.np, 3, M), order="F")
    v2 = np.zeros_like(v1)
    rc, arrs = emu_runner.call("emu_nl_rhs", m.jj, m.rr, m.me, m.np, u, lm, M, v1, v2, -1.0, 0.0, 0.0, 0.0)
    assert rc == 0
    f145, f236 = arrs[-2], arrs[-1]
    g = run(0, m, u, None, None, None, modes, 6)
    w1 = np.zeros((m.np, 3, M), order="F")
    w2 = np.zeros_like(w1)
    rc, arrs = emu_runner.call("emu_rhs", 0, m.jj, m.rr, m.me, m.np, M, g, w1, w2, None, None, None, None, -1.0)
    assert rc == 0
    assert np.array_equal(f145, arrs[3]) and np.array_equal(f236, arrs[4])
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u))
    assert rel(f145, -e1) <= TOL and rel(f236, -e2) <= TOL


@pytest.mark.parametrize("nx,ny,M,with_ff", [(2, 3, 1, True), (3, 2, 3, False), (2, 5, 8, True), (2, 2, 64, True),
                                             (1, 2, 100, False)])
def test_fused_ugradc_to_qs_00_equals_qs_00_of_the_gauss_output(nx, ny, M, with_ff):
    """u.grad c fused to qs_00_gauss (no Gauss array) gives qs_00_gauss of
    smb_ugradc_gauss_fft_par's output bit for bit; and the oracle to 1e-12."""
    from oracle import assembly
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    rng = np.random.default_rng(17)
    u = fourier_field(m.np, 6, M, 15, 2.0)
    c = fourier_field(m.np, 2, M, 16, 2.0)
    hc = rng.uniform(0.5, 2.0, m.me)
    ff = np.asfortranarray(rng.standard_normal((m.np, 2, M))) if with_ff else None
    lm = np.asarray(modes, dtype=np.int32)
    v = np.zeros((m.np, 2, M), order="F")
    rc, arrs = emu_runner.call("emu_nl_qs00", m.jj, m.rr, m.me, m.np, u, c, hc, ff, lm, M, v, -1.0)
    assert rc == 0
    fused = arrs[-1]
    g = run(2, m, u, c, None, hc, modes, 2)
    w = np.zeros((m.np, 2 * M), order="F")
    ffc = None if ff is None else np.asfortranarray(ff.reshape(m.np, 2 * M, order="F"))
    rc, arrs = emu_runner.call("emu_rhs", 1, m.jj, m.rr, m.me, m.np, 2 * M, None, None, None, hc, ffc,
                               np.asfortranarray(g.reshape(m.n_gauss, 2 * M, order="F")), w, -1.0)
    assert rc == 0
    assert np.array_equal(fused.reshape(m.np, 2 * M, order="F"), arrs[-1])
    eg = smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, u, c)
    ffz = np.zeros((m.np, 2, M)) if ff is None else ff
    exp = np.stack([assembly.qs_00_gauss(m.jj, m.rr, hc, ffz[:, k, mm], eg[:, k, mm])
                    for mm in range(M) for k in range(2)], axis=1).reshape(m.np, 2, M, order="F")
    assert rel(fused, -exp) <= TOL
