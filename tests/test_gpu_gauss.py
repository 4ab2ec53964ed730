"""Fused Gauss-point kernels on the GPU through the C-ABI vs the oracle.

Tolerance: relative max-norm error <= 1e-12 (FP64).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import smb  # noqa: E402
from sfemans_amd import nonlinear as NL  # noqa: E402
from sfemans_amd.fft_parallele import fortran_tensor, fortran_empty  # noqa: E402
from sfemans_amd.mesh import rectangle_p2  # noqa: E402
from sfemans_amd.synth import fourier_field  # noqa: E402

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 1), (3, 2, 3), (5, 4, 8), (4, 7, 13), (6, 5, 16),
                                      (3, 4, 32), (4, 4, 64), (2, 3, 100), (2, 2, 128),
                                      # point tiles (16 <= MP <= 64) with a partial last tile,
                                      # and a single square (2 triangles, 14 Gauss points)
                                      (3, 3, 64), (1, 1, 64), (1, 1, 32), (1, 1, 16), (5, 3, 40)])
def test_cross_curl(nx, ny, M):
    print(f"case cross_curl {nx}x{ny} M={M}", flush=True)
    m = rectangle_p2(nx, ny)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    u = fourier_field(m.np, 6, M, 11, 2.0)
    out = np.zeros((m.n_gauss, 6, M), order="F")
    NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, u, out)
    assert rel(out, smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)) <= TOL


def test_cross_curl_precession_selected_modes_device_path():
    m = rectangle_p2(6, 9)
    mesh = NL.Mesh.from_p2(m)
    modes = np.array([0, 1, 3, 5, 7])
    u = fourier_field(m.np, 6, 5, 12, 2.0, modes)
    du = fortran_tensor(u)
    dout = fortran_empty((m.n_gauss, 6, 5))
    NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, du, dout, True, 0.3, 0.25)
    torch.cuda.synchronize()
    exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u, True, 0.3, 0.25)
    assert rel(dout.cpu().numpy(), exp) <= TOL


@pytest.mark.parametrize("M", [3, 8, 64, 128])
def test_curlh_cross_b(M):
    m = rectangle_p2(5, 6)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    H, B = fourier_field(m.np, 6, M, 21, 2.0), fourier_field(m.np, 6, M, 22, 2.0)
    out = np.zeros((m.n_gauss, 6, M), order="F")
    NL.smb_CurlH_cross_B_gauss_sft_par(None, mesh, modes, H, B, out)
    assert rel(out, smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B)) <= TOL


@pytest.mark.parametrize("M", [1, 4, 8, 32, 64, 100, 128, 256])
def test_ugradc(M):
    m = rectangle_p2(4, 5)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    u, c = fourier_field(m.np, 6, M, 31, 2.0), fourier_field(m.np, 2, M, 32, 2.0)
    hc = np.random.default_rng(3).uniform(0.5, 2.0, m.me)
    out = np.zeros((m.n_gauss, 2, M), order="F")
    NL.smb_ugradc_gauss_fft_par(None, mesh, modes, hc, u, c, out)
    assert rel(out, smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, u, c)) <= TOL


@pytest.mark.parametrize("M", [3, 8, 64, 128])
def test_mhd_rotv(M):
    m = rectangle_p2(5, 4)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    u, H, B = (fourier_field(m.np, 6, M, s, 4.0) for s in (41, 42, 43))
    out = np.zeros((m.n_gauss, 6, M), order="F")
    NL.nl_rotv_mhd(mesh, modes, u, H, B, 0.7, out)
    exp = (smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)
           - 0.7 * smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B))
    assert rel(out, exp) <= TOL


def test_mesh_validation_errors():
    from sfemans_amd._lib import SfxError
    m = rectangle_p2(2, 2)
    jj = m.jj.copy(order="F")
    jj[0, 0] = m.np + 5
    with pytest.raises(SfxError):
        NL.Mesh(jj, m.rr)
    rr = m.rr.copy(order="F")
    rr[:, m.jj[3, 0] - 1] += 0.01      # a curved edge is accepted
    NL.Mesh(m.jj, rr).close()
    rr = m.rr.copy(order="F")
    rr[1, m.jj[:, 0] - 1] = 0.0        # the first element flattened onto a line: det J = 0 everywhere
    with pytest.raises(SfxError) as e:
        NL.Mesh(m.jj, rr)
    assert e.value.code == 1


def test_full_size_mhd_sampled():
    """224x448 mesh (200,704 triangles), m_max = 64: sampled elements vs the oracle."""
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = np.arange(M)
    print("case full-size mhd", flush=True)
    u, H, B = (fourier_field(m.np, 6, M, 0x5F3E + k, 4.0) for k in (2, 3, 4))
    du, dH, dB = fortran_tensor(u), fortran_tensor(H), fortran_tensor(B)
    out = fortran_empty((m.n_gauss, 6, M))
    NL.nl_rotv_mhd(mesh, modes, du, dH, dB, 1.0, out)
    torch.cuda.synchronize()
    sel = np.random.default_rng(0).choice(m.me, 300, replace=False)
    jj_s = m.jj[:, sel]
    exp = (smb.smb_cross_prod_gauss_sft_par(jj_s, m.rr, modes, u)
           - smb.smb_CurlH_cross_B_gauss_sft_par(jj_s, m.rr, modes, H, B))
    gidx = torch.as_tensor((sel[:, None] * 7 + np.arange(7)[None, :]).ravel(), device="cuda")
    got = out[gidx].cpu().numpy()
    assert rel(got, exp) <= TOL


def test_rhs_3x3_and_qs00():
    from oracle import assembly
    m = rectangle_p2(7, 9)
    mesh = NL.Mesh.from_p2(m)
    rng = np.random.default_rng(5)
    f = np.asfortranarray(rng.standard_normal((m.n_gauss, 6)))
    v1 = np.zeros((m.np, 3), order="F"); v2 = np.zeros((m.np, 3), order="F")
    NL.rhs_3x3(mesh, 1, f, v1, v2)
    e1, e2 = assembly.rhs_3x3(m.jj, m.rr, f)
    assert rel(v1, e1) <= 1e-13 and rel(v2, e2) <= 1e-13
    hc = rng.uniform(0.5, 2, m.me)
    ff, fg = rng.standard_normal(m.np), rng.standard_normal(m.n_gauss)
    v = np.zeros(m.np)
    NL.qs_00_gauss(mesh, hc, ff, fg, v)
    assert rel(v, assembly.qs_00_gauss(m.jj, m.rr, hc, ff, fg)) <= 1e-13


def test_qs00_device_many_columns():
    """Device path with ncols independent right-hand sides, each vs the oracle."""
    from oracle import assembly
    m = rectangle_p2(9, 6)
    mesh = NL.Mesh.from_p2(m)
    rng = np.random.default_rng(6)
    nc = 7
    hc = rng.uniform(0.5, 2, m.me)
    ff = np.asfortranarray(rng.standard_normal((m.np, nc)))
    fg = np.asfortranarray(rng.standard_normal((m.n_gauss, nc)))
    v = fortran_empty((m.np, nc))
    NL.qs_00_gauss(mesh, torch.as_tensor(hc, device="cuda"), fortran_tensor(ff), fortran_tensor(fg), v)
    torch.cuda.synchronize()
    got = v.cpu().numpy()
    for c in range(nc):
        exp = assembly.qs_00_gauss(m.jj, m.rr, hc, ff[:, c], fg[:, c])
        assert rel(got[:, c], exp) <= 1e-13, c


def test_rhs_ns_3x3_device():
    """Full NS right-hand side (rhs_ns_gauss_3x3 + rhs_3x3, fused) vs the
    oracle's two steps, relative error <= 1e-12, with every field and with
    absent fields (counted as zero)."""
    from oracle import assembly
    m = rectangle_p2(9, 7)
    mesh = NL.Mesh.from_p2(m)
    modes = [0, 1, 2, 4, 5, 9, 11]
    M = len(modes)
    print(f"case rhs_ns_3x3 {m.me} triangles modes={modes}", flush=True)
    rng = np.random.default_rng(9)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    rotv, V1m, ff = F(m.n_gauss, 6, M), F(m.np, 6, M), F(m.np, 6, M)
    P, Pinc = F(m.np, 2, M), F(m.np, 2, M)
    v1, v2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.rhs_ns_3x3(mesh, modes, fortran_tensor(rotv), fortran_tensor(V1m), fortran_tensor(ff),
                  fortran_tensor(P), fortran_tensor(Pinc), v1, v2)
    torch.cuda.synchronize()
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, np.asarray(modes), V1m, ff, P, Pinc, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL
    # no time-derivative term and no pressure increment
    NL.rhs_ns_3x3(mesh, modes, fortran_tensor(rotv), None, fortran_tensor(ff), fortran_tensor(P), None, v1, v2)
    torch.cuda.synchronize()
    z6, z2 = np.zeros_like(V1m), np.zeros_like(P)
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, np.asarray(modes), z6, ff, P, z2, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL


def test_rhs_3x3_device_all_modes():
    """Device path over many modes (element share + gather) vs the oracle."""
    from oracle import assembly
    m = rectangle_p2(8, 5)
    mesh = NL.Mesh.from_p2(m)
    M = 37
    f = np.asfortranarray(np.random.default_rng(7).standard_normal((m.n_gauss, 6, M)))
    v1, v2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.rhs_3x3_all_modes(mesh, fortran_tensor(f), v1, v2, -2.0)
    torch.cuda.synchronize()
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, f)
    assert rel(v1.cpu().numpy(), -2.0 * e1) <= 1e-13 and rel(v2.cpu().numpy(), -2.0 * e2) <= 1e-13


@pytest.mark.parametrize("nx,ny,modes", [(9, 7, [0, 1, 2, 4, 5, 9, 11]), (6, 5, list(range(40)))])
def test_rhs_ns_3x3_with_tensor_device(nx, ny, modes):
    """rhs_3x3's opt_tensor (assembling_rhs.F90:56-121) in the fused NS
    right-hand side vs the oracle, relative error <= 1e-12."""
    from oracle import assembly
    m = rectangle_p2(nx, ny)
    mesh = NL.Mesh.from_p2(m)
    M = len(modes)
    print(f"case rhs_ns_3x3+tensor {m.me} triangles M={M}", flush=True)
    rng = np.random.default_rng(10)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    rotv, V1m, ff = F(m.n_gauss, 6, M), F(m.np, 6, M), F(m.np, 6, M)
    P, Pinc, T = F(m.np, 2, M), F(m.np, 2, M), F(3, m.n_gauss, 6, M)
    v1, v2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.rhs_ns_3x3(mesh, modes, fortran_tensor(rotv), fortran_tensor(V1m), fortran_tensor(ff),
                  fortran_tensor(P), fortran_tensor(Pinc), v1, v2, opt_tensor=fortran_tensor(T))
    torch.cuda.synchronize()
    lm = np.asarray(modes)
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, V1m, ff, P, Pinc, rotv)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g, lm, T)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL
    # the tensor alone
    NL.rhs_ns_3x3(mesh, modes, None, None, None, None, None, v1, v2, opt_tensor=fortran_tensor(T))
    torch.cuda.synchronize()
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, np.zeros((m.n_gauss, 6, M)), lm, T)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL


@pytest.mark.parametrize("M", [3, 16])
def test_nl_rhs_mhd_all_modes(M):
    from oracle import assembly
    m = rectangle_p2(6, 5)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    u, H, B = (fourier_field(m.np, 6, M, s, 4.0) for s in (51, 52, 53))
    work = fortran_empty((m.n_gauss, 6, M))
    v1, v2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.nl_rhs_mhd(mesh, modes, fortran_tensor(u), fortran_tensor(H), fortran_tensor(B), 0.5, work, v1, v2)
    torch.cuda.synchronize()
    rotv = (smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u)
            - 0.5 * smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H, B))
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, -rotv)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL


@pytest.mark.parametrize("M", [5, 32])
def test_tangent_matches_bilinear_form_and_central_difference(M):
    m = rectangle_p2(6, 7)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    U = fourier_field(m.np, 6, M, 71, 2.0)
    D = fourier_field(m.np, 6, M, 72, 2.0)
    out = np.zeros((m.n_gauss, 6, M), order="F")
    NL.nl_tangent(mesh, modes, U, D, out)
    exp = (smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, U, D)
           + smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, D, U))
    assert rel(out, exp) <= TOL
    # central difference of the forward kernel: exact for a bilinear map (eps = 1),
    # rounding-limited at eps = 1e-6
    fp = np.zeros_like(out); fm = np.zeros_like(out)
    for eps, tol in ((1.0, 1e-12), (1e-6, 1e-8)):
        NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, np.asfortranarray(U + eps * D), fp)
        NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, np.asfortranarray(U - eps * D), fm)
        assert rel((fp - fm) / (2 * eps), out) <= tol


@pytest.mark.parametrize("curved", [False, True])
def test_rhs_3x3_all_options_one_mode(curved):
    """rhs_3x3 with opt_tensor and opt_tensor_scaln_bdy (assembling_rhs.F90:5-189),
    one mode through the host entry, vs the oracle; and the device boundary
    term for every mode added after the fused NS right-hand side."""
    from oracle import assembly
    from sfemans_amd.mesh import boundary_sides, sector_p2
    m = sector_p2(4, 6) if curved else rectangle_p2(5, 4)
    mesh = NL.Mesh.from_p2(m, boundary=True)
    jjs = boundary_sides(m.jj)
    mes = jjs.shape[1]
    rng = np.random.default_rng(12)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    mode = 3
    f, T, tb = F(m.n_gauss, 6), F(3, m.n_gauss, 6), F(3 * mes, 6)
    v1 = np.zeros((m.np, 3), order="F"); v2 = np.zeros((m.np, 3), order="F")
    NL.rhs_3x3(mesh, mode, f, v1, v2, opt_tensor=T, opt_tensor_scaln_bdy=tb)
    a1, a2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, f[:, :, None], np.array([mode]), T[..., None])
    b1, b2 = assembly.rhs_3x3_bdy(jjs, m.rr, tb[:, :, None])
    assert rel(v1, (a1 + b1)[:, :, 0]) <= TOL and rel(v2, (a2 + b2)[:, :, 0]) <= TOL
    # the boundary term alone and the tensor alone
    NL.rhs_3x3(mesh, mode, np.zeros_like(f), v1, v2, opt_tensor_scaln_bdy=tb)
    assert rel(v1, b1[:, :, 0]) <= TOL and rel(v2, b2[:, :, 0]) <= TOL
    NL.rhs_3x3(mesh, mode, f, v1, v2, opt_tensor=T)
    assert rel(v1, a1[:, :, 0]) <= TOL and rel(v2, a2[:, :, 0]) <= TOL
    # all modes on the device: the NS right-hand side with the tensor, then the boundary term
    modes = [0, 1, 2, 5, 8]
    M = len(modes)
    rotv, T5, tb5 = F(m.n_gauss, 6, M), F(3, m.n_gauss, 6, M), F(3 * mes, 6, M)
    d1, d2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.rhs_ns_3x3(mesh, modes, fortran_tensor(rotv), None, None, None, None, d1, d2, opt_tensor=fortran_tensor(T5))
    NL.rhs_3x3_bdy(mesh, fortran_tensor(tb5), d1, d2)
    torch.cuda.synchronize()
    a1, a2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, -rotv, np.asarray(modes), T5)
    b1, b2 = assembly.rhs_3x3_bdy(jjs, m.rr, tb5)
    assert rel(d1.cpu().numpy(), a1 + b1) <= TOL and rel(d2.cpu().numpy(), a2 + b2) <= TOL


def test_rhs_3x3_boundary_needs_sides():
    from sfemans_amd._lib import SfxError
    m = rectangle_p2(2, 2)
    mesh = NL.Mesh.from_p2(m)
    z = np.zeros((m.n_gauss, 6), order="F")
    v = np.zeros((m.np, 3), order="F")
    with pytest.raises(SfxError) as e:
        NL.rhs_3x3(mesh, 0, z, v, v.copy(order="F"), opt_tensor_scaln_bdy=np.zeros((24, 6), order="F"))
    assert e.value.code == 1


@pytest.mark.parametrize("modes,tensor,imposed", [([0, 1, 2, 3], False, True), ([2, 0, 5], True, True),
                                                  ([1, 3], True, False), (list(range(16)), False, True)])
def test_rhs_ns_3x3_penalty(modes, tensor, imposed):
    """rhs_ns_gauss_3x3 with if_ns_penalty and if_impose_vel_in_solids
    (computing_rhs_gauss.F90:103-141), assembled by rhs_3x3, vs the oracle:
    relative error <= 1e-12."""
    from oracle import assembly
    m = rectangle_p2(6, 5)
    mesh = NL.Mesh.from_p2(m)
    M = len(modes)
    N = 2 * (3 * M // 2) - 1
    print(f"case rhs_ns_3x3 penalty modes={modes} tensor={tensor} imposed={imposed}", flush=True)
    rng = np.random.default_rng(15)
    F = lambda *s: np.asfortranarray(rng.standard_normal(s))
    rotv, V1m, ff, P, Pinc = F(m.n_gauss, 6, M), F(m.np, 6, M), F(m.np, 6, M), F(m.np, 2, M), F(m.np, 2, M)
    T = F(3, m.n_gauss, 6, M) if tensor else None
    penal = np.asfortranarray(rng.uniform(0.0, 1.0, (N, m.n_gauss)))
    imp = F(m.np, 6) if imposed else None
    d = lambda x: None if x is None else fortran_tensor(x)
    v1, v2 = fortran_empty((m.np, 3, M)), fortran_empty((m.np, 3, M))
    NL.rhs_ns_3x3(mesh, modes, d(rotv), d(V1m), d(ff), d(P), d(Pinc), v1, v2, opt_tensor=d(T),
                  penal_grid=d(penal), imposed_vel=d(imp))
    torch.cuda.synchronize()
    lm = np.asarray(modes)
    g = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, V1m, ff, P, Pinc, rotv, penal_grid=penal, imposed_vel=imp)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, g, lm, T)
    assert rel(v1.cpu().numpy(), e1) <= TOL and rel(v2.cpu().numpy(), e2) <= TOL


@pytest.mark.parametrize("kind,M,prec", [("rect", 1, False), ("rect", 3, False), ("rect", 8, False),
                                         ("rect", 16, False), ("rect", 64, False), ("rect", 100, False),
                                         ("rect", 128, False), ("sector", 5, True)])
def test_fused_nodal_to_rhs(kind, M, prec):
    """nl_rhs_3x3 (no Gauss array) equals smb_cross_prod_gauss_sft_par then
    rhs_3x3 bit for bit, and the oracle to 1e-12."""
    from oracle import assembly
    from sfemans_amd.mesh import sector_p2
    print(f"case fused nodal->rhs {kind} M={M} precession={prec}", flush=True)
    m = rectangle_p2(4, 5) if kind == "rect" else sector_p2(3, 6)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M) if not prec else np.array([0, 1, 3, 5, 7])
    taux, ang = (0.3, 0.25) if prec else (0.0, 0.0)
    u = fourier_field(m.np, 6, len(modes), 14, 2.0)
    du = fortran_tensor(u)
    f145, f236 = fortran_empty((m.np, 3, len(modes))), fortran_empty((m.np, 3, len(modes)))
    NL.nl_rhs_3x3(mesh, modes, du, f145, f236, -1.0, None, prec, taux, ang)
    g = fortran_empty((m.n_gauss, 6, len(modes)))
    NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, du, g, prec, taux, ang)
    w145, w236 = fortran_empty((m.np, 3, len(modes))), fortran_empty((m.np, 3, len(modes)))
    NL.rhs_3x3_all_modes(mesh, g, w145, w236, -1.0)
    torch.cuda.synchronize()
    assert torch.equal(f145, w145) and torch.equal(f236, w236)
    exp_g = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, modes, u, prec, taux, ang)
    e1, e2 = assembly.rhs_3x3_all_modes(m.jj, m.rr, exp_g)
    assert rel(f145.cpu().numpy(), -e1) <= TOL and rel(f236.cpu().numpy(), -e2) <= TOL


@pytest.mark.parametrize("M,with_ff", [(1, True), (3, False), (16, True), (64, True), (100, False), (128, True)])
def test_fused_ugradc_to_qs_00(M, with_ff):
    """nl_qs_00 (no Gauss array) equals smb_ugradc_gauss_fft_par then
    qs_00_gauss bit for bit, and the oracle to 1e-12."""
    from oracle import assembly
    print(f"case fused u.grad c -> qs_00 M={M} ff={with_ff}", flush=True)
    m = rectangle_p2(4, 5)
    mesh = NL.Mesh.from_p2(m)
    modes = np.arange(M)
    rng = np.random.default_rng(18)
    u = fourier_field(m.np, 6, M, 19, 2.0)
    c = fourier_field(m.np, 2, M, 20, 2.0)
    hc = rng.uniform(0.5, 2.0, m.me)
    ff = np.asfortranarray(rng.standard_normal((m.np, 2, M))) if with_ff else None
    du, dc, dhc = fortran_tensor(u), fortran_tensor(c), torch.as_tensor(hc, device="cuda")
    dff = None if ff is None else fortran_tensor(ff)
    v = fortran_empty((m.np, 2, M))
    NL.nl_qs_00(mesh, modes, dhc, du, dc, v, dff, -1.0)
    g = fortran_empty((m.n_gauss, 2, M))
    NL.smb_ugradc_gauss_fft_par(None, mesh, modes, dhc, du, dc, g)
    w = fortran_empty((m.np, 2, M))
    flat = lambda t, n: torch.as_strided(t, (n, 2 * M), (1, n))
    NL.qs_00_gauss(mesh, dhc, None if dff is None else flat(dff, m.np), flat(g, m.n_gauss), flat(w, m.np))
    torch.cuda.synchronize()
    assert torch.equal(v, -w)
    eg = smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, u, c)
    ffz = np.zeros((m.np, 2, M)) if ff is None else ff
    exp = np.stack([assembly.qs_00_gauss(m.jj, m.rr, hc, ffz[:, k, mm], eg[:, k, mm])
                    for mm in range(M) for k in range(2)], axis=1).reshape(m.np, 2, M, order="F")
    assert rel(v.cpu().numpy(), -exp) <= TOL
