"""dot_product_SF, the mass solve and the adjoint of the projected tangent in
dot_product_SF, on the GPU through the C-ABI.

* dot_product_SF (SOURCES/SOLVER/tn_axi.F90:9-37) vs the oracle, relative
  error <= 1e-12, host and device paths, straight and curved meshes.
* M^{-1}: the mass solve's residual ||M x - b|| <= 1e-12 ||b||.
* A = M^{-1} R T_U and its adjoint in <., .>_SF vs the oracle (which forms
  M^{-1} by a sparse LU), relative error <= 1e-11 (the mass solve stops at
  rtol 1e-14 of the preconditioned residual).
* The dot-product test in <., .>_SF at the 50k-triangle, m_max = 32 size:
  |<A x, y>_SF - <x, A* y>_SF| <= 1e-12 ||A x||_SF ||y||_SF, base flow in
  modes 0-4 with amplitude 0.1 beyond mode 0.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import inner as OI  # noqa: E402
from oracle import linearized  # noqa: E402
from sfemans_amd import inner as SI  # noqa: E402
from sfemans_amd import nonlinear as NL  # noqa: E402
from sfemans_amd.fft_parallele import fortran_empty, fortran_tensor  # noqa: E402
from sfemans_amd.mesh import rectangle_p2, sector_p2  # noqa: E402
from sfemans_amd.synth import fourier_field  # noqa: E402


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


@pytest.mark.parametrize("kind,M,nbf", [("rect", 3, 6), ("sector", 5, 2), ("rect", 64, 6)])
def test_dot_product_sf(kind, M, nbf):
    m = rectangle_p2(7, 9) if kind == "rect" else sector_p2(4, 7)
    mesh = NL.Mesh.from_p2(m)
    modes = np.array([0] + list(range(2, M + 1)))
    v, w = fourier_field(m.np, nbf, M, 51, 2.0), fourier_field(m.np, nbf, M, 52, 2.0)
    exp = OI.dot_product_sf(m.jj, m.rr, modes, v, w)
    got_h = SI.dot_product_SF(None, mesh, modes, v, w)
    got_d = SI.dot_product_SF(None, mesh, modes, fortran_tensor(v), fortran_tensor(w))
    assert abs(got_h - exp) <= 1e-12 * abs(exp) and got_d == got_h


def test_mass_solve_residual():
    m = sector_p2(12, 20)
    mesh = NL.Mesh.from_p2(m)
    rng = np.random.default_rng(53)
    b = np.asfortranarray(rng.standard_normal((m.np, 6, 4)))
    db, x, y = fortran_tensor(b), fortran_empty((m.np, 6, 4)), fortran_empty((m.np, 6, 4))
    it = SI.mass_solve(mesh, db, x)
    SI.mass_apply(mesh, x, y)
    torch.cuda.synchronize()
    assert 0 < it < 1000
    assert np.linalg.norm(y.cpu().numpy() - b) <= 1e-12 * np.linalg.norm(b)
    # the mass matrix itself against the oracle's sparse assembly
    Mm = OI.mass_matrix(m.jj, m.rr)
    exp = np.stack([Mm @ x.cpu().numpy()[:, c, k] for k in range(4) for c in range(6)], axis=1)
    got = y.cpu().numpy().reshape(m.np, -1, order="F")
    assert rel(got, exp) <= 1e-13


@pytest.mark.parametrize("kind", ["rect", "sector"])
def test_projected_tangent_and_adjoint_vs_oracle(kind):
    m = rectangle_p2(2, 3) if kind == "rect" else sector_p2(2, 3)
    mesh = NL.Mesh.from_p2(m)
    modes = np.array([0, 1, 3])
    M = len(modes)
    U, D = fourier_field(m.np, 6, M, 54, 2.0), fourier_field(m.np, 6, M, 55, 2.0)
    Y = fourier_field(m.np, 6, M, 56, 2.0)
    out = fortran_empty((m.np, 6, M))
    SI.nl_tangent_proj(mesh, modes, fortran_tensor(U), fortran_tensor(D), out)
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), OI.projected_tangent(m.jj, m.rr, modes, U, D)) <= 1e-11
    SI.nl_adjoint_sf(mesh, modes, fortran_tensor(U), fortran_tensor(Y), out)
    torch.cuda.synchronize()
    exp = OI.projected_adjoint(m.jj, m.rr, modes, U, Y,
                               lambda g: linearized.adjoint_bruteforce(m.jj, m.rr, modes, U, g))
    assert rel(out.cpu().numpy(), exp) <= 1e-11


def test_dot_product_test_in_sf_inner_product_at_size():
    """50,176 triangles (112 x 224), m_max = 32: <A x, y>_SF = <x, A* y>_SF."""
    m = rectangle_p2(112, 224)
    mesh = NL.Mesh.from_p2(m)
    M = 32
    modes = list(range(M))
    print("case SF dot-product test 112x224 m_max=32", flush=True)
    U = fourier_field(m.np, 6, M, 0x5F3E + 4, 2.0)
    U[:, :, 5:] = 0.0
    U[:, :, 1:5] *= 0.1
    x, y = fourier_field(m.np, 6, M, 57, 2.0), fourier_field(m.np, 6, M, 58, 2.0)
    dU, dx, dy = fortran_tensor(U), fortran_tensor(x), fortran_tensor(y)
    Ax, Asy = fortran_empty((m.np, 6, M)), fortran_empty((m.np, 6, M))
    work = torch.empty(int(SI.lib().sfx_nl_proj_work(mesh.handle, M)), dtype=torch.float64, device="cuda")
    SI.nl_tangent_proj(mesh, modes, dU, dx, Ax, work=work)
    SI.nl_adjoint_sf(mesh, modes, dU, dy, Asy, work=work)
    lhs = SI.dot_product_SF(None, mesh, modes, Ax, dy)
    rhs = SI.dot_product_SF(None, mesh, modes, dx, Asy)
    nAx = SI.dot_product_SF(None, mesh, modes, Ax, Ax) ** 0.5
    ny = SI.dot_product_SF(None, mesh, modes, dy, dy) ** 0.5
    assert abs(lhs - rhs) <= 1e-12 * nAx * ny, (lhs, rhs, nAx * ny)
