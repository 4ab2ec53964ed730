"""Pins for the oracle's restatement of rhs_ns_gauss_3x3 (CPU).

SOURCES/SOLVER/computing_rhs_gauss.F90:7-146 without the penalty options:
rhs = (ft + fp_inc + fs - rotv_v) + fp.  Quadratic nodal fields are
reproduced exactly by P2, so every term has a closed form at the Gauss
points: ft and fs are the fields' values, fp the pressure gradient in
(cos, sin) form, (-d_r P_c, -d_r P_s, -m P_s/r, m P_c/r, -d_z P_c, -d_z P_s).
"""
import numpy as np

from oracle import assembly
from sfemans_amd.mesh import rectangle_p2
from tests.test_oracle_gauss import eval_quad, gauss_coords, quad_field


def _zeros(m, M):
    return (np.zeros((m.np, 6, M)), np.zeros((m.np, 6, M)), np.zeros((m.np, 2, M)),
            np.zeros((m.np, 2, M)), np.zeros((m.n_gauss, 6, M)))


def test_only_rotv_gives_minus_rotv():
    m = rectangle_p2(3, 2)
    M = 3
    V1m, ff, P, Pi, R = _zeros(m, M)
    R[:] = np.random.default_rng(1).standard_normal(R.shape)
    out = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, np.arange(M), V1m, ff, P, Pi, R)
    assert np.array_equal(out, -R)


def test_velocity_and_source_terms_are_their_values():
    m = rectangle_p2(3, 4)
    M = 4
    _, _, P, Pi, R = _zeros(m, M)
    V1m, cv = quad_field(m.rr, M, 3)
    ff, cf = quad_field(m.rr, M, 4)
    rg, zg = gauss_coords(m)
    out = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, np.arange(M), V1m, ff, P, Pi, R)
    ex = eval_quad(cv, rg, zg)[0] + eval_quad(cf, rg, zg)[0]
    assert np.allclose(out, ex, atol=1e-12 * np.max(np.abs(ex)))


def test_pressure_terms_are_the_gradient_in_cos_sin_form():
    m = rectangle_p2(4, 3)
    M = 5
    modes = np.array([0, 1, 3, 4, 7])
    V1m, ff, _, _, R = _zeros(m, M)
    Q6, c6 = quad_field(m.rr, M, 5)
    P = np.asfortranarray(Q6[:, :2, :])                # (np, 2, M): columns cos, sin
    rg, zg = gauss_coords(m)
    v, vr, vz = eval_quad(c6[:2], rg, zg)               # (G, 2, M)
    mr = modes[None, :] / rg[:, None]
    ex = np.empty((m.n_gauss, 6, M))
    ex[:, 0] = -vr[:, 0]
    ex[:, 1] = -vr[:, 1]
    ex[:, 2] = -v[:, 1] * mr
    ex[:, 3] = v[:, 0] * mr
    ex[:, 4] = -vz[:, 0]
    ex[:, 5] = -vz[:, 1]
    for p_, pi_ in ((P, np.zeros_like(P)), (np.zeros_like(P), P)):
        out = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, p_, pi_, R)
        assert np.allclose(out, ex, atol=1e-11 * np.max(np.abs(ex)))


def test_linear_in_all_inputs():
    m = rectangle_p2(2, 3)
    M = 3
    rng = np.random.default_rng(6)
    a = [rng.standard_normal(x.shape) for x in _zeros(m, M)]
    b = [rng.standard_normal(x.shape) for x in _zeros(m, M)]
    lm = np.arange(M)
    fa = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, *a)
    fb = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, *b)
    fab = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, lm, *[x + 2.0 * y for x, y in zip(a, b)])
    assert np.allclose(fab, fa + 2.0 * fb, atol=1e-12 * np.max(np.abs(fab)))


def _quad_tensor(m, M, seed):
    """T(3, G, 6, M) of per-(row, column, mode) quadratics at the Gauss points,
    its value and r/z derivatives there, each (3, G, 6, M)."""
    rg, zg = gauss_coords(m)
    rng = np.random.default_rng(seed)
    vals = [eval_quad(rng.standard_normal((6, M, 6)), rg, zg) for _ in range(3)]
    T, Tr, Tz = (np.stack([v[i] for v in vals]) for i in range(3))
    return T, Tr, Tz, rg


def _minus_div(T, Tr, Tz, r, modes):
    """-div T in (cos, sin) form, divergence over the column index, in
    cylindrical coordinates:
      (div T)_r = d_r T_rr + (1/r) d_th T_rth + d_z T_rz + (T_rr - T_thth)/r
      (div T)_th = d_r T_thr + (1/r) d_th T_thth + d_z T_thz + (T_thr + T_rth)/r
      (div T)_z = d_r T_zr + (1/r) d_th T_zth + d_z T_zz + T_zr/r
    with (1/r) d_th (A cos + B sin) = (m B/r) cos - (m A/r) sin."""
    mr = modes[None, :] / r[:, None]
    ir = 1.0 / r[:, None]
    d = np.empty(T.shape[1:])
    for k in range(3):
        for s in range(2):               # s = 0: cos, 1: sin
            th = mr * T[k, :, 3] if s == 0 else -mr * T[k, :, 2]
            d[:, 2 * k + s] = Tr[k, :, s] + th + Tz[k, :, 4 + s]
    d[:, 0] += (T[0, :, 0] - T[1, :, 2]) * ir
    d[:, 1] += (T[0, :, 1] - T[1, :, 3]) * ir
    d[:, 2] += (T[1, :, 0] + T[0, :, 2]) * ir
    d[:, 3] += (T[1, :, 1] + T[0, :, 3]) * ir
    d[:, 4] += T[2, :, 0] * ir
    d[:, 5] += T[2, :, 1] * ir
    return -d


def test_tensor_term_is_the_weak_form_of_minus_the_divergence():
    """Integration by parts pins opt_tensor: with T quadratic, every integrand
    is a polynomial of degree <= 4, so the 7-point rule is exact and, at
    every node whose test function vanishes on the boundary r = 1, z = +-1
    (the r = 0 term carries the factor r), rhs_3x3 with opt_tensor = T
    equals rhs_3x3 of rhs_in = -div T."""
    m = rectangle_p2(4, 5)
    modes = np.array([0, 1, 2, 5])
    M = len(modes)
    T, Tr, Tz, rg = _quad_tensor(m, M, 8)
    zero = np.zeros((m.n_gauss, 6, M))
    a145, a236 = assembly.rhs_3x3_all_modes(m.jj, m.rr, zero, modes, T)
    b145, b236 = assembly.rhs_3x3_all_modes(m.jj, m.rr, _minus_div(T, Tr, Tz, rg, modes))
    inner = (m.rr[0] < 1 - 1e-12) & (np.abs(m.rr[1]) < 1 - 1e-12)
    scale = np.max(np.abs(b145[inner])) + np.max(np.abs(b236[inner]))
    assert np.max(np.abs(a145[inner] - b145[inner])) <= 1e-12 * scale
    assert np.max(np.abs(a236[inner] - b236[inner])) <= 1e-12 * scale
    # on the boundary the two differ by the flux term
    assert np.max(np.abs(a145[~inner] - b145[~inner])) > 1e-3 * scale


def test_tensor_term_adds_to_rhs_in():
    m = rectangle_p2(3, 2)
    modes = np.array([0, 3])
    rng = np.random.default_rng(9)
    f = rng.standard_normal((m.n_gauss, 6, 2))
    T = rng.standard_normal((3, m.n_gauss, 6, 2))
    a = assembly.rhs_3x3_all_modes(m.jj, m.rr, f, modes, T)
    b = assembly.rhs_3x3_all_modes(m.jj, m.rr, f)
    c = assembly.rhs_3x3_all_modes(m.jj, m.rr, np.zeros_like(f), modes, T)
    for x, y, z in zip(a, b, c):
        assert np.allclose(x, y + z, atol=1e-13 * np.max(np.abs(x)))


def _normals_at_side_points(m, jjs):
    """Outward unit normals of the rectangle's boundary sides at their three
    Gauss points, and the physical Gauss points (r, z), each (3, mes)."""
    w, d, p = assembly.element_1d_p2()
    rs = m.rr[:, jjs - 1]                                   # (2, 3, mes)
    rg = np.einsum("nm,nl->lm", rs[0], w)
    zg = np.einsum("nm,nl->lm", rs[1], w)
    nr = np.where(np.isclose(rg, 1.0), 1.0, np.where(np.isclose(rg, 0.0), -1.0, 0.0))
    nz = np.where(np.isclose(zg, 1.0) & (nr == 0), 1.0, np.where(np.isclose(zg, -1.0) & (nr == 0), -1.0, 0.0))
    return nr, nz, rg, zg


def test_boundary_term_closed_form_on_straight_sides():
    """Constant data c on every side: the end nodes of a side of length L at
    radius R receive c R L/6, its midpoint c R 2L/3 (the three-point rule is
    exact); sides on the axis r = 0 add nothing (points with ray < 1e-10)."""
    from sfemans_amd.mesh import boundary_sides
    m = rectangle_p2(2, 3)
    jjs = boundary_sides(m.jj)
    mes = jjs.shape[1]
    c = np.array([1.0, -2.0, 0.5, 3.0, -1.5, 0.25])
    T = np.tile(c[None, :, None], (3 * mes, 1, 1))
    v145, v236 = assembly.rhs_3x3_bdy(jjs, m.rr, T)
    # right side r = 1: three sides of length 2/3 along z
    right = np.nonzero(np.isclose(m.rr[0], 1.0))[0]
    vert = right[np.isclose((m.rr[1, right] + 1) * 1.5, np.round((m.rr[1, right] + 1) * 1.5))]
    inner_vert = vert[np.abs(m.rr[1, vert]) < 1 - 1e-12]
    L = 2.0 / 3.0
    # an inner vertex of the right side belongs to two sides: 2 * L/6
    assert np.allclose(v145[inner_vert, 0, 0], c[0] * 2 * L / 6, atol=1e-15)
    assert np.allclose(v236[inner_vert, 1, 0], c[2] * 2 * L / 6, atol=1e-15)
    mids = np.setdiff1d(right, vert)
    assert np.allclose(v145[mids, 2, 0], c[4] * 2 * L / 3, atol=1e-15)
    assert np.allclose(v236[mids, 2, 0], c[5] * 2 * L / 3, atol=1e-15)
    # nodes only on the axis get nothing
    axis = np.nonzero(np.isclose(m.rr[0], 0.0) & (np.abs(m.rr[1]) < 1 - 1e-12))[0]
    assert np.all(v145[axis] == 0.0) and np.all(v236[axis] == 0.0)


def test_boundary_term_is_the_flux_of_the_tensor():
    """Integration by parts at every node: rhs_3x3 with opt_tensor = T, minus
    rhs_3x3 of rhs_in = -div T, equals the boundary term with
    opt_tensor_scaln_bdy = T.n (the normal flux at the sides' Gauss points),
    exactly (every integrand is a polynomial the rules integrate exactly)."""
    from sfemans_amd.mesh import boundary_sides
    m = rectangle_p2(4, 5)
    modes = np.array([0, 1, 2, 5])
    M = len(modes)
    T, Tr, Tz, rg = _quad_tensor(m, M, 8)
    zero = np.zeros((m.n_gauss, 6, M))
    a145, a236 = assembly.rhs_3x3_all_modes(m.jj, m.rr, zero, modes, T)
    b145, b236 = assembly.rhs_3x3_all_modes(m.jj, m.rr, _minus_div(T, Tr, Tz, rg, modes))
    # the same quadratic tensor evaluated at the sides' Gauss points
    jjs = boundary_sides(m.jj)
    nr, nz, sr, sz = _normals_at_side_points(m, jjs)
    rng = np.random.default_rng(8)
    coefs = [rng.standard_normal((6, M, 6)) for _ in range(3)]
    Ts = np.stack([eval_quad(cf, sr.T.ravel(), sz.T.ravel())[0] for cf in coefs])   # (3, 3 mes, 6, M)
    tb = np.empty((3 * jjs.shape[1], 6, M))
    n_r, n_z = nr.T.ravel()[:, None], nz.T.ravel()[:, None]
    for k in range(3):
        for s in range(2):
            tb[:, 2 * k + s] = Ts[k, :, s] * n_r + Ts[k, :, 4 + s] * n_z
    c145, c236 = assembly.rhs_3x3_bdy(jjs, m.rr, tb)
    scale = np.max(np.abs(a145)) + np.max(np.abs(a236))
    assert np.max(np.abs(a145 - b145 - c145)) <= 1e-12 * scale
    assert np.max(np.abs(a236 - b236 - c236)) <= 1e-12 * scale
    assert np.max(np.abs(c145)) > 1e-3 * scale


def _ns_fields(m, M, seed):
    rng = np.random.default_rng(seed)
    F = lambda *s: rng.standard_normal(s)
    return F(m.np, 6, M), F(m.np, 6, M), F(m.np, 2, M), F(m.np, 2, M), F(m.n_gauss, 6, M)


def test_penalty_one_is_no_penalty():
    """penal = 1 everywhere: the product with the penalty is the identity on
    band-limited fields, except that it drops the sine columns of the mode-0
    slot (fft_parallel.F90:241, 296-299); fp is added after it."""
    m = rectangle_p2(3, 2)
    modes = np.array([0, 1, 2, 3])
    M = len(modes)
    V1m, ff, P, Pinc, rotv = _ns_fields(m, M, 21)
    N = 2 * (3 * M // 2) - 1
    z6, z2, zg = np.zeros_like(V1m), np.zeros_like(P), np.zeros_like(rotv)
    rest = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, z2, Pinc, rotv)
    rest[:, 1::2, 0] = 0.0
    fp = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, z6, z6, P, z2, zg)
    b = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, P, Pinc, rotv, penal_grid=np.ones((N, m.n_gauss)))
    assert np.max(np.abs((rest + fp) - b)) <= 1e-13 * np.max(np.abs(b))


def test_penalty_constant_scales_all_but_fp():
    """penal = c: every term but fp is multiplied by c (computing_rhs_gauss.F90:124-145)."""
    m = rectangle_p2(2, 3)
    modes = np.array([0, 2, 5])
    M = len(modes)
    V1m, ff, P, Pinc, rotv = _ns_fields(m, M, 22)
    N = 2 * (3 * M // 2) - 1
    z6, z2, zg = np.zeros_like(V1m), np.zeros_like(P), np.zeros_like(rotv)
    c = 0.37
    got = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, P, Pinc, rotv, penal_grid=np.full((N, m.n_gauss), c))
    rest = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, z2, Pinc, rotv)
    rest[:, 1::2, 0] = 0.0          # slot 0 is mode 0: its sine columns are dropped
    fp = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, z6, z6, P, z2, zg)
    assert np.max(np.abs(got - (c * rest + fp))) <= 1e-13 * np.max(np.abs(got))


def test_imposed_velocity_in_a_solid():
    """penal = 0 (a solid everywhere) with an imposed velocity linear in (r, z):
    the mode-0 slot is the imposed velocity's Gauss values (closed form) plus
    fp, the other slots fp alone (:105-121, :132-138)."""
    m = rectangle_p2(3, 3, r0=0.5)
    modes = np.array([1, 0, 4])
    M = len(modes)
    V1m, ff, P, Pinc, rotv = _ns_fields(m, M, 23)
    N = 2 * (3 * M // 2) - 1
    coef = np.random.default_rng(24).standard_normal((6, 3))
    imp = np.stack([coef[k, 0] * m.rr[0] + coef[k, 1] * m.rr[1] + coef[k, 2] for k in range(6)], axis=1)
    got = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, V1m, ff, P, Pinc, rotv,
                                    penal_grid=np.zeros((N, m.n_gauss)), imposed_vel=imp)
    z6, z2, zg = np.zeros_like(V1m), np.zeros_like(P), np.zeros_like(rotv)
    fp = assembly.rhs_ns_gauss_3x3(m.jj, m.rr, modes, z6, z6, P, z2, zg)
    rg, zgp = gauss_coords(m)
    ig = np.stack([coef[k, 0] * rg + coef[k, 1] * zgp + coef[k, 2] for k in range(6)], axis=1)   # (G, 6)
    exp = fp.copy()
    exp[:, :, 1] += ig
    assert np.max(np.abs(got - exp)) <= 1e-13 * np.max(np.abs(exp))
