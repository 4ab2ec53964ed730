"""Pins for the oracle's non-polynomial FFT_PAR_* products (CPU).

These products are evaluated on exactly N = 2*m_max_pad-1 angles, so every
pin below fixes the whole chain (packing, c2r, pointwise operation, r2c,
scaling, truncation) through a closed form that does not use the oracle:

* division and pb = 2 scalar-vector: the quotient of q1*q2 by q2 is q1 when
  q1*q2 is band-limited (its coefficients by the product-to-sum rule);
* square root: the root of q**2 is q for 0 <= q <= 1; clamping at 0 and 1;
* Heaviside: regul at its knots, and a level set whose smoothed step is
  0 or 1 at every grid angle, whose coefficients are Dirichlet sums;
* tension: a constant second field gives a tensor linear in the first;
* velocity maximum: constant fields and single modes;
* no-overshoot products: inactive bounds give the band-limited product or
  quotient, active bounds a constant factor; the [0, 1] clamp is the
  identity inside and a constant outside;
* entropy viscosity: with c_in_real = coeff - s on the grid (s band-limited,
  its grid values by explicit sums) the vector part is the product s V; with
  0 <= c <= 1 the scalar part is c - c*c; it clamps at zero above 1; the
  momentum-gradient form likewise gives s V_a; the diffusion products are
  scalar-vector products; the compression form reduces to the entropy form
  without compression and has a closed form for constant fields; the full
  entropy viscosity has closed forms with either bound of its MIN active
  (the element and angle maxima through direct cosine sums); the momentum
  form for fields constant in angle, per element, for l_G = 7 and 3.

The pocketfft engine is also checked against the long-double direct sum.
"""
import numpy as np
import pytest

from oracle import fft_par as F


def times_alpha_beta_cos(C, alpha, beta):
    """Coefficients (np, 2, M) of (alpha + beta cos t) * q, q with coefficients C
    (cos column 0, sin column 1), by cos kt cos t = (cos(k+1)t + cos(k-1)t)/2 and
    sin kt cos t = (sin(k+1)t + sin(k-1)t)/2; C must vanish in mode M-1."""
    npt, _, M = C.shape
    assert np.all(C[:, :, M - 1] == 0)
    out = alpha[:, None, None] * C
    b = beta[:, None]
    for k in range(M - 1):
        a_k, s_k = C[:, 0, k], C[:, 1, k] if k > 0 else 0.0 * C[:, 0, k]
        if k == 0:
            out[:, 0, 1] += b[:, 0] * a_k
            continue
        out[:, 0, k + 1] += 0.5 * b[:, 0] * a_k
        out[:, 1, k + 1] += 0.5 * b[:, 0] * s_k
        if k == 1:
            out[:, 0, 0] += 0.5 * b[:, 0] * a_k       # cos 0 = 1; sin 0 = 0
        else:
            out[:, 0, k - 1] += 0.5 * b[:, 0] * a_k
            out[:, 1, k - 1] += 0.5 * b[:, 0] * s_k
    return out


def band_limited(rng, npt, M, ncol=2):
    C = rng.standard_normal((npt, ncol, M))
    C[:, :, M - 1] = 0.0
    C[:, 1::2, 0] = 0.0
    return C


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_product_rule_helper_against_sampled_values():
    """The helper's closed form, checked by sampling both sides at angles."""
    rng = np.random.default_rng(0)
    M = 5
    C = band_limited(rng, 3, M)
    al, be = rng.uniform(2, 3, 3), rng.uniform(-1, 1, 3)
    P = times_alpha_beta_cos(C, al, be)
    t = np.linspace(0, 2 * np.pi, 17)
    m = np.arange(M)[:, None]
    ev = lambda X: np.einsum("pm,mt->pt", X[:, 0], np.cos(m * t)) + np.einsum("pm,mt->pt", X[:, 1], np.sin(m * t))
    assert np.allclose(ev(P), (al[:, None] + be[:, None] * np.cos(t)) * ev(C), atol=1e-13)


@pytest.mark.parametrize("M,mp", [(4, 6), (8, 12), (16, 24), (7, 11)])
def test_div_of_a_product_is_the_factor(M, mp):
    rng = np.random.default_rng(M)
    npt = 40
    q1 = band_limited(rng, npt, M)
    al, be = rng.uniform(2.0, 3.0, npt), rng.uniform(-1.5, 1.5, npt)
    q2 = np.zeros((npt, 2, M))
    q2[:, 0, 0], q2[:, 0, 1] = al, be
    c1 = times_alpha_beta_cos(q1, al, be)
    out = F.fft_par_div_dcl(c1, q2, mp)
    assert rel(out, q1) <= 1e-13


def test_scalar_vect_pb2_divides_every_component():
    rng = np.random.default_rng(3)
    M, mp, npt = 8, 12, 30
    W = band_limited(rng, npt, M, 6)
    al, be = rng.uniform(2.0, 3.0, npt), rng.uniform(-1.5, 1.5, npt)
    s = np.zeros((npt, 2, M))
    s[:, 0, 0], s[:, 0, 1] = al, be
    V = np.concatenate([times_alpha_beta_cos(W[:, 2 * c:2 * c + 2], al, be) for c in range(3)], axis=1)
    out = F.fft_scalar_vect_dcl(V, s, 2, mp)
    assert rel(out, W) <= 1e-13


def test_squareroot_of_a_square_and_clamping():
    M, mp, npt = 6, 9, 4
    c = np.zeros((npt, 2, M))
    # q = 0.5 + 0.3 cos t in [0.2, 0.8]; q^2 = 0.295 + 0.3 cos t + 0.045 cos 2t
    c[:, 0, 0], c[:, 0, 1], c[:, 0, 2] = 0.25 + 0.045, 0.3, 0.045
    out = F.fft_squareroot_scalar(c, mp)
    ex = np.zeros_like(c)
    ex[:, 0, 0], ex[:, 0, 1] = 0.5, 0.3
    assert np.max(np.abs(out - ex)) <= 1e-14
    hi = np.zeros((npt, 2, M)); hi[:, 0, 0], hi[:, 0, 1] = 2.0, 0.5
    lo = np.zeros((npt, 2, M)); lo[:, 0, 0], lo[:, 1, 3] = -2.0, 0.5
    one = np.zeros_like(hi); one[:, 0, 0] = 1.0
    assert np.max(np.abs(F.fft_squareroot_scalar(hi, mp) - one)) <= 1e-15
    assert np.max(np.abs(F.fft_squareroot_scalar(lo, mp))) <= 1e-15


def test_regul_knots():
    eps = 0.04
    x = np.array([0.5 - 2 * eps, 0.5 - eps, 0.5, 0.5 + eps / 2, 0.5 + eps, 0.5 + 3 * eps])
    r = F.regul(x, eps)
    # x = eps/2: (1 + (eps/2)(eps^2/4 - 3 eps^2)/(-2 eps^3))/2 = (1 + 11/16)/2 = 27/32
    assert np.allclose(r, [0.0, 0.0, 0.5, 27 / 32, 1.0, 1.0], atol=1e-15)


def test_heaviside_of_constant_level_sets():
    M, mp, npt, eps = 4, 6, 3, 0.05
    V = np.zeros((2, npt, 2, M))
    V[0, :, 0, 0] = 0.5               # H1 = 1/2
    V[1, :, 0, 0] = 0.5 + eps / 2     # H2 = 27/32
    vals = np.array([1.0, 3.0, 10.0])
    out = F.fft_heaviside_dcl(V, vals, mp, eps)
    h1, h2 = 0.5, 27 / 32
    rho = (vals[0] * (1 - h1) + vals[1] * h1) * (1 - h2) + vals[2] * h2
    ex = np.zeros((npt, 2, M)); ex[:, 0, 0] = rho
    assert np.max(np.abs(out - ex)) <= 1e-14


@pytest.mark.parametrize("mp", [12, 16])
def test_heaviside_step_is_a_dirichlet_sum(mp):
    """phi = 0.5 + cos t: with eps = 0.01 no grid angle has |cos t_j| < eps,
    so H(t_j) = 1 exactly for |j| <= J and 0 elsewhere, and
    a_0 = v1 + (v2-v1)(2J+1)/N, a_m = (2/N)(v2-v1) sin((2J+1) m pi/N)/sin(m pi/N), b_m = 0."""
    N = 2 * mp - 1
    M = 8
    t = 2 * np.pi * np.arange(N) / N
    assert np.min(np.abs(np.cos(t))) > 0.01
    J = int(np.sum(np.cos(t) > 0) // 2)
    V = np.zeros((1, 2, 2, M))
    V[0, :, 0, 0], V[0, :, 0, 1] = 0.5, 1.0
    v1, v2 = 2.0, 7.0
    out = F.fft_heaviside_dcl(V, [v1, v2], mp, 0.01)
    m = np.arange(1, M)
    ex = np.zeros((2, 2, M))
    ex[:, 0, 0] = v1 + (v2 - v1) * (2 * J + 1) / N
    ex[:, 0, 1:] = 2.0 / N * (v2 - v1) * np.sin((2 * J + 1) * m * np.pi / N) / np.sin(m * np.pi / N)
    assert np.max(np.abs(out - ex)) <= 1e-13


def test_tension_with_a_constant_second_field():
    rng = np.random.default_rng(5)
    M, mp, npt = 8, 12, 20
    U = band_limited(rng, npt, M, 6)
    g = rng.standard_normal((npt, 3))
    G = np.zeros((npt, 6, M))
    G[:, 0::2, 0] = g
    out = F.fft_tensor_dcl(U, G, mp, opt_tension=True)          # (3, np, 6, M)
    nrm = np.sqrt(np.sum(g ** 2, axis=1))
    ex = np.empty_like(out)
    for a in range(3):
        for b in range(3):
            ex[a, :, 2 * b:2 * b + 2] = U[:, 2 * a:2 * a + 2] * (g[:, b] / (nrm + 1e-14))[:, None, None]
            if a == b:
                ex[a, :, 2 * b, 0] -= nrm
    assert rel(out, ex) <= 1e-13


def test_max_vel():
    M, mp = 5, 8
    V = np.zeros((4, 6, M))
    V[:, 0::2, 0] = [[1, 2, 2], [0, 3, 4], [1, 1, 1], [0, 0, 0]]     # |v| = 3, 5, sqrt 3, 0
    assert abs(F.fft_max_vel_dcl(V, mp) - 5.0) <= 1e-14
    W = np.zeros((3, 2, M))
    W[1, 0, 1] = -2.5                                                 # max |c| at t = 0
    assert abs(F.fft_max_vel_dcl(W, mp) - 2.5) <= 1e-14


@pytest.mark.parametrize("op", ["div", "sqrt", "heaviside", "tension", "sv2"])
def test_fft_engine_matches_direct_long_double_sum(op):
    rng = np.random.default_rng(7)
    M, mp, npt = 6, 9, 12
    c = rng.standard_normal((npt, 2, M)) * 0.1
    pos = c.copy(); pos[:, 0, 0] = 3.0
    V = rng.standard_normal((npt, 6, M))
    if op == "div":
        f = lambda e: F.fft_par_div_dcl(c, pos, mp, e)
    elif op == "sqrt":
        cc = c.copy(); cc[:, 0, 0] = 0.5
        f = lambda e: F.fft_squareroot_scalar(cc, mp, e)
    elif op == "heaviside":
        L = rng.standard_normal((2, npt, 2, M)) * 0.05 + 0.5
        f = lambda e: F.fft_heaviside_dcl(L, [1.0, 2.0, 5.0], mp, 0.1, e)
    elif op == "tension":
        f = lambda e: F.fft_tensor_dcl(V, V[:, ::-1].copy(), mp, e, opt_tension=True)
    else:
        f = lambda e: F.fft_scalar_vect_dcl(V, pos, 2, mp, e)
    a, b = f("fft"), f("direct")
    assert rel(a, b) <= 1e-13


def test_no_overshoot_product_with_inactive_bounds_is_the_bilinear_product():
    """Bounds containing every grid value of the scalar: the clamp is the
    identity, so pb = 1 gives the band-limited (grid-independent) product."""
    rng = np.random.default_rng(11)
    M, mp, npt = 8, 12, 25
    V = band_limited(rng, npt, M, 6)
    s = band_limited(rng, npt, M) * 0.05
    s[:, 0, 0] = 1.0
    out = F.fft_scalar_vect_no_overshoot([0.1, 5.0, 2.0], V, s, 1, mp)
    assert rel(out, F.fft_scalar_vect_dcl(V, s, 1, mp)) <= 1e-13
    al, be = s[:, 0, 0].copy(), np.zeros(npt)
    W = band_limited(rng, npt, M, 6)
    P = np.concatenate([times_alpha_beta_cos(W[:, 2 * c:2 * c + 2], al, be) for c in range(3)], axis=1)
    s0 = np.zeros((npt, 2, M)); s0[:, 0, 0] = al
    assert rel(F.fft_scalar_vect_no_overshoot([0.5, 2.0], P, s0, 2, mp), W) <= 1e-13


def test_no_overshoot_active_bounds_clamp_to_a_constant():
    rng = np.random.default_rng(12)
    M, mp, npt = 6, 9, 10
    V = band_limited(rng, npt, M, 6)
    s = np.zeros((npt, 2, M)); s[:, 0, 0] = 10.0; s[:, 0, 1] = 1.0      # s in [9, 11] > max bound
    assert rel(F.fft_scalar_vect_no_overshoot([0.5, 3.0], V, s, 1, mp), 3.0 * V) <= 1e-13
    assert rel(F.fft_scalar_vect_no_overshoot([0.5, 3.0], V, s, 2, mp), V / 3.0) <= 1e-13


def test_no_overshoot_level_set():
    M, mp, npt = 6, 9, 4
    c = np.zeros((npt, 2, M)); c[:, 0, 0], c[:, 0, 2] = 0.5, 0.3       # within [0.2, 0.8]: identity
    assert np.max(np.abs(F.fft_no_overshoot_level_set(c, mp) - c)) <= 1e-15
    hi = np.zeros_like(c); hi[:, 0, 0], hi[:, 1, 1] = 1.5, 0.4
    one = np.zeros_like(c); one[:, 0, 0] = 1.0
    assert np.max(np.abs(F.fft_no_overshoot_level_set(hi, mp) - one)) <= 1e-15
    assert np.max(np.abs(F.fft_no_overshoot_level_set(-hi, mp))) <= 1e-15
    rng = np.random.default_rng(13)
    r = rng.standard_normal((npt, 2, M)) * 0.5 + 0.5
    assert rel(F.fft_no_overshoot_level_set(r, mp), F.fft_no_overshoot_level_set(r, mp, "direct")) <= 1e-13


def grid_values(C, mp):
    """Values on the N = 2*mp-1 grid angles of fields C (np, 2, M) by explicit
    sums a_0 + sum a_m cos(m t) + b_m sin(m t): (N, np)."""
    N = 2 * mp - 1
    t = 2 * np.pi * np.arange(N) / N
    m = np.arange(C.shape[2])
    cs, sn = np.cos(np.outer(t, m)), np.sin(np.outer(t, m))
    sn[:, 0] = 0.0
    return cs @ C[:, 0, :].T + sn @ C[:, 1, :].T


def test_entro_visc_vector_part_is_a_band_limited_product():
    """c_in_real = coeff - s on the grid, s band-limited: V_out is the
    coefficients of s*V, which the scalar-vector product gives."""
    rng = np.random.default_rng(14)
    M, mp, npt = 8, 12, 20
    V = band_limited(rng, npt, M, 6)
    S = band_limited(rng, npt, M)
    coeff = 0.7
    creal = coeff - grid_values(S, mp)
    c = band_limited(rng, npt, M) * 0.05
    c[:, 0, 0] = 0.5
    Vo, co = F.fft_par_entro_visc_dcl(V, c, creal, coeff, mp)
    assert rel(Vo, F.fft_scalar_vect_dcl(V, S, 1, mp)) <= 1e-12
    # 0 <= c <= 1 on the grid: MAX(0, c(1-c)) = c - c*c, a band-limited product
    assert np.min(grid_values(c, mp)) > 0 and np.max(grid_values(c, mp)) < 1
    assert rel(co, c - F.fft_par_prod_dcl(c, c, mp)) <= 1e-12


def test_entro_visc_scalar_part_clamps_at_zero():
    M, mp, npt = 6, 9, 5
    V = np.zeros((npt, 6, M))
    c = np.zeros((npt, 2, M)); c[:, 0, 0], c[:, 0, 1] = 2.0, 0.5      # c in [1.5, 2.5]: c(1-c) < 0
    _, co = F.fft_par_entro_visc_dcl(V, c, np.zeros((2 * mp - 1, npt)), 1.0, mp)
    assert np.max(np.abs(co)) == 0.0


def test_grad_mom_entropy_viscosity_is_a_band_limited_product():
    """c_in_real = coeff + s on the grid, s band-limited: each output is the
    scalar-vector product s V_a."""
    rng = np.random.default_rng(15)
    M, mp, npt = 8, 12, 12
    Vs = [band_limited(rng, npt, M, 6) for _ in range(3)]
    S = band_limited(rng, npt, M)
    coeff = 0.4
    out = F.fft_compute_entropy_visc_grad_mom(*Vs, coeff + grid_values(S, mp), coeff, mp)
    for a in range(3):
        assert rel(out[a], F.fft_scalar_vect_dcl(Vs[a], S, 1, mp)) <= 1e-12


def test_diffu_mom_is_two_scalar_vector_products():
    rng = np.random.default_rng(16)
    M, mp, npt = 6, 9, 10
    V1, V2, s = band_limited(rng, npt, M, 6), band_limited(rng, npt, M, 6), band_limited(rng, npt, M)
    o1, o2 = F.fft_compute_diffu_mom(V1, V2, s, mp)
    assert rel(o1, F.fft_scalar_vect_dcl(V1, s, 1, mp)) <= 1e-14
    assert rel(o2, F.fft_scalar_vect_dcl(V2, s, 1, mp)) <= 1e-14
    assert rel(o1, F.fft_compute_diffu_mom(V1, V2, s, mp, "direct")[0]) <= 1e-13


def test_compression_viscosity_reduces_to_entropy_viscosity_and_constant_case():
    rng = np.random.default_rng(17)
    M, mp, npt = 8, 12, 10
    N = 2 * mp - 1
    V1, V2 = band_limited(rng, npt, M, 6), band_limited(rng, npt, M, 6)
    c = band_limited(rng, npt, M) * 0.05
    c[:, 0, 0] = 0.5
    cr = rng.standard_normal((N, npt))
    a = F.fft_par_compr_entro_visc_dcl(V1, V2, c, cr, 0.3, 0.0, mp)
    b = F.fft_par_entro_visc_dcl(V1, c, cr, 0.3, mp)
    assert rel(a[0], b[0]) <= 1e-13 and rel(a[1], b[1]) <= 1e-13
    # constant grad(phi_reg) = g, phi = phi0, viscosity x0: a constant compression term
    g = rng.standard_normal((npt, 3))
    G = np.zeros((npt, 6, M)); G[:, 0::2, 0] = g
    phi0, x0, comp, coeff = 0.3, 0.8, 1.7, 0.2
    C0 = np.zeros((npt, 2, M)); C0[:, 0, 0] = phi0
    Vo, co = F.fft_par_compr_entro_visc_dcl(V1, G, C0, np.full((N, npt), x0), coeff, comp, mp)
    vc = x0 * comp * phi0 * (1 - phi0) / (np.sqrt(np.sum(g ** 2, axis=1)) + 1e-14)
    ex = (coeff - x0) * V1
    ex[:, 0::2, 0] += vc[:, None] * g
    assert rel(Vo, ex) <= 1e-13
    assert np.allclose(co[:, 0, 0], phi0 * (1 - phi0), atol=1e-15) and np.max(np.abs(co[:, :, 1:])) < 1e-15


def test_entropy_visc_with_the_velocity_bound_active():
    """LES_coeff2 huge: visco = -c1 + c4 max|v|, the maximum over the element's
    points and the angles j-1, j, j+1.  With v = (a cos t, 0, 0) the same at
    every point of an element, max|v|(t_j) = |a| max(|cos t_{j-1}|, |cos t_j|,
    |cos t_{j+1}|); with a constant V3 the output is the direct cosine sum of
    visco(t_j) V3."""
    M, mp, lg, ne = 6, 9, 7, 3
    N = 2 * mp - 1
    npt = lg * ne
    rng = np.random.default_rng(18)
    amp = rng.uniform(0.5, 2.0, ne)
    V1 = np.zeros((npt, 6, M)); V1[:, 0, 1] = np.repeat(amp, lg)
    V2 = rng.standard_normal((npt, 6, M))
    c3 = rng.standard_normal(npt)
    V3 = np.zeros((npt, 6, M)); V3[:, 0, 0] = c3
    Z = np.zeros((npt, 6, M))
    h = rng.uniform(0.1, 1.0, npt)
    c1, c4 = 0.3, 0.8
    o1, o2, o3 = F.fft_compute_entropy_visc(V1, V2, V3, Z, Z, h, 1.0, lg, c1, 1e30, c4, mp)
    t = 2 * np.pi * np.arange(N) / N
    ac = np.abs(np.cos(t))
    mx = np.maximum(np.maximum(np.roll(ac, 1), ac), np.roll(ac, -1))
    vis = -c1 + c4 * np.repeat(amp, lg)[:, None] * mx[None, :]                # (np, N)
    m = np.arange(M)
    a = (vis * c3[:, None]) @ np.cos(np.outer(t, m)) * (2.0 / N)
    a[:, 0] /= 2.0
    b = (vis * c3[:, None]) @ np.sin(np.outer(t, m)) * (2.0 / N)
    b[:, 0] = 0.0
    assert np.max(np.abs(o1[:, 0] - a)) <= 1e-13 and np.max(np.abs(o1[:, 1] - b)) <= 1e-13
    assert np.max(np.abs(o2)) == 0.0 and np.max(np.abs(o3)) == 0.0


def test_entropy_visc_with_the_residual_bound_active():
    """LES_coeff4 huge: visco = -c1 + c2 h |v . res|/(res_norm + 1e-14); with v
    and res constant in angle the viscosity is constant per point."""
    M, mp, lg, npt = 5, 8, 7, 14
    rng = np.random.default_rng(19)
    v, r = rng.standard_normal((npt, 3)), rng.standard_normal((npt, 3))
    V1 = np.zeros((npt, 6, M)); V1[:, 0::2, 0] = v
    V2 = np.zeros((npt, 6, M)); V2[:, 0::2, 0] = r
    Vs = [band_limited(rng, npt, M, 6) for _ in range(3)]
    h = rng.uniform(0.1, 1.0, npt)
    c1, c2, res = 0.1, 0.7, 2.5
    out = F.fft_compute_entropy_visc(V1, V2, *Vs, h, res, lg, c1, c2, 1e30, mp)
    vis = -c1 + c2 * h * np.abs(np.sum(v * r, axis=1)) / (res + 1e-14)
    for k in range(3):
        assert rel(out[k], vis[:, None, None] * Vs[k]) <= 1e-13


@pytest.mark.parametrize("lg", [7, 3])
def test_entropy_visc_mom_for_fields_constant_in_angle(lg):
    """Mode-0 fields: every grid quantity is constant in angle, so the element
    and angle maxima are maxima over the element's points and the outputs
    have a closed form per element."""
    M, mp, ne = 5, 8, 4
    N = 2 * mp - 1
    npt = lg * ne
    rng = np.random.default_rng(20 + lg)
    u, m, r = (rng.standard_normal((npt, 3)) for _ in range(3))
    cc = rng.standard_normal(npt)
    V1, V2, V3 = (np.zeros((npt, 6, M)) for _ in range(3))
    V1[:, 0::2, 0], V2[:, 0::2, 0], V3[:, 0::2, 0] = u, m, r
    C = np.zeros((npt, 2, M)); C[:, 0, 0] = cc
    h = rng.uniform(0.1, 1.0, npt)
    c2, c4 = 0.6, 0.9
    o1, o2 = F.fft_compute_entropy_visc_mom(V1, V2, V3, C, h, lg, c2, c4, mp, want_c2=True)
    uu = np.sum(u ** 2, axis=1)
    raw = c2 * h * np.maximum(np.abs(np.sum(u * r, axis=1)), np.abs(cc * uu))
    emax = lambda x: np.repeat(x.reshape(ne, lg).max(axis=1), lg)
    nv, nm, vis = emax(np.sqrt(uu)), emax(np.sqrt(np.sum(m ** 2, axis=1))), emax(raw)
    vmax = np.max(np.sqrt(uu))
    q = vis / (nv * nm + 1e-30)
    if lg == 3:
        ex1 = np.minimum(4 * c4 * vmax, 16 * q)
        assert o2 is None
    else:
        ex1 = np.minimum(c4 * nv, q)
        assert np.allclose(o2, np.minimum(c4 * vmax, q)[None, :], rtol=1e-13, atol=0)
    assert o1.shape == (N, npt)
    assert np.allclose(o1, ex1[None, :], rtol=1e-13, atol=0)


def test_entropy_visc_mom_engines_agree():
    rng = np.random.default_rng(22)
    M, mp, lg, npt = 6, 9, 7, 14
    Vs = [rng.standard_normal((npt, 6, M)) for _ in range(3)]
    C = rng.standard_normal((npt, 2, M))
    h = rng.uniform(0.1, 1.0, npt)
    a = F.fft_compute_entropy_visc_mom(*Vs, C, h, lg, 0.6, 0.9, mp, True)
    b = F.fft_compute_entropy_visc_mom(*Vs, C, h, lg, 0.6, 0.9, mp, True, "direct")
    assert rel(a[0], b[0]) <= 1e-12 and rel(a[1], b[1]) <= 1e-12


@pytest.mark.parametrize("vscale", [1.0, 5.0])
def test_compression_level_set_for_fields_constant_in_angle(vscale):
    """phi, grad phi constant in angle; velocity v_r = b + c cos t (b, c >= 0),
    so max |v| = max(b + c), reached at t = 0, a grid angle; the last point sits
    past the last whole element.  Per point the product is the constant
    K min(1, vmax) s (MAX(1-(2phi-1)^2, 0)/(2h) - |g|), s = 2 regul(phi) - 1:
    s = 1 above 0.51, -1 below 0.49, 0 at 0.5 and 11/16 at 0.505."""
    M, mp, lg = 6, 9, 3
    phi = np.array([0.8, 0.2, 0.505, 1.3, 0.5, 0.3, 0.9])
    s = np.array([1.0, -1.0, 11 / 16, 1.0, 0.0, -1.0, 1.0])
    npt = phi.size
    rng = np.random.default_rng(23)
    g = rng.standard_normal((npt, 3))
    G = np.zeros((npt, 6, M)); G[:, 0::2, 0] = g
    C = np.zeros((npt, 2, M)); C[:, 0, 0] = phi
    b = np.array([0.1, 0.2, 0.05, 0.0, 0.3, 0.1, 0.4]) * vscale
    c = np.array([0.2, 0.1, 0.0, 0.1, 0.2, 0.0, 0.3]) * vscale
    V = np.zeros((npt, 6, M)); V[:, 0, 0] = b; V[:, 0, 1] = c
    vmax = min(1.0, float(np.max(b + c)))
    comp, c4, h = 1.3, 0.7, 0.05
    out = F.fft_compression_level_set_dcl(G, V, C, lg, comp, c4, h, mp)
    q = np.maximum(1 - (2 * phi - 1) ** 2, 0.0)
    ex = np.zeros((npt, 2, M))
    ex[:, 0, 0] = comp * 4 * c4 * vmax * s * (q / (2 * h) - np.sqrt(np.sum(g ** 2, axis=1)))
    assert rel(out, ex) <= 1e-13


@pytest.mark.parametrize("mp", [12, 16])
def test_compression_level_set_step_by_explicit_sums(mp):
    """phi = 0.5 + 0.3 cos t: no grid angle has |0.3 cos t_j| <= 0.01, so
    s_j = sign(cos t_j) exactly; grad phi = (a cos t, a sin t, 0) has |g| = |a|;
    v = (0.4 + 0.3 cos t, 0, 0) gives vmax = 0.7.  The grid values are
    f_j = K 0.7 s_j ((1 - 0.36 cos^2 t_j)/(2h) - |a|); the coefficients are
    the sums a_0 = sum f_j / N, a_m = 2 sum f_j cos(m t_j) / N,
    b_m = 2 sum f_j sin(m t_j) / N."""
    N = 2 * mp - 1
    M = 8
    t = 2 * np.pi * np.arange(N) / N
    assert np.min(np.abs(0.3 * np.cos(t))) > 0.01
    a_ = -0.6
    G = np.zeros((2, 6, M)); G[:, 0, 1] = a_; G[:, 3, 1] = a_
    C = np.zeros((2, 2, M)); C[:, 0, 0] = 0.5; C[:, 0, 1] = 0.3
    V = np.zeros((2, 6, M)); V[:, 0, 0] = 0.4; V[:, 0, 1] = 0.3
    comp, c4, h = 2.0, 0.25, 0.1
    out = F.fft_compression_level_set_dcl(G, V, C, 2, comp, c4, h, mp)
    f = comp * 4 * c4 * 0.7 * np.sign(np.cos(t)) * ((1 - 0.36 * np.cos(t) ** 2) / (2 * h) - abs(a_))
    m = np.arange(M)[:, None]
    ex = np.zeros((2, 2, M))
    ex[:, 0] = (2.0 / N) * (np.cos(m * t) @ f)
    ex[:, 0, 0] /= 2
    ex[:, 1] = (2.0 / N) * (np.sin(m * t) @ f)
    ex[:, 1, 0] = 0.0
    assert rel(out, ex) <= 1e-13


def test_compression_level_set_engines_agree():
    rng = np.random.default_rng(24)
    M, mp, npt = 6, 9, 10
    G, V = rng.standard_normal((npt, 6, M)), rng.standard_normal((npt, 6, M))
    C = rng.standard_normal((npt, 2, M)) * 0.2
    C[:, 0, 0] = 0.5
    a = F.fft_compression_level_set_dcl(G, V, C, 2, 1.1, 0.6, 0.07, mp)
    b = F.fft_compression_level_set_dcl(G, V, C, 2, 1.1, 0.6, 0.07, mp, "direct")
    assert rel(a, b) <= 1e-12


def test_max_min_vel_constant_fields_and_zero():
    """Constant |v| per point: max and the smallest positive value; zero
    points are skipped; an all-zero field gives (0, 0)."""
    M, mp = 4, 6
    V = np.zeros((4, 6, M))
    V[:, 0::2, 0] = [[1, 2, 2], [0, 0, 0], [0, 3, 4], [2, 0, 0]]      # |v| = 3, 0, 5, 2
    vmax, vmin = F.fft_max_min_vel_dcl(V, mp)
    assert abs(vmax - 5.0) <= 1e-14 and abs(vmin - 2.0) <= 1e-14
    assert F.fft_max_min_vel_dcl(np.zeros((3, 2, M)), mp) == (0.0, 0.0)
    C = np.zeros((2, 4, M)); C[:, 0, 0] = [-1.5, 0.5]; C[:, 2, 0] = 9.0   # nb 4: first field only
    assert F.fft_max_min_vel_dcl(C, mp) == (1.5, 0.5)


def test_max_min_vel_by_direct_sums():
    """Random fields: max and min of |v| on the grid by explicit cosine/sine
    sums (every value is positive, so the minimum is the plain minimum)."""
    rng = np.random.default_rng(25)
    M, mp, npt = 5, 8, 6
    V = rng.standard_normal((npt, 6, M))
    V[:, 0, 0] += 10.0
    N = 2 * mp - 1
    t = 2 * np.pi * np.arange(N) / N
    m = np.arange(1, M)[:, None]
    vals = [V[:, 2 * f, :1] + V[:, 2 * f, 1:] @ np.cos(m * t) + V[:, 2 * f + 1, 1:] @ np.sin(m * t) for f in range(3)]
    nrm = np.sqrt(sum(v ** 2 for v in vals))
    vmax, vmin = F.fft_max_min_vel_dcl(V, mp)
    assert abs(vmax - nrm.max()) <= 1e-13 * nrm.max() and abs(vmin - nrm.min()) <= 1e-13 * nrm.max()


def test_check_interface():
    M, mp, npt = 4, 6, 3
    def ls(*c0, c1=0.0):
        V = np.zeros((len(c0), npt, 2, M))
        for i, c in enumerate(c0):
            V[i, :, 0, 0] = c
        V[0, :, 0, 1] = c1
        return V
    assert F.fft_check_interface(ls(0.5, 0.7), 3, mp) == 0          # 0.5 in [0.1, 0.9] below 0.7
    assert F.fft_check_interface(ls(0.5, 0.3), 3, mp) == 1
    assert F.fft_check_interface(ls(0.95, 0.99), 3, mp) == 1        # outside [0.1, 0.9]
    assert F.fft_check_interface(ls(0.05, 0.99), 3, mp) == 1
    assert F.fft_check_interface(ls(0.6, 0.5, 0.55), 4, mp) == 0    # second pair
    assert F.fft_check_interface(ls(0.6, 0.5, 0.4), 4, mp) == 1
    assert F.fft_check_interface(ls(0.5, 0.5, c1=0.3), 3, mp) == 0  # 0.5 + 0.3 cos t < 0.5 where cos t < 0
    assert F.fft_check_interface(ls(0.2), 2, mp) == 1


@pytest.mark.parametrize("nb", [2, 6])
def test_var_eta_prod_closed_forms(nb):
    """eta constant per point: eta * c; eta = alpha + beta cos t: the product
    rule of times_alpha_beta_cos, per (cos, sin) column pair."""
    rng = np.random.default_rng(26)
    M, mp, npt = 6, 9, 5
    N = 2 * mp - 1
    C = band_limited(rng, npt, M, nb)
    e0 = rng.uniform(0.5, 2.0, npt)
    out = F.fft_par_var_eta_prod(np.tile(e0, (N, 1)), C, mp)
    assert rel(out, e0[:, None, None] * C) <= 1e-13
    al, be = rng.uniform(1, 2, npt), rng.uniform(-1, 1, npt)
    t = 2 * np.pi * np.arange(N) / N
    out = F.fft_par_var_eta_prod(al[None, :] + be[None, :] * np.cos(t)[:, None], C, mp)
    ex = np.concatenate([times_alpha_beta_cos(C[:, 2 * k:2 * k + 2], al, be) for k in range(nb // 2)], axis=1)
    assert rel(out, ex) <= 1e-13


def test_scal_funct_closed_forms():
    """funct(x) = 2x + 1: 2c + 1 in mode 0; funct(x) = x^2 with c = a + b cos t:
    a^2 + b^2/2 + 2ab cos t + (b^2/2) cos 2t; columns 3-4 untouched."""
    M, mp, npt = 5, 8, 3
    rng = np.random.default_rng(27)
    C = band_limited(rng, npt, M, 4)
    out = F.fft_par_scal_funct(C, lambda x: 2 * x + 1, mp)
    ex = C.copy(); ex[:, 0:2] *= 2; ex[:, 0, 0] += 1
    assert rel(out, ex) <= 1e-13
    a, b = np.array([1.0, -0.5, 2.0]), np.array([0.3, 1.0, -2.0])
    D = np.zeros((npt, 2, M)); D[:, 0, 0] = a; D[:, 0, 1] = b
    out = F.fft_par_scal_funct(D, lambda x: x * x, mp)
    ex = np.zeros_like(D)
    ex[:, 0, 0] = a * a + b * b / 2; ex[:, 0, 1] = 2 * a * b; ex[:, 0, 2] = b * b / 2
    assert np.max(np.abs(out - ex)) <= 1e-14 * 10
