"""Pins for the oracle's P2 tables, Gauss preparers and RHS scatter (CPU).

P2 reproduces quadratics exactly, so values, derivatives and curls of
quadratic fields have closed forms at the Gauss points; Stroud's rule is
exact to degree 5, so the scattered right-hand side of a quadratic source
integrates ``f r`` exactly.
"""
import numpy as np
import pytest

from oracle import gauss_p2, smb, assembly
from sfemans_amd.mesh import rectangle_p2


def test_reference_element():
    w, d, p = gauss_p2.element_2d_p2()
    assert np.isclose(p.sum(), 0.5, atol=1e-15)
    assert np.allclose(w.sum(axis=0), 1.0, atol=1e-14)
    assert np.allclose(d.sum(axis=1), 0.0, atol=1e-13)
    # rule exact for degree 5 on the reference triangle: int x^a y^b = a! b! / (a+b+2)!
    from math import factorial
    xx = np.array([1/3, (6-np.sqrt(15))/21, (6-np.sqrt(15))/21, (9+2*np.sqrt(15))/21,
                   (6+np.sqrt(15))/21, (6+np.sqrt(15))/21, (9-2*np.sqrt(15))/21])
    yy = np.array([1/3, (6-np.sqrt(15))/21, (9+2*np.sqrt(15))/21, (6-np.sqrt(15))/21,
                   (6+np.sqrt(15))/21, (9-2*np.sqrt(15))/21, (6+np.sqrt(15))/21])
    for a in range(6):
        for b in range(6 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert np.isclose(np.sum(p * xx**a * yy**b), exact, rtol=1e-13, atol=1e-16)


def test_mesh_numbering_and_tables():
    m = rectangle_p2(3, 4)
    assert m.me == 24 and m.np == 4 * 5 + (3 * 5 + 4 * 4 + 12)
    jj = m.jj - 1
    assert np.all(jj[0] < jj[1]) and np.all(jj[1] < jj[2])
    # node 3+k is the midpoint of the edge opposite vertex k
    for k, (a, b) in enumerate(((1, 2), (0, 2), (0, 1))):
        mid = 0.5 * (m.rr[:, jj[a]] + m.rr[:, jj[b]])
        assert np.allclose(m.rr[:, jj[3 + k]], mid, atol=1e-15)
    ww, dw, rj, ray = gauss_p2.gauss_tables(m.jj, m.rr)
    assert np.isclose(rj.sum(), 2.0, atol=1e-13)          # area of [0,1]x[-1,1]
    assert np.allclose(dw.sum(axis=1), 0.0, atol=1e-12)


def quad_field(rr, M, seed):
    """Nodal values of per-(column, mode) quadratics; returns (V, coef)."""
    c = np.random.default_rng(seed).standard_normal((6, M, 6))   # 1, r, z, r2, rz, z2
    r, z = rr[0], rr[1]
    basis = np.stack([np.ones_like(r), r, z, r * r, r * z, z * z])  # (6, np)
    V = np.einsum("kmb,bn->nkm", c, basis)
    return np.asfortranarray(V), c


def eval_quad(c, r, z):
    b = np.stack([np.ones_like(r), r, z, r * r, r * z, z * z])
    dr = np.stack([0 * r, 1 + 0 * r, 0 * r, 2 * r, z, 0 * r])
    dz = np.stack([0 * r, 0 * r, 1 + 0 * r, 0 * r, r, 2 * z])
    return (np.einsum("kmb,bg->gkm", c, b), np.einsum("kmb,bg->gkm", c, dr),
            np.einsum("kmb,bg->gkm", c, dz))


def gauss_coords(m):
    ww, dw, rj, ray = gauss_p2.gauss_tables(m.jj, m.rr)
    j0 = m.jj - 1
    rg = np.einsum("nm,nl->ml", m.rr[0][j0], ww).ravel()
    zg = np.einsum("nm,nl->ml", m.rr[1][j0], ww).ravel()
    return rg, zg


def test_curl_of_quadratic_field_is_exact():
    m = rectangle_p2(3, 5)
    M = 4
    modes = np.arange(M)
    V, c = quad_field(m.rr, M, 1)
    W, R = smb.gauss_velocity_curl(m.jj, m.rr, modes, V)
    rg, zg = gauss_coords(m)
    u, ur, uz = eval_quad(c, rg, zg)
    mr = modes[None, :] / rg[:, None]
    ex = np.empty_like(u)
    ex[:, 0] = mr * u[:, 5] - uz[:, 2]
    ex[:, 1] = -mr * u[:, 4] - uz[:, 3]
    ex[:, 2] = uz[:, 0] - ur[:, 4]
    ex[:, 3] = uz[:, 1] - ur[:, 5]
    ex[:, 4] = u[:, 2] / rg[:, None] + ur[:, 2] - mr * u[:, 1]
    ex[:, 5] = u[:, 3] / rg[:, None] + ur[:, 3] + mr * u[:, 0]
    assert np.allclose(W, u, atol=1e-12)
    assert np.allclose(R, ex, atol=1e-10 * np.max(np.abs(ex)))


def test_ugradc_of_quadratic_fields_is_exact():
    m = rectangle_p2(2, 3)
    M = 3
    modes = np.arange(M)
    V, cv = quad_field(m.rr, M, 2)
    C, cc = quad_field(m.rr, M, 3)
    C = np.asfortranarray(C[:, :2])
    V[:, 1::2, 0] = 0
    C[:, 1, 0] = 0
    hc = np.random.default_rng(4).uniform(0.5, 2, m.me)
    out = smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, modes, hc, V, C)
    # direct physical-space evaluation on a fine azimuthal grid
    rg, zg = gauss_coords(m)
    u, ur, uz = eval_quad(cv, rg, zg)
    t, tr, tz = eval_quad(cc[:2], rg, zg)
    u[:, 1::2, 0] = ur[:, 1::2, 0] = uz[:, 1::2, 0] = 0
    t[:, 1, 0] = tr[:, 1, 0] = tz[:, 1, 0] = 0
    Nth = 64
    th = 2 * np.pi * np.arange(Nth) / Nth
    cs = np.cos(np.outer(modes, th)); sn = np.sin(np.outer(modes, th))
    phys = lambda a, col: a[:, col] @ cs + a[:, col + 1] @ sn
    dth = lambda a, col: (a[:, col] * modes) @ (-sn) + (a[:, col + 1] * modes) @ cs
    ur_, ut_, uz_ = phys(u, 0), phys(u, 2), phys(u, 4)
    T = phys(t, 0)
    dTr, dTz, dTt = phys(tr, 0), phys(tz, 0), dth(t, 0)
    div = phys(ur, 0) + ur_ / rg[:, None] + dth(u, 2) / rg[:, None] + phys(uz, 4)
    hcg = np.repeat(hc, 7)[:, None]
    f = hcg * (ur_ * dTr + ut_ * dTt / rg[:, None] + uz_ * dTz + div * T)
    a0 = f.mean(axis=1)
    am = 2 * f @ cs.T / Nth
    bm = 2 * f @ sn.T / Nth
    assert np.allclose(out[:, 0, 0], a0, atol=1e-10)
    assert np.allclose(out[:, 0, 1:], am[:, 1:], atol=1e-10)
    assert np.allclose(out[:, 1, 1:], bm[:, 1:], atol=1e-10)


def test_rhs_3x3_integrates_quadratics_and_packs_145_236():
    m = rectangle_p2(4, 6)
    rg, zg = gauss_coords(m)
    f = np.zeros((m.n_gauss, 6))
    coef = [(1.0, 0, 0), (0.5, 1, 0), (-2.0, 0, 2), (3.0, 2, 0), (1.5, 1, 1), (0.25, 0, 1)]
    exact = []
    for k, (a, p, q) in enumerate(coef):
        f[:, k] = a * rg**p * zg**q
        # int_0^1 int_-1^1 a r^p z^q r dz dr
        zi = (1 - (-1) ** (q + 1)) / (q + 1)
        exact.append(a * zi / (p + 2))
    v145, v236 = assembly.rhs_3x3(m.jj, m.rr, f)
    got = [v145[:, 0].sum(), v236[:, 0].sum(), v236[:, 1].sum(), v145[:, 1].sum(),
           v145[:, 2].sum(), v236[:, 2].sum()]
    assert np.allclose(got, exact, atol=1e-13)
    allm = assembly.rhs_3x3_all_modes(m.jj, m.rr, np.stack([f, 2 * f], axis=2))
    assert np.allclose(allm[0][:, :, 1], 2 * v145, atol=1e-14)


def test_qs_00_gauss_integrates():
    m = rectangle_p2(3, 3)
    rg, zg = gauss_coords(m)
    hc = np.full(m.me, 2.0)
    ff = m.rr[0] ** 2                   # nodal quadratic, interpolated exactly
    fg = zg.copy()
    v = assembly.qs_00_gauss(m.jj, m.rr, hc, ff, fg)
    # int (2 r^2 + z) r dr dz = 2 * 2 * 1/4 + 0
    assert np.isclose(v.sum(), 1.0, atol=1e-13)
