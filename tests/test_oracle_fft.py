"""Pins for the oracle's FFT_PAR_* restatement (CPU only).

Each check is chosen so that a plausible mistake fails it: the long-double
direct DFT pins the fast engine; single-mode closed forms pin the 1/2 factors,
the sign of the sine column and the truncation; the TEST1 known answer of the
reference pins the cross-product component order; invariants pin symmetry,
bilinearity, Parseval and grid independence.
"""
import numpy as np
import pytest

from oracle import fft_par as F
from tests.known_answers import ka_test1_fields, ka_test1_rot_cross_v


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("M", [1, 2, 3, 4, 8])
def test_fast_engine_matches_long_double_direct(M):
    V1 = rnd((7, 6, M), 1)
    V2 = rnd((7, 6, M), 2)
    mp = 3 * M // 2 if M > 1 else 1
    for op in ("cross", "dot", "tensor"):
        a = F.fft_product([V1, V2], op, max(mp, M), engine="fft")
        b = F.fft_product([V1, V2], op, max(mp, M), engine="direct")
        assert np.max(np.abs(a - b)) <= 1e-13 * max(1.0, np.max(np.abs(b)))


def single(npt, M, col, m, val=1.0):
    V = np.zeros((npt, 2, M))
    V[:, col, m] = val
    return V


@pytest.mark.parametrize("M", [4, 5, 8])
def test_prod_closed_forms(M):
    mp = 3 * M // 2
    for m in range(M):
        for n in range(M):
            # cos m * cos n = (cos(m-n) + cos(m+n)) / 2
            out = F.fft_par_prod_dcl(single(3, M, 0, m), single(3, M, 0, n), mp)
            exp = np.zeros((3, 2, M))
            for k in (abs(m - n), m + n):
                if k < M:
                    exp[:, 0, k] += 0.5 if not (m == 0 or n == 0) else 0.0
            if m == 0 or n == 0:
                exp = np.zeros((3, 2, M))
                exp[:, 0, max(m, n)] = 1.0
            assert np.allclose(out, exp, atol=1e-14), (m, n)
            if m >= 1 and n >= 1:
                # sin m * cos n = (sin(m+n) + sin(m-n)) / 2
                out = F.fft_par_prod_dcl(single(3, M, 1, m), single(3, M, 0, n), mp)
                exp = np.zeros((3, 2, M))
                if m + n < M:
                    exp[:, 1, m + n] += 0.5
                d = m - n
                if d > 0:
                    exp[:, 1, d] += 0.5
                elif d < 0:
                    exp[:, 1, -d] -= 0.5
                assert np.allclose(out, exp, atol=1e-14), ("sc", m, n)


def test_mode0_sine_is_ignored_and_zero_on_output():
    V1 = rnd((5, 6, 4), 3)
    V2 = rnd((5, 6, 4), 4)
    out = F.fft_par_cross_prod_dcl(V1, V2, 6)
    W1 = V1.copy()
    W1[:, 1::2, 0] = 123.0
    out2 = F.fft_par_cross_prod_dcl(W1, V2, 6)
    assert np.array_equal(out, out2)
    assert np.all(out[:, 1::2, 0] == 0.0)


def test_known_answer_test1_cross_product():
    pts = np.random.default_rng(5).uniform([0.05, -1.0], [1.0, 1.0], size=(11, 2))
    r, z = pts[:, 0], pts[:, 1]
    V, Rot = ka_test1_fields(r, z)                 # mode-1 coefficients (np, 6)
    M = 3
    V1 = np.zeros((len(r), 6, M))
    V2 = np.zeros((len(r), 6, M))
    V1[:, :, 1] = Rot
    V2[:, :, 1] = V
    out = F.fft_par_cross_prod_dcl(V1, V2, 3 * M // 2)
    exp0, exp2 = ka_test1_rot_cross_v(V, Rot)
    scale = np.max(np.abs(exp2))
    assert np.max(np.abs(out[:, :, 0] - exp0)) <= 1e-13 * scale
    assert np.max(np.abs(out[:, :, 2] - exp2)) <= 1e-13 * scale
    assert np.max(np.abs(out[:, :, 1])) <= 1e-13 * scale


def test_cross_antisymmetric_dot_symmetric_and_bilinear():
    a, b, c = rnd((9, 6, 8), 6), rnd((9, 6, 8), 7), rnd((9, 6, 8), 8)
    for x in (a, b, c):
        x[:, 1::2, 0] = 0
    mp = 12
    ab = F.fft_par_cross_prod_dcl(a, b, mp)
    assert np.allclose(ab, -F.fft_par_cross_prod_dcl(b, a, mp), atol=1e-13)
    assert np.allclose(F.fft_par_dot_prod_dcl(a, b, mp), F.fft_par_dot_prod_dcl(b, a, mp), atol=1e-13)
    lhs = F.fft_par_cross_prod_dcl(2.5 * a + c, b, mp)
    rhs = 2.5 * ab + F.fft_par_cross_prod_dcl(c, b, mp)
    assert np.allclose(lhs, rhs, atol=1e-12)


def test_parseval_mode0_and_grid_independence():
    M = 8
    a, b = rnd((6, 2, M), 9), rnd((6, 2, M), 10)
    a[:, 1, 0] = b[:, 1, 0] = 0
    out = F.fft_par_prod_dcl(a, b, 3 * M // 2)
    mean = a[:, 0, 0] * b[:, 0, 0] + 0.5 * np.sum(a[:, 0, 1:] * b[:, 0, 1:] + a[:, 1, 1:] * b[:, 1, 1:], axis=1)
    assert np.allclose(out[:, 0, 0], mean, atol=1e-13)
    # exact de-aliasing for any grid with N >= 3M-2
    ref = F.fft_product([a, b], "prod", 3 * M // 2)
    for mp in (3 * M // 2 + 1, 2 * M, 3 * M):
        assert np.allclose(F.fft_product([a, b], "prod", mp), ref, atol=1e-13)
    # an aliased grid (N < 3M-2) is detectably different
    assert not np.allclose(F.fft_product([a, b], "prod", M), ref, atol=1e-8)


def test_tensor_trace_is_dot_and_layout():
    a, b = rnd((4, 6, 5), 11), rnd((4, 6, 5), 12)
    T = F.fft_tensor_dcl(a, b, 7)
    d = F.fft_par_dot_prod_dcl(a, b, 7)
    tr = T[0, :, 0:2] + T[1, :, 2:4] + T[2, :, 4:6]
    assert np.allclose(tr, d, atol=1e-13)
    # V_out(a, :, 2b-1:2b, :) = V1_a * V2_b
    x = np.zeros_like(a); y = np.zeros_like(b)
    x[:, 2, 1] = 1.0     # V1_theta = cos(theta)
    y[:, 4, 0] = 3.0     # V2_z = 3
    T = F.fft_tensor_dcl(x, y, 7)
    assert np.allclose(T[1, :, 4, 1], 3.0) and np.isclose(np.abs(T).sum(), 3.0 * 4)


def test_scalar_vect_both_problems():
    v, s = rnd((5, 6, 4), 13), rnd((5, 2, 4), 14)
    s[:, :, :] = 0
    s[:, 0, 0] = 2.0
    out1 = F.fft_scalar_vect_dcl(v, s, 1, 6)
    out2 = F.fft_scalar_vect_dcl(v, s, 2, 6)
    vv = v.copy(); vv[:, 1::2, 0] = 0
    assert np.allclose(out1, 2 * vv, atol=1e-13)
    assert np.allclose(out2, vv / 2, atol=1e-13)
    with pytest.raises(ValueError):
        F.fft_scalar_vect_dcl(v, s, 3, 6)


def test_error_convention():
    with pytest.raises(ValueError):
        F.fft_par_cross_prod_dcl(np.zeros((3, 5, 2)), np.zeros((3, 5, 2)), 3)
    with pytest.raises(ValueError):
        F.fft_par_dot_prod_dcl(np.zeros((3, 6, 0)), np.zeros((3, 6, 0)), 3)


def test_fft_par_real_planes():
    M = 4
    V = np.zeros((2, 2, M))
    V[:, 0, 0] = 1.0
    V[:, 1, 3] = 2.0      # 1 + 2 sin(3 theta)
    R = F.fft_par_real(V)
    N = 2 * M - 1
    th = 2 * np.pi * np.arange(N) / N
    assert np.allclose(R[:, 0, 0], 1 + 2 * np.sin(3 * th), atol=1e-13)


@pytest.mark.parametrize("mp", [4, 9, 40])
def test_fft_par_real_on_a_finer_grid(mp):
    """FFT_PAR_REAL with opt_nb_plane = 2*mp-1 > 2M-1: the values of
    a_0 + sum a_m cos(m t) + b_m sin(m t) at t_j = 2 pi j/N, N = 2*mp-1, summed
    here directly (the mode-0 sine is ignored)."""
    rng = np.random.default_rng(mp)
    M, npt = 4, 3
    V = rng.standard_normal((npt, 4, M))
    R = F.fft_par_real(V, mp)
    N = 2 * mp - 1
    t = 2 * np.pi * np.arange(N) / N
    m = np.arange(1, M)[:, None]
    for f in range(2):
        a, b = V[:, 2 * f], V[:, 2 * f + 1]
        ex = a[:, :1] + a[:, 1:] @ np.cos(m * t) + b[:, 1:] @ np.sin(m * t)
        assert R.shape == (N, 2, npt)
        assert np.max(np.abs(R[:, f, :].T - ex)) <= 1e-13 * np.max(np.abs(ex))
