"""Adjoint kernel on the CPU emulator: brute-force transpose and dot-product test."""
import numpy as np
import pytest

from oracle import linearized
from sfemans_amd.mesh import rectangle_p2
from sfemans_amd.synth import fourier_field
from tests import emu_runner


def emu_adjoint(m, U, Y, modes):
    out = np.zeros((m.np, 6, len(modes)), order="F")
    rc, arrs = emu_runner.call("emu_adjoint", m.jj, m.rr, m.me, m.np, U, Y, np.asarray(modes, np.int32),
                               len(modes), out, None, 1.0, None)
    assert rc == 0
    return arrs[-1]


def emu_adjoint_lorentz(m, H0, B0, cl, Y, modes):
    """(dHbar, dBbar) of the transpose of -cl [(curl H0) x W dB + (curl dH) x B0]."""
    dB = np.zeros((m.np, 6, len(modes)), order="F")
    dH = np.zeros((m.np, 6, len(modes)), order="F")
    rc, arrs = emu_runner.call("emu_adjoint", m.jj, m.rr, m.me, m.np, H0, Y, np.asarray(modes, np.int32),
                               len(modes), dB, B0, -cl, dH)
    assert rc == 0
    return arrs[-1], arrs[-3]


def emu_tangent(m, U, D, modes):
    out = np.zeros((m.n_gauss, 6, len(modes)), order="F")
    rc, arrs = emu_runner.call("emu_gauss", 4, m.jj, m.rr, m.me, m.np, U, 6, D, 6, None, None,
                               np.asarray(modes, np.int32), len(modes), out, 6, 0.0, 0.0, 0.0, 0.0)
    assert rc == 0
    return arrs[-1]


@pytest.mark.parametrize("M", [1, 3])
def test_adjoint_equals_bruteforce_transpose(M):
    m = rectangle_p2(1, 2)
    modes = np.arange(M)
    U = fourier_field(m.np, 6, M, 81, 2.0)
    Y = np.asfortranarray(np.random.default_rng(82).standard_normal((m.n_gauss, 6, M)))
    got = emu_adjoint(m, U, Y, modes)
    exp = linearized.adjoint_bruteforce(m.jj, m.rr, modes, U, Y)
    assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp))


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 5), (3, 3, 16), (2, 3, 20), (2, 2, 64), (1, 2, 100)])
def test_dot_product(nx, ny, M):
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    U = fourier_field(m.np, 6, M, 83, 2.0)
    D = fourier_field(m.np, 6, M, 84, 2.0)
    Y = np.asfortranarray(np.random.default_rng(85).standard_normal((m.n_gauss, 6, M)))
    TD = emu_tangent(m, U, D, modes)
    TtY = emu_adjoint(m, U, Y, modes)
    lhs, rhs = np.sum(TD * Y), np.sum(D * TtY)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(TD) * np.linalg.norm(Y)


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 5), (2, 2, 16)])
def test_lorentz_adjoint_dot_product(nx, ny, M):
    """<T_L(dH, dB), Y> = <dH, dHbar> + <dB, dBbar> with T_L the linearized Lorentz term."""
    from oracle import smb
    m = rectangle_p2(nx, ny)
    modes = np.arange(M)
    H0, B0, dH, dB = (fourier_field(m.np, 6, M, s, 2.0) for s in (86, 87, 88, 89))
    cl = 0.7
    Y = np.asfortranarray(np.random.default_rng(90).standard_normal((m.n_gauss, 6, M)))
    TL = -cl * (smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, H0, dB)
                + smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, modes, dH, B0))
    dHbar, dBbar = emu_adjoint_lorentz(m, H0, B0, cl, Y, modes)
    lhs, rhs = np.sum(TL * Y), np.sum(dH * dHbar) + np.sum(dB * dBbar)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(TL) * np.linalg.norm(Y)


def test_lorentz_adjoint_separates_the_two_fields():
    """dB only enters through W (values), dH only through the curl: with B0 = 0
    the dH adjoint vanishes, with H0 = 0 the dB adjoint vanishes (up to the
    rounding of complex signals that pack a zero and a non-zero field)."""
    m = rectangle_p2(2, 2)
    M = 4
    modes = np.arange(M)
    H0, B0 = (fourier_field(m.np, 6, M, s, 2.0) for s in (91, 92))
    Z = np.zeros_like(H0)
    Y = np.asfortranarray(np.random.default_rng(93).standard_normal((m.n_gauss, 6, M)))
    dHbar, dBbar = emu_adjoint_lorentz(m, H0, Z, 1.0, Y, modes)
    assert np.max(np.abs(dHbar)) <= 1e-12 * np.max(np.abs(dBbar))
    dHbar, dBbar = emu_adjoint_lorentz(m, Z, B0, 1.0, Y, modes)
    assert np.max(np.abs(dBbar)) <= 1e-12 * np.max(np.abs(dHbar))
