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
                               len(modes), out)
    assert rc == 0
    return arrs[-1]


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


@pytest.mark.parametrize("nx,ny,M", [(2, 3, 5), (3, 3, 16), (2, 2, 64)])
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
