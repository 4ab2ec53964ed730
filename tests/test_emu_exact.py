"""Exact-grid kernel bodies on the CPU emulator vs the oracle (CPU only).

The non-polynomial FFT_PAR_* products run on N = 2*m_max_pad-1 angles as
real DFT sums (sfx_exact_body.h); the oracle evaluates the same products
with pocketfft on the same grid.  Relative max-norm error <= 1e-12.
"""
import numpy as np
import pytest

from oracle import fft_par as F
from tests import emu_runner

EX_DIV, EX_SV2, EX_SQRT, EX_HEAVI, EX_TENSION, EX_MAXVEL = range(6)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def emu(op, np_, M, mp, in1, in2=None, values=None, eps=0.0, nb=2, out=None):
    nv = 0 if values is None else len(values)
    vals = None if values is None else np.asarray(values, dtype=np.float64)
    rc, arrs = emu_runner.call("emu_exact", op, np_, M, mp, in1, in2, vals, nv, eps, nb, out)
    assert rc == 0
    return arrs[-1]


def fields(rng, npt, nb, M, scale=1.0):
    V = rng.standard_normal((npt, nb, M)) * scale * np.exp(-np.arange(M) / 4.0)
    return np.asfortranarray(V)


CASES = [(5, 4, 6), (17, 8, 12), (9, 13, 20), (33, 16, 24), (3, 1, 1), (4, 32, 48), (2, 64, 96)]


@pytest.mark.parametrize("npt,M,mp", CASES)
def test_div(npt, M, mp):
    rng = np.random.default_rng(M)
    c1 = fields(rng, npt, 2, M)
    c2 = fields(rng, npt, 2, M, 0.1)
    c2[:, 0, 0] = 2.0 + np.abs(c2[:, 0, 0])
    out = emu(EX_DIV, npt, M, mp, c1, c2, out=np.zeros((npt, 2, M), order="F"))
    assert rel(out, F.fft_par_div_dcl(c1, c2, mp)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp", CASES[:5])
def test_scalar_vect_pb2(npt, M, mp):
    rng = np.random.default_rng(10 + M)
    V = fields(rng, npt, 6, M)
    s = fields(rng, npt, 2, M, 0.1)
    s[:, 0, 0] = 1.5
    out = emu(EX_SV2, npt, M, mp, V, s, out=np.zeros((npt, 6, M), order="F"))
    assert rel(out, F.fft_scalar_vect_dcl(V, s, 2, mp)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp,nb", [(7, 5, 8, 2), (11, 16, 24, 4), (3, 64, 96, 2)])
def test_squareroot(npt, M, mp, nb):
    rng = np.random.default_rng(20 + M)
    c = fields(rng, npt, nb, M, 0.4)
    c[:, 0, 0] = 0.5           # values in and out of [0, 1]: the clamp is active
    out = emu(EX_SQRT, npt, M, mp, c, nb=nb, out=np.asfortranarray(c.copy()))
    exp = c.copy()
    exp[:, 0:2] = F.fft_squareroot_scalar(c, mp)
    assert rel(out, exp) <= 1e-12


@pytest.mark.parametrize("npt,M,mp,n_int", [(6, 4, 6, 1), (9, 8, 12, 2), (5, 16, 24, 3), (4, 32, 48, 4)])
def test_heaviside(npt, M, mp, n_int):
    rng = np.random.default_rng(30 + M)
    V = rng.standard_normal((n_int, npt, 2, M)) * 0.3 * np.exp(-np.arange(M) / 3.0)
    V[:, :, 0, 0] += 0.5
    V = np.asfortranarray(V)
    vals = list(np.linspace(1.0, 5.0, n_int + 1))
    eps = 0.2
    out = emu(EX_HEAVI, npt, M, mp, V, values=vals, eps=eps, out=np.zeros((npt, 2, M), order="F"))
    assert rel(out, F.fft_heaviside_dcl(V, vals, mp, eps)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp", CASES[:4])
def test_tension(npt, M, mp):
    rng = np.random.default_rng(40 + M)
    U = fields(rng, npt, 6, M)
    G = fields(rng, npt, 6, M)
    out = emu(EX_TENSION, npt, M, mp, U, G, out=np.zeros((3, npt, 6, M), order="F"))
    assert rel(out, F.fft_tensor_dcl(U, G, mp, opt_tension=True)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp,nb", [(13, 8, 12, 6), (7, 16, 24, 2), (300, 5, 8, 6)])
def test_max_vel(npt, M, mp, nb):
    rng = np.random.default_rng(50 + M)
    V = fields(rng, npt, nb, M)
    out = emu(EX_MAXVEL, npt, M, mp, V, nb=nb, out=np.zeros(1))
    assert abs(out[0] - F.fft_max_vel_dcl(V, mp)) <= 1e-13 * F.fft_max_vel_dcl(V, mp)


EX_SVCLAMP, EX_CLAMP01 = 6, 7


@pytest.mark.parametrize("npt,M,mp,pb", [(9, 8, 12, 1), (7, 16, 24, 2), (3, 64, 96, 1)])
def test_scalar_vect_no_overshoot(npt, M, mp, pb):
    rng = np.random.default_rng(70 + M)
    V = fields(rng, npt, 6, M)
    s = fields(rng, npt, 2, M, 0.4)
    s[:, 0, 0] = 1.5
    bounds = [1.2, 1.8]
    out = emu(EX_SVCLAMP, npt, M, mp, V, s, values=bounds, nb=pb, out=np.zeros((npt, 6, M), order="F"))
    assert rel(out, F.fft_scalar_vect_no_overshoot(bounds, V, s, pb, mp)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp,nb", [(7, 5, 8, 2), (11, 32, 48, 4)])
def test_no_overshoot_level_set(npt, M, mp, nb):
    rng = np.random.default_rng(80 + M)
    c = fields(rng, npt, nb, M, 0.4)
    c[:, 0, 0] = 0.5
    out = emu(EX_CLAMP01, npt, M, mp, c, nb=nb, out=np.asfortranarray(c.copy()))
    exp = c.copy()
    exp[:, 0:2] = F.fft_no_overshoot_level_set(c, mp)
    assert rel(out, exp) <= 1e-12


EX_ENTRO = 8


@pytest.mark.parametrize("npt,M,mp", [(9, 8, 12), (5, 16, 24), (3, 64, 96)])
def test_entro_visc(npt, M, mp):
    rng = np.random.default_rng(90 + M)
    V = fields(rng, npt, 6, M)
    c = fields(rng, npt, 2, M, 0.3)
    c[:, 0, 0] = 0.5
    N = 2 * mp - 1
    creal = rng.standard_normal((N, npt))
    packed = np.concatenate([c.ravel(order="F"), creal.ravel(order="F")])
    out = emu(EX_ENTRO, npt, M, mp, V, packed, values=[0.3], out=np.zeros(npt * 8 * M))
    Vo, co = F.fft_par_entro_visc_dcl(V, c, creal, 0.3, mp)
    assert rel(out[:npt * 6 * M].reshape((npt, 6, M), order="F"), Vo) <= 1e-12
    assert rel(out[npt * 6 * M:].reshape((npt, 2, M), order="F"), co) <= 1e-12


EX_GRADMOM = 9


@pytest.mark.parametrize("npt,M,mp", [(7, 8, 12), (3, 32, 48)])
def test_entropy_visc_grad_mom(npt, M, mp):
    rng = np.random.default_rng(95 + M)
    Vs = [fields(rng, npt, 6, M) for _ in range(3)]
    N = 2 * mp - 1
    creal = rng.standard_normal((N, npt))
    packed = np.concatenate([v.ravel(order="F") for v in Vs] + [creal.ravel(order="F")])
    out = emu(EX_GRADMOM, npt, M, mp, packed, values=[0.25], out=np.zeros((3, npt, 6, M), order="F"))
    assert rel(out, F.fft_compute_entropy_visc_grad_mom(*Vs, creal, 0.25, mp)) <= 1e-12


EX_COMPR = 10


@pytest.mark.parametrize("npt,M,mp", [(7, 8, 12), (3, 32, 48)])
def test_compression_entropy_viscosity(npt, M, mp):
    rng = np.random.default_rng(97 + M)
    V1, V2 = fields(rng, npt, 6, M), fields(rng, npt, 6, M)
    c = fields(rng, npt, 2, M, 0.3)
    c[:, 0, 0] = 0.5
    N = 2 * mp - 1
    creal = rng.standard_normal((N, npt))
    in1 = np.concatenate([V1.ravel(order="F"), V2.ravel(order="F")])
    in2 = np.concatenate([c.ravel(order="F"), creal.ravel(order="F")])
    out = emu(EX_COMPR, npt, M, mp, in1, in2, values=[0.3, 1.7], out=np.zeros(npt * 8 * M))
    Vo, co = F.fft_par_compr_entro_visc_dcl(V1, V2, c, creal, 0.3, 1.7, mp)
    assert rel(out[:npt * 6 * M].reshape((npt, 6, M), order="F"), Vo) <= 1e-12
    assert rel(out[npt * 6 * M:].reshape((npt, 2, M), order="F"), co) <= 1e-12


@pytest.mark.parametrize("ne,M,mp,lg", [(5, 8, 12, 7), (40, 4, 6, 7), (3, 32, 48, 3)])
def test_entropy_visc(ne, M, mp, lg):
    rng = np.random.default_rng(99 + M)
    npt = ne * lg
    Vs = [fields(rng, npt, 6, M) for _ in range(5)]
    h = rng.uniform(0.1, 1.0, npt)
    packed = np.concatenate([v.ravel(order="F") for v in Vs])
    rc, arrs = emu_runner.call("emu_exact_evisc", npt, M, mp, lg, packed, h, 0.3, 0.9, 0.8, 1.7,
                               np.zeros(3 * npt * 6 * M))
    assert rc == 0
    out = arrs[-1]
    exp = F.fft_compute_entropy_visc(*Vs, h, 1.7, lg, 0.3, 0.9, 0.8, mp)
    for k in range(3):
        got = out[k * npt * 6 * M:(k + 1) * npt * 6 * M].reshape((npt, 6, M), order="F")
        assert rel(got, exp[k]) <= 1e-12


@pytest.mark.parametrize("ne,M,mp,lg,c2out", [(5, 8, 12, 7, True), (40, 4, 6, 7, False), (6, 16, 24, 3, False)])
def test_entropy_visc_mom(ne, M, mp, lg, c2out):
    rng = np.random.default_rng(101 + M)
    npt = ne * lg
    Vs = [fields(rng, npt, 6, M) for _ in range(3)]
    c = fields(rng, npt, 2, M)
    h = rng.uniform(0.1, 1.0, npt)
    N = 2 * mp - 1
    packed = np.concatenate([v.ravel(order="F") for v in Vs])
    o2 = np.zeros((N, npt), order="F") if c2out else None
    rc, arrs = emu_runner.call("emu_exact_emom", npt, M, mp, lg, packed, c, h, 0.6, 0.9,
                               np.zeros((N, npt), order="F"), o2)
    assert rc == 0
    e1, e2 = F.fft_compute_entropy_visc_mom(*Vs, c, h, lg, 0.6, 0.9, mp, c2out)
    assert rel(arrs[3 if not c2out else -2], e1) <= 1e-12
    if c2out:
        assert rel(arrs[-1], e2) <= 1e-12


EX_CLS = 13


@pytest.mark.parametrize("npt,M,mp,vs", [(9, 8, 12, 0.3), (5, 16, 24, 3.0), (3, 64, 96, 0.3)])
def test_compression_level_set(npt, M, mp, vs):
    rng = np.random.default_rng(103 + M)
    G, V = fields(rng, npt, 6, M), fields(rng, npt, 6, M, vs)
    c = fields(rng, npt, 2, M, 0.2)
    c[:, 0, 0] = 0.5
    in1 = np.concatenate([G.ravel(order="F"), V.ravel(order="F")])
    vals = [1.3, 0.7, 0.05]
    out = emu(EX_CLS, npt, M, mp, in1, c, values=vals, out=np.zeros((npt, 2, M), order="F"))
    assert rel(out, F.fft_compression_level_set_dcl(G, V, c, 1, *vals, mp)) <= 1e-12


EX_REAL = 14


@pytest.mark.parametrize("npt,M,mp,nb", [(9, 8, 8, 6), (5, 16, 40, 2), (3, 1, 1, 4), (4, 64, 64, 2)])
def test_fft_par_real(npt, M, mp, nb):
    rng = np.random.default_rng(107 + M)
    V = fields(rng, npt, nb, M)
    N = 2 * mp - 1
    out = emu(EX_REAL, npt, M, mp, V, nb=nb, out=np.zeros((N, nb // 2, npt), order="F"))
    assert rel(out, F.fft_par_real(V, mp)) <= 1e-12


EX_VELSTAT, EX_CHECKIF = 15, 16


@pytest.mark.parametrize("npt,M,mp,nb", [(13, 8, 12, 6), (7, 16, 24, 2), (300, 5, 8, 6), (5, 4, 6, 4)])
def test_max_min_vel(npt, M, mp, nb):
    rng = np.random.default_rng(110 + M)
    V = fields(rng, npt, nb, M)
    out = emu(EX_VELSTAT, npt, M, mp, V, nb=nb, out=np.zeros(2))
    vmax, vmin = F.fft_max_min_vel_dcl(V, mp)
    assert abs(out[0] - vmax) <= 1e-13 * vmax and abs(out[1] - vmin) <= 1e-13 * vmax


@pytest.mark.parametrize("npt,M,mp,nf,shift", [(9, 8, 12, 3, 0.0), (9, 8, 12, 3, 0.4), (6, 16, 24, 5, 0.3),
                                              (4, 4, 6, 2, 0.0)])
def test_check_interface(npt, M, mp, nf, shift):
    rng = np.random.default_rng(120 + M)
    V = rng.standard_normal((nf - 1, npt, 2, M)) * 0.01 * np.exp(-np.arange(M) / 3.0)
    for i in range(nf - 1):
        V[i, :, 0, 0] = 0.8 - 0.2 * i + (shift if i == nf - 2 else 0.0)   # ordered unless shifted up
    V = np.asfortranarray(V)
    out = emu(EX_CHECKIF, npt, M, mp, V, nb=nf, out=np.zeros(1))
    assert out[0] == F.fft_check_interface(V, nf, mp)


EX_ETA, EX_ETA3, EX_FROMREAL = 17, 18, 19


@pytest.mark.parametrize("npt,M,mp,nb", [(9, 8, 12, 2), (7, 16, 24, 6), (3, 64, 96, 6), (4, 1, 1, 2)])
def test_var_eta_prod(npt, M, mp, nb):
    rng = np.random.default_rng(130 + M)
    c = fields(rng, npt, nb, M)
    eta = np.asfortranarray(rng.uniform(0.5, 2.0, (2 * mp - 1, npt)))
    out = emu(EX_ETA if nb == 2 else EX_ETA3, npt, M, mp, c, eta, nb=nb, out=np.zeros((npt, nb, M), order="F"))
    assert rel(out, F.fft_par_var_eta_prod(eta, c, mp)) <= 1e-12


@pytest.mark.parametrize("npt,M,mp", [(9, 8, 12), (5, 32, 48), (2, 3, 3)])
def test_from_real_inverts_real_planes(npt, M, mp):
    """The analysis of FFT_PAR_REAL's planes gives the coefficients back
    (mode-0 sine zero), and the analysis of any values X matches the oracle's
    product of X with the constant 1."""
    rng = np.random.default_rng(140 + M)
    c = fields(rng, npt, 2, M)
    c[:, 1, 0] = 0.0
    R = np.asfortranarray(F.fft_par_real(c, mp)[:, 0, :])
    out = emu(EX_FROMREAL, npt, M, mp, c, R, out=np.zeros((npt, 2, M), order="F"))
    assert rel(out, c) <= 1e-12
    X = np.asfortranarray(rng.standard_normal((2 * mp - 1, npt)))
    out = emu(EX_FROMREAL, npt, M, mp, c, X, out=np.zeros((npt, 2, M), order="F"))
    one = np.zeros((npt, 2, M))
    one[:, 0, 0] = 1.0                       # the constant 1: eta * 1 = eta
    assert rel(out, F.fft_par_var_eta_prod(X, one, mp)) <= 1e-12
