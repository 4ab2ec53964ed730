"""FFT_PAR_* kernel bodies on the CPU emulator vs the oracle (CPU only).

Same kernel source as the GPU library, every index bounds-checked; ragged
point counts exercise partially filled waves.
"""
import numpy as np
import pytest

from oracle import fft_par as F
from tests import emu_runner
from tests.known_answers import ka_test1_fields, ka_test1_rot_cross_v

OPS = {  # op id, nb1, nb2, nbout, oracle
    "cross": (0, 6, 6, 6, lambda a, b, mp: F.fft_par_cross_prod_dcl(a, b, mp)),
    "dot": (1, 6, 6, 2, lambda a, b, mp: F.fft_par_dot_prod_dcl(a, b, mp)),
    "prod": (2, 2, 2, 2, lambda a, b, mp: F.fft_par_prod_dcl(a, b, mp)),
    "svect1": (3, 6, 2, 6, lambda a, b, mp: F.fft_scalar_vect_dcl(a, b, 1, mp)),
    "tensor": (4, 6, 6, 18, lambda a, b, mp: F.fft_tensor_dcl(a, b, mp)),
}


def run(op, V1, V2, wpb=2):
    opid, nb1, nb2, nbo, _ = OPS[op]
    npt, _, M = V1.shape
    out = np.zeros((3, npt, 6, M) if op == "tensor" else (npt, nbo, M), order="F")
    rc, arrs = emu_runner.call("emu_fftprod", opid, V1, V2, out, npt, M, nb1, nb2, nbo, wpb)
    assert rc == 0
    return arrs[2]


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("M,npt", [(1, 3), (3, 7), (8, 9), (13, 5), (64, 6), (128, 3)])
def test_emulated_products_match_oracle(op, M, npt):
    _, nb1, nb2, _, ref = OPS[op]
    rng = np.random.default_rng(M * 31 + npt)
    V1 = np.asfortranarray(rng.standard_normal((npt, nb1, M)))
    V2 = np.asfortranarray(rng.standard_normal((npt, nb2, M)))
    got = run(op, V1, V2)
    exp = ref(V1, V2, max(3 * M // 2, M))
    assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp))


def test_emulated_test1_known_answer():
    pts = np.random.default_rng(7).uniform([0.05, -1.0], [1.0, 1.0], size=(13, 2))
    V, Rot = ka_test1_fields(pts[:, 0], pts[:, 1])
    V1 = np.zeros((13, 6, 3), order="F"); V2 = np.zeros((13, 6, 3), order="F")
    V1[:, :, 1] = Rot
    V2[:, :, 1] = V
    out = run("cross", V1, V2)
    e0, e2 = ka_test1_rot_cross_v(V, Rot)
    s = np.max(np.abs(e2))
    assert np.max(np.abs(out[:, :, 0] - e0)) <= 1e-13 * s
    assert np.max(np.abs(out[:, :, 2] - e2)) <= 1e-13 * s
    assert np.max(np.abs(out[:, :, 1])) <= 1e-13 * s
