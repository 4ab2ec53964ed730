"""Batched persistent launch on the CPU emulator vs the oracle (CPU only).

Every problem of the batch must equal smb_cross_prod_gauss_sft_par of the
oracle (relative max-norm error <= 1e-12) under every scheduling policy; the
scheduler's lists must cover every tile exactly once.
"""
import numpy as np
import pytest

from oracle import smb
from sfemans_amd.mesh import rectangle_p2
from sfemans_amd.synth import fourier_field
from tests import emu_runner

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# (nx, ny, M): uneven meshes and mode counts, several grid sizes in one batch
PROBS = [(2, 3, 4), (1, 2, 13), (3, 2, 8), (2, 2, 33), (1, 1, 1), (2, 1, 64)]


def _pack(probs):
    meshes, fields, outs = [], [], []
    for k, (nx, ny, M) in enumerate(probs):
        m = rectangle_p2(nx, ny)
        meshes.append(m)
        fields.append(fourier_field(m.np, 6, M, 40 + k, 2.0))
        outs.append(np.zeros((m.n_gauss, 6, M), order="F"))

    def flat(arrs, dtype):
        off = np.cumsum([0] + [a.size for a in arrs[:-1]]).astype(np.int64)
        return np.concatenate([np.asarray(a, dtype=dtype).ravel(order="F") for a in arrs]), off

    jj, jj_off = flat([m.jj for m in meshes], np.int32)
    rr, rr_off = flat([m.rr for m in meshes], np.float64)
    U, U_off = flat(fields, np.float64)
    lm, lm_off = flat([np.arange(M) for (_, _, M) in probs], np.int32)
    out, out_off = flat(outs, np.float64)
    me = np.array([m.me for m in meshes], dtype=np.int64)
    npn = np.array([m.np for m in meshes], dtype=np.int64)
    Ms = np.array([M for (_, _, M) in probs], dtype=np.int32)
    return meshes, fields, (jj, jj_off, rr, rr_off, me, npn, U, U_off, lm, lm_off, Ms, out, out_off)


@pytest.mark.parametrize("policy,n_wg", [(0, 3), (1, 4), (2, 5), (0, 1000), (3, 1)])
def test_batch_matches_oracle(policy, n_wg):
    meshes, fields, args = _pack(PROBS)
    used = np.zeros(1, dtype=np.int32)
    rc, arrs = emu_runner.call("emu_batch", len(PROBS), *args, policy, n_wg, used)
    assert rc == 0
    out = arrs[11]
    out_off = args[12]
    for k, (m, u, (_, _, M)) in enumerate(zip(meshes, fields, PROBS)):
        n = m.n_gauss * 6 * M
        got = out[out_off[k]:out_off[k] + n].reshape((m.n_gauss, 6, M), order="F")
        exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, np.arange(M), u)
        assert rel(got, exp) <= TOL, (k, M)


def test_unknown_policy_is_rejected():
    _, _, args = _pack(PROBS[:1])
    used = np.zeros(1, dtype=np.int32)
    rc, _ = emu_runner.call("emu_batch", 1, *args, 7, 2, used)
    assert rc == 1
