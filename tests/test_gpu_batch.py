"""Batched persistent launch on the GPU through the C-ABI.

* every problem of a small batch vs smb_cross_prod_gauss_sft_par of the
  oracle (relative max-norm error <= 1e-12);
* every schedule (persistent lists or one grid per grid size) gives
  bit-identical results (a tile's arithmetic does not
  depend on which work-group runs it);
* hipGraph replays, with eager work in between, equal the eager launch bit
  for bit;
* the batched workload of 32 problems (2k-50k triangles, m_max in
  {4..128}) equals the single-problem kernel bit for bit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import smb  # noqa: E402
from sfemans_amd import batch as BT  # noqa: E402
from sfemans_amd import nonlinear as NL  # noqa: E402
from sfemans_amd._lib import SfxError  # noqa: E402
from sfemans_amd.fft_parallele import fortran_tensor, fortran_empty  # noqa: E402
from sfemans_amd.mesh import rectangle_p2  # noqa: E402
from sfemans_amd.synth import fourier_field  # noqa: E402

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


PROBS = [(2, 3, 4), (1, 2, 13), (3, 2, 8), (4, 5, 33), (1, 1, 1), (2, 1, 64), (6, 7, 128), (3, 3, 20)]


def _small_batch():
    items = []
    for k, (nx, ny, M) in enumerate(PROBS):
        m = rectangle_p2(nx, ny)
        mesh = NL.Mesh.from_p2(m)
        u = fourier_field(m.np, 6, M, 60 + k, 2.0)
        items.append((m, mesh, M, u, fortran_tensor(u), fortran_empty((m.n_gauss, 6, M))))
    return items


def test_batch_vs_oracle_and_schedules():
    items = _small_batch()
    print("case batch small: 8 problems", flush=True)
    b = BT.Batch([(mesh, list(range(M)), du, dout) for (_, mesh, M, _, du, dout) in items])
    b.run()
    torch.cuda.synchronize()
    first = [dout.cpu().numpy() for (*_, dout) in items]
    for (m, _, M, u, _, _), got in zip(items, first):
        exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, np.arange(M), u)
        assert rel(got, exp) <= TOL, M
    for policy in (BT.LPT, BT.CONTIGUOUS, BT.ROUND_ROBIN, BT.BY_SIZE):
        for n_wg in (1, 3, 256, 100000):
            print(f"case batch schedule policy={policy} n_wg={n_wg}", flush=True)
            for (*_, dout) in items:
                dout.fill_(np.nan)
            b.schedule(policy, n_wg)
            b.run()
            torch.cuda.synchronize()
            for (*_, dout), ref in zip(items, first):
                assert np.array_equal(dout.cpu().numpy(), ref)
    b.close()


def test_batch_autotune_and_graph_replays():
    items = _small_batch()
    b = BT.Batch([(mesh, list(range(M)), du, dout) for (_, mesh, M, _, du, dout) in items])
    print("case batch autotune", flush=True)
    b.autotune(reps=2)
    info = b.info()
    assert info["n_wg"] >= 1 and info["tuned_ms"] > 0
    b.run()
    torch.cuda.synchronize()
    eager = [dout.cpu().numpy() for (*_, dout) in items]
    print("case batch graph replays", flush=True)
    b.capture()
    assert b.info()["captured"]
    for rep in range(4):
        for (*_, dout) in items:
            dout.zero_()
        _ = torch.ones(1000, device="cuda", dtype=torch.float64).sum()   # eager work between replays
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for (*_, dout), ref in zip(items, eager):
            assert np.array_equal(dout.cpu().numpy(), ref), rep
    b.close()


def test_batch_argument_errors():
    m = rectangle_p2(2, 2)
    mesh = NL.Mesh.from_p2(m)
    du = fortran_tensor(fourier_field(m.np, 6, 4, 1, 2.0))
    with pytest.raises(ValueError):
        BT.Batch([(mesh, list(range(4)), du, fortran_empty((m.n_gauss, 6, 3)))])
    big = 300
    du2 = fortran_empty((m.np, 6, big))
    with pytest.raises(SfxError):
        BT.Batch([(mesh, list(range(big)), du2, fortran_empty((m.n_gauss, 6, big)))])
    b = BT.Batch([(mesh, list(range(4)), du, fortran_empty((m.n_gauss, 6, 4)))])
    with pytest.raises(SfxError):
        b.schedule(9, 4)
    with pytest.raises(SfxError):
        b.schedule(0, 0)
    b.close()


def uneven_problems(seed=0x5F3E + 3, n=32):
    """32 problems: triangles log-uniform in [2k, 50k] (nx = ny/2), m_max from {4..128}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        me = float(np.exp(rng.uniform(np.log(2000), np.log(50000))))
        ny = max(2, int(round(np.sqrt(me))))
        nx = max(1, ny // 2)
        M = int(rng.choice([4, 8, 16, 32, 64, 128]))
        out.append((nx, ny, M))
    return out


def test_batch_32_uneven_problems_equal_single_problem_kernel():
    probs = uneven_problems()
    items = []
    for k, (nx, ny, M) in enumerate(probs):
        m = rectangle_p2(nx, ny)
        mesh = NL.Mesh.from_p2(m)
        g = torch.Generator(device="cuda").manual_seed(100 + k)
        du = torch.randn((M, 6, m.np), generator=g, device="cuda", dtype=torch.float64).permute(2, 1, 0)
        items.append((m, mesh, M, du, fortran_empty((m.n_gauss, 6, M)), fortran_empty((m.n_gauss, 6, M))))
    print(f"case batch 32 uneven problems: {len(items)} problems, {sum(m.me for m, *_ in items)} triangles", flush=True)
    b = BT.Batch([(mesh, list(range(M)), du, out) for (m, mesh, M, du, out, _) in items])
    b.autotune(reps=1)
    b.run()
    for (m, mesh, M, du, _, ref) in items:
        NL.smb_cross_prod_gauss_sft_par(None, mesh, list(range(M)), du, ref)
    torch.cuda.synchronize()
    for (m, mesh, M, du, out, ref) in items:
        assert torch.equal(out, ref), (m.me, M)
    # the smallest problem against the oracle
    k = int(np.argmin([m.me * M for (m, _, M, *_) in items]))
    m, mesh, M, du, out, _ = items[k]
    u = du.cpu().numpy()
    u[:, 1, 0] = 0.0
    exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, np.arange(M), u)
    assert rel(out.cpu().numpy(), exp) <= TOL
    b.close()
