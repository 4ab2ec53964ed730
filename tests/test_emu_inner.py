"""The dot_product_SF, Gauss-value and column-dot kernel bodies on the CPU
emulator vs the oracle (CPU only), and the sanitizer build of the emulator.

* dot_product_SF (SOURCES/SOLVER/tn_axi.F90:9-37) for work-group counts 1, 3,
  7 and the library's own count: the strided split changes only the order of
  the sum, relative error <= 1e-13 against the oracle.
* The weighted Gauss values R^T z (rj*ray per Gauss point, mode weights per
  column) against oracle.inner.gauss_weighted, straight and curved meshes.
* The sanitizer self-check: the ASan+UBSan build stops on a planted
  out-of-bounds read and a planted signed overflow (host code of the
  emulator, never GPU code), and runs one dot product cleanly.
"""
import numpy as np
import pytest

from oracle import inner as OI
from sfemans_amd.mesh import rectangle_p2, sector_p2
from sfemans_amd.synth import fourier_field
from tests import emu_runner


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _mesh(kind):
    return rectangle_p2(4, 5) if kind == "rect" else sector_p2(3, 5)


def _lib_blocks(me, ncols):
    # the library's count: one work-group per 256 (element, column) items, at most 2048
    items = me * ncols
    return min((items + 255) // 256, 2048)


@pytest.mark.parametrize("kind", ["rect", "sector"])
@pytest.mark.parametrize("nblk", [1, 3, 7, None])
def test_dotsf_any_split(kind, nblk):
    m = _mesh(kind)
    modes = np.array([0, 2, 3, 5], dtype=np.int32)
    M, nc = len(modes), 6
    v, w = fourier_field(m.np, nc, M, 71, 2.0), fourier_field(m.np, nc, M, 72, 2.0)
    nb = _lib_blocks(m.me, nc * M) if nblk is None else nblk
    out = np.zeros(1)
    rc, arrs = emu_runner.call("emu_dotsf", m.jj, m.rr, m.me, m.np, modes, M, nc, v, w, nb, out)
    assert rc == 0
    exp = OI.dot_product_sf(m.jj, m.rr, modes, v, w)
    assert abs(arrs[-1][0] - exp) <= 1e-13 * abs(exp)


def test_dotsf_more_blocks_than_items():
    """Work-groups beyond the items contribute zero partials."""
    m = rectangle_p2(1, 1)
    modes = np.array([0], dtype=np.int32)
    v, w = fourier_field(m.np, 2, 1, 73, 2.0), fourier_field(m.np, 2, 1, 74, 2.0)
    out = np.zeros(1)
    rc, arrs = emu_runner.call("emu_dotsf", m.jj, m.rr, m.me, m.np, modes, 1, 2, v, w, 5, out)
    assert rc == 0
    exp = OI.dot_product_sf(m.jj, m.rr, modes, v, w)
    assert abs(arrs[-1][0] - exp) <= 1e-13 * abs(exp)


@pytest.mark.parametrize("kind", ["rect", "sector"])
def test_weighted_gauss_values(kind):
    m = _mesh(kind)
    modes = np.array([0, 1, 4], dtype=np.int32)
    M, nc = len(modes), 6
    z = fourier_field(m.np, nc, M, 75, 2.0)
    out = np.zeros((m.n_gauss, nc, M), order="F")
    rc, arrs = emu_runner.call("emu_gauss_values", m.jj, m.rr, m.me, m.np, modes, M, nc, 1, 1.0, 0.5, z, out)
    assert rc == 0
    wk = OI.mode_weights(modes)
    exp = OI.gauss_weighted(m.jj, m.rr, z * wk[None, None, :])
    assert rel(arrs[-1], exp) <= 1e-13
    # unweighted, unit scale: the plain Gauss values, whose sum over an element's
    # points times rj*ray is the weighted form
    out2 = np.zeros((m.n_gauss, nc, M), order="F")
    rc, arrs2 = emu_runner.call("emu_gauss_values", m.jj, m.rr, m.me, m.np, modes, M, nc, 0, 1.0, 1.0, z, out2)
    assert rc == 0
    ones = np.ones((m.np, 1))
    wt = OI.gauss_weighted(m.jj, m.rr, ones)[:, 0]
    assert rel(arrs2[-1] * wt[:, None, None], OI.gauss_weighted(m.jj, m.rr, z)) <= 1e-13


def test_coldot():
    rng = np.random.default_rng(76)
    A = np.asfortranarray(rng.standard_normal((1000, 5)))
    B = np.asfortranarray(rng.standard_normal((1000, 5)))
    out = np.zeros(5)
    rc, arrs = emu_runner.call("emu_coldot", A, B, 1000, 5, out)
    assert rc == 0
    assert rel(arrs[-1], np.sum(A * B, axis=0)) <= 1e-13


def test_sanitizer_build_catches_planted_errors():
    rc, _ = emu_runner.call("emu_sanitizer_probe", 0, 8, sanitize=True)
    assert rc == 1
    with pytest.raises(RuntimeError, match="heap-buffer-overflow"):
        emu_runner.call("emu_sanitizer_probe", 1, 8, sanitize=True)
    with pytest.raises(RuntimeError, match="signed integer overflow"):
        emu_runner.call("emu_sanitizer_probe", 2, 8, sanitize=True)
    m = rectangle_p2(2, 2)
    modes = np.array([0, 1], dtype=np.int32)
    v, w = fourier_field(m.np, 2, 2, 77, 2.0), fourier_field(m.np, 2, 2, 78, 2.0)
    rc, arrs = emu_runner.call("emu_dotsf", m.jj, m.rr, m.me, m.np, modes, 2, 2, v, w, 3, np.zeros(1),
                               sanitize=True)
    exp = OI.dot_product_sf(m.jj, m.rr, modes, v, w)
    assert rc == 0 and abs(arrs[-1][0] - exp) <= 1e-13 * abs(exp)
