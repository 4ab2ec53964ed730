"""The allocation-free form of dot_product_SF against the workspace form.

sfx_dot_product_sf_noalloc keeps its partial sums in a fixed device array
instead of a caller's workspace; it runs the same kernels in the same order,
so its result must equal sfx_dot_product_sf_dev's bit for bit.  Two calls in
a row, the second on a smaller input, check that nothing of the first call's
contents leaks into the second.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from sfemans_amd import nonlinear as NL  # noqa: E402
from sfemans_amd._lib import check, lib  # noqa: E402
from sfemans_amd.fft_parallele import fortran_tensor  # noqa: E402
from sfemans_amd.mesh import rectangle_p2  # noqa: E402
from sfemans_amd.synth import fourier_field  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


def _both(mesh, m, modes, nbf, seed):
    M = len(modes)
    v, w = fourier_field(m.np, nbf, M, seed, 2.0), fourier_field(m.np, nbf, M, seed + 1, 2.0)
    dv, dw = fortran_tensor(v), fortran_tensor(w)
    lm = mesh.list_mode_device(modes)
    L = lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    res = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    work = torch.empty(int(L.sfx_dot_product_sf_work(mesh.handle, M, nbf)), dtype=torch.float64, device="cuda")
    check(L.sfx_dot_product_sf_dev(mesh.handle, lm.data_ptr(), M, nbf, dv.data_ptr(), dw.data_ptr(),
                                   res.data_ptr(), work.data_ptr(), stream))
    check(L.sfx_dot_product_sf_noalloc(mesh.handle, lm.data_ptr(), M, nbf, dv.data_ptr(), dw.data_ptr(),
                                       res.data_ptr() + 8, stream))
    torch.cuda.synchronize()
    a, b = res.cpu().numpy()
    return a, b


def test_noalloc_matches_workspace_form_twice():
    big = rectangle_p2(96, 160)           # 30,720 triangles, m_max = 16: 2048 partial sums, strided
    mesh_b = NL.Mesh.from_p2(big)
    a, b = _both(mesh_b, big, list(range(16)), 6, 61)
    assert np.isfinite(a) and a == b
    small = rectangle_p2(5, 3)            # second call on a smaller input: 2 partial sums
    mesh_s = NL.Mesh.from_p2(small)
    a2, b2 = _both(mesh_s, small, [0, 2, 3], 2, 62)
    assert np.isfinite(a2) and a2 == b2 and a2 != a
