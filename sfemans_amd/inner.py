"""Python mirror of the inner product of SFEMaNS fields and of the adjoint in it.

``dot_product_SF`` follows the reference's ``dot_product_SF``
(SOURCES/SOLVER/tn_axi.F90:9-37), the inner product LightKrylov uses on
SFEMaNS fields (SOURCES/TOOLS/def_type_field.F90:176-198); ``communicator``
must be None (one rank).  ``nl_tangent_proj`` and ``nl_adjoint_sf`` are the
L2-projected tangent of the convective term and its adjoint in that inner
product (no reference counterpart).  Device arrays are Fortran-strided CUDA
tensors (np, nb_field, m); work buffers are allocated when omitted.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .fft_parallele import fdims, is_device, ptr, stream_ptr

__all__ = ["dot_product_SF", "mass_apply", "mass_solve", "nl_tangent_proj", "nl_adjoint_sf"]


def dot_product_SF(communicator, mesh, list_mode, v, w):
    """2 pi SUM_k w_k SUM_c (v, w)_{r dr dz} over nodal fields (np, nb_field, m)."""
    if communicator is not None and getattr(communicator, "size", 1) != 1:
        raise ValueError("one GPU: the communicators must have a single rank")
    npts, nbf, M = fdims(v)
    L = lib()
    if is_device(v):
        import torch
        work = torch.empty(int(L.sfx_dot_product_sf_work(mesh.handle, M, nbf)), dtype=torch.float64, device="cuda")
        res = torch.empty(1, dtype=torch.float64, device="cuda")
        check(L.sfx_dot_product_sf_dev(mesh.handle, mesh.list_mode_device(list_mode).data_ptr(), M, nbf, ptr(v),
                                       ptr(w), res.data_ptr(), work.data_ptr(), stream_ptr()))
        return float(res.item())
    lm = np.ascontiguousarray(list_mode, dtype=np.int32)
    out = C.c_double()
    check(L.sfx_dot_product_sf(mesh.handle, lm.ctypes.data, M, nbf, ptr(np.asfortranarray(v)),
                               ptr(np.asfortranarray(w)), C.byref(out)))
    return out.value


def mass_apply(mesh, x, y, work_elem=None):
    """y = M x, M the r-weighted P2 mass matrix, for the columns of (np, ...) device arrays."""
    import torch
    nc = x.numel() // mesh.np
    if work_elem is None:
        work_elem = torch.empty(6 * mesh.me * nc, dtype=torch.float64, device="cuda")
    check(lib().sfx_mass_apply_dev(mesh.handle, nc, ptr(x), ptr(y), work_elem.data_ptr(), stream_ptr()))
    return y


def mass_solve(mesh, b, x, rtol=1e-14, max_iter=0, work=None):
    """x = M^{-1} b column by column (device arrays); returns the iteration count."""
    import torch
    nc = b.numel() // mesh.np
    L = lib()
    if work is None:
        work = torch.empty(int(L.sfx_mass_solve_work(mesh.handle, nc)), dtype=torch.float64, device="cuda")
    it = C.c_int32()
    check(L.sfx_mass_solve_dev(mesh.handle, nc, ptr(b), ptr(x), float(rtol), int(max_iter), work.data_ptr(),
                               C.byref(it), stream_ptr()))
    return it.value


def _proj_work(mesh, M, work):
    import torch
    if work is None:
        work = torch.empty(int(lib().sfx_nl_proj_work(mesh.handle, M)), dtype=torch.float64, device="cuda")
    return work


def nl_tangent_proj(mesh, list_mode, U, dU, out, rtol=1e-14, work=None):
    """A dU = M^{-1} R T_U dU (nodal (np, 6, m) device arrays)."""
    M = len(list_mode)
    work = _proj_work(mesh, M, work)
    check(lib().sfx_nl_tangent_proj_dev(mesh.handle, mesh.list_mode_device(list_mode).data_ptr(), M, ptr(U), ptr(dU),
                                        ptr(out), float(rtol), work.data_ptr(), stream_ptr()))
    return out


def nl_adjoint_sf(mesh, list_mode, U, ybar, out, rtol=1e-14, work=None):
    """A* ybar, the adjoint of :func:`nl_tangent_proj` in dot_product_SF."""
    M = len(list_mode)
    work = _proj_work(mesh, M, work)
    check(lib().sfx_nl_adjoint_sf_dev(mesh.handle, mesh.list_mode_device(list_mode).data_ptr(), M, ptr(U), ptr(ybar),
                                      ptr(out), float(rtol), work.data_ptr(), stream_ptr()))
    return out
