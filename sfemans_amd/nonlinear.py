"""Python mirror of the reference's Gauss-point nonlinear-term routines.

Routine names and argument order follow SOURCES/SIMULATION:
``smb_cross_prod_gauss_sft_par`` and ``smb_CurlH_cross_B_gauss_sft_par``
(subroutine_ns_with_velocity.F90:514, 632) and ``smb_ugradc_gauss_fft_par``
(subroutine_temperature_with_T.F90:247).  ``communicator`` must be ``None``
(one rank).  ``mesh`` is a :class:`Mesh` registered once from the reference
arrays ``jj(6,me)`` and ``rr(2,np)``.  Outputs are INTENT(OUT) arrays filled
in place: Fortran-ordered numpy arrays (host path) or Fortran-strided CUDA
tensors (device path, launched on torch's current stream).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib, check
from .fft_parallele import fdims, is_device, ptr, stream_ptr

__all__ = ["Mesh", "smb_cross_prod_gauss_sft_par", "nl_rhs_3x3", "nl_qs_00", "smb_CurlH_cross_B_gauss_sft_par",
           "smb_ugradc_gauss_fft_par", "nl_rotv_mhd", "rhs_3x3", "rhs_3x3_all_modes", "rhs_3x3_bdy",
           "qs_00_gauss", "nl_rhs_mhd", "nl_tangent", "nl_tangent_mhd", "nl_adjoint", "nl_adjoint_lorentz",
           "nl_adjoint_mhd"]


class Mesh:
    """A P2 meridian mesh registered with libsfx (device copies of jj, rr)."""

    def __init__(self, jj, rr):
        jj = np.asfortranarray(jj, dtype=np.int32)
        rr = np.asfortranarray(rr, dtype=np.float64)
        if jj.ndim != 2 or jj.shape[0] != 6 or rr.ndim != 2 or rr.shape[0] != 2:
            raise ValueError("expected jj(6, me) and rr(2, np)")
        self.me, self.np = int(jj.shape[1]), int(rr.shape[1])
        h = C.c_void_p()
        check(lib().sfx_mesh_create(jj.ctypes.data, rr.ctypes.data, 6, 7, self.me, self.np, C.byref(h)))
        self._h = h
        self._lm_cache = {}

    @classmethod
    def from_p2(cls, m, boundary=False):
        """Mesh of a P2Mesh; with ``boundary`` its boundary sides are registered too."""
        mesh = cls(m.jj, m.rr)
        if boundary:
            from .mesh import boundary_sides
            mesh.set_boundary(boundary_sides(m.jj))
        return mesh

    def set_boundary(self, jjs):
        """Register the boundary sides jjs(3, mes) (mesh%jjs), for rhs_3x3's boundary term."""
        jjs = np.asfortranarray(jjs, dtype=np.int32)
        if jjs.ndim != 2 or jjs.shape[0] != 3:
            raise ValueError("expected jjs(3, mes)")
        check(lib().sfx_mesh_set_boundary(self._h, jjs.ctypes.data, int(jjs.shape[1])))
        self.mes = int(jjs.shape[1])

    @property
    def handle(self):
        return self._h

    @property
    def n_gauss(self):
        return 7 * self.me

    def list_mode_device(self, list_mode):
        """Device int32 copy of list_mode (cached)."""
        import torch
        key = tuple(int(x) for x in list_mode)
        t = self._lm_cache.get(key)
        if t is None:
            t = torch.tensor(key, dtype=torch.int32, device="cuda")
            self._lm_cache[key] = t
        return t

    def close(self):
        if getattr(self, "_h", None):
            lib().sfx_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _lm(list_mode):
    return np.ascontiguousarray(np.asarray(list_mode, dtype=np.int32))


def _check_comm(communicator):
    if communicator is not None and getattr(communicator, "size", 1) != 1:
        raise ValueError("one GPU: the Fourier communicator must have a single rank")


def smb_cross_prod_gauss_sft_par(communicator, mesh: Mesh, list_mode, V_in, V_out, precession_in=False,
                                 taux_precession=0.0, angle_s_pi=0.0):
    """V_out(G,6,m) = (curl V_in + 2 Omega) x V_in at the Gauss points."""
    _check_comm(communicator)
    M = len(list_mode)
    fdims(V_in), fdims(V_out)
    L = lib()
    if is_device(V_in):
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_smb_cross_prod_gauss_sft_par_dev(mesh.handle, lm.data_ptr(), M, ptr(V_in), ptr(V_out),
                                                     int(bool(precession_in)), float(taux_precession),
                                                     float(angle_s_pi), stream_ptr()))
    else:
        lm = _lm(list_mode)
        check(L.sfx_smb_cross_prod_gauss_sft_par(mesh.handle, lm.ctypes.data, M, ptr(V_in), ptr(V_out),
                                                 int(bool(precession_in)), float(taux_precession),
                                                 float(angle_s_pi)))
    return V_out


def smb_CurlH_cross_B_gauss_sft_par(communicator, mesh: Mesh, list_mode, V_in, W_in, V_out):
    """V_out(G,6,m) = (curl V_in) x W_in at the Gauss points."""
    _check_comm(communicator)
    M = len(list_mode)
    L = lib()
    if is_device(V_in):
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_smb_curlh_cross_b_gauss_sft_par_dev(mesh.handle, lm.data_ptr(), M, ptr(V_in), ptr(W_in),
                                                        ptr(V_out), stream_ptr()))
    else:
        lm = _lm(list_mode)
        check(L.sfx_smb_curlh_cross_b_gauss_sft_par(mesh.handle, lm.ctypes.data, M, ptr(V_in), ptr(W_in),
                                                    ptr(V_out)))
    return V_out


def smb_ugradc_gauss_fft_par(communicator, mesh: Mesh, list_mode, heat_capa_in, V_in, c_in, c_out):
    """c_out(G,2,m) = (hc grad c).u + (div u)(hc c) at the Gauss points."""
    _check_comm(communicator)
    M = len(list_mode)
    L = lib()
    if is_device(V_in):
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_smb_ugradc_gauss_fft_par_dev(mesh.handle, lm.data_ptr(), M, ptr(heat_capa_in), ptr(V_in),
                                                 ptr(c_in), ptr(c_out), stream_ptr()))
    else:
        lm = _lm(list_mode)
        hc = np.ascontiguousarray(heat_capa_in, dtype=np.float64)
        check(L.sfx_smb_ugradc_gauss_fft_par(mesh.handle, lm.ctypes.data, M, hc.ctypes.data, ptr(V_in),
                                             ptr(c_in), ptr(c_out)))
    return c_out


def nl_rotv_mhd(mesh: Mesh, list_mode, u, H, B, coeff_lorentz, rotv_v):
    """rotv_v(G,6,m) = (curl u) x u - coeff_lorentz (curl H) x B in one launch."""
    M = len(list_mode)
    L = lib()
    if is_device(u):
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_nl_rotv_mhd_dev(mesh.handle, lm.data_ptr(), M, ptr(u), ptr(H), ptr(B),
                                    float(coeff_lorentz), ptr(rotv_v), stream_ptr()))
    else:
        lm = _lm(list_mode)
        check(L.sfx_nl_rotv_mhd(mesh.handle, lm.ctypes.data, M, ptr(u), ptr(H), ptr(B),
                                float(coeff_lorentz), ptr(rotv_v)))
    return rotv_v


def rhs_3x3(mesh: Mesh, mode, rhs_in, vb_145, vb_236, opt_tensor=None, opt_tensor_scaln_bdy=None):
    """One mode (SOURCES/SOLVER/assembling_rhs.F90:5-189): rhs_in(G,6) -> vb_145(np,3), vb_236(np,3).

    Host arrays.  ``opt_tensor`` (3,G,6) adds the weak divergence of the
    tensor (its 1/r terms use ``mode``); ``opt_tensor_scaln_bdy`` (3*mes,6)
    the boundary term (the mesh's boundary sides must be registered).
    """
    fdims(rhs_in), fdims(vb_145), fdims(vb_236)
    if opt_tensor is None and opt_tensor_scaln_bdy is None:
        check(lib().sfx_rhs_3x3(mesh.handle, int(mode), ptr(rhs_in), ptr(vb_145), ptr(vb_236)))
    else:
        opt = lambda x: None if x is None else ptr(np.asfortranarray(x, dtype=np.float64))
        t = None if opt_tensor is None else np.asfortranarray(opt_tensor, dtype=np.float64)
        b = None if opt_tensor_scaln_bdy is None else np.asfortranarray(opt_tensor_scaln_bdy, dtype=np.float64)
        check(lib().sfx_rhs_3x3_mode(mesh.handle, int(mode), ptr(rhs_in), None if t is None else ptr(t),
                                     None if b is None else ptr(b), ptr(vb_145), ptr(vb_236)))
    return vb_145, vb_236


def rhs_3x3_bdy(mesh: Mesh, tensor_bdy, vb_145, vb_236, scale=1.0):
    """Add rhs_3x3's boundary term (assembling_rhs.F90:129-185) for every mode
    to the device arrays vb_145(np,3,m), vb_236(np,3,m): tensor_bdy(3*mes,6,m)."""
    M = fdims(tensor_bdy)[2]
    check(lib().sfx_rhs_3x3_bdy_dev(mesh.handle, M, ptr(tensor_bdy), ptr(vb_145), ptr(vb_236), float(scale),
                                    stream_ptr()))
    return vb_145, vb_236


def _work_elem(mesh: Mesh, M, work_elem):
    if work_elem is None:
        import torch
        work_elem = torch.empty(36 * mesh.me * M, dtype=torch.float64, device="cuda")
    return work_elem


def rhs_3x3_all_modes(mesh: Mesh, rhs_gauss, vb_145, vb_236, scale=1.0, work_elem=None):
    """All modes on the device: rhs_gauss(G,6,m) -> vb_145(np,3,m), vb_236(np,3,m).

    ``work_elem``: (36*me*m) float64 CUDA buffer (allocated when omitted).
    """
    M = fdims(rhs_gauss)[2]
    work_elem = _work_elem(mesh, M, work_elem)
    check(lib().sfx_rhs_3x3_dev(mesh.handle, M, ptr(rhs_gauss), ptr(vb_145), ptr(vb_236), float(scale),
                                work_elem.data_ptr(), stream_ptr()))
    return vb_145, vb_236


def nl_rhs_3x3(mesh: Mesh, list_mode, V_in, vb_145, vb_236, scale=-1.0, work_elem=None, precession_in=False,
               taux_precession=0.0, angle_s_pi=0.0):
    """smb_cross_prod_gauss_sft_par then rhs_3x3 of its output, fused (no Gauss
    array; device arrays): vb_145(np,3,m), vb_236(np,3,m) = scale *
    rhs_3x3((curl V_in + 2 Omega) x V_in), scale = -1 for the -rotv_v share of
    the NS right-hand side (computing_rhs_gauss.F90:100).  Equal, bit for bit,
    to the two calls."""
    M = len(list_mode)
    fdims(V_in), fdims(vb_145), fdims(vb_236)
    lm = mesh.list_mode_device(list_mode)
    work_elem = _work_elem(mesh, M, work_elem)
    check(lib().sfx_nl_rhs_3x3_dev(mesh.handle, lm.data_ptr(), M, ptr(V_in), ptr(vb_145), ptr(vb_236),
                                   float(scale), int(bool(precession_in)), float(taux_precession),
                                   float(angle_s_pi), work_elem.data_ptr(), stream_ptr()))
    return vb_145, vb_236


def nl_qs_00(mesh: Mesh, list_mode, heat_capa, V_in, c_in, vect, ff=None, scale=-1.0, work_elem=None):
    """smb_ugradc_gauss_fft_par then qs_00_gauss of its output, fused (no Gauss
    array; device arrays): vect(np,2,m) = scale * qs_00_gauss(heat_capa, ff,
    (hc grad c).u + (div u)(hc c)), ff (np,2,m) the nodal source or None.
    Equal, bit for bit, to the two calls."""
    import torch
    M = len(list_mode)
    fdims(V_in), fdims(c_in), fdims(vect)
    lm = mesh.list_mode_device(list_mode)
    if work_elem is None:
        work_elem = torch.empty(12 * mesh.me * M, dtype=torch.float64, device="cuda")
    check(lib().sfx_nl_qs_00_dev(mesh.handle, lm.data_ptr(), M, ptr(heat_capa), ptr(V_in), ptr(c_in),
                                 None if ff is None else ptr(ff), ptr(vect), float(scale), work_elem.data_ptr(),
                                 stream_ptr()))
    return vect


def rhs_ns_3x3(mesh: Mesh, list_mode, rotv_v, V1m, ff, P, P_inc, vb_145, vb_236, work_elem=None,
               opt_tensor=None, penal_grid=None, imposed_vel=None, work_gauss=None):
    """Full explicit NS right-hand side assembled for every mode (device arrays).

    rhs_ns_gauss_3x3 without the penalty options
    (SOURCES/SOLVER/computing_rhs_gauss.F90:7-146) followed by rhs_3x3
    (SOURCES/SOLVER/assembling_rhs.F90:5-126), fused.  ``P``, ``P_inc``: the
    pressure and its increment on the velocity mesh, (np, 2, m);
    ``opt_tensor``: rhs_3x3's tensor at the Gauss points, (3, G, 6, m).  Any
    of ``rotv_v``, ``V1m``, ``ff``, ``P``, ``P_inc``, ``opt_tensor`` may be
    None (zero).  ``penal_grid`` (2*m_max_pad-1, G), m_max_pad = 3m/2: the
    penalty function on the grid angles at the Gauss points
    (``if_ns_penalty``, computing_rhs_gauss.F90:103-141); ``imposed_vel``
    (np, 6): 3*imposed_velocity/(2 dt) (``if_impose_vel_in_solids``);
    ``work_gauss``: (2*G*6*m) buffer (allocated when omitted).
    """
    M = len(list_mode)
    lm = mesh.list_mode_device(list_mode)
    work_elem = _work_elem(mesh, M, work_elem)
    opt = lambda x: 0 if x is None else ptr(x)
    if penal_grid is not None:
        import torch
        if work_gauss is None:
            work_gauss = torch.empty(2 * mesh.n_gauss * 6 * M, dtype=torch.float64, device="cuda")
        zero = [i for i, md in enumerate(list_mode) if int(md) == 0]
        check(lib().sfx_rhs_ns_3x3_penal_dev(mesh.handle, lm.data_ptr(), M, zero[0] if zero else -1, opt(rotv_v),
                                             opt(V1m), opt(ff), opt(P), opt(P_inc), ptr(penal_grid),
                                             opt(imposed_vel), opt(opt_tensor), work_gauss.data_ptr(),
                                             work_elem.data_ptr(), ptr(vb_145), ptr(vb_236), stream_ptr()))
        return vb_145, vb_236
    check(lib().sfx_rhs_ns_3x3_dev(mesh.handle, lm.data_ptr(), M, opt(rotv_v), opt(V1m), opt(ff), opt(P),
                                   opt(P_inc), opt(opt_tensor), work_elem.data_ptr(), ptr(vb_145), ptr(vb_236),
                                   stream_ptr()))
    return vb_145, vb_236


def qs_00_gauss(mesh: Mesh, heat_capa, ff, ff_gauss, vect, work_elem=None):
    """Scalar RHS (SOURCES/SOLVER/fem_rhs_axi.F90:940): vect(np) (host) or vect(np,ncols) (device).

    Device path: ``work_elem`` is a (6*me*ncols) float64 CUDA buffer (allocated when omitted).
    """
    if is_device(vect):
        import torch
        nc = vect.shape[1] if vect.dim() > 1 else 1
        if work_elem is None:
            work_elem = torch.empty(6 * mesh.me * nc, dtype=torch.float64, device="cuda")
        check(lib().sfx_qs_00_gauss_dev(mesh.handle, nc, ptr(heat_capa), ptr(ff), ptr(ff_gauss),
                                        work_elem.data_ptr(), ptr(vect), stream_ptr()))
    else:
        hc = np.ascontiguousarray(heat_capa, dtype=np.float64)
        ffa = np.ascontiguousarray(ff, dtype=np.float64)
        fga = np.ascontiguousarray(ff_gauss, dtype=np.float64)
        check(lib().sfx_qs_00_gauss(mesh.handle, hc.ctypes.data, ffa.ctypes.data, fga.ctypes.data, ptr(vect)))
    return vect


def nl_rhs_mhd(mesh: Mesh, list_mode, u, H, B, coeff_lorentz, work_gauss, vb_145, vb_236, work_elem=None):
    """Nonlinear share of the NS right-hand sides of an MHD step, all modes (device)."""
    M = len(list_mode)
    lm = mesh.list_mode_device(list_mode)
    work_elem = _work_elem(mesh, M, work_elem)
    check(lib().sfx_nl_rhs_mhd_dev(mesh.handle, lm.data_ptr(), M, ptr(u), ptr(H), ptr(B), float(coeff_lorentz),
                                   ptr(work_gauss), work_elem.data_ptr(), ptr(vb_145), ptr(vb_236), stream_ptr()))
    return vb_145, vb_236


def nl_tangent(mesh: Mesh, list_mode, U, dU, out):
    """Tangent of smb_cross_prod_gauss_sft_par around U: B(U,dU) + B(dU,U), B(a,b) = P[(curl a) x b]."""
    M = len(list_mode)
    L = lib()
    if is_device(U):
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_nl_tangent_dev(mesh.handle, lm.data_ptr(), M, ptr(U), ptr(dU), ptr(out), stream_ptr()))
    else:
        lm = _lm(list_mode)
        check(L.sfx_nl_tangent(mesh.handle, lm.ctypes.data, M, ptr(U), ptr(dU), ptr(out)))
    return out


def nl_tangent_mhd(mesh: Mesh, list_mode, U, H0, B0, dU, dH, dB, coeff_lorentz, out):
    """Linearized convective + Lorentz terms around (U, H0, B0) (device arrays)."""
    M = len(list_mode)
    lm = mesh.list_mode_device(list_mode)
    check(lib().sfx_nl_tangent_mhd_dev(mesh.handle, lm.data_ptr(), M, ptr(U), ptr(H0), ptr(B0), ptr(dU), ptr(dH),
                                       ptr(dB), float(coeff_lorentz), ptr(out), stream_ptr()))
    return out


def nl_adjoint(mesh: Mesh, list_mode, U, ybar, dbar, work_elem=None):
    """Exact transpose of :func:`nl_tangent`: dbar(np,6,m) = T_U^T ybar(G,6,m).

    Device path: ``work_elem`` is a (6*me*6*m) float64 CUDA buffer (allocated
    when omitted).
    """
    M = len(list_mode)
    L = lib()
    if is_device(U):
        import torch
        if work_elem is None:
            work_elem = torch.empty(6 * mesh.me * 6 * M, dtype=torch.float64, device="cuda")
        lm = mesh.list_mode_device(list_mode)
        check(L.sfx_nl_adjoint_dev(mesh.handle, lm.data_ptr(), M, ptr(U), ptr(ybar), work_elem.data_ptr(), ptr(dbar),
                                   stream_ptr()))
    else:
        lm = _lm(list_mode)
        check(L.sfx_nl_adjoint(mesh.handle, lm.ctypes.data, M, ptr(U), ptr(ybar), ptr(dbar)))
    return dbar


def nl_adjoint_lorentz(mesh: Mesh, list_mode, H0, B0, coeff_lorentz, ybar, dHbar, dBbar, work_elem=None):
    """Transpose of -coeff_lorentz [(curl H0) x dB + (curl dH) x B0] (device arrays).

    ``work_elem``: (2*36*me*m) float64 CUDA buffer (allocated when omitted).
    """
    import torch
    M = len(list_mode)
    if work_elem is None:
        work_elem = torch.empty(2 * 36 * mesh.me * M, dtype=torch.float64, device="cuda")
    lm = mesh.list_mode_device(list_mode)
    check(lib().sfx_nl_adjoint_lorentz_dev(mesh.handle, lm.data_ptr(), M, ptr(H0), ptr(B0), float(coeff_lorentz),
                                           ptr(ybar), work_elem.data_ptr(), ptr(dHbar), ptr(dBbar), stream_ptr()))
    return dHbar, dBbar


def nl_adjoint_mhd(mesh: Mesh, list_mode, U, H0, B0, coeff_lorentz, ybar, dUbar, dHbar, dBbar, work_elem=None):
    """Transpose of :func:`nl_tangent_mhd` around (U, H0, B0) (device arrays)."""
    import torch
    M = len(list_mode)
    if work_elem is None:
        work_elem = torch.empty(2 * 36 * mesh.me * M, dtype=torch.float64, device="cuda")
    lm = mesh.list_mode_device(list_mode)
    check(lib().sfx_nl_adjoint_mhd_dev(mesh.handle, lm.data_ptr(), M, ptr(U), ptr(H0), ptr(B0),
                                       float(coeff_lorentz), ptr(ybar), work_elem.data_ptr(), ptr(dUbar),
                                       ptr(dHbar), ptr(dBbar), stream_ptr()))
    return dUbar, dHbar, dBbar
