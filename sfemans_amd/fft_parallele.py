"""Python mirror of the reference module ``fft_parallele`` (one GPU).

Same routine names, argument order and meaning as
SOURCES/POST_PROCESSING/fft_parallel.F90 (``communicator`` must be ``None``:
one rank; ``V_out`` is INTENT(OUT) and is filled in place).  Arrays are
either Fortran-ordered numpy arrays (host path: staged to the GPU and back)
or CUDA tensors with Fortran strides (device path: no copies, launched on
torch's current stream).  ``fortran_tensor`` builds such tensors.

Errors raise ``SfxError`` where the reference would STOP.
"""
from __future__ import annotations

import numpy as np

from ._lib import lib, check

__all__ = [
    "FFT_PAR_CROSS_PROD_DCL", "FFT_PAR_DOT_PROD_DCL", "FFT_PAR_PROD_DCL",
    "FFT_SCALAR_VECT_DCL", "FFT_TENSOR_DCL", "FFT_PAR_DIV_DCL", "FFT_HEAVISIDE_DCL",
    "FFT_SQUAREROOT_SCALAR", "FFT_MAX_VEL_DCL", "FFT_SCALAR_VECT_NO_OVERSHOOT",
    "FFT_NO_OVERSHOOT_LEVEL_SET", "FFT_PAR_ENTRO_VISC_DCL", "FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM",
    "FFT_COMPUTE_DIFFU_MOM", "FFT_PAR_COMPR_ENTRO_VISC_DCL", "FFT_COMPUTE_ENTROPY_VISC",
    "FFT_COMPUTE_ENTROPY_VISC_MOM",
    "fortran_tensor", "fortran_empty",
    "is_device", "ptr", "fdims", "stream_ptr",
]


def is_device(x) -> bool:
    return hasattr(x, "is_cuda") and x.is_cuda


def fdims(x):
    """Fortran dimensions of an array/tensor with column-major layout."""
    if is_device(x):
        shape, st = tuple(x.shape), tuple(x.stride())
        exp = 1
        for d, s in zip(shape, st):
            if d > 1 and s != exp:
                raise ValueError(f"tensor is not Fortran-contiguous: shape {shape} strides {st}")
            exp *= d
        if x.dtype.__str__() != "torch.float64":
            raise TypeError("tensors must be float64")
        return shape
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.f_contiguous):
        raise TypeError("arrays must be Fortran-contiguous float64 numpy arrays")
    return x.shape


def ptr(x):
    if x is None:
        return None
    if is_device(x):
        return x.data_ptr()
    return x.ctypes.data


def stream_ptr():
    import torch
    return torch.cuda.current_stream().cuda_stream


def fortran_tensor(a, device="cuda"):
    """Copy a numpy array to a CUDA tensor with the same column-major layout."""
    import torch
    a = np.asfortranarray(a, dtype=np.float64)
    t = torch.as_tensor(np.ascontiguousarray(a.T), device=device)
    return t.permute(*reversed(range(a.ndim)))


def fortran_empty(shape, device="cuda"):
    import torch
    t = torch.empty(tuple(reversed(shape)), dtype=torch.float64, device=device)
    return t.permute(*reversed(range(len(shape))))


def _defaults(np_, M, nb_procs, bloc_size, m_max_pad):
    if bloc_size is None:
        bloc_size = np_ // nb_procs + 1
    if m_max_pad is None:
        m_max_pad = (3 * M * nb_procs) // 2
    return bloc_size, m_max_pad


def _single_rank(communicator):
    if communicator is not None and getattr(communicator, "size", 1) != 1:
        raise ValueError("one GPU: the Fourier communicator must have a single rank")


def _temps_ptr(temps):
    if temps is None:
        return None
    if not (isinstance(temps, np.ndarray) and temps.dtype == np.float64 and temps.size >= 3):
        raise TypeError("temps must be a float64 array of 3 entries")
    return temps.ctypes.data


def _binary(name, communicator, V1_in, V2_in, V_out, nb_procs, bloc_size, m_max_pad, temps,
            extra=()):
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    fdims(V2_in), fdims(V_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        dev = getattr(L, f"sfx_{name}_dev")
        if name == "fft_scalar_vect_dcl":
            check(dev(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, M, extra[1], m_max_pad, stream_ptr()))
        elif name == "fft_tensor_dcl":
            check(dev(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, M, m_max_pad, int(bool(extra[0])),
                      stream_ptr()))
        elif name == "fft_par_div_dcl":
            check(dev(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, M, m_max_pad, stream_ptr()))
        else:
            check(dev(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, M, m_max_pad, stream_ptr()))
        return V_out
    fn = getattr(L, f"sfx_{name}")
    if name in ("fft_par_prod_dcl", "fft_par_div_dcl"):
        check(fn(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, M, nb_procs, bloc_size, m_max_pad, _temps_ptr(temps)))
    elif name == "fft_scalar_vect_dcl":
        nb2 = fdims(V2_in)[1]
        check(fn(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, nb, nb2, M, extra[1], nb_procs, bloc_size,
                 m_max_pad, _temps_ptr(temps)))
    elif name == "fft_tensor_dcl":
        check(fn(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, nb, M, nb_procs, bloc_size, m_max_pad,
                 int(bool(extra[0])), _temps_ptr(temps)))
    else:
        check(fn(ptr(V1_in), ptr(V2_in), ptr(V_out), np_, nb, M, nb_procs, bloc_size, m_max_pad,
                 _temps_ptr(temps)))
    return V_out


def FFT_PAR_CROSS_PROD_DCL(communicator, V1_in, V2_in, V_out, nb_procs=1, bloc_size=None,
                           m_max_pad=None, temps=None):
    """V_out(np,6,m) = V1 x V2 (fft_parallel.F90:147)."""
    return _binary("fft_par_cross_prod_dcl", communicator, V1_in, V2_in, V_out, nb_procs,
                   bloc_size, m_max_pad, temps)


def FFT_PAR_DOT_PROD_DCL(communicator, V1_in, V2_in, c_out, nb_procs=1, bloc_size=None,
                         m_max_pad=None, temps=None):
    """c_out(np,2,m) = V1 . V2 (fft_parallel.F90:330)."""
    return _binary("fft_par_dot_prod_dcl", communicator, V1_in, V2_in, c_out, nb_procs,
                   bloc_size, m_max_pad, temps)


def FFT_PAR_PROD_DCL(communicator, c1_in, c2_in, c_out, nb_procs=1, bloc_size=None,
                     m_max_pad=None, temps=None):
    """c_out(np,2,m) = c1 * c2 (fft_parallel.F90:502)."""
    return _binary("fft_par_prod_dcl", communicator, c1_in, c2_in, c_out, nb_procs,
                   bloc_size, m_max_pad, temps)


def FFT_SCALAR_VECT_DCL(communicator, V1_in, V2_in, V_out, pb, nb_procs=1, bloc_size=None,
                        m_max_pad=None, temps=None):
    """V_out(np,6,m) = s * v (pb = 1) or v / s (pb = 2, on the grid of
    2*m_max_pad-1 angles) (fft_parallel.F90:1021)."""
    return _binary("fft_scalar_vect_dcl", communicator, V1_in, V2_in, V_out, nb_procs,
                   bloc_size, m_max_pad, temps, extra=(None, pb))


def FFT_TENSOR_DCL(communicator, V1_in, V2_in, V_out, nb_procs=1, bloc_size=None,
                   m_max_pad=None, temps=None, opt_tension=None):
    """V_out(3,np,6,m) = V1 (x) V2, or with ``opt_tension`` the surface-tension
    tensor V1_a V2_b/(|V2| + 1e-14) - delta_ab |V2| on the grid of
    2*m_max_pad-1 angles (fft_parallel.F90:1543, :1688-1706)."""
    return _binary("fft_tensor_dcl", communicator, V1_in, V2_in, V_out, nb_procs,
                   bloc_size, m_max_pad, temps, extra=(opt_tension, None))


def FFT_PAR_DIV_DCL(communicator, c1_in, c2_in, c_out, nb_procs=1, bloc_size=None,
                    m_max_pad=None, temps=None):
    """c_out(np,2,m) = c1 / c2 on the grid of 2*m_max_pad-1 angles (fft_parallel.F90:667)."""
    return _binary("fft_par_div_dcl", communicator, c1_in, c2_in, c_out, nb_procs,
                   bloc_size, m_max_pad, temps)


def FFT_HEAVISIDE_DCL(communicator, V1_in, values, V_out, nb_procs=1, bloc_size=None,
                      m_max_pad=None, coeff_tanh=None, temps=None):
    """Density of level sets V1_in(n_int, np, 2, m) with ``values`` (n_int+1)
    and regul of half-width ``coeff_tanh`` (fft_parallel.F90:832): V_out(np,2,m)."""
    _single_rank(communicator)
    n_int, np_, nb, M = fdims(V1_in)
    fdims(V_out)
    vals = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
    if vals.size != n_int + 1:
        raise ValueError("SIZE(values) must be SIZE(V1_in,1) + 1")
    if coeff_tanh is None:
        raise TypeError("coeff_tanh is required")
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_heaviside_dcl_dev(ptr(V1_in), vals.ctypes.data, ptr(V_out), vals.size, np_, M,
                                          m_max_pad, float(coeff_tanh), stream_ptr()))
    else:
        check(L.sfx_fft_heaviside_dcl(ptr(V1_in), vals.ctypes.data, ptr(V_out), vals.size, np_, nb, M,
                                      nb_procs, bloc_size, m_max_pad, float(coeff_tanh), _temps_ptr(temps)))
    return V_out


def _inplace(name, communicator, c1_inout, nb_procs, bloc_size, m_max_pad, temps):
    _single_rank(communicator)
    np_, nb, M = fdims(c1_inout)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(c1_inout):
        check(getattr(L, f"sfx_{name}_dev")(ptr(c1_inout), np_, nb, M, m_max_pad, stream_ptr()))
    else:
        check(getattr(L, f"sfx_{name}")(ptr(c1_inout), np_, nb, M, nb_procs, bloc_size, m_max_pad,
                                        _temps_ptr(temps)))
    return c1_inout


def FFT_SQUAREROOT_SCALAR(communicator, c1_inout, nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """c1_inout(np,nb,m): its first scalar becomes SQRT(MAX(0, MIN(1, c))) on the
    grid of 2*m_max_pad-1 angles (fft_parallel.F90:4601), in place."""
    return _inplace("fft_squareroot_scalar", communicator, c1_inout, nb_procs, bloc_size, m_max_pad, temps)


def FFT_NO_OVERSHOOT_LEVEL_SET(communicator, c1_inout, nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """c1_inout(np,nb,m): its first scalar becomes MAX(0, MIN(1, c)) on the grid
    of 2*m_max_pad-1 angles (fft_parallel.F90:3649), in place."""
    return _inplace("fft_no_overshoot_level_set", communicator, c1_inout, nb_procs, bloc_size, m_max_pad, temps)


def FFT_SCALAR_VECT_NO_OVERSHOOT(communicator, scalar_bounds, V1_in, V2_in, V_out, pb, nb_procs=1,
                                 bloc_size=None, m_max_pad=None, temps=None):
    """FFT_SCALAR_VECT_DCL with the scalar clamped to [MINVAL(scalar_bounds),
    MAXVAL(scalar_bounds)] on the grid of 2*m_max_pad-1 angles (fft_parallel.F90:1222)."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    nb2 = fdims(V2_in)[1]
    fdims(V_out)
    bnd = np.ascontiguousarray(np.atleast_1d(np.asarray(scalar_bounds, dtype=np.float64)).ravel())
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_scalar_vect_no_overshoot_dev(bnd.ctypes.data, bnd.size, ptr(V1_in), ptr(V2_in), ptr(V_out),
                                                     np_, M, pb, m_max_pad, stream_ptr()))
    else:
        check(L.sfx_fft_scalar_vect_no_overshoot(bnd.ctypes.data, bnd.size, ptr(V1_in), ptr(V2_in), ptr(V_out), np_,
                                                 nb, nb2, M, pb, nb_procs, bloc_size, m_max_pad, _temps_ptr(temps)))
    return V_out


def FFT_MAX_VEL_DCL(communicator, V1_in, nb_procs=1, bloc_size=None, m_max_pad=None, V_out=None):
    """Maximum over points and the 2*m_max_pad-1 angles of |v| (nb_field 6) or |c|
    (nb_field 2) (fft_parallel.F90:1426).  Host arrays: returns a float.  CUDA
    tensors: writes the maximum into ``V_out`` (a 1-element float64 tensor,
    allocated when omitted) on the current stream and returns it."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        import torch
        if V_out is None:
            V_out = torch.empty(1, dtype=torch.float64, device=V1_in.device)
        check(L.sfx_fft_max_vel_dcl_dev(ptr(V1_in), ptr(V_out), np_, nb, M, m_max_pad, stream_ptr()))
        return V_out
    out = np.zeros(1)
    check(L.sfx_fft_max_vel_dcl(ptr(V1_in), out.ctypes.data, np_, nb, M, nb_procs, bloc_size, m_max_pad))
    return float(out[0])


def _grid_angles(m_max_pad):
    n = 2 * m_max_pad - 1
    return 2 * np.pi * np.arange(n) / n


def _eta_prod(name, communicator, eta, H_mesh, c_in, c_out, nb_procs, bloc_size, m_max_pad, time, temps):
    _single_rank(communicator)
    np_, nb, M = fdims(c_in)
    fdims(c_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    N = 2 * m_max_pad - 1
    if callable(eta):
        # the reference's ETA(mesh, angles, nb_angles, nb, ne, time), points nb..ne (1-based)
        eg = eta(H_mesh, _grid_angles(m_max_pad), N, 1, np_, time)
    else:
        eg = eta
    if tuple(fdims(eg) if is_device(eg) else np.shape(eg)) != (N, np_):
        raise ValueError(f"{name}: eta must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(c_in):
        if not is_device(eg):
            eg = fortran_tensor(np.asarray(eg, dtype=np.float64), device=c_in.device)
        check(L.sfx_fft_par_var_eta_prod_dcl_dev(ptr(eg), ptr(c_in), ptr(c_out), np_, nb, M, m_max_pad,
                                                 stream_ptr()))
    else:
        eg = np.asfortranarray(eg, dtype=np.float64)
        check(L.sfx_fft_par_var_eta_prod_dcl(ptr(eg), ptr(c_in), ptr(c_out), np_, nb, M, nb_procs, bloc_size,
                                             m_max_pad, _temps_ptr(temps)))
    return c_out


def FFT_PAR_VAR_ETA_PROD_T_DCL(communicator, eta, H_mesh, c_in, c_out, nb_procs=1, bloc_size=None, m_max_pad=None,
                               time=0.0, temps=None):
    """c_out = coefficients of eta * c_in (nb_field 2 or 6) on the grid of
    2*m_max_pad-1 angles (fft_parallel.F90:1798).  eta is the reference's
    user function eta(H_mesh, angles, nb_angles, nb, ne, time) -> (nb_angles,
    ne-nb+1) values at the nodes, called once for all points, or its values
    as an (2*m_max_pad-1, np) array."""
    return _eta_prod("FFT_PAR_VAR_ETA_PROD_T_DCL", communicator, eta, H_mesh, c_in, c_out, nb_procs, bloc_size,
                     m_max_pad, time, temps)


def FFT_PAR_VAR_ETA_PROD_GAUSS_DCL(communicator, eta, H_mesh, c_in, c_out, nb_procs=1, bloc_size=None,
                                   m_max_pad=None, time=0.0, temps=None):
    """As FFT_PAR_VAR_ETA_PROD_T_DCL with eta at the Gauss points
    (fft_parallel.F90:2015)."""
    return _eta_prod("FFT_PAR_VAR_ETA_PROD_GAUSS_DCL", communicator, eta, H_mesh, c_in, c_out, nb_procs,
                     bloc_size, m_max_pad, time, temps)


def FFT_PAR_FROM_REAL(real_in, c_out, m_max_pad=None):
    """Coefficients (np, 2, m) of values given on the 2*m_max_pad-1 grid
    angles, real_in (2*m_max_pad-1, np): the analysis every product ends with."""
    np_, _, M = fdims(c_out)
    rshape = tuple(fdims(real_in)) if is_device(real_in) else np.shape(real_in)
    if m_max_pad is None:
        m_max_pad = (rshape[0] + 1) // 2
    if tuple(rshape) != (2 * m_max_pad - 1, np_):
        raise ValueError("real_in must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(c_out):
        check(L.sfx_fft_par_from_real_dev(ptr(real_in), ptr(c_out), np_, M, m_max_pad, stream_ptr()))
    else:
        check(L.sfx_fft_par_from_real(ptr(np.asfortranarray(real_in, dtype=np.float64)), ptr(c_out), np_, M,
                                      m_max_pad))
    return c_out


def FFT_PAR_SCAL_FUNCT(communicator, c1_inout, funct, nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """Columns 1-2 of c1_inout replaced by the coefficients of funct(first
    field) on the grid of 2*m_max_pad-1 angles (fft_parallel.F90:4294).  The
    synthesis and the analysis run on the GPU; funct runs here, on the real
    planes: it is called once on the whole array (numpy for host data, a
    torch tensor for CUDA data) and must act value by value; a function that
    does not accept an array is applied to each value."""
    _single_rank(communicator)
    np_, nb, M = fdims(c1_inout)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    N = 2 * m_max_pad - 1
    first = c1_inout[:, 0:2, :]
    dev = is_device(c1_inout)
    if dev:
        import torch
        R = FFT_PAR_REAL(communicator, first.contiguous().permute(2, 1, 0).contiguous().permute(2, 1, 0),
                         opt_nb_plane=N)
        vals = funct(R[:, 0, :])
        out = fortran_empty((np_, 2, M), device=c1_inout.device)
        FFT_PAR_FROM_REAL(vals.contiguous().t().contiguous().t(), out, m_max_pad)
        c1_inout[:, 0:2, :] = out
        return c1_inout
    R = FFT_PAR_REAL(communicator, np.asfortranarray(first), opt_nb_plane=N)[:, 0, :]
    try:
        vals = np.asarray(funct(R), dtype=np.float64)
        if vals.shape != R.shape:
            raise ValueError
    except Exception:
        vals = np.vectorize(funct, otypes=[np.float64])(R)
    out = np.zeros((np_, 2, M), order="F")
    FFT_PAR_FROM_REAL(vals, out, m_max_pad)
    c1_inout[:, 0:2, :] = out
    return c1_inout


def FFT_MAX_MIN_VEL_DCL(communicator, V1_in, nb_procs=1, bloc_size=None, m_max_pad=None, V_out=None):
    """(max, min) of |v| over the points and the 2*m_max_pad-1 angles
    (fft_parallel.F90:4467); the minimum as the reference computes it: of the
    positive values, or 0 when |v| = 0 on the whole last angle.  Host arrays:
    returns a (2,) array.  CUDA tensors: writes into ``V_out`` (2 float64,
    allocated when omitted) on the current stream and returns it."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        import torch
        if V_out is None:
            V_out = torch.empty(2, dtype=torch.float64, device=V1_in.device)
        check(L.sfx_fft_max_min_vel_dcl_dev(ptr(V1_in), ptr(V_out), np_, nb, M, m_max_pad, stream_ptr()))
        return V_out
    out = np.zeros(2)
    check(L.sfx_fft_max_min_vel_dcl(ptr(V1_in), out.ctypes.data, np_, nb, M, nb_procs, bloc_size, m_max_pad))
    return out


def FFT_CHECK_INTERFACE(communicator, V1_in, nb_fluids, nb_procs=1, bloc_size=None, m_max_pad=None, temps=None,
                        interface_ok=None):
    """interface_ok (fft_parallel.F90:2267): 0 if at some point and grid angle
    a level set in [0.1, 0.9] lies below the next one, else 1.  V1_in is
    (nb_fluids-1, np, 2, m).  Host arrays: returns an int.  CUDA tensors:
    writes into ``interface_ok`` (1 int32, allocated when omitted) on the
    current stream and returns it."""
    _single_rank(communicator)
    nl, np_, nb, M = fdims(V1_in)
    if nl != nb_fluids - 1:
        raise ValueError("V1_in must hold nb_fluids-1 level sets")
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        import torch
        if interface_ok is None:
            interface_ok = torch.empty(1, dtype=torch.int32, device=V1_in.device)
        check(L.sfx_fft_check_interface_dev(ptr(V1_in), int(nb_fluids), ptr(interface_ok), np_, nb, M, m_max_pad,
                                            stream_ptr()))
        return interface_ok
    ok = np.zeros(1, dtype=np.int32)
    check(L.sfx_fft_check_interface(ptr(V1_in), int(nb_fluids), ok.ctypes.data, np_, nb, M, nb_procs, bloc_size,
                                    m_max_pad, _temps_ptr(temps)))
    return int(ok[0])


def FFT_PAR_ENTRO_VISC_DCL(communicator, V1_in, c1_in, c_in_real, hloc_gauss, coeff1_in_level, V_out, c_out,
                           nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """V_out(np,6,m) = (coeff1_in_level - c_in_real) V1 and c_out(np,2,m) =
    MAX(0, c (1 - c)) on the grid of 2*m_max_pad-1 angles, c_in_real given on
    that grid, shape (2*m_max_pad-1, np) (fft_parallel.F90:4055)."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    nb2 = fdims(c1_in)[1]
    fdims(V_out), fdims(c_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    nr, npr = fdims(c_in_real)
    if nr != 2 * m_max_pad - 1 or npr != np_:
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_par_entro_visc_dcl_dev(ptr(V1_in), ptr(c1_in), ptr(c_in_real), float(coeff1_in_level),
                                               ptr(V_out), ptr(c_out), np_, M, m_max_pad, stream_ptr()))
    else:
        check(L.sfx_fft_par_entro_visc_dcl(ptr(V1_in), ptr(c1_in), ptr(c_in_real), ptr(hloc_gauss),
                                           float(coeff1_in_level), ptr(V_out), ptr(c_out), np_, nb, nb2, M,
                                           nb_procs, bloc_size, m_max_pad, _temps_ptr(temps)))
    return V_out, c_out


def FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM(communicator, V1_in, V2_in, V3_in, c_in_real, V_out, les_coeff1_mom,
                                      nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """V_out(3,np,6,m): coefficients of (c_in_real - les_coeff1_mom) V_a on the grid
    of 2*m_max_pad-1 angles (fft_parallel.F90:3419); ``les_coeff1_mom`` is the
    reference's inputs%LES_coeff1_mom, passed explicitly."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    fdims(V2_in), fdims(V3_in), fdims(V_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    if tuple(fdims(c_in_real)) != (2 * m_max_pad - 1, np_):
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_compute_entropy_visc_grad_mom_dev(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(c_in_real),
                                                          float(les_coeff1_mom), ptr(V_out), np_, M, m_max_pad,
                                                          stream_ptr()))
    else:
        check(L.sfx_fft_compute_entropy_visc_grad_mom(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(c_in_real),
                                                      float(les_coeff1_mom), ptr(V_out), np_, nb, M, nb_procs,
                                                      bloc_size, m_max_pad, _temps_ptr(temps)))
    return V_out


def FFT_COMPUTE_DIFFU_MOM(communicator, V1_in, V2_in, V3_in, V1_out, V2_out, nb_procs=1, bloc_size=None,
                          m_max_pad=None, temps=None):
    """V1_out = s V1, V2_out = s V2 with s = V3_in (fft_parallel.F90:2939)."""
    _single_rank(communicator)
    np_, nb1, M = fdims(V1_in)
    nb2, nb3 = fdims(V2_in)[1], fdims(V3_in)[1]
    fdims(V1_out), fdims(V2_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_compute_diffu_mom_dev(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(V1_out), ptr(V2_out), np_, M,
                                              m_max_pad, stream_ptr()))
    else:
        check(L.sfx_fft_compute_diffu_mom(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(V1_out), ptr(V2_out), np_, nb1,
                                          nb2, nb3, M, nb_procs, bloc_size, m_max_pad, _temps_ptr(temps)))
    return V1_out, V2_out


def FFT_PAR_COMPR_ENTRO_VISC_DCL(communicator, V1_in, V2_in, c1_in, c_in_real, hloc_gauss, coeff1_in_level,
                                 V_out, c_out, level_set_comp_factor, nb_procs=1, bloc_size=None, m_max_pad=None,
                                 temps=None):
    """Compressive entropy viscosity of a level set (fft_parallel.F90:3160):
    V_out(np,6,m), c_out(np,2,m); ``level_set_comp_factor`` is the reference's
    inputs%level_set_comp_factor, passed explicitly."""
    _single_rank(communicator)
    np_, nb1, M = fdims(V1_in)
    nb2, nb3 = fdims(V2_in)[1], fdims(c1_in)[1]
    fdims(V_out), fdims(c_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    if tuple(fdims(c_in_real)) != (2 * m_max_pad - 1, np_):
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_par_compr_entro_visc_dcl_dev(ptr(V1_in), ptr(V2_in), ptr(c1_in), ptr(c_in_real),
                                                     float(coeff1_in_level), float(level_set_comp_factor),
                                                     ptr(V_out), ptr(c_out), np_, M, m_max_pad, stream_ptr()))
    else:
        check(L.sfx_fft_par_compr_entro_visc_dcl(ptr(V1_in), ptr(V2_in), ptr(c1_in), ptr(c_in_real), ptr(hloc_gauss),
                                                 float(coeff1_in_level), float(level_set_comp_factor), ptr(V_out),
                                                 ptr(c_out), np_, nb1, nb2, nb3, M, nb_procs, bloc_size, m_max_pad,
                                                 _temps_ptr(temps)))
    return V_out, c_out


def FFT_PAR_REAL(communicator, V1_in, V_out=None, opt_nb_plane=None):
    """Real planes of a Fourier field (fft_parallel.F90:28): V1_in (np, nb_field,
    m_max_c) -> V_out (N, nb_field/2, np), the values on the N angles 2 pi j/N,
    N = 2*m_max_pad - 1 with m_max_pad = (opt_nb_plane+1)//2 when opt_nb_plane
    > 2*m_max_c - 1, else m_max_c.  The reference allocates V_out; here it is
    allocated when not given (on the device of V1_in) and returned."""
    _single_rank(communicator)
    np_, nb, M = fdims(V1_in)
    L = lib()
    onp = 0 if opt_nb_plane is None else int(opt_nb_plane)
    N = L.sfx_fft_par_real_n(M, onp)
    shape = (N, nb // 2, np_)
    if is_device(V1_in):
        if V_out is None:
            V_out = fortran_empty(shape)
        if tuple(fdims(V_out)) != shape:
            raise ValueError(f"V_out must be {shape}")
        check(L.sfx_fft_par_real_dev(ptr(V1_in), ptr(V_out), np_, nb, M, onp, stream_ptr()))
    else:
        if V_out is None:
            V_out = np.zeros(shape, order="F")
        if tuple(fdims(V_out)) != shape:
            raise ValueError(f"V_out must be {shape}")
        check(L.sfx_fft_par_real(ptr(V1_in), ptr(V_out), np_, nb, M, 1, onp))
    return V_out


def FFT_COMPRESSION_LEVEL_SET_DCL(communicator_F, communicator_S, V1_in, V2_in, c_in, c_out, hloc_gauss, l_G,
                                  level_set_comp_factor, les_coeff4, h_min_distance, nb_procs=1, bloc_size=None,
                                  m_max_pad=None, temps=None):
    """Level-set compression term (fft_parallel.F90:3821): V1_in the level-set
    gradient, V2_in the velocity (np,6,m), c_in the level set (np,2,m) ->
    c_out (np,2,m).  level_set_comp_factor, les_coeff4 and h_min_distance are
    the reference's inputs%level_set_comp_factor, LES_coeff4, h_min_distance,
    passed explicitly; hloc_gauss and l_G do not change the result."""
    _single_rank(communicator_F)
    _single_rank(communicator_S)
    np_, nb, M = fdims(V1_in)
    fdims(V2_in), fdims(c_in), fdims(c_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    args = (float(level_set_comp_factor), float(les_coeff4), float(h_min_distance))
    if is_device(V1_in):
        check(L.sfx_fft_compression_level_set_dcl_dev(ptr(V1_in), ptr(V2_in), ptr(c_in), ptr(c_out), np_, M,
                                                      m_max_pad, *args, stream_ptr()))
    else:
        check(L.sfx_fft_compression_level_set_dcl(ptr(V1_in), ptr(V2_in), ptr(c_in), ptr(c_out), None, int(l_G),
                                                  np_, nb, M, nb_procs, bloc_size, m_max_pad, *args,
                                                  _temps_ptr(temps)))
    return c_out


def FFT_COMPUTE_ENTROPY_VISC(communicator, communicator_S, V1_in, V2_in, V3_in, V4_in, V5_in, hloc_gauss,
                             V1_out, V2_out, V3_out, residual_normalization, l_G, les_coeff1, les_coeff2,
                             les_coeff4, nb_procs=1, bloc_size=None, m_max_pad=None, temps=None):
    """Entropy viscosity times three tensor rows (fft_parallel.F90:2418); the
    LES coefficients are the reference's inputs%LES_coeff1/2/4, passed
    explicitly.  Points are grouped l_G per element."""
    _single_rank(communicator)
    _single_rank(communicator_S)
    np_, nb, M = fdims(V1_in)
    for x in (V2_in, V3_in, V4_in, V5_in, V1_out, V2_out, V3_out):
        fdims(x)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    L = lib()
    args = (float(residual_normalization), int(l_G), float(les_coeff1), float(les_coeff2), float(les_coeff4))
    if is_device(V1_in):
        check(L.sfx_fft_compute_entropy_visc_dev(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(V4_in), ptr(V5_in),
                                                 ptr(hloc_gauss), ptr(V1_out), ptr(V2_out), ptr(V3_out), np_, M,
                                                 m_max_pad, *args, stream_ptr()))
    else:
        check(L.sfx_fft_compute_entropy_visc(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(V4_in), ptr(V5_in),
                                             ptr(np.ascontiguousarray(hloc_gauss, dtype=np.float64)), ptr(V1_out),
                                             ptr(V2_out), ptr(V3_out), np_, nb, M, nb_procs, bloc_size, m_max_pad,
                                             *args, _temps_ptr(temps)))
    return V1_out, V2_out, V3_out


def FFT_COMPUTE_ENTROPY_VISC_MOM(communicator, communicator_S, V1_in, V2_in, V3_in, c1_in, hloc_gauss,
                                 c1_real_out, l_G, les_coeff2, les_coeff4, nb_procs=1, bloc_size=None,
                                 m_max_pad=None, opt_c2_real_out=None, temps=None):
    """Entropy viscosities of the momentum (and level-set) equations on the grid
    (fft_parallel.F90:2701); outputs (2*m_max_pad-1, np).  The LES coefficients
    are the reference's inputs%LES_coeff2/4, passed explicitly."""
    _single_rank(communicator)
    _single_rank(communicator_S)
    np_, nb, M = fdims(V1_in)
    fdims(V2_in), fdims(V3_in), fdims(c1_in), fdims(c1_real_out)
    bloc_size, m_max_pad = _defaults(np_, M, nb_procs, bloc_size, m_max_pad)
    if tuple(fdims(c1_real_out)) != (2 * m_max_pad - 1, np_):
        raise ValueError("c1_real_out must be (2*m_max_pad-1, np)")
    L = lib()
    if is_device(V1_in):
        check(L.sfx_fft_compute_entropy_visc_mom_dev(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(c1_in), ptr(hloc_gauss),
                                                     ptr(c1_real_out), ptr(opt_c2_real_out), np_, M, m_max_pad,
                                                     int(l_G), float(les_coeff2), float(les_coeff4), stream_ptr()))
    else:
        check(L.sfx_fft_compute_entropy_visc_mom(ptr(V1_in), ptr(V2_in), ptr(V3_in), ptr(c1_in),
                                                 ptr(np.ascontiguousarray(hloc_gauss, dtype=np.float64)),
                                                 ptr(c1_real_out), ptr(opt_c2_real_out), np_, nb, M, nb_procs,
                                                 bloc_size, m_max_pad, int(l_G), float(les_coeff2),
                                                 float(les_coeff4), _temps_ptr(temps)))
    return c1_real_out, opt_c2_real_out
