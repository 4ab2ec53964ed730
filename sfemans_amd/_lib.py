"""ctypes binding of the C-ABI in ``include/sfx.h`` (``libsfx.so``).

The library is built in-tree (``sfemans_amd/libsfx.so``) by
``sfemans_amd/build.py``.  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

__all__ = ["lib", "LIB_PATH", "SfxError", "check", "SIGNATURES"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsfx.so")

SFX_OK = 0


class SfxError(RuntimeError):
    """A non-zero status returned by libsfx (message from sfx_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"sfx error {code}: {msg}")
        self.code = code


_d = C.c_double
_pd = C.POINTER(C.c_double)
_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_pi32 = C.POINTER(C.c_int32)

# name -> argument types (restype is int unless listed in _RESTYPE)
SIGNATURES = {
    "sfx_last_error": [],
    "sfx_version": [],
    "sfx_init": [_i32],
    "sfx_stream_release": [_vp],
    "sfx_fft_par_cross_prod_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_par_dot_prod_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_par_prod_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_scalar_vect_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_tensor_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32, _i32, _vp],
    "sfx_fft_par_cross_prod_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_par_dot_prod_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_par_prod_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_scalar_vect_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_tensor_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_par_div_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_par_div_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_heaviside_dcl": [_vp, _vp, _vp, _i32, _i64, _i32, _i32, _i32, _i64, _i32, _d, _vp],
    "sfx_fft_heaviside_dcl_dev": [_vp, _vp, _vp, _i32, _i64, _i32, _i32, _d, _vp],
    "sfx_fft_squareroot_scalar": [_vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_squareroot_scalar_dev": [_vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_max_vel_dcl": [_vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32],
    "sfx_fft_scalar_vect_no_overshoot": [_vp, _i32, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i64,
                                         _i32, _vp],
    "sfx_fft_scalar_vect_no_overshoot_dev": [_vp, _i32, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_no_overshoot_level_set": [_vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_par_entro_visc_dcl": [_vp, _vp, _vp, _vp, _d, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i64, _i32,
                                   _vp],
    "sfx_fft_par_entro_visc_dcl_dev": [_vp, _vp, _vp, _d, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_compute_entropy_visc_grad_mom": [_vp, _vp, _vp, _vp, _d, _vp, _i64, _i32, _i32, _i32, _i64, _i32,
                                              _vp],
    "sfx_fft_compute_entropy_visc_grad_mom_dev": [_vp, _vp, _vp, _vp, _d, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_par_var_eta_prod_dcl": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_par_var_eta_prod_dcl_dev": [_vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_par_from_real": [_vp, _vp, _i64, _i32, _i32],
    "sfx_fft_par_from_real_dev": [_vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_max_min_vel_dcl": [_vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32],
    "sfx_fft_max_min_vel_dcl_dev": [_vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_check_interface": [_vp, _i32, _vp, _i64, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_check_interface_dev": [_vp, _i32, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_par_real_n": [_i32, _i32],
    "sfx_fft_par_real": [_vp, _vp, _i64, _i32, _i32, _i32, _i32],
    "sfx_fft_par_real_dev": [_vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_compression_level_set_dcl": [_vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, _i32, _i32, _i64, _i32, _d,
                                          _d, _d, _vp],
    "sfx_fft_compression_level_set_dcl_dev": [_vp, _vp, _vp, _vp, _i64, _i32, _i32, _d, _d, _d, _vp],
    "sfx_fft_par_compr_entro_visc_dcl": [_vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp, _i64, _i32, _i32, _i32, _i32,
                                         _i32, _i64, _i32, _vp],
    "sfx_fft_par_compr_entro_visc_dcl_dev": [_vp, _vp, _vp, _vp, _d, _d, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_compute_entropy_visc": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64,
                                     _i32, _d, _i32, _d, _d, _d, _vp],
    "sfx_fft_compute_entropy_visc_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _d, _i32,
                                         _d, _d, _d, _vp],
    "sfx_fft_compute_entropy_visc_mom": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i64, _i32,
                                         _i32, _d, _d, _vp],
    "sfx_fft_compute_entropy_visc_mom_dev": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _d, _d,
                                             _vp],
    "sfx_fft_compute_diffu_mom": [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i64, _i32, _vp],
    "sfx_fft_compute_diffu_mom_dev": [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _vp],
    "sfx_fft_no_overshoot_level_set_dev": [_vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_fft_max_vel_dcl_dev": [_vp, _vp, _i64, _i32, _i32, _i32, _vp],
    "sfx_mesh_create": [_vp, _vp, _i32, _i32, _i64, _i64, _vp],
    "sfx_mesh_destroy": [_vp],
    "sfx_mesh_info": [_vp, _vp, _vp],
    "sfx_smb_cross_prod_gauss_sft_par": [_vp, _vp, _i32, _vp, _vp, _i32, _d, _d],
    "sfx_smb_cross_prod_gauss_sft_par_dev": [_vp, _vp, _i32, _vp, _vp, _i32, _d, _d, _vp],
    "sfx_nl_rhs_3x3_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _i32, _d, _d, _vp, _vp],
    "sfx_nl_qs_00_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_smb_curlh_cross_b_gauss_sft_par": [_vp, _vp, _i32, _vp, _vp, _vp],
    "sfx_smb_curlh_cross_b_gauss_sft_par_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp],
    "sfx_smb_ugradc_gauss_fft_par": [_vp, _vp, _i32, _vp, _vp, _vp, _vp],
    "sfx_smb_ugradc_gauss_fft_par_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "sfx_nl_rotv_mhd": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp],
    "sfx_nl_rotv_mhd_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_nl_tangent": [_vp, _vp, _i32, _vp, _vp, _vp],
    "sfx_nl_tangent_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp],
    "sfx_nl_tangent_mhd_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_nl_adjoint": [_vp, _vp, _i32, _vp, _vp, _vp],
    "sfx_nl_adjoint_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "sfx_nl_adjoint_lorentz_dev": [_vp, _vp, _i32, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp],
    "sfx_nl_adjoint_mhd_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp],
    "sfx_rhs_3x3": [_vp, _i32, _vp, _vp, _vp],
    "sfx_rhs_3x3_mode": [_vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "sfx_rhs_3x3_bdy_dev": [_vp, _i32, _vp, _vp, _vp, _d, _vp],
    "sfx_mesh_set_boundary": [_vp, _vp, _i64],
    "sfx_rhs_3x3_dev": [_vp, _i32, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_rhs_ns_3x3_penal_dev": [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _vp],
    "sfx_rhs_ns_3x3_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sfx_qs_00_gauss": [_vp, _vp, _vp, _vp, _vp],
    "sfx_dot_product_sf_work": [_vp, _i32, _i32],
    "sfx_dot_product_sf_dev": [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp],
    "sfx_dot_product_sf_noalloc": [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp],
    "sfx_dot_product_sf": [_vp, _vp, _i32, _i32, _vp, _vp, _vp],
    "sfx_mass_apply_dev": [_vp, _i32, _vp, _vp, _vp, _vp],
    "sfx_mass_solve_work": [_vp, _i32],
    "sfx_mass_solve_dev": [_vp, _i32, _vp, _vp, _d, _i32, _vp, _vp, _vp],
    "sfx_nl_proj_work": [_vp, _i32],
    "sfx_nl_tangent_proj_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_nl_adjoint_sf_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp, _vp],
    "sfx_qs_00_gauss_dev": [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "sfx_nl_rhs_mhd_dev": [_vp, _vp, _i32, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp],
    "sfx_batch_create": [_vp, _i32, _vp],
    "sfx_batch_schedule": [_vp, _i32, _i32],
    "sfx_batch_autotune": [_vp, _i32, _vp],
    "sfx_batch_capture": [_vp],
    "sfx_batch_run": [_vp, _vp],
    "sfx_batch_info": [_vp, _vp, _vp, _vp, _vp, _vp],
    "sfx_batch_destroy": [_vp],
}
_RESTYPE = {"sfx_last_error": C.c_char_p, "sfx_version": C.c_char_p, "sfx_dot_product_sf_work": C.c_int64,
            "sfx_mass_solve_work": C.c_int64, "sfx_nl_proj_work": C.c_int64}

_lib = None


def lib():
    """Load libsfx.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m sfemans_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def check(rc: int):
    if rc != SFX_OK:
        msg = lib().sfx_last_error()
        raise SfxError(rc, msg.decode() if msg else "")
