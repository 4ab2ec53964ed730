"""Gauss-point preparers of the nonlinear terms (restatement).

* ``gauss_velocity_curl`` - the Gauss loop of ``smb_cross_prod_gauss_sft_par``
  (SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:559-611): ``W`` (the
  field) and ``RotV`` (its curl in cos/sin form, with ``mode/ray`` terms).
* ``smb_cross_prod_gauss_sft_par`` (:514-630), including the precession
  term ``RotV += 2*Om`` (:536-551, 613-620), then
  ``FFT_PAR_CROSS_PROD_DCL(RotV, W)`` with ``m_max_pad = 3*M/2`` (:624-627).
* ``smb_CurlH_cross_B_gauss_sft_par`` (:632-727): curl of ``V_in`` crossed
  with the Gauss values of ``W_in``.
* ``smb_ugradc_gauss_fft_par`` (SOURCES/SIMULATION/subroutine_temperature_with_T.F90:247-350):
  ``DOT(heat_capa*grad c, u) + PROD(div u, heat_capa*c)``.

Gauss point ``index = (m-1)*l_G + l`` (element-major), so arrays are
(G, ncol, M) with ``G = 7*me``.
"""
from __future__ import annotations

import numpy as np

from .gauss_p2 import gauss_tables, element_2d_p2
from .fft_par import (fft_par_cross_prod_dcl, fft_par_dot_prod_dcl,
                      fft_par_prod_dcl)

__all__ = ["interp", "gauss_velocity_curl", "smb_cross_prod_gauss_sft_par",
           "smb_CurlH_cross_B_gauss_sft_par", "smb_ugradc_gauss_fft_par",
           "precession_omega", "m_max_pad_of"]


def m_max_pad_of(M: int, nb_procs: int = 1) -> int:
    """``m_max_pad = 3*SIZE(list_mode)*nb_procs/2`` (subroutine_ns_with_velocity.F90:626)."""
    return (3 * M * nb_procs) // 2


def interp(jj, rr, V_in):
    """Values and physical r/z derivatives of nodal data at Gauss points.

    Returns ``(W, Dr, Dz, ray)``, each of W/Dr/Dz shaped (me, 7, ncol, M),
    ``ray`` (me, 7).
    """
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = np.asarray(jj, dtype=np.int64) - 1
    Vs = np.asarray(V_in)[j0.T]                       # (me, 6, ncol, M)
    W = np.einsum("enki,nl->elki", Vs, ww)
    Dr = np.einsum("enki,nle->elki", Vs, dw[0])
    Dz = np.einsum("enki,nle->elki", Vs, dw[1])
    return W, Dr, Dz, ray.T


def _gauss_flat(X):
    me, l, k, M = X.shape
    return np.ascontiguousarray(X.reshape(me * l, k, M))


def gauss_velocity_curl(jj, rr, list_mode, V_in):
    """``(W, RotV)`` at Gauss points, (G, 6, M) each (:577-608)."""
    W, Dr, Dz, ray = interp(jj, rr, V_in)
    mode = np.asarray(list_mode, dtype=np.float64)[None, None, :]
    r = ray[:, :, None]
    mr = mode / r                                      # (me, 7, M)
    R = np.empty_like(W)
    R[:, :, 0] = mr * W[:, :, 5] - Dz[:, :, 2]
    R[:, :, 3] = Dz[:, :, 1] - Dr[:, :, 5]
    R[:, :, 4] = 1 / r * W[:, :, 2] + Dr[:, :, 2] - mr * W[:, :, 1]
    R[:, :, 1] = -mr * W[:, :, 4] - Dz[:, :, 3]
    R[:, :, 2] = Dz[:, :, 0] - Dr[:, :, 4]
    R[:, :, 5] = 1 / r * W[:, :, 3] + Dr[:, :, 3] + mr * W[:, :, 0]
    return _gauss_flat(W), _gauss_flat(R)


def precession_omega(list_mode, taux_precession, angle_s_pi):
    """Nodal precession vector per mode (subroutine_ns_with_velocity.F90:536-551).

    Returns a (6, M) array of the constant nodal values.
    """
    PI = np.arccos(-1.0)
    om = np.zeros((6, len(list_mode)))
    for i, mode in enumerate(list_mode):
        if mode == 1:
            om[0, i] = taux_precession * np.sin(angle_s_pi * PI)
            om[3, i] = -taux_precession * np.sin(angle_s_pi * PI)
        elif mode == 0:
            om[4, i] = taux_precession * np.cos(angle_s_pi * PI)
    return om


def smb_cross_prod_gauss_sft_par(jj, rr, list_mode, V_in, precession_in=False,
                                 taux_precession=0.0, angle_s_pi=0.0, engine="fft"):
    """``V_out = P_M[(curl V + 2 Om) x V]`` at Gauss points, (G, 6, M)."""
    W, RotV = gauss_velocity_curl(jj, rr, list_mode, V_in)
    if precession_in:
        ww = element_2d_p2()[0]
        om = precession_omega(list_mode, taux_precession, angle_s_pi)
        # Om(index,k,i) = SUM(Omega_s(:,k)*ww(:,l)) with a constant nodal Omega
        Om = om[None, :, :] * ww.sum(axis=0)[:, None, None]     # (7, 6, M)
        RotV = RotV + 2.0 * np.tile(Om, (np.shape(jj)[1], 1, 1))
    return fft_par_cross_prod_dcl(RotV, W, m_max_pad_of(len(list_mode)), engine)


def smb_CurlH_cross_B_gauss_sft_par(jj, rr, list_mode, V_in, W_in, engine="fft"):
    """``V_out = P_M[(curl V_in) x W_in]`` at Gauss points, (G, 6, M)."""
    _, RotV = gauss_velocity_curl(jj, rr, list_mode, V_in)
    Wg, _, _, _ = interp(jj, rr, W_in)
    return fft_par_cross_prod_dcl(RotV, _gauss_flat(Wg), m_max_pad_of(len(list_mode)), engine)


def smb_ugradc_gauss_fft_par(jj, rr, list_mode, heat_capa_in, V_in, c_in, engine="fft"):
    """``c_out = u.grad(c) + (div u) c`` with per-element ``heat_capa``, (G, 2, M)."""
    W, Dr, Dz, ray = interp(jj, rr, V_in)
    C, Cr, Cz, _ = interp(jj, rr, c_in)
    mode = np.asarray(list_mode, dtype=np.float64)[None, None, :]
    r = ray[:, :, None]
    hc = np.asarray(heat_capa_in, dtype=np.float64)[:, None, None]
    Div = np.empty(W.shape[:2] + (2,) + W.shape[3:])
    Div[:, :, 0] = Dr[:, :, 0] + W[:, :, 0] / r + (mode / r) * W[:, :, 3] + Dz[:, :, 4]
    Div[:, :, 1] = Dr[:, :, 1] + W[:, :, 1] / r - (mode / r) * W[:, :, 2] + Dz[:, :, 5]
    G = np.empty(W.shape)
    G[:, :, 0] = Cr[:, :, 0]
    G[:, :, 1] = Cr[:, :, 1]
    G[:, :, 2] = mode / r * C[:, :, 1]
    G[:, :, 3] = -mode / r * C[:, :, 0]
    G[:, :, 4] = Cz[:, :, 0]
    G[:, :, 5] = Cz[:, :, 1]
    G = hc[:, :, :, None] * G
    Cg = hc[:, :, :, None] * C
    mp = m_max_pad_of(len(list_mode))
    c_out = fft_par_dot_prod_dcl(_gauss_flat(G), _gauss_flat(W), mp, engine)
    cint = fft_par_prod_dcl(_gauss_flat(Div), _gauss_flat(Cg), mp, engine)
    return c_out + cint
