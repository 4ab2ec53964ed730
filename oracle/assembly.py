"""Element-to-node right-hand-side assembly (restatement).

``rhs_3x3`` follows SOURCES/SOLVER/assembling_rhs.F90:5-126, 185-188 without
``opt_tensor``/``opt_tensor_scaln_bdy``: for each element and Gauss point
``smb = rhs_in(index,:)*ray*rj(l,m)`` (:71-73); components (1,4,5) go to
``vb_145`` and (2,3,6) to ``vb_236`` for velocity components ki = 1,2,3
(:75-95), added at row ``loc_to_glob(ki, jj(ni,m))``.  On one rank
``loc_to_glob(ki, i) = (ki-1)*np + i`` (SOURCES/INITIALISATION/st_matrix.F90:101),
so the dense vectors here are (np, 3), column ki.

``qs_00_gauss`` follows SOURCES/SOLVER/fem_rhs_axi.F90:940-987:
``fl = (heat_capa(m)*SUM(ff_loc*ww(:,l)) + ff_gauss(index))*rj(l,m)*ray``.

``rhs_ns_gauss_nl`` is the nonlinear share ``-rotv_v`` of ``rhs_ns_gauss_3x3``
(SOURCES/SOLVER/computing_rhs_gauss.F90:100) with
``rotv_v = (curl u) x u - coeff_lorentz*(curl H) x B``
(SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:227, 246-248).
"""
from __future__ import annotations

import numpy as np

from .gauss_p2 import gauss_tables

__all__ = ["rhs_3x3", "rhs_3x3_all_modes", "qs_00_gauss"]

_C145 = (0, 3, 4)
_C236 = (1, 2, 5)


def _nodal_sum(jj, npts, contrib):
    """Sum per-element nodal contributions (me, 6, ...) into (np, ...)."""
    j0 = (np.asarray(jj, dtype=np.int64) - 1).T          # (me, 6)
    out = np.zeros((npts,) + contrib.shape[2:])
    np.add.at(out, j0.ravel(), contrib.reshape((-1,) + contrib.shape[2:]))
    return out


def rhs_3x3(jj, rr, rhs_in):
    """One mode: ``rhs_in`` (G, 6) -> ``(vb_145, vb_236)``, each (np, 3)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    me = np.shape(jj)[1]
    f = np.asarray(rhs_in, dtype=np.float64).reshape(me, 7, 6)
    smb = f * (ray.T * rj.T)[:, :, None]                   # (me, 7, 6)
    loc = np.einsum("nl,elk->enk", ww, smb)                # (me, 6 nodes, 6 comps)
    npts = np.shape(rr)[1]
    v145 = _nodal_sum(jj, npts, loc[:, :, list(_C145)])
    v236 = _nodal_sum(jj, npts, loc[:, :, list(_C236)])
    return v145, v236


def rhs_3x3_all_modes(jj, rr, rhs_gauss):
    """All modes: (G, 6, M) -> ``(vb_145, vb_236)`` each (np, 3, M)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    me = np.shape(jj)[1]
    M = rhs_gauss.shape[2]
    f = np.asarray(rhs_gauss, dtype=np.float64).reshape(me, 7, 6, M)
    smb = f * (ray.T * rj.T)[:, :, None, None]
    loc = np.einsum("nl,elkm->enkm", ww, smb)              # (me, 6, 6, M)
    npts = np.shape(rr)[1]
    v145 = _nodal_sum(jj, npts, loc[:, :, list(_C145), :])
    v236 = _nodal_sum(jj, npts, loc[:, :, list(_C236), :])
    return v145, v236


def qs_00_gauss(jj, rr, heat_capa, ff, ff_gauss):
    """Scalar RHS (np,) from nodal ``ff`` (np,) and Gauss ``ff_gauss`` (G,)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = (np.asarray(jj, dtype=np.int64) - 1).T
    me = j0.shape[0]
    ffl = np.asarray(ff, dtype=np.float64)[j0]             # (me, 6)
    fg = np.asarray(ff_gauss, dtype=np.float64).reshape(me, 7)
    hc = np.asarray(heat_capa, dtype=np.float64)[:, None]
    fl = (hc * (ffl @ ww) + fg) * rj.T * ray.T             # (me, 7)
    loc = fl @ ww.T                                        # (me, 6)
    return _nodal_sum(jj, np.shape(rr)[1], loc)
