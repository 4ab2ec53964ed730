"""Element-to-node right-hand-side assembly (restatement).

``rhs_3x3`` follows SOURCES/SOLVER/assembling_rhs.F90:5-126, 185-188 without
``opt_tensor_scaln_bdy``: for each element and Gauss point
``smb = (rhs_in(index,:) + f_tensor)*ray*rj(l,m)`` (:71-73); components
(1,4,5) go to ``vb_145`` and (2,3,6) to ``vb_236`` for velocity components
ki = 1,2,3 (:75-95), added at row ``loc_to_glob(ki, jj(ni,m))``.  On one rank
``loc_to_glob(ki, i) = (ki-1)*np + i`` (SOURCES/INITIALISATION/st_matrix.F90:101),
so the dense vectors here are (np, 3), column ki.  With ``opt_tensor``
T(3, G, 6) per mode (row k = velocity component, column = the 6 (cos, sin)
columns of the derivative direction r, theta, z), ``f_tensor`` holds its
1/r and mode terms (:56-68) and every node adds
``(dw(1,ni)*T(ki, c) + dw(2,ni)*T(ki, c+4))*ray*rj`` with c = 1 for the
cos slot and 2 for the sin slot of row ki (:97-121): the weak form
``int T : grad(phi) r dr dz`` of the tensor divergence.

``qs_00_gauss`` follows SOURCES/SOLVER/fem_rhs_axi.F90:940-987:
``fl = (heat_capa(m)*SUM(ff_loc*ww(:,l)) + ff_gauss(index))*rj(l,m)*ray``.

``rhs_3x3_bdy`` is the ``opt_tensor_scaln_bdy`` term of rhs_3x3
(assembling_rhs.F90:129-185): for every boundary side ``ms`` (nodes
``jjs(1:3, ms)``, the two ends then the midpoint) and each of its three
Gauss points, ``ray = SUM(rr(1,jjs(:,ms))*wws(:,ls))`` (points with
``ray < 1e-10`` skipped), ``smb = T_bdy(indexs,:)*ray*rjs(ls,ms)`` with
``rjs = |d rr/ds| pps(ls)`` (gauss_points_2d.F90:182-190) and the quadratic
edge element of GP_2D_p2.F90:176-212; then the (1,4,5)/(2,3,6) packing of
the volume term.

``rhs_ns_gauss_3x3`` follows SOURCES/SOLVER/computing_rhs_gauss.F90:7-146,
the penalty options included (below); without them, per mode i and Gauss
point, ``rhs_gauss = (ft + fp_inc + fs - rotv_v) + fp`` (:100, :145) with
``ft = SUM(V1m*ww)``, ``fs = SUM(ff*ww)`` (``ff`` the nodal momentum
source), and the pressure terms of the pressure injected onto the velocity
mesh (``P``, ``P_inc``, both (np, 2, M)) (:69-97):
``fp = (-d_r P_c, -d_r P_s, -m P_s/r, m P_c/r, -d_z P_c, -d_z P_s)``, where
``rotv_v = (curl u) x u - coeff_lorentz*(curl H) x B``
(SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:227, 246-248).
"""
from __future__ import annotations

import numpy as np

from .gauss_p2 import gauss_tables

__all__ = ["rhs_3x3", "rhs_3x3_all_modes", "qs_00_gauss", "rhs_ns_gauss_3x3", "element_1d_p2", "rhs_3x3_bdy"]

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


def rhs_3x3_all_modes(jj, rr, rhs_gauss, list_mode=None, opt_tensor=None):
    """All modes: (G, 6, M) -> ``(vb_145, vb_236)`` each (np, 3, M).

    ``opt_tensor``: (3, G, 6, M) or None; ``list_mode`` (M,) is needed with it.
    """
    ww, dw, rj, ray = gauss_tables(jj, rr)
    me = np.shape(jj)[1]
    M = rhs_gauss.shape[2]
    f = np.asarray(rhs_gauss, dtype=np.float64).reshape(me, 7, 6, M)
    wgt = (ray.T * rj.T)[:, :, None, None]                 # (me, 7, 1, 1)
    if opt_tensor is not None:
        T = np.asarray(opt_tensor, dtype=np.float64).reshape(3, me, 7, 6, M)
        mode = np.asarray(list_mode, dtype=np.float64)[None, None, :]
        r = ray.T[:, :, None]
        ft = np.empty((me, 7, 6, M))
        ft[:, :, 0] = (-mode * T[0, :, :, 3] + T[1, :, :, 2]) / r
        ft[:, :, 1] = (mode * T[0, :, :, 2] + T[1, :, :, 3]) / r
        ft[:, :, 2] = (-T[0, :, :, 2] - mode * T[1, :, :, 3]) / r
        ft[:, :, 3] = (-T[0, :, :, 3] + mode * T[1, :, :, 2]) / r
        ft[:, :, 4] = -T[2, :, :, 3] * mode / r
        ft[:, :, 5] = T[2, :, :, 2] * mode / r
        f = f + ft
    loc = np.einsum("nl,elkm->enkm", ww, f * wgt)          # (me, 6, 6, M)
    if opt_tensor is not None:
        # component c: row c // 2 of T, its cos (c even) or sin (c odd) columns
        for c in range(6):
            k, p = c // 2, c % 2
            g = (np.einsum("nle,elm->enlm", dw[0], T[k, :, :, p])
                 + np.einsum("nle,elm->enlm", dw[1], T[k, :, :, 4 + p]))   # (me, 6, 7, M)
            loc[:, :, c, :] += np.einsum("enlm,el->enm", g, ray.T * rj.T)
    npts = np.shape(rr)[1]
    v145 = _nodal_sum(jj, npts, loc[:, :, list(_C145), :])
    v236 = _nodal_sum(jj, npts, loc[:, :, list(_C236), :])
    return v145, v236


def element_1d_p2():
    """Quadratic edge element with three Gauss points (GP_2D_p2.F90:176-212):
    ``w (3, 3)`` values, ``d (3, 3)`` derivatives of the shape functions of the
    end nodes (x = -1, +1) and of the midpoint (x = 0) at the Gauss points
    x = -sqrt(3/5), 0, sqrt(3/5), and the weights ``p (3,)``."""
    xx = np.array([-np.sqrt(3.0 / 5.0), 0.0, np.sqrt(3.0 / 5.0)])
    w = np.stack([(xx - 1) * xx / 2, (xx + 1) * xx / 2, (xx + 1) * (1 - xx)])
    d = np.stack([(2 * xx - 1) / 2, (2 * xx + 1) / 2, -2 * xx])
    p = np.array([5.0, 8.0, 5.0]) / 9.0
    return w, d, p


def rhs_3x3_bdy(jjs, rr, tbdy):
    """Boundary term of rhs_3x3: ``tbdy`` (3*mes, 6, M) at the sides' Gauss
    points -> ``(vb_145, vb_236)``, each (np, 3, M)."""
    w, d, p = element_1d_p2()
    js = np.asarray(jjs, dtype=np.int64) - 1                # (3, mes)
    mes = js.shape[1]
    rs = np.asarray(rr)[:, js]                             # (2, 3, mes)
    ray = np.einsum("nm,nl->lm", rs[0], w)                 # (3 ls, mes)
    drs = np.einsum("knm,nl->klm", rs, d)                  # (2, ls, mes)
    rjs = np.sqrt(drs[0] ** 2 + drs[1] ** 2) * p[:, None]
    wgt = np.where(ray < 1e-10, 0.0, ray * rjs)            # skipped points add nothing
    T = np.asarray(tbdy, dtype=np.float64).reshape(mes, 3, 6, -1)
    loc = np.einsum("nl,mlkq->mnkq", w, T * wgt.T[:, :, None, None])   # (mes, 3 nodes, 6, M)
    npts = np.shape(rr)[1]
    v145 = _nodal_sum(js + 1, npts, loc[:, :, list(_C145), :])
    v236 = _nodal_sum(js + 1, npts, loc[:, :, list(_C236), :])
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


def rhs_ns_gauss_3x3(jj, rr, list_mode, V1m, ff, P, P_inc, rotv_v, penal_grid=None, imposed_vel=None):
    """``rhs_gauss`` (G, 6, M) of computing_rhs_gauss.F90:7-146.

    ``V1m``, ``ff``: nodal (np, 6, M); ``P``, ``P_inc``: pressure on the
    velocity mesh, (np, 2, M); ``rotv_v``: (G, 6, M).  With ``penal_grid``
    (``if_ns_penalty``): the values (N, G) of penal_in_real_space on the
    grid N = 2*m_max_pad - 1, m_max_pad = 3*M/2 (:124-129), at the Gauss
    points; ``imposed_vel`` (``if_impose_vel_in_solids``): the nodal (np, 6)
    array 3*imposed_velocity_by_penalty/(2 dt) of :107, subtracted from the
    mode-0 slot before the product with the penalty and added back after it
    (:105-121, :132-138); then fp is added (:145).
    """
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = (np.asarray(jj, dtype=np.int64) - 1).T            # (me, 6)
    me = j0.shape[0]
    M = np.shape(V1m)[2]
    mode = np.asarray(list_mode, dtype=np.float64)[None, None, :]
    r = ray.T[:, :, None]                                  # (me, 7, 1)

    def val(X):   # (np, c, M) -> (me, 7, c, M)
        return np.einsum("enkm,nl->elkm", np.asarray(X, dtype=np.float64)[j0], ww)

    def grad(X, d):   # physical derivative d (0: r, 1: z)
        return np.einsum("enkm,nle->elkm", np.asarray(X, dtype=np.float64)[j0], dw[d])

    def fpress(Q):
        Qv, Qr, Qz = val(Q), grad(Q, 0), grad(Q, 1)
        f = np.empty((me, 7, 6, M))
        f[:, :, 0] = -Qr[:, :, 0]
        f[:, :, 1] = -Qr[:, :, 1]
        f[:, :, 2] = -Qv[:, :, 1] / r * mode
        f[:, :, 3] = Qv[:, :, 0] / r * mode
        f[:, :, 4] = -Qz[:, :, 0]
        f[:, :, 5] = -Qz[:, :, 1]
        return f
    rv = np.asarray(rotv_v, dtype=np.float64).reshape(me, 7, 6, M)
    if penal_grid is None:
        out = (val(V1m) + fpress(P_inc) + val(ff) - rv) + fpress(P)
        return np.ascontiguousarray(out.reshape(me * 7, 6, M))
    from .fft_par import fft_par_var_eta_prod
    g = val(V1m) + fpress(P_inc) + val(ff) - rv                       # (me, 7, 6, M)
    zero = [i for i, md in enumerate(np.asarray(list_mode)) if md == 0]
    if imposed_vel is not None and zero:
        imp = val(np.asarray(imposed_vel, dtype=np.float64)[:, :, None])[:, :, :, 0]   # (me, 7, 6)
        for i in zero:
            g[:, :, :, i] = g[:, :, :, i] - imp
    gp = fft_par_var_eta_prod(penal_grid, g.reshape(me * 7, 6, M), (3 * M) // 2).reshape(me, 7, 6, M)
    if imposed_vel is not None and zero:
        for i in zero:
            gp[:, :, :, i] = gp[:, :, :, i] + imp
    out = gp + fpress(P)
    return np.ascontiguousarray(out.reshape(me * 7, 6, M))

