"""P2 reference element and per-element Gauss data.

``element_2d_p2`` restates SOURCES/GAUSS_POINTS/GP_2D_p2.F90:9-88: the six
quadratic shape functions on the triangle (0,0),(1,0),(0,1) with node order
"3 / 5 4 / 1 6 2", their derivatives, and Stroud's 7-point degree-5 rule.

``gauss_tables`` restates the volume loop of ``create_gauss_points_2d``
(SOURCES/GAUSS_POINTS/gauss_points_2d.F90:102-125): the Jacobian
``dr(k,k1) = sum_n rr(k, jj(n,m)) * dd(k1,n,l)``, the physical derivatives
``dw = J^{-T} dd`` and ``rj = |det J| * p``.  ``ray`` is the radius of each
Gauss point, ``SUM(rr(1,jj(:,m))*ww(:,l))`` (subroutine_ns_with_velocity.F90:574).
"""
from __future__ import annotations

import numpy as np

__all__ = ["element_2d_p2", "gauss_tables"]


def element_2d_p2():
    """Return ``w(6,7)``, ``d(2,6,7)``, ``p(7)`` (GP_2D_p2.F90:9-88)."""
    half, one, two, three, four = 0.5, 1.0, 2.0, 3.0, 4.0
    area = one / two
    sq = np.sqrt(15.0)
    r = one / three
    a = area * 9.0 / 40
    s1 = (6 - sq) / 21
    t1 = (9 + 2 * sq) / 21
    b1 = area * (155 - sq) / 1200
    s2 = (6 + sq) / 21
    t2 = (9 - 2 * sq) / 21
    b2 = area * (155 + sq) / 1200
    xx = np.array([r, s1, s1, t1, s2, s2, t2])
    yy = np.array([r, s1, t1, s1, s2, t2, s2])
    p = np.array([a, b1, b1, b1, b2, b2, b2])
    x, y = xx, yy
    w = np.empty((6, 7))
    d = np.empty((2, 6, 7))
    w[0] = (half - x - y) * (one - x - y) * two
    w[1] = x * (x - half) * two
    w[2] = y * (y - half) * two
    w[3] = x * y * four
    w[4] = y * (one - x - y) * four
    w[5] = x * (one - x - y) * four
    d[0, 0] = -three + four * (x + y)
    d[0, 1] = (two * x - half) * two
    d[0, 2] = 0.0
    d[0, 3] = y * four
    d[0, 4] = -y * four
    d[0, 5] = (one - two * x - y) * four
    d[1, 0] = -three + four * (x + y)
    d[1, 1] = 0.0
    d[1, 2] = (two * y - half) * two
    d[1, 3] = x * four
    d[1, 4] = (one - x - two * y) * four
    d[1, 5] = -x * four
    return w, d, p


def gauss_tables(jj: np.ndarray, rr: np.ndarray):
    """Per-element data for a P2 mesh.

    ``jj`` (6, me) 1-based, ``rr`` (2, np).  Returns ``ww (6,7)``,
    ``dw (2,6,7,me)``, ``rj (7,me)`` and ``ray (7,me)``.
    """
    ww, dd, pp = element_2d_p2()
    j0 = np.asarray(jj, dtype=np.int64) - 1                 # (6, me)
    rloc = np.asarray(rr)[:, j0]                            # (2, 6, me)
    # dr(k, k1, l, m) = sum_n rloc(k, n, m) * dd(k1, n, l)
    dr = np.einsum("knm,jnl->kjlm", rloc, dd)
    rjac = dr[0, 0] * dr[1, 1] - dr[0, 1] * dr[1, 0]        # (7, me)
    dw = np.empty((2, 6, 7, j0.shape[1]))
    dw[0] = (dd[0][:, :, None] * dr[1, 1][None] - dd[1][:, :, None] * dr[1, 0][None]) / rjac[None]
    dw[1] = (-dd[0][:, :, None] * dr[0, 1][None] + dd[1][:, :, None] * dr[0, 0][None]) / rjac[None]
    rj = np.abs(rjac) * pp[:, None]
    ray = np.einsum("nm,nl->lm", rloc[0], ww)
    return ww, dw, rj, ray
