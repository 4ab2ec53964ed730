"""The inner product of SFEMaNS fields and the adjoint of the linearized
convective term in it (restatement).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``.

* ``dot_product_sf`` follows ``dot_product_SF``
  (SOURCES/SOLVER/tn_axi.F90:9-37) through ``dot_product_champ``
  (SOURCES/SOLVER/fem_tn_ns_mhd.F90:92-135) and the element loop
  ``dot_product`` (SOURCES/SOLVER/fem_tn_axi.F90:8-48): for every mode slot
  k and column c, SUM over elements and Gauss points of
  ``SUM(v*ww) * SUM(w*ww) * rj * ray``, halved for ``list_mode(k) /= 0``,
  summed and multiplied by 2*pi.  LightKrylov's ``dot`` on SFEMaNS fields is
  this product (SOURCES/TOOLS/def_type_field.F90:176-198).  In matrix form
  <v, w>_W = 2 pi SUM_k w_k SUM_c v(:,c,k)^T M w(:,c,k) with M the
  r-weighted P2 mass matrix and w_k = 1 for mode 0, 1/2 otherwise.
* ``mass_matrix`` is M, assembled sparse from the same quadrature.
* ``projected_tangent`` is A x = M^{-1} R T_U x: the L2(r dr dz) projection
  onto the P2 space, mode by mode and column by column, of the tangent of
  the convective term (``linearized.tangent``), with R the Galerkin
  right-hand side SUM_l ww(i,l) f(l) rj ray of ``qs_00_gauss``.
* ``projected_adjoint`` is its adjoint in <., .>_W:
  A* y = M^{-1} w^{-1} T_U^T R^T (w y), where T_U^T is the Euclidean
  transpose (``linearized.adjoint_bruteforce``, or the kernel under test)
  and R^T z = the Gauss values of z times rj ray.

Pins (tests/test_oracle_inner.py): <1, 1>_W = 2 pi times the domain's
integral of r (closed form on rectangles and per mode weight), quadratic
fields integrated exactly; the adjoint equals W^{-1} A^T W formed from the
dense matrices of A and W on a tiny mesh.
"""
from __future__ import annotations

import numpy as np

from .gauss_p2 import gauss_tables

__all__ = ["dot_product_sf", "mode_weights", "mass_matrix", "galerkin_rhs", "gauss_weighted",
           "projected_tangent", "projected_adjoint"]


def mode_weights(list_mode) -> np.ndarray:
    """1 for mode 0, 1/2 otherwise (fem_tn_ns_mhd.F90:115-118)."""
    return np.where(np.asarray(list_mode) == 0, 1.0, 0.5)


def dot_product_sf(jj, rr, list_mode, v, w) -> float:
    """2 pi SUM_k w_k SUM_c SUM_(e,l) v_g w_g rj ray for nodal v, w (np, nc, M)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = np.asarray(jj, dtype=np.int64) - 1                 # (6, me)
    vg = np.einsum("nekm,nl->lekm", np.asarray(v)[j0], ww)   # (7, me, nc, M)
    wg = np.einsum("nekm,nl->lekm", np.asarray(w)[j0], ww)
    s = np.einsum("lekm,lekm,le->m", vg, wg, rj * ray)      # per mode slot
    return float(2.0 * np.pi * np.sum(mode_weights(list_mode) * s))


def mass_matrix(jj, rr):
    """Sparse r-weighted P2 mass matrix M_ij = SUM_(e,l) ww(i,l) ww(j,l) rj ray."""
    import scipy.sparse as sp
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = np.asarray(jj, dtype=np.int64) - 1
    wt = rj * ray                                           # (7, me)
    loc = np.einsum("il,jl,le->eij", ww, ww, wt)            # (me, 6, 6)
    rows = np.repeat(j0.T, 6, axis=1).ravel()
    cols = np.tile(j0.T, (1, 6)).ravel()
    npts = np.shape(rr)[1]
    return sp.csr_matrix((loc.ravel(), (rows, cols)), shape=(npts, npts))


def galerkin_rhs(jj, rr, f):
    """R f: (R f)_i = SUM_(e,l) ww(i,l) f(7e+l) rj ray, f (G, ...) -> (np, ...)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = (np.asarray(jj, dtype=np.int64) - 1).T             # (me, 6)
    me = j0.shape[0]
    fg = np.asarray(f, dtype=np.float64).reshape((me, 7) + np.shape(f)[1:])
    wt = (rj * ray).T.reshape((me, 7) + (1,) * (fg.ndim - 2))
    loc = np.einsum("il,el...->ei...", ww, fg * wt)
    out = np.zeros((np.shape(rr)[1],) + fg.shape[2:])
    np.add.at(out, j0.ravel(), loc.reshape((-1,) + fg.shape[2:]))
    return out


def gauss_weighted(jj, rr, z):
    """R^T z: the Gauss values of nodal z (np, ...) times rj ray, (G, ...)."""
    ww, dw, rj, ray = gauss_tables(jj, rr)
    j0 = np.asarray(jj, dtype=np.int64) - 1
    zg = np.einsum("ne...,nl->el...", np.asarray(z)[j0], ww)   # (me, 7, ...)
    wt = (rj * ray).T.reshape(zg.shape[:2] + (1,) * (zg.ndim - 2))
    return (zg * wt).reshape((-1,) + zg.shape[2:])


def _solve(jj, rr, b):
    from scipy.sparse.linalg import splu
    lu = splu(mass_matrix(jj, rr).tocsc())
    flat = np.asarray(b).reshape(b.shape[0], -1)
    return lu.solve(flat).reshape(b.shape)


def projected_tangent(jj, rr, list_mode, U, D):
    """A D = M^{-1} R T_U D, nodal (np, 6, M)."""
    from .linearized import tangent
    return _solve(jj, rr, galerkin_rhs(jj, rr, tangent(jj, rr, list_mode, U, D)))


def projected_adjoint(jj, rr, list_mode, U, Y, euclidean_adjoint=None):
    """A* Y = M^{-1} w^{-1} T_U^T R^T (w Y), nodal (np, 6, M); the Euclidean
    transpose T_U^T is ``linearized.adjoint_bruteforce`` unless given."""
    from .linearized import adjoint_bruteforce
    wk = mode_weights(list_mode)
    g = gauss_weighted(jj, rr, np.asarray(Y) * wk[None, None, :])
    tt = adjoint_bruteforce(jj, rr, list_mode, U, g) if euclidean_adjoint is None else euclidean_adjoint(g)
    return _solve(jj, rr, tt / wk[None, None, :])
