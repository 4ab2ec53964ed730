"""Synthetic P2 meridian meshes with the reference's node numbering.

The reference reads a P1 triangulation, sorts the three vertices of every cell
by increasing global index (SOURCES/MESH/load_mesh.F90:101-112) and builds the
P2 grid by appending one node per edge after the vertices
(SOURCES/MESH/refine_mesh.F90:310-365, 423-428).  Local node ``3+k`` is the
midpoint of the edge opposite local vertex ``k``
(SOURCES/GAUSS_POINTS/GP_2D_p2.F90:15-17) and its coordinates are computed as
``rr(start) + (rr(end) - rr(start))/2`` from the lower-index end
(refine_mesh.F90:341-343).

``rectangle_p2`` produces the same structure for an ``nx x ny`` grid of squares
on ``[r0,r1] x [z0,z1]``, each square split into two triangles, elements
numbered row by row and edges numbered in order of first encounter.  The
arrays use the Fortran layout the C-ABI receives: ``jj(6, me)`` (1-based,
column-major) and ``rr(2, np)``.
"""
from __future__ import annotations

import numpy as np

__all__ = ["P2Mesh", "rectangle_p2"]


class P2Mesh:
    """Connectivity ``jj`` (6, me) int32 1-based and coordinates ``rr`` (2, np)."""

    def __init__(self, jj: np.ndarray, rr: np.ndarray, nx: int = 0, ny: int = 0):
        self.jj = np.asfortranarray(jj, dtype=np.int32)
        self.rr = np.asfortranarray(rr, dtype=np.float64)
        self.nx, self.ny = nx, ny

    @property
    def me(self) -> int:
        return int(self.jj.shape[1])

    @property
    def np(self) -> int:
        return int(self.rr.shape[1])

    @property
    def n_w(self) -> int:
        return 6

    @property
    def l_G(self) -> int:
        return 7

    @property
    def n_gauss(self) -> int:
        return 7 * self.me

    def __repr__(self) -> str:
        return f"P2Mesh(me={self.me}, np={self.np}, nx={self.nx}, ny={self.ny})"


def rectangle_p2(nx: int, ny: int, r0: float = 0.0, r1: float = 1.0,
                 z0: float = -1.0, z1: float = 1.0) -> P2Mesh:
    """Structured P2 mesh of ``2*nx*ny`` triangles on ``[r0,r1] x [z0,z1]``.

    Vertices are numbered row by row (r fastest).  Square (i, j) is split
    along its (i,j)-(i+1,j+1) diagonal.  The vertex triplet of every cell is
    sorted ascending, as the reference does when it loads a mesh.
    """
    if nx < 1 or ny < 1:
        raise ValueError("nx and ny must be >= 1")
    nvr = nx + 1
    nv = nvr * (ny + 1)
    ir, iz = np.meshgrid(np.arange(nvr), np.arange(ny + 1), indexing="xy")
    vr = r0 + (r1 - r0) * ir.ravel() / nx
    vz = z0 + (z1 - z0) * iz.ravel() / ny

    # cells, vertices 0-based, two per square, rows of squares in z
    i = np.arange(nx)
    j = np.arange(ny)
    I, J = np.meshgrid(i, j, indexing="xy")
    v00 = (J * nvr + I).ravel()
    v10 = v00 + 1
    v01 = v00 + nvr
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    tri = np.empty((2 * nx * ny, 3), dtype=np.int64)
    tri[0::2] = t1
    tri[1::2] = t2
    tri.sort(axis=1)
    me = tri.shape[0]

    # edge opposite local vertex k joins the two other vertices (k = 0,1,2)
    opp = np.stack([np.stack([tri[:, 1], tri[:, 2]], 1),
                    np.stack([tri[:, 0], tri[:, 2]], 1),
                    np.stack([tri[:, 0], tri[:, 1]], 1)], axis=1)  # (me, 3, 2)
    keys = opp[:, :, 0] * nv + opp[:, :, 1]                        # (me, 3)
    flat = keys.ravel()                                            # order m, then k
    uniq, first_idx, inv = np.unique(flat, return_index=True, return_inverse=True)
    # renumber edges in order of first encounter
    order = np.argsort(first_idx, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    edge_id = rank[inv].reshape(me, 3)
    nedge = uniq.size
    npts = nv + nedge

    rr = np.empty((2, npts), dtype=np.float64)
    rr[0, :nv] = vr
    rr[1, :nv] = vz
    ekeys = uniq[order]
    es = ekeys // nv
    ee = ekeys % nv
    rr[0, nv:] = vr[es] + (vr[ee] - vr[es]) / 2.0
    rr[1, nv:] = vz[es] + (vz[ee] - vz[es]) / 2.0

    jj = np.empty((6, me), dtype=np.int32)
    jj[0:3, :] = tri.T + 1
    jj[3:6, :] = (nv + edge_id).T + 1
    return P2Mesh(jj, rr, nx, ny)
