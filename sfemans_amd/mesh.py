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
numbered row by row and edges numbered in order of first encounter.
``sector_p2`` meshes an annular sector ``rho0 <= |x| <= rho1``,
``theta0 <= atan2(z, r) <= theta1`` with curved P2 elements: the midpoint of
every edge on one of the two arcs is moved radially onto the circle, as the
reference places the midpoints of sides on a spherical boundary
(SOURCES/MESH/refine_mesh.F90:342-353, 1043-1047).
``load_fem_p2`` reads the P1 section of a formatted SFEMaNS ``.FEM`` mesh file
(SOURCES/MESH/load_mesh.F90:59-93: ``np nw me nws mes``, then per cell
``jj(nw) neigh(3) i_d``, per boundary side ``jjs(nws) neighs sides``, per node
``rr(2)``), keeps the cells of the listed domains, numbers their nodes in order
of first appearance (:246-259) and builds the P2 grid as above.  The arrays
use the Fortran layout the C-ABI receives: ``jj(6, me)`` (1-based,
column-major) and ``rr(2, np)``.
"""
from __future__ import annotations

import numpy as np

__all__ = ["P2Mesh", "rectangle_p2", "sector_p2", "p2_from_p1", "read_fem_p1", "load_fem_p2",
           "edge_ends", "boundary_sides"]


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
    return p2_from_p1(tri, vr, vz, nx, ny)


def edge_ends(jj: np.ndarray) -> np.ndarray:
    """(3, 2, me) 1-based end nodes of the edge of each local midpoint node
    3+k: the two vertices other than local vertex k."""
    jj = np.asarray(jj)
    return np.stack([np.stack([jj[1], jj[2]]), np.stack([jj[0], jj[2]]), np.stack([jj[0], jj[1]])])


def boundary_sides(jj: np.ndarray) -> np.ndarray:
    """Boundary sides of a P2 mesh, ``jjs(3, mes)`` 1-based as mesh%jjs holds
    them: the two end vertices, then the midpoint node.  A side is an edge of
    a single element; sides are listed in order of their element, then of the
    edge opposite local vertex 0, 1, 2."""
    jj = np.asarray(jj)
    ends = edge_ends(jj)                                        # (3, 2, me)
    mids = jj[3:6]                                              # (3, me)
    flat_mid = mids.T.ravel()                                   # element-major
    uniq, cnt = np.unique(flat_mid, return_counts=True)
    single = set(uniq[cnt == 1].tolist())
    cols = []
    for e in range(jj.shape[1]):
        for k in range(3):
            if int(mids[k, e]) in single:
                cols.append((ends[k, 0, e], ends[k, 1, e], mids[k, e]))
    return np.asfortranarray(np.array(cols, dtype=np.int32).reshape(-1, 3).T)


def sector_p2(nrho: int, ntheta: int, rho0: float = 0.5, rho1: float = 1.0, theta0: float = -np.pi / 3,
              theta1: float = np.pi / 3, curved: bool = True) -> P2Mesh:
    """P2 mesh of ``2*nrho*ntheta`` triangles on the annular sector
    ``rho0 <= sqrt(r^2+z^2) <= rho1``, ``theta0 <= atan2(z, r) <= theta1``
    (r > 0 when |theta| < pi/2).  Vertices of an (rho, theta) grid, squares
    split as in ``rectangle_p2``, then, with ``curved``, the midpoint of every
    edge whose two ends lie on the inner or the outer arc rescaled onto that
    circle about the origin: rr = rr_ref * radius / |rr_ref|
    (refine_mesh.F90:1043-1047)."""
    if nrho < 1 or ntheta < 1:
        raise ValueError("nrho and ntheta must be >= 1")
    base = rectangle_p2(nrho, ntheta, rho0, rho1, theta0, theta1)
    nv = (nrho + 1) * (ntheta + 1)
    rho, th = base.rr[0, :nv].copy(), base.rr[1, :nv].copy()
    vr, vz = rho * np.cos(th), rho * np.sin(th)
    tri = base.jj[0:3].T.astype(np.int64) - 1
    m = p2_from_p1(tri, vr, vz, nrho, ntheta)
    if curved:
        ends = edge_ends(m.jj) - 1                                   # (3, 2, me) 0-based
        mids = m.jj[3:6] - 1                                         # (3, me)
        on_in = np.isclose(rho, rho0, rtol=0, atol=1e-12 * rho1)
        on_out = np.isclose(rho, rho1, rtol=0, atol=1e-12 * rho1)
        for ring, radius in ((on_in, rho0), (on_out, rho1)):
            sel = ring[ends[:, 0]] & ring[ends[:, 1]]                # both ends on this arc
            nodes = np.unique(mids[sel])
            rr = m.rr[:, nodes]
            m.rr[:, nodes] = rr * (radius / np.sqrt(np.sum(rr * rr, axis=0)))
    return m


def p2_from_p1(tri: np.ndarray, vr: np.ndarray, vz: np.ndarray, nx: int = 0, ny: int = 0) -> P2Mesh:
    """P2 mesh from P1 cells ``tri`` (me, 3), 0-based, each triplet in the
    reference's local vertex order, and vertex coordinates ``vr``, ``vz``:
    one node per edge after the vertices, edges numbered in order of first
    encounter (cell by cell, edge opposite local vertex 0, 1, 2), as the
    single-process mesh extraction numbers them
    (SOURCES/MESH/mesh_metis_distribution.F90:728-745)."""
    tri = np.asarray(tri, dtype=np.int64)
    nv = int(np.asarray(vr).size)
    me = tri.shape[0]
    # edge opposite local vertex k joins the two other vertices (k = 0,1,2)
    opp = np.stack([np.stack([tri[:, 1], tri[:, 2]], 1),
                    np.stack([tri[:, 0], tri[:, 2]], 1),
                    np.stack([tri[:, 0], tri[:, 1]], 1)], axis=1)  # (me, 3, 2)
    lo, hi = np.minimum(opp[:, :, 0], opp[:, :, 1]), np.maximum(opp[:, :, 0], opp[:, :, 1])
    keys = lo * nv + hi                                            # (me, 3)
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
    # midpoint from the end with the lower local index in the cell that
    # meets the edge first (refine_mesh.F90:323-329, 341-343)
    first = opp.reshape(-1, 2)[first_idx[order]]
    es, ee = first[:, 0], first[:, 1]
    rr[0, nv:] = vr[es] + (vr[ee] - vr[es]) / 2.0
    rr[1, nv:] = vz[es] + (vz[ee] - vz[es]) / 2.0

    jj = np.empty((6, me), dtype=np.int32)
    jj[0:3, :] = tri.T + 1
    jj[3:6, :] = (nv + edge_id).T + 1
    return P2Mesh(jj, rr, nx, ny)


def read_fem_p1(path: str):
    """First (P1) section of a formatted ``.FEM`` file:
    ``(jj (3, me) 1-based as in the file, i_d (me,), rr (2, np))``."""
    with open(path) as f:
        head = f.readline().split()
        npt, nw, me, nws, mes = (int(x) for x in head[:5])
        if nw != 3 or nws != 2:
            raise ValueError(f"{path}: first section is not a P1 triangulation (nw={nw}, nws={nws})")
        cells = np.array([f.readline().split()[:7] for _ in range(me)], dtype=np.int64)
        for _ in range(mes):
            f.readline()
        rr = np.array([f.readline().replace("D", "E").split()[:2] for _ in range(npt)], dtype=np.float64)
    return np.asfortranarray(cells[:, :3].T), cells[:, 6].copy(), np.asfortranarray(rr.T)


def load_fem_p2(path: str, list_dom=None) -> P2Mesh:
    """P2 mesh of the cells of ``list_dom`` (all domains, ascending, if None)
    from the P1 section of a formatted ``.FEM`` file.  The vertex triplet of
    every cell is sorted by file index (load_mesh.F90:101-112), the kept
    cells are taken domain by domain in file order and their nodes numbered
    in order of first appearance (:186-196, :246-259); a single process keeps
    that order (mesh_metis_distribution.F90:62-65, 689-727).  Edges are then
    added as in ``p2_from_p1``.  Nodes on the listed domains' common sides
    are shared: the duplicated nodes of interfaces declared as real
    interfaces (stat 3, :233-236) are not produced."""
    jj1, i_d, rr1 = read_fem_p1(path)
    tri = np.sort(jj1.T - 1, axis=1)
    doms = np.unique(i_d) if list_dom is None else np.asarray(list_dom)
    new = -np.ones(rr1.shape[1], dtype=np.int64)
    cells, n = [], 0
    for d in doms:
        for m in np.nonzero(i_d == d)[0]:
            for v in tri[m]:
                if new[v] < 0:
                    new[v] = n
                    n += 1
            cells.append(m)
    if not cells:
        raise ValueError(f"{path}: no cell in domains {list(doms)}")
    tri_new = new[tri[cells]]
    old = np.empty(n, dtype=np.int64)
    old[new[new >= 0]] = np.nonzero(new >= 0)[0]
    return p2_from_p1(tri_new, rr1[0, old], rr1[1, old])

