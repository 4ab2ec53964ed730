"""Tangent and adjoint of the convective term (no reference counterpart).

The reference's LightKrylov glue binds rmatvec to the forward matvec
(SOURCES/LIGHTKRYLOV_OBJECTS/exponential_propagator.F90:61-62), so nothing
in the reference pins these operators.  They are defined from the forward
map of ``smb_cross_prod_gauss_sft_par``
(SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514-630): with
B(a, b) = P[(curl a) x b] (bilinear in the nodal coefficients),

* tangent around U: T_U d = B(U, d) + B(d, U);
* adjoint: T_U^T in the Euclidean inner product of the coefficient arrays.

Pins: the tangent equals the central difference of the forward map with
step 1 (exact for a bilinear map); the adjoint equals the transpose of the
dense matrix of T_U assembled column by column (brute force, tiny inputs).
"""
from __future__ import annotations

import numpy as np

from .smb import smb_CurlH_cross_B_gauss_sft_par

__all__ = ["tangent", "tangent_matrix", "adjoint_bruteforce"]


def tangent(jj, rr, list_mode, U, D, engine="fft"):
    return (smb_CurlH_cross_B_gauss_sft_par(jj, rr, list_mode, U, D, engine)
            + smb_CurlH_cross_B_gauss_sft_par(jj, rr, list_mode, D, U, engine))


def tangent_matrix(jj, rr, list_mode, U):
    """Dense matrix (G*6*M, np*6*M) of T_U in Fortran-order flattening."""
    npts, nc, M = U.shape
    n = npts * nc * M
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        D = e.reshape((npts, nc, M), order="F")
        cols.append(tangent(jj, rr, list_mode, U, D).ravel(order="F"))
    return np.stack(cols, axis=1)


def adjoint_bruteforce(jj, rr, list_mode, U, Y):
    """T_U^T Y through the dense matrix (tiny inputs only)."""
    A = tangent_matrix(jj, rr, list_mode, U)
    return (A.T @ np.asarray(Y).ravel(order="F")).reshape(U.shape, order="F")
