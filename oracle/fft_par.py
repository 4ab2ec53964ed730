"""Restatement of the ``FFT_PAR_*`` pseudo-spectral products, ``nb_procs = 1``.

Reference: SOURCES/POST_PROCESSING/fft_parallel.F90.  With one rank the two
MPI_ALLTOALL transposes are identity copies, so each routine reduces to, per
point (``fft_parallel.F90:225-324`` for the cross product):

1. pack ``c_0 = a_0``, ``c_m = (a_m - i b_m)/2``, zero-padded to ``m_max_pad``
   modes (``:227-241``);
2. unnormalised complex-to-real DFT of length ``N = 2*m_max_pad-1``
   (``:254-267``);
3. pointwise operation on the real grid (``:270-274``);
4. real-to-complex DFT, scaled by ``1/N`` (``:277-288``);
5. keep modes ``0..m_max-1``: mode 0 as is, mode m>=1 as ``2*conj`` (``:295-299``),
   unpacked as (real, imaginary) = (cos, sin) columns (``:311-324``).

Non-polynomial products (division, square root, the smoothed Heaviside of
level sets, the surface-tension tensor, the velocity maximum) depend on the
grid size, so they are evaluated on exactly ``N = 2*m_max_pad-1`` points as
the reference does: ``fft_par_div_dcl`` (:667-830), ``fft_scalar_vect_dcl``
with pb = 2 (:1021-1220), ``fft_heaviside_dcl`` with ``regul`` (:832-1019,
:2234-2246), ``fft_squareroot_scalar`` (:4601-4771), ``fft_tensor_dcl`` with
``opt_tension`` (:1688-1706), ``fft_max_vel_dcl`` (:1426-1541),
``fft_scalar_vect_no_overshoot`` (:1222-1424),
``fft_no_overshoot_level_set`` (:3649-3819), ``fft_par_entro_visc_dcl``
(:4055-4292), ``fft_par_compr_entro_visc_dcl`` (:3160-3417) and
``fft_compute_entropy_visc_grad_mom`` (:3419-3647), ``fft_compute_entropy_visc``
(:2418-2699), ``fft_compute_entropy_visc_mom`` (:2701-2937),
``fft_compression_level_set_dcl`` (:3821-4053), ``fft_max_min_vel_dcl``
(:4467-4599), ``fft_check_interface`` (:2267-2416), ``fft_par_var_eta_prod``
(:1798-2232) and ``fft_par_scal_funct`` (:4294-4465);
``fft_compute_diffu_mom`` (:2939-3158) is bilinear.

Two DFT engines are available: ``"fft"`` (numpy's pocketfft, the fast path
used for larger parity cases and the CPU baseline) and ``"direct"`` (a direct
sum in 80-bit long double, the pin for the fast engine on small inputs).
Every routine processes points in chunks so that large inputs stay bounded
in memory; points are independent, so chunking does not change the result.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "n_grid", "pack", "c2r", "r2c", "unpack", "fft_product",
    "fft_par_real", "fft_par_cross_prod_dcl", "fft_par_dot_prod_dcl",
    "fft_par_prod_dcl", "fft_scalar_vect_dcl", "fft_tensor_dcl",
    "fft_par_div_dcl", "fft_heaviside_dcl", "fft_squareroot_scalar", "fft_max_vel_dcl",
    "fft_scalar_vect_no_overshoot", "fft_no_overshoot_level_set", "fft_par_entro_visc_dcl",
    "fft_compute_entropy_visc_grad_mom", "fft_compute_diffu_mom", "fft_par_compr_entro_visc_dcl",
    "fft_compute_entropy_visc", "fft_compute_entropy_visc_mom", "fft_compression_level_set_dcl",
    "fft_max_min_vel_dcl", "fft_check_interface", "fft_par_var_eta_prod", "fft_par_scal_funct",
    "regul", "OPS",
]

_PI_LD = np.longdouble("3.141592653589793238462643383279502884197")
_CHUNK = 16384


def n_grid(m_max_pad: int) -> int:
    """Azimuthal grid size ``N_r_pad = 2*m_max_pad-1`` (fft_parallel.F90:185)."""
    return 2 * m_max_pad - 1


def pack(V: np.ndarray, m_max_pad: int, dtype=np.complex128) -> np.ndarray:
    """(np, nb_field, M) cos/sin columns -> (np, nb_field/2, m_max_pad) complex.

    fft_parallel.F90:236-241: ``cu = 0.5*CMPLX(a, -b)``, then ``cu(1) = 2*REAL(cu(1))``.
    """
    npt, nb, M = V.shape
    cu = np.zeros((npt, nb // 2, m_max_pad), dtype=dtype)
    mm = min(M, m_max_pad)
    cu[:, :, :mm] = 0.5 * (V[:, 0::2, :mm] - 1j * V[:, 1::2, :mm])
    cu[:, :, 0] = 2 * cu[:, :, 0].real
    return cu


def _angles(N: int, rows: int, cols: int, sign: int) -> np.ndarray:
    k = np.arange(rows, dtype=np.int64)[:, None]
    j = np.arange(cols, dtype=np.int64)[None, :]
    t = (2 * _PI_LD * np.asarray((k * j) % N, dtype=np.longdouble)) / N
    return np.cos(t) + sign * 1j * np.sin(t)


def c2r(cu: np.ndarray, N: int, engine: str = "fft") -> np.ndarray:
    """Unnormalised Hermitian complex-to-real DFT along the last axis (FFTW c2r)."""
    K = cu.shape[-1]
    if engine == "fft":
        return np.fft.irfft(cu, n=N, axis=-1) * N
    if engine == "direct":
        E = _angles(N, K, N, +1)                       # (K, N), e^{+2 pi i m j / N}
        c = cu.astype(np.clongdouble)
        out = c[..., :1].real * np.ones(N, dtype=np.longdouble)
        if K > 1:
            out = out + 2 * np.real(c[..., 1:] @ E[1:, :])
        return out
    raise ValueError(engine)


def r2c(ru: np.ndarray, K: int, engine: str = "fft") -> np.ndarray:
    """Unnormalised real-to-complex DFT, first ``K`` outputs (FFTW r2c)."""
    N = ru.shape[-1]
    if engine == "fft":
        return np.fft.rfft(ru, axis=-1)[..., :K]
    if engine == "direct":
        E = _angles(N, N, K, -1)                       # (N, K), e^{-2 pi i j m / N}
        return ru.astype(np.longdouble) @ E
    raise ValueError(engine)


def unpack(prod_cu: np.ndarray, N: int, m_max: int) -> np.ndarray:
    """(np, nf, >=m_max) spectrum -> (np, 2*nf, m_max) cos/sin columns.

    fft_parallel.F90:291-299 (scale 1/N, mode 0 as is, ``2*CONJG`` for m>=1)
    and :311-324 (real -> odd column, imaginary -> even column).
    """
    pc = prod_cu[..., :m_max] * (1.0 / N)
    pc[..., 1:] = 2 * np.conj(pc[..., 1:])
    npt, nf, _ = pc.shape
    out = np.empty((npt, 2 * nf, m_max), dtype=np.float64)
    out[:, 0::2, :] = np.real(pc)
    out[:, 1::2, :] = np.imag(pc)
    return out


def _check(nb_field: int, m_max_c: int, name: str):
    # fft_parallel.F90:188-191 (WRITE + STOP on odd nb_field or empty mode range)
    if nb_field % 2 != 0 or m_max_c == 0:
        raise ValueError(f"BUG in {name}: MOD(nb_field,2)/=0 .OR. m_max_c==0")


def _op_cross(ru):
    a, b = ru[:, 0:3], ru[:, 3:6]
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def _op_dot(ru):
    return (ru[:, 0] * ru[:, 3] + ru[:, 1] * ru[:, 4] + ru[:, 2] * ru[:, 5])[:, None]


def _op_prod(ru):
    return (ru[:, 0] * ru[:, 1])[:, None]


def _op_scalar_vect_1(ru):
    return ru[:, 3:4] * ru[:, 0:3]


def _op_scalar_vect_2(ru):
    return ru[:, 0:3] / ru[:, 3:4]


def _op_tensor(ru):
    # prod_ru_a(:, b) = ru(:, a) * ru(:, 3+b)   (fft_parallel.F90:1674-1686)
    return (ru[:, 0:3, None, :] * ru[:, None, 3:6, :]).reshape(ru.shape[0], 9, ru.shape[-1])


def _op_div(ru):
    # prod_ru = ru(:,1,:)/ru(:,2,:)   (fft_parallel.F90:790)
    return ru[:, 0:1] / ru[:, 1:2]


def _op_sqrt(ru):
    # SQRT(MAX(0, MIN(1, ru(:,1,n))))   (fft_parallel.F90:4724)
    return np.sqrt(np.maximum(0.0, np.minimum(1.0, ru[:, 0:1])))


def _op_tension(ru):
    # prod_ru_a(:, b) = ru(:,a)*ru(:,3+b)/(|g| + 1e-14) - delta_ab |g|   (fft_parallel.F90:1688-1706)
    g = ru[:, 3:6]
    nrm = np.sqrt(g[:, 0] ** 2 + g[:, 1] ** 2 + g[:, 2] ** 2)
    t = ru[:, 0:3, None, :] * g[:, None, :, :] / (nrm + 1e-14)[:, None, None, :]
    for a in range(3):
        t[:, a, a, :] -= nrm
    return t.reshape(ru.shape[0], 9, ru.shape[-1])


def regul(phi, eps):
    """Smoothed Heaviside of a level set (fft_parallel.F90:2234-2246):
    0 below 0.5-eps, 1 above 0.5+eps, (1 + x(x^2-3eps^2)/(-2eps^3))/2 between."""
    x = np.asarray(phi, dtype=np.float64) - 0.5
    mid = (1 + x * (x ** 2 - 3 * eps ** 2) / (-2 * eps ** 3)) / 2
    return np.where(x <= -eps, 0.0, np.where(x <= eps, mid, 1.0))


OPS = {
    "div": _op_div,
    "sqrt": _op_sqrt,
    "tension": _op_tension,
    "cross": _op_cross,
    "dot": _op_dot,
    "prod": _op_prod,
    "scalar_vect_1": _op_scalar_vect_1,
    "scalar_vect_2": _op_scalar_vect_2,
    "tensor": _op_tensor,
}


def fft_product(inputs, op, m_max_pad: int, engine: str = "fft") -> np.ndarray:
    """Generic de-aliased product: concatenate input columns, c2r, ``op``, r2c.

    ``inputs`` is a list of (np, nb_k, M) arrays; ``op`` maps the real fields
    (np, nfields, N) to outputs (np, nout, N).  Returns (np, 2*nout, M).
    """
    V = np.concatenate([np.asarray(x, dtype=np.float64) for x in inputs], axis=1)
    npt, nb, M = V.shape
    N = n_grid(m_max_pad)
    f = OPS[op] if isinstance(op, str) else op
    outs = []
    for s in range(0, npt, _CHUNK):
        cu = pack(V[s:s + _CHUNK], m_max_pad)
        ru = c2r(cu, N, engine)
        pr = f(ru)
        pc = r2c(pr, m_max_pad, engine)
        outs.append(unpack(np.asarray(pc, dtype=np.complex128), N, M))
    return np.concatenate(outs, axis=0) if outs else np.zeros((0, 0, M))


def fft_par_real(V1_in, m_max_pad=None, engine="fft"):
    """FFT_PAR_REAL (fft_parallel.F90:28-145): real planes (N, nb_field/2, np).

    ``m_max_pad`` defaults to ``m_max`` (no ``opt_nb_plane``).
    """
    V = np.asarray(V1_in, dtype=np.float64)
    npt, nb, M = V.shape
    _check(nb, M, "FFT_PAR_REAL")
    if m_max_pad is None:
        m_max_pad = M
    N = n_grid(m_max_pad)
    ru = c2r(pack(V, m_max_pad), N, engine)
    return np.transpose(np.asarray(ru, dtype=np.float64), (2, 1, 0))


def fft_par_cross_prod_dcl(V1_in, V2_in, m_max_pad, engine="fft"):
    """FFT_PAR_CROSS_PROD_DCL (fft_parallel.F90:147-328): V1 x V2, (np,6,M)."""
    nb = np.shape(V1_in)[1]
    _check(nb, np.shape(V1_in)[2], "FFT_PAR_CROSS_PROD_DCL")
    if nb != 6:
        raise ValueError("cross product needs nb_field == 6 (fft_parallel.F90:270)")
    return fft_product([V1_in, V2_in], "cross", m_max_pad, engine)


def fft_par_dot_prod_dcl(V1_in, V2_in, m_max_pad, engine="fft"):
    """FFT_PAR_DOT_PROD_DCL (fft_parallel.F90:330-500): V1 . V2, (np,2,M)."""
    nb = np.shape(V1_in)[1]
    _check(nb, np.shape(V1_in)[2], "FFT_PAR_DOT_PROD_DCL")
    if nb != 6:
        raise ValueError("dot product needs nb_field == 6 (fft_parallel.F90:448)")
    return fft_product([V1_in, V2_in], "dot", m_max_pad, engine)


def fft_par_prod_dcl(c1_in, c2_in, m_max_pad, engine="fft"):
    """FFT_PAR_PROD_DCL (fft_parallel.F90:502-665): c1 * c2, (np,2,M)."""
    if np.shape(c1_in)[2] == 0:
        raise ValueError("BUG FFT_PAR_PROD_DCL: m_max_c==0")
    return fft_product([np.asarray(c1_in)[:, 0:2], np.asarray(c2_in)[:, 0:2]],
                       "prod", m_max_pad, engine)


def fft_scalar_vect_dcl(V1_in, V2_in, pb, m_max_pad, engine="fft"):
    """FFT_SCALAR_VECT_DCL (fft_parallel.F90:1021-1220): pb=1 s*v, pb=2 v/s."""
    nb1, nb2 = np.shape(V1_in)[1], np.shape(V2_in)[1]
    if nb1 % 2 or nb2 % 2 or np.shape(V1_in)[2] == 0:
        raise ValueError("BUG in FFT_SCALAR_VECT_DCL")
    if not (nb1 == 6 and nb2 == 2):
        raise ValueError("error in dimension of INPUT variables while calling FFT_SCALAR_VECT_DCL")
    if pb not in (1, 2):
        raise ValueError("error in problem type while calling FFT_SCALAR_VECT_DCL")
    return fft_product([V1_in, V2_in], f"scalar_vect_{pb}", m_max_pad, engine)


def fft_tensor_dcl(V1_in, V2_in, m_max_pad, engine="fft", opt_tension=False):
    """FFT_TENSOR_DCL (fft_parallel.F90:1543-1796).

    Returns V_out(3, np, 6, M) with ``V_out(a, n, 2b-1:2b, m)`` the (cos, sin)
    coefficients of ``V1_a * V2_b``, or with ``opt_tension`` of
    ``V1_a V2_b/(|V2| + 1e-14) - delta_ab |V2|`` (:1688-1706).
    """
    nb = np.shape(V1_in)[1]
    _check(nb, np.shape(V1_in)[2], "FFT_TENSOR_DCL")
    if nb != 6:
        raise ValueError("tensor product needs nb_field == 6")
    out = fft_product([V1_in, V2_in], "tension" if opt_tension else "tensor", m_max_pad, engine)
    npt, _, M = out.shape
    # column 2*(3a+b) / 2*(3a+b)+1 -> V_out(a, n, 2b, m) / (a, n, 2b+1, m)
    res = np.empty((3, npt, 6, M))
    for a in range(3):
        for b in range(3):
            res[a, :, 2 * b, :] = out[:, 2 * (3 * a + b), :]
            res[a, :, 2 * b + 1, :] = out[:, 2 * (3 * a + b) + 1, :]
    return res


def fft_par_div_dcl(c1_in, c2_in, m_max_pad, engine="fft"):
    """FFT_PAR_DIV_DCL (fft_parallel.F90:667-830): c1 / c2 on the grid, (np,2,M)."""
    if np.shape(c1_in)[2] == 0:
        raise ValueError("BUG in FFT_PAR_DIV_DCL: m_max_c==0")
    return fft_product([np.asarray(c1_in)[:, 0:2], np.asarray(c2_in)[:, 0:2]], "div", m_max_pad, engine)


def fft_squareroot_scalar(c1_in, m_max_pad, engine="fft"):
    """FFT_SQUAREROOT_SCALAR (fft_parallel.F90:4601-4771): SQRT(MAX(0, MIN(1, c)))
    of the first scalar; the reference overwrites c1_inout(:, 1:2, :)."""
    c = np.asarray(c1_in, dtype=np.float64)
    _check(c.shape[1], c.shape[2], "FFT_SQUAREROOT_SCALAR")
    return fft_product([c[:, 0:2]], "sqrt", m_max_pad, engine)


def fft_heaviside_dcl(V1_in, values, m_max_pad, coeff_tanh, engine="fft"):
    """FFT_HEAVISIDE_DCL (fft_parallel.F90:832-1019): V1_in(n_int, np, 2, M)
    level sets, ``values`` (n_int+1); on the grid
    ``rho = values(1)``, then for each interface i
    ``rho = rho*(1 - regul(phi_i)) + values(i+1)*regul(phi_i)``; (np, 2, M)."""
    V = np.asarray(V1_in, dtype=np.float64)
    vals = np.asarray(values, dtype=np.float64)
    n_int = vals.size - 1
    _check(V.shape[2], V.shape[3], "FFT_HEAVISIDE_DCL")

    def op(ru):
        rho = np.full((ru.shape[0], 1, ru.shape[-1]), vals[0])
        for i in range(n_int):
            h = regul(ru[:, i:i + 1], coeff_tanh)
            rho = rho * (1.0 - h) + vals[i + 1] * h
        return rho
    # the fields of all interfaces side by side: (np, 2*n_int, M)
    fields = np.concatenate([V[i, :, 0:2, :] for i in range(n_int)], axis=1)
    return fft_product([fields], op, m_max_pad, engine)


def fft_max_vel_dcl(V1_in, m_max_pad, engine="fft"):
    """FFT_MAX_VEL_DCL (fft_parallel.F90:1426-1541): max over points and grid
    angles of |v| (nb_field 6) or |c| (nb_field 2)."""
    V = np.asarray(V1_in, dtype=np.float64)
    npt, nb, M = V.shape
    _check(nb, M, "FFT_MAX_VEL_DCL")
    N = n_grid(m_max_pad)
    best = 0.0
    for s in range(0, npt, _CHUNK):
        ru = np.asarray(c2r(pack(V[s:s + _CHUNK], m_max_pad), N, engine), dtype=np.float64)
        if nb == 6:
            nrm = np.sqrt(ru[:, 0] ** 2 + ru[:, 1] ** 2 + ru[:, 2] ** 2)
        else:
            nrm = np.sqrt(ru[:, 0] ** 2)
        best = max(best, float(np.max(nrm)))
    return best


def _grid_norms(V, m_max_pad, engine):
    """|v| (nb_field 6: first three fields; else |first field|) on the grid,
    (np, N), as FFT_MAX_VEL_DCL and FFT_MAX_MIN_VEL_DCL form it."""
    npt, nb, M = V.shape
    N = n_grid(m_max_pad)
    out = np.empty((npt, N))
    for s in range(0, npt, _CHUNK):
        ru = np.asarray(c2r(pack(V[s:s + _CHUNK], m_max_pad), N, engine), dtype=np.float64)
        if nb == 6:
            out[s:s + _CHUNK] = np.sqrt(ru[:, 0] ** 2 + ru[:, 1] ** 2 + ru[:, 2] ** 2)
        else:
            out[s:s + _CHUNK] = np.sqrt(ru[:, 0] ** 2)
    return out


def fft_max_min_vel_dcl(V1_in, m_max_pad, engine="fft"):
    """FFT_MAX_MIN_VEL_DCL (fft_parallel.F90:4467-4599), one rank: (max, min) of
    |v| on the grid (:4559-4562).  The minimum follows the reference's loops
    (:4563-4576): the first loop's EXIT leaves only the inner loop over
    points, so its start value is that of the last angle, the first positive
    |v| there (or the last point's value, 0, when there is none); the second
    loop lowers it to any smaller positive value."""
    V = np.asarray(V1_in, dtype=np.float64)
    npt, nb, M = V.shape
    if nb % 2 != 0 or M == 0:
        raise ValueError(" BUG in FFT_MAX_MIN_VEL_DCL ")
    nrm = _grid_norms(V, m_max_pad, engine)
    vmax = float(np.max(nrm))
    last = nrm[:, -1]
    pos = np.nonzero(last > 0.0)[0]
    start = float(last[pos[0]]) if pos.size else float(last[-1])
    lower = nrm[(nrm < start) & (nrm > 0.0)]
    return vmax, (float(np.min(lower)) if lower.size else start)


def fft_par_var_eta_prod(eta_grid, c_in, m_max_pad, engine="fft"):
    """FFT_PAR_VAR_ETA_PROD_T_DCL (fft_parallel.F90:1798-2013) and
    FFT_PAR_VAR_ETA_PROD_GAUSS_DCL (:2015-2232), one rank: coefficients of
    eta * c on the grid (:1954-1962), eta_grid (N, np) the user function's
    values on the angles 2 pi j/N, c (np, 2 or 6, m)."""
    c = np.asarray(c_in, dtype=np.float64)
    eta = np.asarray(eta_grid, dtype=np.float64)
    npt, nb, M = c.shape
    if M == 0:
        raise ValueError(" BUG in FFT_PAR_VAR_ETA_PROD_T_DCL ")
    if nb not in (2, 6):
        raise ValueError("error in nb_field of c_in while calling FFT_PAR_VAR_ETA_PROD_T_DCL")
    N = n_grid(m_max_pad)
    if eta.shape != (N, npt):
        raise ValueError("eta must be (2*m_max_pad-1, np)")
    outs = []
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        ru = np.asarray(c2r(pack(c[sl], m_max_pad), N, engine), dtype=np.float64)
        pr = ru * eta[:, sl].T[:, None, :]
        outs.append(unpack(np.asarray(r2c(pr, m_max_pad, engine), dtype=np.complex128), N, M))
    return np.concatenate(outs, axis=0)


def fft_par_scal_funct(c1_inout, funct, m_max_pad, engine="fft"):
    """FFT_PAR_SCAL_FUNCT (fft_parallel.F90:4294-4465), one rank: columns 1-2
    replaced by the coefficients of funct(first field) on the grid
    (:4412-4416), the other columns unchanged.  funct is applied value by
    value."""
    c = np.array(c1_inout, dtype=np.float64)
    npt, nb, M = c.shape
    _check(nb, M, "FFT_PAR_SCAL_FUNCT")
    N = n_grid(m_max_pad)
    f = np.vectorize(funct, otypes=[np.float64])
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        ru = np.asarray(c2r(pack(c[sl, 0:2], m_max_pad), N, engine), dtype=np.float64)
        pr = f(ru[:, 0])[:, None, :]
        c[sl, 0:2] = unpack(np.asarray(r2c(pr, m_max_pad, engine), dtype=np.complex128), N, M)
    return c


def fft_check_interface(V1_in, nb_fluids, m_max_pad, engine="fft"):
    """FFT_CHECK_INTERFACE (fft_parallel.F90:2267-2416), one rank: V1_in
    (nb_fluids-1, np, 2, m) level sets; 0 if at some point and grid angle a
    level set in [0.1, 0.9] is below the next one (:2388-2404), else 1."""
    V = np.asarray(V1_in, dtype=np.float64)
    n_int, npt, nb, M = V.shape
    if nb % 2 != 0 or M == 0:
        raise ValueError(" BUG in FFT_CHECK_INTERFACE ")
    if nb != 2:
        raise ValueError("BUG in FFT_CHECK_INTERFACE: pb with V1_in dimensions")
    if nb_fluids == 2:
        return 1
    N = n_grid(m_max_pad)
    ru = [np.asarray(c2r(pack(V[i], m_max_pad), N, engine), dtype=np.float64)[:, 0] for i in range(n_int)]
    for i in range(nb_fluids - 2):
        a, b = ru[i], ru[i + 1]
        if np.any((0.1 <= a) & (a <= 0.9) & (a < b)):
            return 0
    return 1


def fft_scalar_vect_no_overshoot(scalar_bounds, V1_in, V2_in, pb, m_max_pad, engine="fft"):
    """FFT_SCALAR_VECT_NO_OVERSHOOT (fft_parallel.F90:1222-1424): as
    fft_scalar_vect_dcl with the scalar clamped on the grid to
    [MINVAL(scalar_bounds), MAXVAL(scalar_bounds)] (:1352-1361)."""
    b = np.atleast_1d(np.asarray(scalar_bounds, dtype=np.float64))
    lo, hi = float(np.min(b)), float(np.max(b))
    nb1, nb2 = np.shape(V1_in)[1], np.shape(V2_in)[1]
    if nb1 % 2 or nb2 % 2 or np.shape(V1_in)[2] == 0:
        raise ValueError("BUG in FFT_SCALAR_VECT_NO_OVERSHOOT")
    if not (nb1 == 6 and nb2 == 2) or pb not in (1, 2):
        raise ValueError("error in dimension or problem type while calling FFT_SCALAR_VECT_NO_OVERSHOOT")

    def op(ru):
        s = np.maximum(lo, np.minimum(ru[:, 3:4], hi))
        return s * ru[:, 0:3] if pb == 1 else ru[:, 0:3] / s
    return fft_product([V1_in, V2_in], op, m_max_pad, engine)


def fft_no_overshoot_level_set(c1_in, m_max_pad, engine="fft"):
    """FFT_NO_OVERSHOOT_LEVEL_SET (fft_parallel.F90:3649-3819): MAX(0, MIN(1, c))
    of the first scalar on the grid (:3771); the reference overwrites c1_inout(:, 1:2, :)."""
    c = np.asarray(c1_in, dtype=np.float64)
    _check(c.shape[1], c.shape[2], "FFT_NO_OVERSHOOT_LEVEL_SET")
    return fft_product([c[:, 0:2]], lambda ru: np.maximum(0.0, np.minimum(1.0, ru[:, 0:1])), m_max_pad, engine)


def fft_par_entro_visc_dcl(V1_in, c1_in, c_in_real, coeff1_in_level, m_max_pad, engine="fft"):
    """FFT_PAR_ENTRO_VISC_DCL (fft_parallel.F90:4055-4292): with c_in_real
    (N, np) given on the grid, V_out = coefficients of -visc_entro * V1 with
    visc_entro = -coeff1_in_level + c_in_real (:4212-4219), and c_out of
    MAX(0, c (1 - c)) (:4209).  Returns (V_out (np,6,M), c_out (np,2,M))."""
    V = np.asarray(V1_in, dtype=np.float64)
    c = np.asarray(c1_in, dtype=np.float64)
    cr = np.asarray(c_in_real, dtype=np.float64)
    npt, nb1, M = V.shape
    if nb1 != 6 or c.shape[1] != 2 or M == 0:
        raise ValueError("error in problem type while calling FFT_PAR_ENTRO_VISC_DCL")
    N = n_grid(m_max_pad)
    if cr.shape != (N, npt):
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    outs_v, outs_c = [], []
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        cu = pack(np.concatenate([V[sl], c[sl]], axis=1), m_max_pad)
        ru = np.asarray(c2r(cu, N, engine), dtype=np.float64)        # (n, 4, N)
        vis = -coeff1_in_level + cr[:, sl].T                             # (n, N)
        pv = -vis[:, None, :] * ru[:, 0:3]
        pc = np.maximum(0.0, ru[:, 3:4] * (1.0 - ru[:, 3:4]))
        outs_v.append(unpack(np.asarray(r2c(pv, m_max_pad, engine), dtype=np.complex128), N, M))
        outs_c.append(unpack(np.asarray(r2c(pc, m_max_pad, engine), dtype=np.complex128), N, M))
    return np.concatenate(outs_v, axis=0), np.concatenate(outs_c, axis=0)


def fft_compute_entropy_visc_grad_mom(V1_in, V2_in, V3_in, c_in_real, les_coeff1_mom, m_max_pad, engine="fft"):
    """FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM (fft_parallel.F90:3419-3647): with
    visco_entro = -LES_coeff1_mom + c_in_real on the grid (:3553), V_out(a, n,
    2f-1:2f, m) = coefficients of visco_entro * V_a,f (:3560-3568, :3635-3640)."""
    Vs = [np.asarray(x, dtype=np.float64) for x in (V1_in, V2_in, V3_in)]
    cr = np.asarray(c_in_real, dtype=np.float64)
    npt, nb, M = Vs[0].shape
    _check(nb, M, "FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM")
    N = n_grid(m_max_pad)
    if cr.shape != (N, npt):
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    res = np.empty((3, npt, 6, M))
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        vis = -les_coeff1_mom + cr[:, sl].T                              # (n, N)
        for a in range(3):
            ru = np.asarray(c2r(pack(Vs[a][sl], m_max_pad), N, engine), dtype=np.float64)
            pc = r2c(vis[:, None, :] * ru, m_max_pad, engine)
            res[a, sl] = unpack(np.asarray(pc, dtype=np.complex128), N, M)
    return res


def fft_compute_diffu_mom(V1_in, V2_in, V3_in, m_max_pad, engine="fft"):
    """FFT_COMPUTE_DIFFU_MOM (fft_parallel.F90:2939-3158): (s V1, s V2) with
    s = V3_in (:3082-3087)."""
    if np.shape(V1_in)[1] != 6 or np.shape(V2_in)[1] != 6 or np.shape(V3_in)[1] != 2:
        raise ValueError("error on inputs while calling FFT_COMPUTE_DIFFU_MOM")
    return (fft_product([V1_in, V3_in], "scalar_vect_1", m_max_pad, engine),
            fft_product([V2_in, V3_in], "scalar_vect_1", m_max_pad, engine))


def fft_par_compr_entro_visc_dcl(V1_in, V2_in, c1_in, c_in_real, coeff1_in_level, level_set_comp_factor,
                                 m_max_pad, engine="fft"):
    """FFT_PAR_COMPR_ENTRO_VISC_DCL (fft_parallel.F90:3160-3417), on the grid
    (:3316-3330): c2 = MAX(0, phi (1 - phi)), |grad phi_reg| + 1e-14,
    visc_entro = -coeff1_in_level + c_in_real, visc_comp = c_in_real *
    level_set_comp_factor * c2 / |grad phi_reg|; V_out = coefficients of
    -visc_entro grad phi + visc_comp grad phi_reg, c_out of c2."""
    V1 = np.asarray(V1_in, dtype=np.float64)
    V2 = np.asarray(V2_in, dtype=np.float64)
    c = np.asarray(c1_in, dtype=np.float64)
    cr = np.asarray(c_in_real, dtype=np.float64)
    npt, _, M = V1.shape
    if V1.shape[1] != 6 or V2.shape[1] != 6 or c.shape[1] != 2 or M == 0:
        raise ValueError("error in problem type while calling FFT_PAR_COMPR_ENTRO_VISC_DCL")
    N = n_grid(m_max_pad)
    if cr.shape != (N, npt):
        raise ValueError("c_in_real must be (2*m_max_pad-1, np)")
    outs_v, outs_c = [], []
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        ru = np.asarray(c2r(pack(np.concatenate([V1[sl], V2[sl], c[sl]], axis=1), m_max_pad), N, engine),
                        dtype=np.float64)                                  # (n, 7, N)
        c2 = np.maximum(0.0, ru[:, 6] * (1.0 - ru[:, 6]))
        ng = np.sqrt(ru[:, 3] ** 2 + ru[:, 4] ** 2 + ru[:, 5] ** 2) + 1e-14
        x = cr[:, sl].T
        ve = -coeff1_in_level + x
        vc = x * level_set_comp_factor * c2 / ng
        pv = -ve[:, None, :] * ru[:, 0:3] + vc[:, None, :] * ru[:, 3:6]
        outs_v.append(unpack(np.asarray(r2c(pv, m_max_pad, engine), dtype=np.complex128), N, M))
        outs_c.append(unpack(np.asarray(r2c(c2[:, None, :], m_max_pad, engine), dtype=np.complex128), N, M))
    return np.concatenate(outs_v, axis=0), np.concatenate(outs_c, axis=0)


def fft_compression_level_set_dcl(V1_in, V2_in, c_in, l_G, level_set_comp_factor, les_coeff4, h_min_distance,
                                  m_max_pad, engine="fft"):
    """FFT_COMPRESSION_LEVEL_SET_DCL (fft_parallel.F90:3821-4053), one rank:
    grad phi = V1 components 1-3, velocity = V2 components 1-3, phi = c.
    |v| is maximised over each element (l_G consecutive points) and the angles
    j-1, j, j+1 (:3960-3984; points past the last whole element keep |v|),
    vmax = MIN(1, max over all) (:3987-3998); the product is
    level_set_comp_factor*4*les_coeff4*vmax*(2 regul(phi, 0.01) - 1)
    *(MAX(1-(2phi-1)^2, 0)/(2 h_min_distance) - |grad phi|) (:4001-4005)."""
    V1 = np.asarray(V1_in, dtype=np.float64)
    V2 = np.asarray(V2_in, dtype=np.float64)
    c = np.asarray(c_in, dtype=np.float64)
    npt, nb, M = V1.shape
    _check(nb, M, "FFT_COMPRESSION_LEVEL_SET_DCL")
    N = n_grid(m_max_pad)
    nv = np.empty((npt, N))
    for s0 in range(0, npt, _CHUNK):
        sl = slice(s0, s0 + _CHUNK)
        ru = np.asarray(c2r(pack(V2[sl, 0:6], m_max_pad), N, engine), dtype=np.float64)
        nv[sl] = np.sqrt(np.sum(ru[:, 0:3] ** 2, axis=1))
    ne = npt // l_G
    if ne:
        nv[:ne * l_G] = _element_angle_max(nv[:ne * l_G], l_G)
    vmax = min(1.0, float(np.max(nv))) if npt else 1.0
    coeff = level_set_comp_factor * 4 * les_coeff4 * vmax

    def op(ru):
        ng = np.sqrt(ru[:, 0] ** 2 + ru[:, 1] ** 2 + ru[:, 2] ** 2)
        phi = ru[:, 3]
        return (coeff * (2.0 * regul(phi, 0.01) - 1.0)
                * (np.maximum(1.0 - (2 * phi - 1.0) ** 2, 0.0) / (2 * h_min_distance) - ng))[:, None]
    return fft_product([V1[:, 0:6], c[:, 0:2]], op, m_max_pad, engine)


def fft_compute_entropy_visc(V1_in, V2_in, V3_in, V4_in, V5_in, hloc_gauss, residual_normalization, l_G,
                             les_coeff1, les_coeff2, les_coeff4, m_max_pad, engine="fft"):
    """FFT_COMPUTE_ENTROPY_VISC (fft_parallel.F90:2418-2699), one rank: |v| on
    the grid (:2571), its maximum over the l_G points of each element
    (:2572-2577) and over the angles j-1, j, j+1, periodic (:2578-2591);
    visco = -LES_coeff1 + MIN(LES_coeff4 |v|max, LES_coeff2 h |v . res| /
    (residual_normalization + 1e-14)) (:2601-2607); outputs the coefficients
    of visco V3, visco V4, visco V5 (:2612-2620).  Returns three (np, 6, M)."""
    Vs = [np.asarray(x, dtype=np.float64) for x in (V1_in, V2_in, V3_in, V4_in, V5_in)]
    h = np.asarray(hloc_gauss, dtype=np.float64)
    npt, nb, M = Vs[0].shape
    _check(nb, M, "FFT_COMPUTE_ENTROPY_VISC")
    if npt % l_G:
        raise ValueError("np must be a multiple of l_G")
    N = n_grid(m_max_pad)
    chunk = l_G * max(1, _CHUNK // l_G)
    outs = [[], [], []]
    for s0 in range(0, npt, chunk):
        sl = slice(s0, s0 + chunk)
        ru = [np.asarray(c2r(pack(V[sl], m_max_pad), N, engine), dtype=np.float64) for V in Vs]   # (n, 3, N)
        n = ru[0].shape[0]
        nv = np.sqrt(np.sum(ru[0] ** 2, axis=1))                                                  # (n, N)
        ne = nv.reshape(n // l_G, l_G, N).max(axis=1)                                             # (ne, N)
        if N >= 3:
            ne = np.maximum(np.maximum(np.roll(ne, 1, axis=1), ne), np.roll(ne, -1, axis=1))
        sm = np.repeat(ne, l_G, axis=0)                                                           # (n, N)
        dot = np.abs(np.sum(ru[0] * ru[1], axis=1))
        vis = -les_coeff1 + np.minimum(les_coeff4 * sm,
                                       les_coeff2 * h[sl][:, None] * dot / (residual_normalization + 1e-14))
        for k in range(3):
            pc = r2c(vis[:, None, :] * ru[2 + k], m_max_pad, engine)
            outs[k].append(unpack(np.asarray(pc, dtype=np.complex128), N, M))
    return tuple(np.concatenate(o, axis=0) for o in outs)


def _element_angle_max(x, l_G):
    """(n, N) per point and angle -> per point: the maximum over the point's
    element (l_G consecutive points) and the angles j-1, j, j+1 (periodic)."""
    n, N = x.shape
    e = x.reshape(n // l_G, l_G, N).max(axis=1)
    if N >= 3:
        e = np.maximum(np.maximum(np.roll(e, 1, axis=1), e), np.roll(e, -1, axis=1))
    return np.repeat(e, l_G, axis=0)


def fft_compute_entropy_visc_mom(V1_in, V2_in, V3_in, c1_in, hloc_gauss, l_G, les_coeff2, les_coeff4,
                                 m_max_pad, want_c2=False, engine="fft"):
    """FFT_COMPUTE_ENTROPY_VISC_MOM (fft_parallel.F90:2701-2937), one rank: |u|,
    |m| on the grid (:2839-2840) and the raw viscosity LES_coeff2 h
    MAX(|u . res|, |c |u|^2|) (:2890-2893), each maximised over the element
    and the angles j-1, j, j+1 (:2841-2870, :2896-2916); vmax = max |u|
    (:2872-2876); c1_real_out = MIN(4 c4 vmax, 16 vis/(nv nm + 1e-30)) when
    l_G = 3, else MIN(c4 nv, vis/(nv nm + 1e-30)) and opt_c2_real_out =
    MIN(c4 vmax, vis/(nv nm + 1e-30)) (:2918-2931).  Returns (N, np) arrays."""
    V1, V2, V3 = (np.asarray(x, dtype=np.float64) for x in (V1_in, V2_in, V3_in))
    c = np.asarray(c1_in, dtype=np.float64)
    h = np.asarray(hloc_gauss, dtype=np.float64)
    npt, nb, M = V1.shape
    _check(nb, M, "FFT_COMPUTE_ENTROPY_VISC_MOM")
    if npt % l_G:
        raise ValueError("np must be a multiple of l_G")
    N = n_grid(m_max_pad)
    nv, nm, vis = (np.empty((npt, N)) for _ in range(3))
    chunk = l_G * max(1, _CHUNK // l_G)
    for s0 in range(0, npt, chunk):
        sl = slice(s0, s0 + chunk)
        ru = np.asarray(c2r(pack(np.concatenate([V1[sl], V2[sl], V3[sl], c[sl, 0:2]], axis=1), m_max_pad), N,
                            engine), dtype=np.float64)                       # (n, 10, N)
        uu = np.sum(ru[:, 0:3] ** 2, axis=1)
        raw = les_coeff2 * h[sl][:, None] * np.maximum(np.abs(np.sum(ru[:, 0:3] * ru[:, 6:9], axis=1)),
                                                       np.abs(ru[:, 9] * uu))
        nv[sl] = _element_angle_max(np.sqrt(uu), l_G)
        nm[sl] = _element_angle_max(np.sqrt(np.sum(ru[:, 3:6] ** 2, axis=1)), l_G)
        vis[sl] = _element_angle_max(raw, l_G)
    vmax = float(np.max(nv)) if npt else 0.0
    q = vis / (nv * nm + 1e-30)
    if l_G == 3:
        return np.minimum(4 * les_coeff4 * vmax, 16 * q).T.copy(), None
    out2 = np.minimum(les_coeff4 * vmax, q).T.copy() if want_c2 else None
    return np.minimum(les_coeff4 * nv, q).T.copy(), out2
