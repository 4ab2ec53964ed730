"""Seeded synthetic Fourier fields (host side).

Coefficients are ``N(0,1) * exp(-m/decay)`` per (node, column, mode), drawn
from SplitMix64 + Box-Muller, in Fortran order of the reference array
``V(np, nb_field, m_max_c)`` (node index fastest).  The sine column of mode 0
is set to zero because the reference ignores it on input
(SOURCES/POST_PROCESSING/fft_parallel.F90:236-241).
"""
from __future__ import annotations

import numpy as np

__all__ = ["splitmix64", "normal", "fourier_field", "level_set_density", "SEED_BASE"]

SEED_BASE = 0x5F3E
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """``n`` consecutive SplitMix64 outputs starting from state ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normal(seed: int, n: int) -> np.ndarray:
    """``n`` standard normal doubles (Box-Muller on 53-bit uniforms)."""
    m = (n + 1) // 2
    u = splitmix64(seed, 2 * m)
    f = (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = 1.0 - f[0::2]            # (0, 1]
    u2 = f[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2.0 * np.pi * u2)
    z[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return z[:n]


def fourier_field(npts: int, nb_field: int, m_max: int, seed: int,
                  decay: float = 2.0, list_mode=None) -> np.ndarray:
    """Array ``(npts, nb_field, m_max)`` Fortran-ordered, spectral decay exp(-m/decay)."""
    if list_mode is None:
        list_mode = np.arange(m_max)
    z = normal(seed, npts * nb_field * m_max).reshape((npts, nb_field, m_max), order="F")
    scale = np.exp(-np.asarray(list_mode, dtype=np.float64) / decay)
    v = np.asfortranarray(z * scale[None, None, :])
    for i, mode in enumerate(list_mode):
        if mode == 0:
            v[:, 1::2, i] = 0.0
    return v


def level_set_density(torch, rr, m_max: int, x0: float = 0.3, radius: float = 0.25, eps: float = 0.05):
    """Variable density of a level set, as (cos, sin) coefficients on the device.

    rho = 1 + 0.5 tanh(phi/eps) with phi = |x - x0| - radius, the signed
    distance to a sphere centred at (x0, 0, 0), sampled at every node on 2*m_max
    azimuthal points and transformed once into the coefficients of modes
    0..m_max-1; returns a Fortran-strided CUDA tensor (np, 2, m_max).
    """
    r = torch.tensor(rr[0], device="cuda", dtype=torch.float64)[:, None]
    z = torch.tensor(rr[1], device="cuda", dtype=torch.float64)[:, None]
    n = 2 * m_max
    th = torch.arange(n, device="cuda", dtype=torch.float64)[None, :] * (2 * np.pi / n)
    out = torch.empty((m_max, 2, r.shape[0]), device="cuda", dtype=torch.float64)
    for i0 in range(0, r.shape[0], 65536):
        rs, zs = r[i0:i0 + 65536], z[i0:i0 + 65536]
        phi = torch.sqrt((rs * torch.cos(th) - x0) ** 2 + (rs * torch.sin(th)) ** 2 + zs ** 2) - radius
        rho = 1.0 + 0.5 * torch.tanh(phi / eps)
        c = torch.fft.rfft(rho, dim=1)[:, :m_max] / n      # rho = sum_m c_m e^{i m theta}
        a = 2.0 * c.real
        b = -2.0 * c.imag
        a[:, 0] = c[:, 0].real
        b[:, 0] = 0.0
        out[:, 0, i0:i0 + 65536] = a.T
        out[:, 1, i0:i0 + 65536] = b.T
    return out.permute(2, 1, 0)
