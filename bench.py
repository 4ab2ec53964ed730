"""Benchmark of the pseudo-spectral nonlinear-term path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line.  ``value`` is Gauss-pt*modes per second over the timed
steps (inputs resident in HBM).  A step is a list of phases (one library call
each); every phase is bracketed by HIP events on the stream the kernels are
launched on (torch's current stream, passed to every entry point), so each
phase's average duration is measured live inside the timed region.
``roofline`` reports the dominant phase: its algorithmic HBM bytes per launch
divided by its average duration, against MI355X's 8 TB/s.  ``cpu_baseline``
times the CPU restatement of the reference algorithm (the oracle's C
restatement, on one core and on all cores) on a bounded sample.

Workloads (all FP64, synthetic seeded inputs, N(0,1)*exp(-m/decay) nodal
coefficients, mode-0 sine set to zero):

* ``nl`` (default): the nonlinear term and its adjoint on the 224x448 P2
  mesh (200,704 triangles, 1,404,928 Gauss points), m_max = 64.  Forward:
  smb_cross_prod_gauss_sft_par, (curl u) x u at the Gauss points.  Adjoint:
  the exact transpose of its linearization around u applied to a Gauss
  field.  units = 2 * G * m_max per step.
* ``mhd``: MHD nonlinear terms of one explicit step on the same mesh:
  rotv_v = (curl u) x u - (curl H) x B, the full per-mode NS right-hand
  sides built from it (rhs_ns_gauss_3x3 with the BDF term, a source and
  the pressure gradients, assembled by rhs_3x3), and u x B at the nodes.
  units = G * m_max.
* ``fftprod``: FFT_PAR_CROSS_PROD_DCL on Gauss-point arrays of that size.
* ``tangent``: tangent and adjoint around a base flow (modes 0-4, amplitude
  0.1) on a 50k-triangle mesh (158x316), m_max = 32.
* ``batch``: 32 independent problems (2k-50k triangles, m_max in {4..128})
  in one persistent launch replayed from a hipGraph, schedule autotuned.
* ``nsrhs``: the full per-mode NS right-hand side with rhs_3x3's tensor
  (LES or variable viscosity) on the 224x448 mesh, m_max = 64.
* ``exact``: the six non-polynomial FFT_PAR products (Heaviside of a level
  set, pb = 2 scalar-vector, division, square root, tension tensor, velocity
  maximum) on the reference grid N = 191 at 1,404,928 points, m_max = 64.
* ``entropy``: the six entropy-viscosity products at 1,404,928 Gauss points
  (200,704 elements of 7), m_max = 64, grid N = 191.
* ``full``: the explicit nonlinear step at m_max = 128 on the 224x448 mesh:
  (curl u) x u -> rhs_3x3, u x B, u.grad T -> qs_00_gauss, rho u and
  (rho u) (x) u with a level-set density, and the convective adjoint.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
SEED = 0x5F3E


def mp_of(M: int) -> int:
    mp = 4
    while mp < M:
        mp *= 2
    return mp


def fft_flops_per_point(M: int, n_in: int, n_out: int, n_prod: int) -> float:
    """Flops per point of the azimuthal part: one real transform of length
    N = 3*MP per real field in and out at 2.5 N log2 N each, plus n_prod
    flops per grid angle for the pointwise product."""
    N = 3 * mp_of(M)
    return (n_in + n_out) * 2.5 * N * np.log2(N) + n_prod * N


def b2_flops(G: int, M: int) -> float:
    """Flops of smb_cross_prod_gauss_sft_par: interpolation and curl (168 per
    Gauss-pt*mode) plus six real fields in, three out and the cross product
    (9 flops) per grid angle."""
    return G * M * 168.0 + G * fft_flops_per_point(M, 6, 3, 9)


def _fields(torch, npts, M, decay, seeds, ncol=6):
    """Seeded nodal Fourier fields generated on the GPU, Fortran strides (npts, ncol, M)."""
    out = []
    dec = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / decay)
    for sd in seeds:
        g = torch.Generator(device="cuda").manual_seed(int(sd))
        t = torch.randn((M, ncol, npts), generator=g, device="cuda", dtype=torch.float64) * dec[:, None, None]
        t[0, 1::2, :] = 0.0
        out.append(t.permute(2, 1, 0))
    return out


class Phase:
    """One library call of a step.  ``bytes_`` are its compulsory HBM bytes
    (inputs read once, outputs written once, mesh once); ``overhead`` the
    bytes of work arrays it writes and reads back, reported apart and never
    counted as achieved traffic."""

    def __init__(self, name, fn, bytes_, flops, kernel, overhead=0):
        self.name, self.fn, self.bytes, self.flops, self.kernel = name, fn, bytes_, flops, kernel
        self.overhead = overhead


# ---------------------------------------------------------------- workloads
def wl_nl(torch, args):
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = list(range(M))
    G, npn = m.n_gauss, m.np
    u, = _fields(torch, npn, M, 4.0, (SEED + 2,))
    ybar, = _fields(torch, G, M, 4.0, (SEED + 5,))
    fwd_out = FP.fortran_empty((G, 6, M))
    dbar = FP.fortran_empty((npn, 6, M))
    work = torch.empty(6 * m.me * 6 * M, dtype=torch.float64, device="cuda")
    mesh_bytes = m.me * 24 + npn * 16
    # compulsory bytes: every input read once, every output written once, the mesh once
    fwd_bytes = npn * 6 * M * 8 + G * 6 * M * 8 + mesh_bytes
    # adjoint: ybar and U read, dbar written (the element work array is re-traffic, not credited)
    adj_bytes = G * 6 * M * 8 + npn * 6 * M * 8 + npn * 6 * M * 8 + mesh_bytes
    f_fwd = b2_flops(G, M)
    f_adj = 2 * f_fwd
    phases = [
        Phase("forward", lambda: NL.smb_cross_prod_gauss_sft_par(None, mesh, modes, u, fwd_out),
              fwd_bytes, f_fwd, "gauss_pt_kernel<64,CROSS_CURL>"),
        Phase("adjoint", lambda: NL.nl_adjoint(mesh, modes, u, ybar, dbar, work),
              adj_bytes, f_adj, "adjoint_kernel<64> + adjoint_gather_kernel",
              overhead=2 * (6 * m.me * 6 * M * 8)),
    ]
    cfg = {"workload": "nonlinear term (curl u) x u at Gauss points + adjoint of its linearization, "
                       "224x448 P2 mesh (200,704 triangles, 1,404,928 Gauss points), m_max=64",
           "me": m.me, "np": npn, "gauss_points": G, "m_max": M, "grid": 3 * mp_of(M),
           "units_per_step": "2 * G * m_max (forward + adjoint)"}
    return phases, 2 * G * M, cfg, ("nl", m, M)


def wl_mhd(torch, args):
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = list(range(M))
    G, npn = m.n_gauss, m.np
    u, H, B = _fields(torch, npn, M, 4.0, (SEED + 2, SEED + 3, SEED + 4))
    # BDF term, momentum source and pressures on the velocity mesh of the NS RHS
    v1m, fsrc = _fields(torch, npn, M, 4.0, (SEED + 6, SEED + 7))
    pr, pinc = _fields(torch, npn, M, 4.0, (SEED + 8, SEED + 9), ncol=2)
    work = FP.fortran_empty((G, 6, M))
    v145 = FP.fortran_empty((npn, 3, M))
    v236 = FP.fortran_empty((npn, 3, M))
    uxb = FP.fortran_empty((npn, 6, M))
    welem = torch.empty(36 * m.me * M, dtype=torch.float64, device="cuda")
    mesh_bytes = m.me * 24 + npn * 16
    rotv_bytes = 3 * npn * 6 * M * 8 + G * 6 * M * 8 + mesh_bytes
    phases = [
        Phase("rotv_mhd", lambda: NL.nl_rotv_mhd(mesh, modes, u, H, B, 1.0, work), rotv_bytes,
              2 * b2_flops(G, M), "gauss_kernel<64,PAIRS> (2 passes)"),
        # full NS RHS (rhs_ns_gauss_3x3 + rhs_3x3): rotv_v read once, V1m, ff (6 columns) and
        # P, P_inc (2 columns) read, vb written, mesh read; the element contributions
        # written and read back are re-traffic
        Phase("rhs_ns_3x3", lambda: NL.rhs_ns_3x3(mesh, modes, work, v1m, fsrc, pr, pinc, v145, v236, welem),
              G * 6 * M * 8 + npn * 16 * M * 8 + npn * 6 * M * 8 + m.me * 7 * 8 + 12 * 6 * m.me + npn * 16,
              G * M * (6 * 2 * 6 + 16 * 6 * 2), "rhs_ns_elem_kernel + rhs_gather_kernel",
              overhead=2 * 36 * m.me * M * 8),
        Phase("u_x_B_nodes", lambda: FP.FFT_PAR_CROSS_PROD_DCL(None, u, B, uxb), 3 * npn * 6 * M * 8,
              npn * fft_flops_per_point(M, 6, 3, 9), "fftprod_kernel<64,CROSS>"),
    ]
    cfg = {"workload": "MHD nonlinear terms of one step, 224x448 P2 mesh (200,704 triangles), m_max=64: "
                       "rotv_v=(curl u)xu-(curl H)xB at Gauss points -> full per-mode NS RHS (rhs_ns_gauss_3x3 + rhs_3x3), "
                       "u x B at nodes",
           "me": m.me, "np": npn, "gauss_points": G, "m_max": M, "grid": 3 * mp_of(M)}
    return phases, G * M, cfg, ("mhd", m, M)


def wl_nsrhs(torch, args):
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = list(range(M))
    G, npn = m.n_gauss, m.np
    v1m, fsrc = _fields(torch, npn, M, 4.0, (SEED + 6, SEED + 7))
    pr, pinc = _fields(torch, npn, M, 4.0, (SEED + 8, SEED + 9), ncol=2)
    gen = torch.Generator(device="cuda").manual_seed(SEED + 10)
    rotv = torch.randn((M, 6, G), dtype=torch.float64, device="cuda", generator=gen).permute(2, 1, 0)
    tens = torch.randn((M, 6, G, 3), dtype=torch.float64, device="cuda", generator=gen).permute(3, 2, 1, 0)
    v145 = FP.fortran_empty((npn, 3, M))
    v236 = FP.fortran_empty((npn, 3, M))
    welem = torch.empty(36 * m.me * M, dtype=torch.float64, device="cuda")
    # rotv_v and the tensor read once, V1m, ff (6 columns) and P, P_inc (2 columns) read,
    # vb written, mesh read (the element contributions written and read back are re-traffic)
    rest = G * 6 * M * 8 + npn * 16 * M * 8 + npn * 6 * M * 8 + m.me * 7 * 8 + 12 * 6 * m.me + npn * 16
    phases = [
        Phase("rhs_ns_3x3_tensor",
              lambda: NL.rhs_ns_3x3(mesh, modes, rotv, v1m, fsrc, pr, pinc, v145, v236, welem, opt_tensor=tens),
              rest + G * 18 * M * 8, G * M * (6 * 2 * 6 + 16 * 6 * 2 + 6 * 6 * 6 + 20),
              "rhs_ns_elem_kernel<true> + rhs_gather_kernel", overhead=2 * 36 * m.me * M * 8),
    ]
    cfg = {"workload": "full per-mode NS right-hand side with rhs_3x3's tensor (LES / variable viscosity), "
                       "224x448 P2 mesh (200,704 triangles, 1,404,928 Gauss points), m_max=64: "
                       "rhs_ns_gauss_3x3 (BDF term, source, pressure gradients, -rotv_v) + rhs_3x3 with "
                       "opt_tensor(3,G,6,m)",
           "me": m.me, "np": npn, "gauss_points": G, "m_max": M}
    return phases, G * M, cfg, ("nsrhs", m, M)


def wl_b3(torch, args):
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = list(range(M))
    G, npn = m.n_gauss, m.np
    u, = _fields(torch, npn, M, 4.0, (SEED + 2,))
    v145 = FP.fortran_empty((npn, 3, M))
    v236 = FP.fortran_empty((npn, 3, M))
    welem = torch.empty(36 * m.me * M, dtype=torch.float64, device="cuda")
    # compulsory: u read, vb_145 and vb_236 written, mesh (jj, rr, rj*ray) once; the
    # element contributions written and read back by the gather are re-traffic
    b3_bytes = npn * 6 * M * 8 + npn * 6 * M * 8 + m.me * 24 + npn * 16 + m.me * 7 * 8
    phases = [
        Phase("nl_rhs_3x3", lambda: NL.nl_rhs_3x3(mesh, modes, u, v145, v236, -1.0, welem),
              b3_bytes, b2_flops(G, M) + 72.0 * G * M, "gauss_rhs_kernel<64> + rhs_gather_kernel",
              overhead=2 * 36 * m.me * M * 8),
    ]
    cfg = {"workload": "nonlinear term fused to the per-mode NS right-hand side (nodal u in, rhs_3x3's "
                       "vb_145/vb_236 of -(curl u) x u out, no Gauss-point array), 224x448 P2 mesh "
                       "(200,704 triangles, 1,404,928 Gauss points), m_max=64",
           "me": m.me, "np": npn, "gauss_points": G, "m_max": M, "grid": 3 * mp_of(M)}
    return phases, G * M, cfg, ("nl", m, M)


def _exact_ops(FP, torch, npt, M, seed):
    """Inputs and calls of the six non-polynomial products on npt points."""
    dec = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / 4.0)
    gen = torch.Generator(device="cuda").manual_seed(seed)

    def fld(nb, scale=1.0, mean=None):
        t = torch.randn((M, nb, npt), generator=gen, device="cuda", dtype=torch.float64) * dec[:, None, None] * scale
        t[0, 1::2, :] = 0.0
        if mean is not None:
            t[0, 0, :] = mean
        return t.permute(2, 1, 0)
    phi = torch.randn((M, 2, npt, 1), generator=gen, device="cuda", dtype=torch.float64) * 0.3 * dec[:, None, None, None]
    phi[0, 0] += 0.5
    phi[0, 1] = 0.0
    phi = phi.permute(3, 2, 1, 0)                                   # (1, np, 2, M)
    mom, rho, c, d = fld(6), fld(2, 0.1, 1.5), fld(2), fld(2, 0.1, 2.0)
    u, g = fld(6), fld(6)
    rho_out, u_out = FP.fortran_empty((npt, 2, M)), FP.fortran_empty((npt, 6, M))
    q_out, tens = FP.fortran_empty((npt, 2, M)), FP.fortran_empty((3, npt, 6, M))
    s_io = FP.fortran_empty((npt, 2, M))
    vmax = torch.empty(1, dtype=torch.float64, device="cuda")
    vals = [1.0, 3.0]
    K = 3 * M // 2
    f_t = 4.0 * M * K                                               # flops of one real DFT per point
    b = 8 * M * npt
    return [
        ("heaviside", lambda: FP.FFT_HEAVISIDE_DCL(None, phi, vals, rho_out, coeff_tanh=0.05), 4 * b, 2 * f_t * npt,
         "exact_grid_kernel<HEAVI,1> + exact_r2c_kernel"),
        ("scalar_vect_pb2", lambda: FP.FFT_SCALAR_VECT_DCL(None, mom, rho, u_out, 2), 14 * b, 7 * f_t * npt,
         "exact_grid_kernel<SV2,4> + exact_r2c_kernel"),
        ("div", lambda: FP.FFT_PAR_DIV_DCL(None, c, d, q_out), 6 * b, 3 * f_t * npt,
         "exact_grid_kernel<DIV,2> + exact_r2c_kernel"),
        ("squareroot", lambda: (s_io.copy_(rho_out), FP.FFT_SQUAREROOT_SCALAR(None, s_io)), 6 * b, 2 * f_t * npt,
         "copy + exact_grid_kernel<SQRT,1> + exact_r2c_kernel"),
        ("tension", lambda: FP.FFT_TENSOR_DCL(None, u, g, tens, opt_tension=True), 30 * b, 15 * f_t * npt,
         "exact_grid_kernel<TENSION,6> + exact_r2c_kernel"),
        ("max_vel", lambda: FP.FFT_MAX_VEL_DCL(None, u, V_out=vmax), 6 * b, 3 * f_t * npt,
         "exact_grid_kernel<MAXVEL,3> + exact_max_kernel"),
    ]


def wl_exact(torch, args):
    from sfemans_amd import fft_parallele as FP
    npt, M = 1404928, 64
    phases = [Phase(n, f, b, fl, k) for n, f, b, fl, k in _exact_ops(FP, torch, npt, M, SEED + 11)]
    cfg = {"workload": "non-polynomial FFT_PAR products on the reference grid N=2*m_max_pad-1=191 at 1,404,928 "
                       "points, m_max=64: HEAVISIDE (one level set), SCALAR_VECT pb=2, DIV, SQUAREROOT, "
                       "TENSOR with tension, MAX_VEL",
           "points": npt, "m_max": M, "grid": 2 * (3 * M // 2) - 1}
    return phases, npt * M, cfg, ("exact", None, M)


def wl_grid(torch, args):
    """The grid-function and reduction FFT_PAR routines at 1.4M points, m_max = 64."""
    from sfemans_amd import fft_parallele as FP
    npt, M = 1404928, 64
    mp = 3 * M // 2
    N = 2 * mp - 1
    dec = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / 4.0)
    gen = torch.Generator(device="cuda").manual_seed(SEED + 13)

    def fld(nb, scale=1.0, mean=None):
        t = torch.randn((M, nb, npt), generator=gen, device="cuda", dtype=torch.float64) * dec[:, None, None] * scale
        t[0, 1::2, :] = 0.0
        if mean is not None:
            t[0, 0, :] = mean
        return t.permute(2, 1, 0)
    u, gphi = fld(6), fld(6)
    phi = fld(2, 0.2, 0.5)
    ls = torch.randn((M, 2, npt, 2), generator=gen, device="cuda", dtype=torch.float64) * 0.01 * dec[:, None, None, None]
    ls[0, 0, :, 0] += 0.8
    ls[0, 0, :, 1] += 0.6
    ls[0, 1] = 0.0
    ls = ls.permute(3, 2, 1, 0)                                     # (2, np, 2, M): three fluids
    eta = torch.rand((npt, N), generator=gen, device="cuda", dtype=torch.float64).t() + 0.5   # (N, np) column-major
    planes = FP.fortran_empty((N, 3, npt))
    back = FP.fortran_empty((npt, 2, M))
    c_out, v_out = FP.fortran_empty((npt, 2, M)), FP.fortran_empty((npt, 6, M))
    mm = torch.empty(2, dtype=torch.float64, device="cuda")
    ok = torch.empty(1, dtype=torch.int32, device="cuda")
    f_t = 4.0 * M * mp                                              # flops of one real DFT per point
    b = 8 * M * npt                                                 # bytes of one (cos, sin) column pair
    r = 8 * N * npt                                                 # bytes of one field on the grid
    phases = [
        Phase("par_real", lambda: FP.FFT_PAR_REAL(None, u, planes, opt_nb_plane=N), 6 * b + 3 * r, 3 * f_t * npt,
              "exact_real_kernel x3"),
        Phase("from_real", lambda: FP.FFT_PAR_FROM_REAL(eta, back, mp), r + 2 * b, f_t * npt,
              "exact_grid_kernel<FROMREAL,1> + exact_r2c_kernel"),
        Phase("var_eta_prod", lambda: FP.FFT_PAR_VAR_ETA_PROD_T_DCL(None, eta, None, u, v_out), r + 12 * b,
              6 * f_t * npt, "exact_grid_kernel<ETA3,3> + exact_r2c_kernel"),
        Phase("compression_level_set", lambda: FP.FFT_COMPRESSION_LEVEL_SET_DCL(None, None, gphi, u, phi, c_out,
                                                                               None, 7, 1.0, 0.5, 0.01),
              16 * b, 8 * f_t * npt, "exact_grid_kernel<MAXVEL,3> + exact_max_kernel + exact_grid_kernel<CLS,4> "
                                     "+ exact_r2c_kernel"),
        Phase("max_min_vel", lambda: FP.FFT_MAX_MIN_VEL_DCL(None, u, V_out=mm), 6 * b, 3 * f_t * npt,
              "exact_velstat_kernel<3> + exact_velstat_final_kernel"),
        Phase("check_interface", lambda: FP.FFT_CHECK_INTERFACE(None, ls, 3, interface_ok=ok), 4 * b, 2 * f_t * npt,
              "exact_grid_kernel<CHECKIF,2> + exact_max_kernel"),
    ]
    cfg = {"workload": "grid-function and reduction FFT_PAR routines on the reference grid N=2*m_max_pad-1=191 at "
                       "1,404,928 points, m_max=64: PAR_REAL (3 planes), the analysis of values on the grid, "
                       "VAR_ETA_PROD (vector), COMPRESSION_LEVEL_SET, MAX_MIN_VEL, CHECK_INTERFACE (3 fluids)",
           "points": npt, "m_max": M, "grid": N}
    return phases, npt * M, cfg, ("grid", None, M)


def wl_entropy(torch, args):
    from sfemans_amd import fft_parallele as FP
    npt, M, lg = 1404928, 64, 7
    mp = 3 * M // 2
    N = 2 * mp - 1
    dec = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / 4.0)
    gen = torch.Generator(device="cuda").manual_seed(SEED + 12)

    def fld(nb, mean=None):
        t = torch.randn((M, nb, npt), generator=gen, device="cuda", dtype=torch.float64) * dec[:, None, None]
        t[0, 1::2, :] = 0.0
        if mean is not None:
            t[0, 0, :] = mean
        return t.permute(2, 1, 0)
    u, mom, res, g1, g2, g3 = (fld(6) for _ in range(6))
    c, rmass = fld(2, 0.5), fld(2)
    creal = torch.rand((npt, N), generator=gen, device="cuda", dtype=torch.float64).t()   # (N, np), Fortran
    h = torch.rand(npt, generator=gen, device="cuda", dtype=torch.float64) * 0.01
    v6 = [FP.fortran_empty((npt, 6, M)) for _ in range(3)]
    c2o = FP.fortran_empty((npt, 2, M))
    t3 = FP.fortran_empty((3, npt, 6, M))
    r1, r2 = FP.fortran_empty((N, npt)), FP.fortran_empty((N, npt))
    f_t = 4.0 * M * mp
    b = 8 * M * npt
    ops = [
        ("par_entro_visc", lambda: FP.FFT_PAR_ENTRO_VISC_DCL(None, u, c, creal, None, 0.3, v6[0], c2o),
         8 * b + 8 * N * npt, 6 * f_t * npt, "exact_grid_kernel<ENTRO,4> + exact_r2c_kernel"),
        ("compr_entro_visc", lambda: FP.FFT_PAR_COMPR_ENTRO_VISC_DCL(None, u, g1, c, creal, None, 0.3, v6[0], c2o,
                                                                      1.5),
         14 * b + 8 * N * npt, 11 * f_t * npt, "exact_grid_kernel<COMPR,7> + exact_r2c_kernel"),
        ("entropy_visc_grad_mom", lambda: FP.FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM(None, g1, g2, g3, creal, t3, 0.2),
         36 * b + 8 * N * npt, 18 * f_t * npt, "exact_grid_kernel<GRADMOM,9> + exact_r2c_kernel"),
        ("entropy_visc", lambda: FP.FFT_COMPUTE_ENTROPY_VISC(None, None, u, res, g1, g2, g3, h, *v6, 1.0, lg, 0.3,
                                                             0.9, 0.8),
         48 * b + 8 * npt, 24 * f_t * npt,
         "exact_enorm_kernel + exact_grid_kernel<EVISCV,6> + exact_grid_kernel<GRADMOM,9> + exact_r2c_kernel"),
        ("entropy_visc_mom", lambda: FP.FFT_COMPUTE_ENTROPY_VISC_MOM(None, None, u, mom, res, rmass, h, r1, lg, 0.6,
                                                                     0.9, opt_c2_real_out=r2),
         20 * b + 8 * npt + 16 * N * npt, 10 * f_t * npt, "exact_emom_kernel + exact_emom_out_kernel"),
        ("diffu_mom", lambda: FP.FFT_COMPUTE_DIFFU_MOM(None, g1, g2, c, v6[1], v6[2]), 26 * b,
         npt * fft_flops_per_point(M, 4, 3, 3) * 2, "fftprod_kernel<64,SVECT1> x 2"),
    ]
    phases = [Phase(n, f, bb, fl, k) for n, f, bb, fl, k in ops]
    cfg = {"workload": "entropy-viscosity FFT_PAR products on the reference grid N=191 at 1,404,928 Gauss "
                       "points (200,704 elements of 7), m_max=64: FFT_PAR_ENTRO_VISC_DCL, "
                       "FFT_PAR_COMPR_ENTRO_VISC_DCL, FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM, FFT_COMPUTE_ENTROPY_VISC, "
                       "FFT_COMPUTE_ENTROPY_VISC_MOM, FFT_COMPUTE_DIFFU_MOM",
           "points": npt, "m_max": M, "grid": N}
    return phases, npt * M, cfg, ("entropy", None, M)


def wl_fftprod(torch, args):
    from sfemans_amd import fft_parallele as FP
    npt, M = 1404928, 64
    a, b = _fields(torch, npt, M, 4.0, (SEED + 2, SEED + 3))
    out = FP.fortran_empty((npt, 6, M))
    phases = [Phase("cross_prod", lambda: FP.FFT_PAR_CROSS_PROD_DCL(None, a, b, out), 3 * npt * 6 * M * 8,
                    npt * fft_flops_per_point(M, 6, 3, 9), "fftprod_kernel<64,CROSS>")]
    cfg = {"workload": "FFT_PAR_CROSS_PROD_DCL on Gauss-point arrays (1,404,928 points), m_max=64",
           "points": npt, "m_max": M, "grid": 3 * mp_of(M)}
    return phases, npt * M, cfg, ("fftprod", None, M)


def wl_tangent(torch, args):
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(158, 316)
    mesh = NL.Mesh.from_p2(m)
    M = 32
    modes = list(range(M))
    G, npn = m.n_gauss, m.np
    U, = _fields(torch, npn, M, 2.0, (SEED + 4,))
    U[:, :, 5:] = 0.0
    U[:, :, 1:5] *= 0.1
    d, = _fields(torch, npn, M, 2.0, (SEED + 6,))
    ybar, = _fields(torch, G, M, 2.0, (SEED + 7,))
    tout = FP.fortran_empty((G, 6, M))
    dbar = FP.fortran_empty((npn, 6, M))
    work = torch.empty(6 * m.me * 6 * M, dtype=torch.float64, device="cuda")
    tb = 2 * npn * 6 * M * 8 + G * 6 * M * 8
    # adjoint: ybar and U read, dbar written; its element work array is re-traffic
    ab = G * 6 * M * 8 + npn * 6 * M * 8 + npn * 6 * M * 8
    phases = [
        Phase("tangent", lambda: NL.nl_tangent(mesh, modes, U, d, tout), tb,
              2 * b2_flops(G, M), "gauss_kernel<32,PAIRS> (2 passes)"),
        Phase("adjoint", lambda: NL.nl_adjoint(mesh, modes, U, ybar, dbar, work), ab,
              2 * b2_flops(G, M), "adjoint_kernel<32> + adjoint_gather_kernel",
              overhead=2 * (6 * m.me * 6 * M * 8)),
    ]
    cfg = {"workload": "tangent + adjoint of the convective term around a base flow (modes 0-4, amplitude 0.1), "
                       "158x316 P2 mesh (99,856 triangles), m_max=32",
           "me": m.me, "np": npn, "gauss_points": G, "m_max": M}
    return phases, 2 * G * M, cfg, ("tangent", m, M)


def wl_batch(torch, args):
    from sfemans_amd import batch as BT
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    rng = np.random.default_rng(SEED + 3)
    items, units, byts, flops = [], 0, 0, 0.0
    for k in range(32):
        me = float(np.exp(rng.uniform(np.log(2000), np.log(50000))))
        ny = max(2, int(round(np.sqrt(me))))
        nx = max(1, ny // 2)
        M = int(rng.choice([4, 8, 16, 32, 64, 128]))
        m = rectangle_p2(nx, ny)
        mesh = NL.Mesh.from_p2(m)
        u, = _fields(torch, m.np, M, 2.0, (SEED + 100 + k,))
        out = FP.fortran_empty((m.n_gauss, 6, M))
        items.append((mesh, list(range(M)), u, out))
        units += m.n_gauss * M
        byts += (m.np + m.n_gauss) * 6 * M * 8 + m.me * 24 + m.np * 16
        flops += b2_flops(m.n_gauss, M)
    b = BT.Batch(items)
    b.autotune(reps=2)
    b.capture()
    info = b.info()
    phases = [Phase("batch", b.run, byts, flops, "batch_kernel (persistent, hipGraph replay)")]
    cfg = {"workload": "32 independent (curl u) x u problems, 2k-50k triangles, m_max in {4..128}, "
                       "one persistent launch replayed from a hipGraph",
           "problems": 32, "triangles": int(sum(it[0].me for it in items)), "schedule": info,
           "_keep": b}
    return phases, units, cfg, ("batch", None, None)


def wl_full(torch, args):
    """Full explicit nonlinear step, forward and adjoint, 200,704 triangles, m_max = 128."""
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    from sfemans_amd.synth import level_set_density
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 128
    modes = list(range(M))
    G, npn, me = m.n_gauss, m.np, m.me
    u, B = _fields(torch, npn, M, 4.0, (SEED + 5, SEED + 6))
    T, = _fields(torch, npn, M, 4.0, (SEED + 7,), ncol=2)
    rho = level_set_density(torch, m.rr, M)
    ybar, = _fields(torch, G, M, 4.0, (SEED + 8,))
    hc = torch.ones(me, dtype=torch.float64, device="cuda")
    v145, v236 = FP.fortran_empty((npn, 3, M)), FP.fortran_empty((npn, 3, M))
    welem = torch.empty(36 * me * M, dtype=torch.float64, device="cuda")
    uxb = FP.fortran_empty((npn, 6, M))
    ff0 = torch.zeros((2 * M, npn), dtype=torch.float64, device="cuda").T
    vT = FP.fortran_empty((npn, 2 * M))
    wq = torch.empty(6 * me * 2 * M, dtype=torch.float64, device="cuda")
    rhou = FP.fortran_empty((npn, 6, M))
    mten = torch.empty((M, 6, npn, 3), dtype=torch.float64, device="cuda").permute(3, 2, 1, 0)
    dbar = FP.fortran_empty((npn, 6, M))
    nod6, gau6 = npn * 6 * M * 8, G * 6 * M * 8
    # the two right-hand sides in their fused forms (no Gauss array), the same
    # values bit for bit as smb_cross_prod_gauss_sft_par then rhs_3x3 and
    # smb_ugradc_gauss_fft_par then qs_00_gauss
    vTf = torch.as_strided(vT, (npn, 2, M), (1, npn, 2 * npn))
    ff0f = torch.as_strided(ff0, (npn, 2, M), (1, npn, 2 * npn))
    phases = [
        Phase("ns_nl_rhs", lambda: NL.nl_rhs_3x3(mesh, modes, u, v145, v236, -1.0, welem),
              2 * nod6 + me * 80 + npn * 16, b2_flops(G, M) + G * M * 72, "gauss_rhs_kernel<128,CROSS_CURL> + rhs_gather_kernel",
              overhead=2 * 36 * me * M * 8),
        Phase("u_x_B_nodes", lambda: FP.FFT_PAR_CROSS_PROD_DCL(None, u, B, uxb), 3 * nod6,
              npn * fft_flops_per_point(M, 6, 3, 9), "fftprod_kernel<128,CROSS>"),
        Phase("temperature_nl_rhs", lambda: NL.nl_qs_00(mesh, modes, hc, u, T, vTf, ff0f, -1.0, wq),
              nod6 + 3 * (nod6 // 3) + me * 80 + npn * 16,
              G * M * 100 + G * fft_flops_per_point(M, 8, 1, 8) + G * 2 * M * 30,
              "gauss_rhs_kernel<128,UGRADC> + qs00_gather_kernel", overhead=2 * 6 * me * 2 * M * 8),
        Phase("rho_u", lambda: FP.FFT_SCALAR_VECT_DCL(None, u, rho, rhou, 1), 2 * nod6 + nod6 // 3,
              npn * fft_flops_per_point(M, 4, 3, 3), "fftprod_kernel<128,SVECT1>"),
        Phase("m_tensor_u", lambda: FP.FFT_TENSOR_DCL(None, rhou, u, mten), 5 * nod6,
              npn * fft_flops_per_point(M, 6, 9, 9), "fftprod_kernel<128,TENSOR>"),
        Phase("adjoint", lambda: NL.nl_adjoint(mesh, modes, u, ybar, dbar, welem),
              gau6 + 2 * nod6, 2 * b2_flops(G, M),
              "adjoint_kernel<128> + adjoint_gather_kernel", overhead=2 * 36 * me * M * 8),
    ]
    cfg = {"workload": "full explicit nonlinear step, forward + adjoint, 224x448 P2 mesh (200,704 triangles), "
                       "m_max=128: (curl u)xu -> rhs_3x3 and u.grad T -> qs_00_gauss (fused, no Gauss arrays), u x B, "
                       "rho u and (rho u) (x) u "
                       "with rho = 1 + 0.5 tanh(phi/0.05) of a sphere, adjoint of the convective term",
           "me": me, "np": npn, "gauss_points": G, "m_max": M, "grid": 3 * mp_of(M)}
    return phases, G * M, cfg, ("full", m, M)


WORKLOADS = {"nl": wl_nl, "mhd": wl_mhd, "fftprod": wl_fftprod, "tangent": wl_tangent, "batch": wl_batch,
             "full": wl_full, "nsrhs": wl_nsrhs, "b3": wl_b3, "exact": wl_exact, "entropy": wl_entropy,
             "grid": wl_grid}


# ------------------------------------------------------------- CPU baseline
def _cpu_threads() -> int:
    """Host threads for the all-core CPU timing: the cores this process may
    use, at most 16 (a one-GPU share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(kind, mesh, M, sample_s: float):
    """The CPU restatement of the reference algorithm on a bounded sample of
    the same workload: the compiled C restatement (OpenMP over element
    blocks) timed on one core and on all cores, or the numpy one on one core
    when the C library is not built."""
    from oracle import cpu as OC
    from oracle import fft_par as F
    from sfemans_amd.synth import fourier_field
    if kind == "exact":
        # the oracle's six products (pocketfft on the same grid) on 2,048 points, same m_max
        rng = np.random.default_rng(3)
        npt = 2048
        dec = np.exp(-np.arange(M) / 4.0)
        fld = lambda nb, sc=1.0: np.asfortranarray(rng.standard_normal((npt, nb, M)) * dec * sc)
        phi = np.asfortranarray(rng.standard_normal((1, npt, 2, M)) * 0.3 * dec)
        phi[0, :, 0, 0] += 0.5
        mom, rho, c, d, u, g = fld(6), fld(2, 0.1), fld(2), fld(2, 0.1), fld(6), fld(6)
        rho[:, 0, 0], d[:, 0, 0] = 1.5, 2.0
        mp = 3 * M // 2

        def run():
            F.fft_heaviside_dcl(phi, [1.0, 3.0], mp, 0.05)
            F.fft_scalar_vect_dcl(mom, rho, 2, mp)
            F.fft_par_div_dcl(c, d, mp)
            F.fft_squareroot_scalar(rho, mp)
            F.fft_tensor_dcl(u, g, mp, opt_tension=True)
            F.fft_max_vel_dcl(u, mp)
        run()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() < t0 + sample_s or n == 0:
            run()
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n * npt * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
                "sample": f"the six non-polynomial FFT_PAR products restated in numpy (pocketfft on the grid "
                          f"N={2 * mp - 1}), {n} passes over {npt} points x m_max={M}, 1 thread, {dt:.1f} s"}
    if kind == "grid":
        # the oracle's grid-function and reduction routines on 2,048 points, same m_max
        rng = np.random.default_rng(5)
        npt = 2048
        mp = 3 * M // 2
        N = 2 * mp - 1
        dec = np.exp(-np.arange(M) / 4.0)
        fld = lambda nb, sc=1.0: np.asfortranarray(rng.standard_normal((npt, nb, M)) * dec * sc)
        u, gphi, phi = fld(6), fld(6), fld(2, 0.2)
        phi[:, 0, 0] = 0.5
        ls = np.asfortranarray(rng.standard_normal((2, npt, 2, M)) * 0.01 * dec)
        ls[0, :, 0, 0] += 0.8
        ls[1, :, 0, 0] += 0.6
        eta = rng.random((N, npt)) + 0.5
        one = np.zeros((npt, 2, M))
        one[:, 0, 0] = 1.0

        def run():
            R = F.fft_par_real(u, mp)
            F.fft_par_var_eta_prod(R[:, 0, :], one, mp)
            F.fft_par_var_eta_prod(eta, u, mp)
            F.fft_compression_level_set_dcl(gphi, u, phi, 7, 1.0, 0.5, 0.01, mp)
            F.fft_max_min_vel_dcl(u, mp)
            F.fft_check_interface(ls, 3, mp)
        run()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() < t0 + sample_s or n == 0:
            run()
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n * npt * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
                "sample": f"the six grid-function and reduction FFT_PAR routines restated in numpy (pocketfft on "
                          f"the grid N={N}; the analysis as the product with 1), {n} passes over {npt} points x "
                          f"m_max={M}, 1 thread, {dt:.1f} s"}
    if kind == "entropy":
        # the oracle's six entropy-viscosity products on 2,058 points (294 elements of 7), same m_max
        rng = np.random.default_rng(4)
        npt, lg = 2058, 7
        mp = 3 * M // 2
        N = 2 * mp - 1
        dec = np.exp(-np.arange(M) / 4.0)
        fld = lambda nb: np.asfortranarray(rng.standard_normal((npt, nb, M)) * dec)
        u, mom, res, g1, g2, g3 = (fld(6) for _ in range(6))
        c, rmass = fld(2), fld(2)
        c[:, 0, 0] = 0.5
        creal = rng.random((N, npt))
        h = rng.random(npt) * 0.01

        def run():
            F.fft_par_entro_visc_dcl(u, c, creal, 0.3, mp)
            F.fft_par_compr_entro_visc_dcl(u, g1, c, creal, 0.3, 1.5, mp)
            F.fft_compute_entropy_visc_grad_mom(g1, g2, g3, creal, 0.2, mp)
            F.fft_compute_entropy_visc(u, res, g1, g2, g3, h, 1.0, lg, 0.3, 0.9, 0.8, mp)
            F.fft_compute_entropy_visc_mom(u, mom, res, rmass, h, lg, 0.6, 0.9, mp, True)
            F.fft_compute_diffu_mom(g1, g2, c, mp)
        run()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() < t0 + sample_s or n == 0:
            run()
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n * npt * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
                "sample": f"the six entropy-viscosity FFT_PAR products restated in numpy (pocketfft on the grid "
                          f"N={N}), {n} passes over {npt} points x m_max={M}, 1 thread, {dt:.1f} s"}
    if kind == "nsrhs":
        from oracle import assembly
        from sfemans_amd.mesh import rectangle_p2
        # the oracle's rhs_ns_gauss_3x3 + rhs_3x3 with opt_tensor on a 16x16 P2 mesh, same m_max
        sm = rectangle_p2(16, 16)
        rng = np.random.default_rng(3)
        F = lambda *sh: rng.standard_normal(sh)
        lm = np.arange(M)
        ins = (F(sm.np, 6, M), F(sm.np, 6, M), F(sm.np, 2, M), F(sm.np, 2, M), F(sm.n_gauss, 6, M))
        T = F(3, sm.n_gauss, 6, M)

        def run():
            g = assembly.rhs_ns_gauss_3x3(sm.jj, sm.rr, lm, *ins)
            assembly.rhs_3x3_all_modes(sm.jj, sm.rr, g, lm, T)
        run()
        n, t0 = 0, time.perf_counter()
        t_end = t0 + sample_s
        while time.perf_counter() < t_end or n == 0:
            run()
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n * sm.n_gauss * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
                "sample": f"rhs_ns_gauss_3x3 + rhs_3x3 with opt_tensor restated in numpy, {n} calls on a "
                          f"16x16 P2 mesh ({sm.me} triangles) x m_max={M}, 1 thread, {dt:.1f} s"}
    if kind in ("nl", "mhd", "tangent", "batch", "full") and OC.available():
        from sfemans_amd.mesh import rectangle_p2

        def rate(sm, Mm, th, fft, budget):
            """Median Gauss-pt*modes/s of the C restatement on mesh sm, m_max Mm, th threads."""
            u = fourier_field(sm.np, 6, Mm, 3, 4.0)
            modes = np.arange(Mm, dtype=np.int32)
            out = np.empty((sm.n_gauss, 6, Mm), order="F")
            OC.smb_cross_prod_gauss(sm.jj, sm.rr, modes, u, th, out, fft=fft)          # warm-up
            times = []
            t_end = time.perf_counter() + budget
            while time.perf_counter() < t_end or len(times) < 3:
                t0 = time.perf_counter()
                OC.smb_cross_prod_gauss(sm.jj, sm.rr, modes, u, th, out, fft=fft)
                times.append(time.perf_counter() - t0)
            return sm.n_gauss * Mm / float(np.median(times)), len(times)

        # forward nonlinear term on a 64x64 P2 mesh (8,192 triangles), same m_max, FFT variant
        sm = rectangle_p2(64, 64)
        nth = _cpu_threads()
        v_all, n_all = rate(sm, M, nth, True, 0.35 * sample_s)
        v_1, n_1 = rate(sm, M, 1, True, 0.25 * sample_s) if nth > 1 else (v_all, n_all)
        v_dir, n_dir = rate(sm, M, nth, False, 0.15 * sample_s)
        # the 40x80 rectangle (6,400 triangles) at m_max = 8 on one core (one MPI rank)
        c1 = rectangle_p2(40, 80)
        v_c1, n_c1 = rate(c1, 8, 1, True, 0.25 * sample_s)
        return {"value": v_all, "unit": "Gauss-pt*modes/s", "cores": nth, "kind": "port",
                "value_1core": v_1, "value_direct_reference_grid": v_dir,
                "rectangle_40x80_m8_1core": v_c1,
                "sample": f"smb_cross_prod_gauss_sft_par restated in C: Gauss interpolation + curl, then a "
                          f"mixed-radix (4,2,3) FFT on N={OC.fft_grid(M)} (no FFTW on the host), OpenMP over "
                          f"element blocks; 64x64 P2 mesh ({sm.me} triangles) x m_max={M}, median of {n_all} "
                          f"calls on {nth} threads and of {n_1} calls on 1 thread; value_direct_reference_grid: "
                          f"direct sums on N={2 * (3 * M // 2) - 1}, median of {n_dir} calls on {nth} threads; "
                          f"rectangle_40x80_m8_1core: the 40x80 mesh ({c1.me} triangles) at m_max=8 "
                          f"(N={OC.fft_grid(8)}), median of {n_c1} calls on 1 thread; forward term only"}
    if kind in ("nl", "mhd", "tangent", "batch", "full") and mesh is not None:
        from oracle import smb
        t_end = time.perf_counter() + sample_s
        ne = 512
        jj = np.asfortranarray(mesh.jj[:, :ne])
        used = np.unique(jj)
        remap = np.zeros(mesh.np + 1, dtype=np.int32)
        remap[used] = np.arange(1, used.size + 1)
        jj = np.asfortranarray(remap[jj])
        rr = np.asfortranarray(mesh.rr[:, used - 1])
        u = fourier_field(rr.shape[1], 6, M, 3, 4.0)
        modes = np.arange(M)
        smb.smb_cross_prod_gauss_sft_par(jj, rr, modes, u)   # warm-up
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() < t_end or n == 0:
            smb.smb_cross_prod_gauss_sft_par(jj, rr, modes, u)
            n += 1
        dt = time.perf_counter() - t0
        return {"value": n * 7 * ne * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
                "sample": f"smb_cross_prod_gauss_sft_par restated in numpy (pocketfft, reference grid "
                          f"N={2 * (3 * M // 2) - 1}), {n} calls on {ne} triangles x m_max={M}, 1 thread, "
                          f"{dt:.1f} s; forward term only"}
    t_end = time.perf_counter() + sample_s
    M = M or 64
    npt = 4096
    rng = np.random.default_rng(1)
    V1 = np.asfortranarray(rng.standard_normal((npt, 6, M)))
    V2 = np.asfortranarray(rng.standard_normal((npt, 6, M)))
    mp = 3 * M // 2
    F.fft_par_cross_prod_dcl(V1, V2, mp)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() < t_end or n == 0:
        F.fft_par_cross_prod_dcl(V1, V2, mp)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n * npt * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
            "sample": f"FFT_PAR_CROSS_PROD_DCL restated in numpy (pocketfft, reference grid N={2 * mp - 1}), "
                      f"{n} calls x {npt} points x m_max={M}, 1 thread, {dt:.1f} s"}


def measured_traffic(workload: str, phase: str):
    """HBM bytes per launch of the dominant phase from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, each divided by the factor the
    same counter reads for a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    ph = t.get("phases", {}).get(phase)
    if not ph:
        return None
    return ph["fetch_bytes"] / t["copy_fetch_factor"] + ph["write_bytes"] / t["copy_write_factor"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="nl", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample-s", type=float, default=12.0)
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="run one plain copy kernel (y = x * 1.0) of this size first, for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import torch
    torch.cuda.set_device(0)
    if args.calibrate_copy_gib > 0:
        n = int(args.calibrate_copy_gib * (1 << 30)) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y
    phases, units, cfg, cpu_kind = WORKLOADS[args.workload](torch, args)
    for _ in range(args.warmup):
        for ph in phases:
            ph.fn()
    torch.cuda.synchronize()
    nph = len(phases)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(nph + 1)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        e = evs[k]
        for i, ph in enumerate(phases):
            e[i].record()
            ph.fn()
        e[nph].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ph_ms = [sum(evs[k][i].elapsed_time(evs[k][i + 1]) for k in range(args.steps)) / args.steps for i in range(nph)]
    ms_per_step = wall * 1000.0 / args.steps
    value = units / (ms_per_step / 1000.0)
    dom = int(np.argmax(ph_ms))
    dp = phases[dom]
    achieved = dp.bytes / (ph_ms[dom] / 1000.0) / 1e9

    def roofs(ph, ms):
        hbm = ph.bytes / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS
        fp64 = ph.flops / (ms / 1000.0) / 1e12 / FP64_PEAK_TFLOPS
        # the roof that bounds the kernel: the one its work would take longer to meet at peak
        binding = "HBM" if ph.bytes / (HBM_PEAK_GBS * 1e9) > ph.flops / (FP64_PEAK_TFLOPS * 1e12) else "FP64"
        return hbm, fp64, binding

    phase_info = {}
    for i, ph in enumerate(phases):
        hbm, fp64, binding = roofs(ph, ph_ms[i])
        phase_info[ph.name] = {"ms": ph_ms[i], "kernel": ph.kernel, "GB_per_launch": ph.bytes / 1e9,
                               "GFLOP_per_launch": ph.flops / 1e9, "hbm_frac": hbm, "fp64_frac": fp64,
                               "binding_roof": binding,
                               "binding_frac": hbm if binding == "HBM" else fp64,
                               "work_array_GB": ph.overhead / 1e9}
    cfg["phases"] = phase_info
    cfg.pop("_keep", None)
    kind, mesh, M = cpu_kind
    cpu = cpu_baseline(kind, mesh, M if M else 64, args.cpu_sample_s)
    line = {
        "metric": "Gauss-pt*modes/s for nonlinear term fwd+adjoint (FP64), % HBM roofline",
        "value": value, "unit": "Gauss-pt*modes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded N(0,1)*exp(-m/decay) nodal Fourier coefficients on structured P2 meshes)",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, dp.name),
                     "algorithmic_bytes": dp.bytes, "work_array_bytes": dp.overhead,
                     "kernel": dp.kernel, "phase": dp.name, "kernel_ms": ph_ms[dom],
                     "flops": dp.flops, "fp64_tflops": dp.flops / (ph_ms[dom] / 1000.0) / 1e12,
                     "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                     "fp64_frac": phase_info[dp.name]["fp64_frac"],
                     "binding_roof": phase_info[dp.name]["binding_roof"],
                     "binding_frac": phase_info[dp.name]["binding_frac"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
