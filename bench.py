"""Benchmark of the pseudo-spectral nonlinear-term path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line.  ``value`` is Gauss-pt*modes per second for the timed
workload with inputs resident in HBM; ``roofline`` describes the dominant
kernel (algorithmic bytes / measured kernel time, HIP events on the launch
stream); ``cpu_baseline`` is the CPU oracle (numpy/pocketfft restatement of
the reference algorithm) timed on this host's cores on a bounded sample.

Workloads
* ``fftprod``: FFT_PAR_CROSS_PROD_DCL on the Gauss-point arrays of the
  224x448 P2 mesh (1,404,928 points, m_max = 64), inputs N(0,1)*exp(-m/4).
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


def flops_fft_unit(M: int) -> float:
    """Algorithmic flops per Gauss-pt*mode of one cross product (FFT part).

    9 real transforms of length N = 3*MP (6 inverse, 3 forward) at
    2.5*N*log2(N) flops each, plus 9 flops per grid point for the product.
    """
    MP = 4
    while MP < M:
        MP *= 2
    N = 3 * MP
    return (9 * 2.5 * N * np.log2(N) + 9 * N) / M


def bench_fftprod(args, torch):
    from sfemans_amd import fft_parallele as FP
    npt, M = 1404928, 64
    g = torch.Generator(device="cuda").manual_seed(0x5F3E + 2)
    decay = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / 4.0)
    a = (torch.randn((M, 6, npt), generator=g, device="cuda", dtype=torch.float64) * decay[:, None, None]).permute(2, 1, 0)
    b = (torch.randn((M, 6, npt), generator=g, device="cuda", dtype=torch.float64) * decay[:, None, None]).permute(2, 1, 0)
    out = FP.fortran_empty((npt, 6, M))

    def step():
        FP.FFT_PAR_CROSS_PROD_DCL(None, a, b, out)

    units = npt * M
    bytes_launch = 3 * npt * 6 * M * 8
    flops_launch = flops_fft_unit(M) * units
    cfg = {"workload": "FFT_PAR_CROSS_PROD_DCL at Gauss points, 224x448 P2 mesh",
           "points": npt, "m_max": M, "grid": 3 * 64, "kernel": "fftprod_kernel<64,CROSS>"}
    return step, units, bytes_launch, flops_launch, cfg


def _fields_device(torch, npts, M, decay, seeds):
    """Seeded N(0,1)*exp(-m/decay) nodal fields generated on the GPU (Fortran strides)."""
    out = []
    dec = torch.exp(-torch.arange(M, device="cuda", dtype=torch.float64) / decay)
    for sd in seeds:
        g = torch.Generator(device="cuda").manual_seed(sd)
        t = torch.randn((M, 6, npts), generator=g, device="cuda", dtype=torch.float64) * dec[:, None, None]
        t[0, 1::2, :] = 0.0
        out.append(t.permute(2, 1, 0))
    return out


def bench_mhd(args, torch):
    """MHD nonlinear terms of one explicit step on the 224x448 P2 mesh, m_max = 64."""
    from sfemans_amd import fft_parallele as FP
    from sfemans_amd import nonlinear as NL
    from sfemans_amd.mesh import rectangle_p2
    m = rectangle_p2(224, 448)
    mesh = NL.Mesh.from_p2(m)
    M = 64
    modes = list(range(M))
    u, H, B = _fields_device(torch, m.np, M, 4.0, (0x5F3E + 2, 0x5F3E + 3, 0x5F3E + 4))
    rotv = FP.fortran_empty((m.n_gauss, 6, M))
    uxb = FP.fortran_empty((m.np, 6, M))

    def step():
        NL.nl_rotv_mhd(mesh, modes, u, H, B, 1.0, rotv)
        FP.FFT_PAR_CROSS_PROD_DCL(None, u, B, uxb)

    units = m.n_gauss * M
    # dominant kernel: the fused Gauss-point MHD term (reads u, H, B nodal once, writes rotv_v)
    bytes_launch = (3 * m.np * 6 * M + m.n_gauss * 6 * M) * 8
    flops_launch = 2 * flops_fft_unit(M) * units
    cfg = {"workload": "MHD nonlinear terms, 224x448 P2 mesh (200,704 triangles), m_max=64: "
                       "(curl u)xu - (curl H)xB at Gauss points + u x B at nodes",
           "me": m.me, "np": m.np, "gauss_points": m.n_gauss, "m_max": M, "grid": 192,
           "kernel": "gauss_kernel<64,MHD>"}
    return step, units, bytes_launch, flops_launch, cfg


def cpu_baseline_fftprod(sample_s: float):
    from oracle import fft_par as F
    M = 64
    rng = np.random.default_rng(1)
    npt = 4096
    V1 = np.asfortranarray(rng.standard_normal((npt, 6, M)))
    V2 = np.asfortranarray(rng.standard_normal((npt, 6, M)))
    F.fft_par_cross_prod_dcl(V1, V2, 96)  # warm-up
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < sample_s:
        F.fft_par_cross_prod_dcl(V1, V2, 96)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n * npt * M / dt, "unit": "Gauss-pt*modes/s", "cores": 1, "kind": "port",
            "sample": f"{n} calls x {npt} points x m_max={M} on the reference grid N=191 "
                      f"(numpy/pocketfft restatement, 1 thread, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="fftprod")
    ap.add_argument("--cpu-sample-s", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import torch
    torch.cuda.set_device(0)
    step, units, bytes_launch, flops_launch, cfg = {"fftprod": bench_fftprod, "mhd": bench_mhd}[args.workload](args, torch)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1000.0 / args.steps
    value = units / (ms_per_step / 1000.0)
    achieved = bytes_launch / (kern_ms / 1000.0) / 1e9
    cpu = cpu_baseline_fftprod(args.cpu_sample_s)
    line = {
        "metric": "Gauss-pt*modes/s for nonlinear term fwd+adjoint (FP64), % HBM roofline",
        "value": value, "unit": "Gauss-pt*modes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": kern_ms, "fp64_tflops": flops_launch / (kern_ms / 1000.0) / 1e12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
