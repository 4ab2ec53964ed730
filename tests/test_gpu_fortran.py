"""The Fortran modules fft_parallele and sfx_nonlinear over libsfx.so.

Runs the compiled driver ``sfemans_amd/fortran/sfx_fortran_driver`` in a
child process: it calls every PUBLIC routine of the reference's
fft_parallele (SOURCES/POST_PROCESSING/fft_parallel.F90:8-17) and the
Gauss-point and assembly routines with the reference's argument lists, on
raw float64 files, and each result is checked here against the oracle
(relative max-norm error <= 1e-12; the reductions to 1e-13).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import assembly, smb  # noqa: E402
from oracle import fft_par as F  # noqa: E402
from sfemans_amd.mesh import boundary_sides, rectangle_p2  # noqa: E402
from sfemans_amd.synth import fourier_field  # noqa: E402

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "sfemans_amd", "fortran", "sfx_fortran_driver")
TOL = 1e-12
INPUTS = dict(LES_coeff1=0.3, LES_coeff2=0.9, LES_coeff4=0.8, LES_coeff1_mom=0.2, level_set_comp_factor=1.5,
              h_min_distance=0.01, taux_precession=0.3, angle_s_pi=0.25)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_fortran_modules(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (amdflang absent at build time)")
    np_, M = 7 * 40, 8
    mp = 3 * M // 2
    N = 2 * mp - 1
    rng = np.random.default_rng(31)
    v = [fourier_field(np_, 6, M, 21 + k, 2.0) for k in range(5)]
    s1 = fourier_field(np_, 2, M, 26, 2.0) * 0.1
    s1[:, 0, 0] = 0.5
    s2 = fourier_field(np_, 2, M, 27, 2.0) * 0.1
    s2[:, 0, 0] = 0.8
    creal = rng.uniform(0.0, 1.0, (N, np_))
    h = rng.uniform(0.001, 0.01, np_)
    ls = np.asfortranarray(rng.standard_normal((2, np_, 2, M)) * 0.02)
    ls[0, :, 0, 0] += 0.7
    ls[1, :, 0, 0] += 0.4
    ls[:, :, 1, 0] = 0.0
    rg2 = np.asfortranarray(np.stack([rng.uniform(0.1, 1.0, np_), rng.uniform(-1, 1, np_)]))
    m = rectangle_p2(4, 5)
    jjs = boundary_sides(m.jj)
    mes = jjs.shape[1]
    vn, wn, cn = fourier_field(m.np, 6, M, 28, 2.0), fourier_field(m.np, 6, M, 29, 2.0), \
        fourier_field(m.np, 2, M, 30, 2.0)
    hc = rng.uniform(0.5, 2.0, m.me)
    rhs = rng.standard_normal((m.n_gauss, 6))
    ten = rng.standard_normal((3, m.n_gauss, 6))
    tbd = rng.standard_normal((3 * mes, 6))
    ffn, ffg = rng.standard_normal(m.np), rng.standard_normal(m.n_gauss)
    (tmp_path / "dims.txt").write_text(f"{np_} {M} {m.me} {m.np} {mes}\n")
    files = dict(v1=v[0], v2=v[1], v3=v[2], v4=v[3], v5=v[4], s1=s1, s2=s2, creal=creal, h=h, ls=ls, rg2=rg2,
                 rr=m.rr, vn=vn, wn=wn, cn=cn, hc=hc, rhs=rhs, ten=ten, tbd=tbd, ffn=ffn, ffg=ffg)
    for name, a in files.items():
        np.asfortranarray(a, dtype="<f8").T.tofile(tmp_path / f"{name}.bin")
    for name, a in (("jj", m.jj), ("jjs", jjs)):
        np.asfortranarray(a, dtype="<i4").T.tofile(tmp_path / f"{name}.bin")
    print(f"case fortran driver np={np_} M={M} mesh 4x5 mes={mes}", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def out(name, shape):
        return np.fromfile(tmp_path / f"{name}.out", dtype="<f8").reshape(shape[::-1]).T

    v1, v2, v3, v4, v5 = v
    n6, n2, t3 = (np_, 6, M), (np_, 2, M), (3, np_, 6, M)
    I = INPUTS
    checks = {
        "cross": (n6, F.fft_par_cross_prod_dcl(v1, v2, mp)),
        "dot": (n2, F.fft_par_dot_prod_dcl(v1, v2, mp)),
        "prod": (n2, F.fft_par_prod_dcl(s1, s2, mp)),
        "svect1": (n6, F.fft_scalar_vect_dcl(v1, s2, 1, mp)),
        "tensor": (t3, F.fft_tensor_dcl(v1, v2, mp)),
        "div": (n2, F.fft_par_div_dcl(np.asfortranarray(v1[:, 0:2]), s2, mp)),
        "svect2": (n6, F.fft_scalar_vect_dcl(v1, s2, 2, mp)),
        "svect_no": (n6, F.fft_scalar_vect_no_overshoot([0.2, 0.7, 0.5], v1, s1, 1, mp)),
        "tension": (t3, F.fft_tensor_dcl(v1, v2, mp, opt_tension=True)),
        "sqrt": (n2, F.fft_squareroot_scalar(s1, mp)),
        "no_ls": (n2, F.fft_no_overshoot_level_set(s1, mp)),
        "heavi": (n2, F.fft_heaviside_dcl(ls, [1.0, 3.0, 0.5], mp, 0.05)),
        "real": ((2 * M + 1, 3, np_), F.fft_par_real(v1, M + 1)),
        "scal": (n2, F.fft_par_scal_funct(s2, lambda x: x * x, mp)),
        "ev1": (n6, None), "ev2": (n6, None), "ev3": (n6, None),
        "evm1": ((N, np_), None), "evm2": ((N, np_), None),
        "evgrad": (t3, F.fft_compute_entropy_visc_grad_mom(v1, v2, v3, creal, I["LES_coeff1_mom"], mp)),
        "entro_v": (n6, None), "entro_c": (n2, None),
        "compr_v": (n6, None), "compr_c": (n2, None),
        "cls": (n2, F.fft_compression_level_set_dcl(v1, v2, s1, 7, I["level_set_comp_factor"], I["LES_coeff4"],
                                                    I["h_min_distance"], mp)),
    }
    e1, e2, e3 = F.fft_compute_entropy_visc(v1, v2, v3, v4, v5, h, 1.0, 7, I["LES_coeff1"], I["LES_coeff2"],
                                            I["LES_coeff4"], mp)
    checks["ev1"], checks["ev2"], checks["ev3"] = (n6, e1), (n6, e2), (n6, e3)
    r1, r2 = F.fft_compute_entropy_visc_mom(v1, v2, v3, s2, h, 7, I["LES_coeff2"], I["LES_coeff4"], mp, True)
    checks["evm1"], checks["evm2"] = ((N, np_), r1), ((N, np_), r2)
    d1, d2 = F.fft_compute_diffu_mom(v1, v2, s2, mp)
    checks["diffu1"], checks["diffu2"] = (n6, d1), (n6, d2)
    ev, ec = F.fft_par_entro_visc_dcl(v1, s1, creal, 0.3, mp)
    checks["entro_v"], checks["entro_c"] = (n6, ev), (n2, ec)
    cv, cc = F.fft_par_compr_entro_visc_dcl(v1, v2, s1, creal, 0.3, I["level_set_comp_factor"], mp)
    checks["compr_v"], checks["compr_c"] = (n6, cv), (n2, cc)
    # the user functions, evaluated here as the driver's ETA does
    th = 2 * np.pi * np.arange(N) / N
    npe = min(np_, m.np)
    eta_t = 1 + np.cos(th)[:, None] * m.rr[0, :npe] / 2 + np.sin(th)[:, None] * m.rr[1, :npe] / 4 + 0.5
    checks["eta_t"] = ((npe, 6, M), F.fft_par_var_eta_prod(eta_t, np.asfortranarray(v1[:npe]), mp))
    eta_g = 1 + np.cos(th)[:, None] * rg2[0] / 2 + np.sin(th)[:, None] * rg2[1] / 4 + 0.5
    checks["eta_g"] = (n2, F.fft_par_var_eta_prod(eta_g, s1, mp))
    for name, (shape, exp) in checks.items():
        assert rel(out(name, shape), exp) <= TOL, name
    vmax = F.fft_max_vel_dcl(v1, mp)
    assert abs(out("maxvel", (1,))[0] - vmax) <= 1e-13 * vmax
    vmx, vmn = F.fft_max_min_vel_dcl(v1, mp)
    mm = out("maxmin", (2,))
    assert abs(mm[0] - vmx) <= 1e-13 * vmx and abs(mm[1] - vmn) <= 1e-13 * vmx
    assert out("checkif", (1,))[0] == F.fft_check_interface(ls, 3, mp)
    assert np.allclose(out("regul", (4,)), F.regul(np.array([0.47, 0.6, 0.4, 0.5]), 0.05), rtol=0, atol=1e-15)

    # the Gauss-point terms (precession on) and the assembly on the 4x5 mesh
    lm = np.arange(M)
    exp = smb.smb_cross_prod_gauss_sft_par(m.jj, m.rr, lm, vn, True, I["taux_precession"], I["angle_s_pi"])
    assert rel(out("smb_cross", (m.n_gauss, 6, M)), exp) <= TOL
    assert rel(out("smb_curlh", (m.n_gauss, 6, M)), smb.smb_CurlH_cross_B_gauss_sft_par(m.jj, m.rr, lm, vn, wn)) <= TOL
    assert rel(out("smb_ugradc", (m.n_gauss, 2, M)), smb.smb_ugradc_gauss_fft_par(m.jj, m.rr, lm, hc, vn, cn)) <= TOL
    a145, a236 = assembly.rhs_3x3(m.jj, m.rr, rhs)
    assert rel(out("rhs_a145", (3 * m.np,)), a145.T.ravel()) <= TOL
    assert rel(out("rhs_a236", (3 * m.np,)), a236.T.ravel()) <= TOL
    b145, b236 = assembly.rhs_3x3_all_modes(m.jj, m.rr, rhs[:, :, None], np.array([2]), ten[..., None])
    c145, c236 = assembly.rhs_3x3_bdy(jjs, m.rr, tbd[:, :, None])
    assert rel(out("rhs_b145", (3 * m.np,)), (b145 + c145)[:, :, 0].T.ravel()) <= TOL
    assert rel(out("rhs_b236", (3 * m.np,)), (b236 + c236)[:, :, 0].T.ravel()) <= TOL
    assert rel(out("qs00", (m.np,)), assembly.qs_00_gauss(m.jj, m.rr, hc, ffn, ffg)) <= TOL
