"""Navier-Stokes restart files (sfemans_amd.restart), CPU only.

The reader is checked against a file written by a Fortran program that
issues the same WRITE sequence as write_restart_ns
(SOURCES/INITIALISATION/restart.F90:59-97), compiled here with the image's
Fortran compiler (the test is skipped without one), and against the
reference's reading rules: missing modes, val_init, interpol, the dt
extrapolation.  No restart file ships with the reference, so the layout is
pinned by the compiler's own output, not by a file from a run.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sfemans_amd import restart as R

HERE = os.path.dirname(os.path.abspath(__file__))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


def expected(np_v, np_p, m):
    i = np.arange(1, m + 1)[None, None, :]
    un = (np.arange(1, np_v + 1)[:, None, None] + 10 * np.arange(1, 7)[None, :, None] + 100 * i) + 0.25
    un_m1 = -(un - 0.25) - 0.5
    base_p = np.arange(1, np_p + 1)[:, None, None] + 10 * np.arange(1, 3)[None, :, None] + 100 * i
    ls = (0.125 * np.arange(1, 3)[:, None, None, None] + np.arange(1, np_v + 1)[None, :, None, None]
          + 10 * np.arange(1, 3)[None, None, :, None] + 100 * np.arange(1, m + 1)[None, None, None, :])
    ls_m1 = ls - 0.25 * np.arange(1, 3)[:, None, None, None]
    return dict(un=un, un_m1=un_m1, pn=1000 + base_p, pn_m1=2000 + base_p, incpn=3000 + base_p,
                incpn_m1=4000 + base_p, level_set=ls, level_set_m1=ls_m1)


@pytest.fixture(scope="module")
def fortran_files(tmp_path_factory):
    if FC is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("suite")
    exe = d / "write_suite_ns"
    subprocess.run([FC, "-O0", os.path.join(HERE, "fortran", "write_suite_ns.F90"), "-o", str(exe)],
                   check=True, cwd=d, timeout=300)
    files = {}
    for v in (0, 1, 2, 3):
        f = d / f"suite_ns_{v}"
        subprocess.run([str(exe), str(f), str(v)], check=True, timeout=60)
        files[v] = str(f)
    return files


def test_reads_fortran_written_file(fortran_files):
    r = R.read_restart_ns(fortran_files[0], [0, 3, 5], 7, 4, n_int=2, dt=0.01, nb_procs_S=1)
    ex = expected(7, 4, 3)
    assert r["time"] == 1.5 and r["dt"] == 0.01 and r["m_max_c"] == 3
    for k, v in ex.items():
        assert np.array_equal(r[k], v), k
    assert r["max_vel"] == 2.5 + 3          # the last mode read (restart.F90:270)


def test_mode_order_missing_modes_and_val_init(fortran_files):
    ex = expected(7, 4, 3)
    r = R.read_restart_ns(fortran_files[0], [5, 4, 0], 7, 4, n_int=2, dt=0.01)
    assert np.array_equal(r["un"][..., 0], ex["un"][..., 2])
    assert np.array_equal(r["un"][..., 1], 0 * ex["un"][..., 0])        # mode 4 absent: zeros
    assert np.array_equal(r["pn_m1"][..., 2], ex["pn_m1"][..., 0])
    r = R.read_restart_ns(fortran_files[0], [4], 7, 4, n_int=2, dt=0.01, val_init=-3.0)
    assert np.all(r["incpn"] == -3.0) and np.all(r["level_set_m1"] == -3.0) and r["max_vel"] == -3.0


def test_mono_header_and_interpol(fortran_files):
    ex = expected(7, 4, 2)
    r = R.read_restart_ns(fortran_files[1], [2, 7], 7, 4, mono=True)
    assert (r["np_v"], r["np_p"], r["nb_procs_F"], r["m_max_c"]) == (7, 4, 1, 2)
    assert np.array_equal(r["un"], ex["un"]) and np.array_equal(r["incpn_m1"], ex["incpn_m1"])
    r = R.read_restart_ns(fortran_files[1], [0, 1], 7, 4, mono=True, interpol=True)
    assert r["list_mode"] == [2, 7] and np.array_equal(r["pn"], ex["pn"])


def test_dt_extrapolation(fortran_files):
    ex = expected(7, 4, 3)
    r = R.read_restart_ns(fortran_files[0], [0, 3, 5], 7, 4, n_int=2, dt=0.02)
    ratio = 2.0                                                  # restart.F90:316-335
    assert np.allclose(r["un_m1"], ratio * ex["un_m1"] + (1 - ratio) * ex["un"], rtol=0, atol=1e-12)
    assert np.allclose(r["level_set_m1"], ratio * ex["level_set_m1"] + (1 - ratio) * ex["level_set"], atol=1e-12)
    assert np.array_equal(r["un"], ex["un"])


def test_writer_round_trip_and_names(tmp_path):
    rng = np.random.default_rng(3)
    np_v, np_p, modes = 11, 5, [1, 4]
    f = {k: rng.standard_normal(s) for k, s in (("un", (np_v, 6, 2)), ("un_m1", (np_v, 6, 2)), ("pn", (np_p, 2, 2)),
                                               ("pn_m1", (np_p, 2, 2)), ("incpn", (np_p, 2, 2)),
                                               ("incpn_m1", (np_p, 2, 2)))}
    p = str(tmp_path / R.suite_ns_name("run1", rang_S=2))
    R.write_restart_ns(p, 0.75, modes, **f, nb_procs_S=3)
    r = R.read_restart_ns(p, modes, np_v, np_p, nb_procs_S=3)
    for k, v in f.items():
        assert np.array_equal(r[k], v)
    assert os.path.basename(p) == "suite_ns_S002.run1"
    assert R.suite_ns_name("x", mono=True, it=7) == "suite_ns_I007.x"
    with pytest.raises(ValueError):
        R.read_restart_ns(p, modes, np_v, np_p, nb_procs_S=2)


def test_corrupt_file_is_rejected(tmp_path):
    p = tmp_path / "bad"
    p.write_bytes(b"\x08\x00\x00\x00" + b"\x00" * 8 + b"\x07\x00\x00\x00")
    with pytest.raises(ValueError):
        R.read_records(str(p))


def test_reads_fortran_written_maxwell_files(fortran_files):
    ex = expected(7, 4, 3)
    r = R.read_restart_maxwell(fortran_files[2], [5, 0, 4], 7, 4, dt=0.01, nb_procs_S=1)
    assert np.array_equal(r["Hn"][..., 0], ex["un"][..., 2]) and np.array_equal(r["Hn1"][..., 1], ex["un_m1"][..., 0])
    assert np.array_equal(r["Bn"][..., 0], ex["un"][..., 2] + 0.5)
    assert np.array_equal(r["Bn1"][..., 1], ex["un_m1"][..., 0] - 0.5)
    assert np.array_equal(r["phin"][..., 0], ex["pn"][..., 2]) and np.array_equal(r["phin1"][..., 1], ex["pn_m1"][..., 0])
    assert np.all(r["Hn"][..., 2] == 0.0)                                   # mode 4 absent
    r = R.read_restart_maxwell(fortran_files[2], [0, 3, 5], 7, 4, dt=0.005)
    assert np.allclose(r["Bn1"], 0.5 * (ex["un_m1"] - 0.5) + 0.5 * (ex["un"] + 0.5), atol=1e-12)
    ex2 = expected(7, 4, 2)
    r = R.read_restart_maxwell(fortran_files[3], [7, 2], 7, 0, mono=True)    # empty phi mesh
    assert (r["np_v"], r["np_p"]) == (7, 0) and r["phin"].shape == (0, 2, 2)
    assert np.array_equal(r["Hn"][..., 1], ex2["un"][..., 0]) and np.array_equal(r["Bn"][..., 0], ex2["un"][..., 1] + 0.5)


def test_maxwell_writer_round_trip(tmp_path):
    rng = np.random.default_rng(8)
    f = {"Hn": rng.standard_normal((9, 6, 2)), "Hn1": rng.standard_normal((9, 6, 2)),
         "Bn": rng.standard_normal((9, 6, 2)), "Bn1": rng.standard_normal((9, 6, 2)),
         "phin": rng.standard_normal((3, 2, 2)), "phin1": rng.standard_normal((3, 2, 2))}
    p = str(tmp_path / R.suite_maxwell_name("run", mono=True))
    R.write_restart_maxwell(p, 2.0, [0, 6], **f, mono=True)
    r = R.read_restart_maxwell(p, [6, 0], 9, 3, mono=True)
    for k, v in f.items():
        assert np.array_equal(r[k], v[..., ::-1])
    assert os.path.basename(p) == "suite_maxwell.run"
