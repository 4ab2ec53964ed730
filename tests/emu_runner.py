"""Build and call the CPU emulator of the libsfx kernel bodies (tests only).

The emulator (tests/emu/sfx_emu.cpp) compiles the same kernel-body headers as
the GPU library with SFX_EMULATE, runs each work-group as std::threads and
checks every computed index against its buffer.  A violation aborts, so each
call runs in a child process.

Sanitizer build: with SFX_EMU_SANITIZE=1 in the environment (or
``call(..., sanitize=True)``) the emulator is built with AddressSanitizer and
UndefinedBehaviorSanitizer (host code only; g++, no GPU code) and each call
runs in a fresh Python child with the ASan runtime preloaded; any report
ends the child with a non-zero status, which ``call`` turns into an error.
``SFX_EMU_SANITIZE=1 python -m pytest tests -m "not gpu" -k emu`` runs the
emulator tests under both sanitizers.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os
import pickle
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "sfx_emu.cpp")
OUT_DIR = os.path.join(HERE, "emu", "_build")
LIB = os.path.join(OUT_DIR, "libsfx_emu.so")
LIB_SAN = os.path.join(OUT_DIR, "libsfx_emu_asan_ubsan.so")
CSRC = os.path.join(os.path.dirname(HERE), "sfemans_amd", "csrc")
SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def sanitize_default() -> bool:
    return os.environ.get("SFX_EMU_SANITIZE", "") == "1"


def build_emulator(sanitize: bool = False) -> str:
    lib = LIB_SAN if sanitize else LIB
    deps = [SRC] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if os.path.exists(lib) and os.path.getmtime(lib) >= max(os.path.getmtime(d) for d in deps):
        return lib
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = lib + f".tmp{os.getpid()}"      # one per process: parallel test workers may build at once
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-pthread"]
                   + (SAN_FLAGS if sanitize else []) + ["-o", tmp, SRC], check=True)
    os.replace(tmp, lib)
    return lib


def _load(path: str = LIB):
    L = C.CDLL(path)
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
    L.emu_fftprod.argtypes = [i32, vp, vp, vp, i64, i32, i32, i32, i32, i32]
    L.emu_fftprod.restype = C.c_int
    d = C.c_double
    L.emu_gauss.argtypes = [i32, vp, vp, i64, i64, vp, i32, vp, i32, vp, vp, vp, i32, vp, i32, d, d, d, d]
    L.emu_gauss.restype = C.c_int
    L.emu_rhs.argtypes = [i32, vp, vp, i64, i64, i32, vp, vp, vp, vp, vp, vp, vp, d]
    L.emu_rhs.restype = C.c_int
    L.emu_rhs_ns.argtypes = [vp, vp, i64, i64, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.emu_rhs_ns.restype = C.c_int
    L.emu_rhs_ns_gauss.argtypes = [vp, vp, i64, i64, vp, i32, vp, vp, vp, vp, vp, i32, vp]
    L.emu_rhs_ns_gauss.restype = C.c_int
    L.emu_rhs_bdy.argtypes = [vp, i64, vp, i64, i32, vp, vp, vp, d]
    L.emu_rhs_bdy.restype = C.c_int
    L.emu_adjoint.argtypes = [vp, vp, i64, i64, vp, vp, vp, i32, vp, vp, d, vp]
    L.emu_adjoint.restype = C.c_int
    L.emu_batch.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    L.emu_batch.restype = C.c_int
    L.emu_exact.argtypes = [i32, i64, i32, i32, vp, vp, vp, i32, C.c_double, i32, vp]
    L.emu_exact.restype = C.c_int
    L.emu_exact_evisc.argtypes = [i64, i32, i32, i32, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double, vp]
    L.emu_exact_evisc.restype = C.c_int
    L.emu_exact_emom.argtypes = [i64, i32, i32, i32, vp, vp, vp, C.c_double, C.c_double, vp, vp]
    L.emu_exact_emom.restype = C.c_int
    L.emu_dotsf.argtypes = [vp, vp, i64, i64, vp, i32, i32, vp, vp, i64, vp]
    L.emu_dotsf.restype = C.c_int
    L.emu_gauss_values.argtypes = [vp, vp, i64, i64, vp, i32, i32, i32, d, d, vp, vp]
    L.emu_gauss_values.restype = C.c_int
    L.emu_coldot.argtypes = [vp, vp, i64, i32, vp]
    L.emu_coldot.restype = C.c_int
    L.emu_nl_rhs.argtypes = [vp, vp, i64, i64, vp, vp, i32, vp, vp, d, d, d, d]
    L.emu_nl_rhs.restype = C.c_int
    L.emu_nl_qs00.argtypes = [vp, vp, i64, i64, vp, vp, vp, vp, vp, i32, vp, d]
    L.emu_nl_qs00.restype = C.c_int
    L.emu_sanitizer_probe.argtypes = [i32, i32]
    L.emu_sanitizer_probe.restype = C.c_int
    return L


def _run(L, fname, args):
    arrays = [np.asfortranarray(a) if isinstance(a, np.ndarray) else a for a in args]
    ptrs = [a.ctypes.data if isinstance(a, np.ndarray) else a for a in arrays]
    ptrs = [None if (a is None) else p for a, p in zip(arrays, ptrs)]
    rc = getattr(L, fname)(*ptrs)
    return rc, [a for a in arrays if isinstance(a, np.ndarray)]


def _child(conn, fname, args):
    try:
        conn.send(_run(_load(), fname, args))
    except Exception as e:  # pragma: no cover
        conn.send((repr(e), None))
    conn.close()


def _call_sanitized(fname: str, args, timeout: float):
    """One call of the sanitizer build in a fresh interpreter with the ASan
    runtime preloaded (it must come first in the process)."""
    lib = build_emulator(sanitize=True)
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:detect_stack_use_after_return=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    with tempfile.TemporaryDirectory(prefix="sfx_emu_san") as d:
        fin, fout = os.path.join(d, "in.pkl"), os.path.join(d, "out.pkl")
        with open(fin, "wb") as f:
            pickle.dump((lib, fname, list(args)), f)
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--child", fin, fout], env=env,
                           capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"sanitized emulator call {fname} exited with {p.returncode}:\n{p.stderr[-4000:]}")
        with open(fout, "rb") as f:
            return pickle.load(f)


def call(fname: str, *args, timeout: float = 300.0, sanitize: bool | None = None):
    """Run emulator entry ``fname`` in a child process; returns (rc, arrays)."""
    if sanitize is None:
        sanitize = sanitize_default()
    if sanitize:
        return _call_sanitized(fname, args, timeout)
    build_emulator()
    ctx = mp.get_context("fork")
    a, b = ctx.Pipe()
    p = ctx.Process(target=_child, args=(b, fname, args))
    p.start()
    b.close()   # the child's end: a child that dies then shows as end of file at once
    if not a.poll(timeout):
        p.kill()
        raise TimeoutError(fname)
    try:
        res = a.recv()
    except EOFError:
        res = None
    p.join()
    if p.exitcode not in (0, None):
        raise RuntimeError(f"emulator child exited with {p.exitcode} (out-of-bounds abort?)")
    return res


if __name__ == "__main__" and len(sys.argv) == 4 and sys.argv[1] == "--child":
    # sanitized child: arguments and results pass through files this module wrote
    with open(sys.argv[2], "rb") as f:
        lib_path, fn, call_args = pickle.load(f)
    res = _run(_load(lib_path), fn, call_args)
    with open(sys.argv[3], "wb") as f:
        pickle.dump(res, f)
