"""Build and call the CPU emulator of the libsfx kernel bodies (tests only).

The emulator (tests/emu/sfx_emu.cpp) compiles the same kernel-body headers as
the GPU library with SFX_EMULATE, runs each work-group as std::threads and
checks every computed index against its buffer.  A violation aborts, so each
call runs in a child process.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "sfx_emu.cpp")
OUT_DIR = os.path.join(HERE, "emu", "_build")
LIB = os.path.join(OUT_DIR, "libsfx_emu.so")
CSRC = os.path.join(os.path.dirname(HERE), "sfemans_amd", "csrc")


def build_emulator() -> str:
    deps = [SRC] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(d) for d in deps):
        return LIB
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + f".tmp{os.getpid()}"      # one per process: parallel test workers may build at once
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-pthread",
                    "-o", tmp, SRC], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    L = C.CDLL(LIB)
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
    return L


def _child(conn, fname, args):
    try:
        L = _load()
        arrays = [np.asfortranarray(a) if isinstance(a, np.ndarray) else a for a in args]
        ptrs = [a.ctypes.data if isinstance(a, np.ndarray) else a for a in arrays]
        ptrs = [None if (a is None) else p for a, p in zip(arrays, ptrs)]
        rc = getattr(L, fname)(*ptrs)
        conn.send((rc, [a for a in arrays if isinstance(a, np.ndarray)]))
    except Exception as e:  # pragma: no cover
        conn.send((repr(e), None))
    conn.close()


def call(fname: str, *args, timeout: float = 300.0):
    """Run emulator entry ``fname`` in a child process; returns (rc, arrays)."""
    build_emulator()
    ctx = mp.get_context("fork")
    a, b = ctx.Pipe()
    p = ctx.Process(target=_child, args=(b, fname, args))
    p.start()
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
