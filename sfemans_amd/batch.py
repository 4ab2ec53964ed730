"""Batched persistent launch of many smb_cross_prod_gauss_sft_par problems.

A :class:`Batch` holds a fixed list of independent problems (a registered
:class:`~sfemans_amd.nonlinear.Mesh`, the mode list, the nodal velocity
``V_in(np,6,m)`` and the Gauss output ``V_out(7me,6,m)``, all CUDA tensors)
and runs them as one kernel launch.  Each problem is what the reference
computes with one call of ``smb_cross_prod_gauss_sft_par``
(SOURCES/SIMULATION/subroutine_ns_with_velocity.F90:514).  ``autotune``
picks the schedule (policy and number of persistent work-groups) by timing
the candidates; ``capture`` records the launch in a hipGraph that ``run``
then replays.
"""
from __future__ import annotations

import ctypes as C

from ._lib import lib, check
from .fft_parallele import fdims, ptr, stream_ptr

__all__ = ["Batch", "LPT", "CONTIGUOUS", "ROUND_ROBIN", "BY_SIZE"]

LPT, CONTIGUOUS, ROUND_ROBIN, BY_SIZE = 0, 1, 2, 3


class _Problem(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("list_mode_dev", C.c_void_p), ("m_max_c", C.c_int32),
                ("V_in", C.c_void_p), ("V_out", C.c_void_p)]


class Batch:
    def __init__(self, problems):
        """problems: sequence of (mesh, list_mode, V_in, V_out)."""
        self._keep = []
        arr = (_Problem * len(problems))()
        for i, (mesh, list_mode, V_in, V_out) in enumerate(problems):
            M = len(list_mode)
            np_, nc, m = fdims(V_in)
            go, nco, mo = fdims(V_out)
            if (np_, nc, m) != (mesh.np, 6, M) or (go, nco, mo) != (mesh.n_gauss, 6, M):
                raise ValueError(f"problem {i}: expected V_in({mesh.np},6,{M}) and V_out({mesh.n_gauss},6,{M})")
            lm = mesh.list_mode_device(list_mode)
            self._keep += [mesh, lm, V_in, V_out]
            arr[i] = _Problem(mesh.handle.value, lm.data_ptr(), M, ptr(V_in), ptr(V_out))
        h = C.c_void_p()
        check(lib().sfx_batch_create(C.addressof(arr), len(problems), C.byref(h)))
        self._h = h

    def schedule(self, policy: int, n_wg: int):
        check(lib().sfx_batch_schedule(self._h, int(policy), int(n_wg)))

    def autotune(self, reps: int = 3):
        check(lib().sfx_batch_autotune(self._h, int(reps), stream_ptr()))

    def capture(self):
        check(lib().sfx_batch_capture(self._h))

    def run(self):
        check(lib().sfx_batch_run(self._h, stream_ptr()))

    def info(self) -> dict:
        pol, nwg, nt, cap = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
        ms = C.c_double()
        check(lib().sfx_batch_info(self._h, C.byref(pol), C.byref(nwg), C.byref(nt), C.byref(ms), C.byref(cap)))
        return {"policy": ["lpt", "contiguous", "round_robin", "by_size"][pol.value], "n_wg": nwg.value,
                "n_tiles": nt.value, "tuned_ms": ms.value, "captured": bool(cap.value)}

    def close(self):
        if getattr(self, "_h", None):
            lib().sfx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
