"""Build ``libsfx.so`` in-tree for gfx950 (``python -m sfemans_amd.build``).

Runs ``make`` in ``sfemans_amd/csrc`` (hipcc, --offload-arch=gfx950).  The
Fortran ISO_C_BINDING shim is compiled separately by ``build_fortran`` when
``amdflang`` is available.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "libsfx.so")


def build(jobs: int = 8, force: bool = True) -> str:
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
    env = dict(os.environ)
    subprocess.run(["make", "-C", CSRC, f"-j{jobs}"], check=True, env=env)
    if not os.path.exists(LIB):
        raise RuntimeError("libsfx.so was not produced")
    return LIB


def build_fortran() -> str | None:
    """Compile the Fortran shim module and its test driver (needs amdflang)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None or not os.path.isdir(FSRC):
        return None
    subprocess.run(["make", "-C", FSRC, f"FC={fc}"], check=True)
    return os.path.join(FSRC, "sfx_fortran_driver")


if __name__ == "__main__":
    print(build())
    if "--fortran" in sys.argv:
        print(build_fortran())
