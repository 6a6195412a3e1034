"""Build step: compiles libscarl_hip.so for gfx950 (scarl_amd/csrc/Makefile) and the Fortran
drop-in modules over it (scarl_amd/fortran/Makefile, amdflang), in-tree."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> str:
    subprocess.run(["make", "-s", "-j", str(jobs), "-C", os.path.join(HERE, "csrc")], check=True)
    # Fortran drop-in modules (fftw_ScaRL, normal_ScaRL, constants_ScaRL) and their driver
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "fortran")], check=True)
    return os.path.join(HERE, "libscarl_hip.so")
