"""Build the in-tree native artefacts.

    python -m boxzyhydro_amd.build

* ``boxzyhydro_amd/libboxzy_hip.so`` -- HIP kernels + C-ABI, hipcc for gfx950
* ``boxzyhydro_amd/fortran/boxzy_driver`` -- Fortran driver over the C-ABI (amdflang)
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_hip(jobs: int = 8) -> str:
    """Compile every HIP source (always, also over an existing build: this is
    the does-it-build check) and link the library."""
    subprocess.run(["make", "-B", "-s", f"-j{jobs}", "-C", os.path.join(HERE, "csrc")], check=True)
    return os.path.join(HERE, "libboxzy_hip.so")


def build_fortran() -> bool:
    mk = os.path.join(HERE, "fortran", "Makefile")
    if not os.path.exists(mk) or shutil.which("amdflang") is None:
        return False
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "fortran")], check=True)
    return True


def main() -> None:
    print("built", build_hip())
    if build_fortran():
        print("built Fortran driver")


if __name__ == "__main__":
    main()
