"""Builds tests/fortran/_build/drive_lp: the Fortran caller of libaladyn_gpu.so
through the bind(C) module aladyn_amd/fortran/aladyn_gpu.f90."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
MODULE = ROOT / "aladyn_amd" / "fortran" / "aladyn_gpu.f90"
DRIVER = HERE / "drive_lp.f90"
OUT = HERE / "_build"
EXE = OUT / "drive_lp"


def fortran_compiler() -> str | None:
    for cand in ("/opt/rocm/llvm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if cand and Path(cand).exists():
            return cand
    return None


def build(verbose: bool = False) -> Path:
    fc = fortran_compiler()
    if fc is None:
        raise RuntimeError("no Fortran compiler (amdflang) found")
    OUT.mkdir(exist_ok=True)
    lib_dir = ROOT / "aladyn_amd"
    cmds = [
        [fc, "-O2", "-c", str(MODULE), "-J", str(OUT), "-o", str(OUT / "aladyn_gpu.o")],
        [fc, "-O2", "-c", str(DRIVER), "-I", str(OUT), "-J", str(OUT), "-o", str(OUT / "drive_lp.o")],
        [fc, "-O2", str(OUT / "drive_lp.o"), str(OUT / "aladyn_gpu.o"), "-o", str(EXE),
         f"-L{lib_dir}", "-laladyn_gpu", "-Wl,-rpath,$ORIGIN/../../../aladyn_amd"],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
