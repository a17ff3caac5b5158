"""Compiles fortran_formats.f90 with the image's Fortran compiler (amdflang)
and runs it here, writing fortran_formats.txt: the formatted header records
of the reference's output files for fixed values.  tests/test_outputs.py
checks the Python writers against it."""
import shutil
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent

if __name__ == "__main__":
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    with tempfile.TemporaryDirectory() as t:
        exe = Path(t) / "ff"
        subprocess.run([fc, "-O0", str(HERE / "fortran_formats.f90"), "-o", str(exe)], check=True,
                       cwd=t)
        subprocess.run([str(exe)], check=True, cwd=t)
        shutil.copy(Path(t) / "fortran_formats.txt", HERE / "fortran_formats.txt")
