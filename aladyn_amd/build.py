"""Builds the HIP library libaladyn_gpu.so for gfx950 in-tree.

Run ``python -m aladyn_amd.build`` (or ``__graft_entry__.build()``).  The
library is a plain C-ABI shared object (include/aladyn_gpu.h); it does not
link against PyTorch.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libaladyn_gpu.so"
SOURCES = ["capi.hip", "fields.hip", "tile_kernel.hip", "tile_store.hip", "reductions.hip",
           "envelope.hip", "beam.hip"]
ARCH = os.environ.get("ALADYN_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Wall",
        "-Wno-unused-result",
        # packed FP32 ops have produced silent wrong results on this hardware
        # pool; keep the compiler from emitting them
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        "-I", str(ROOT / "include"),
    ]


def build(verbose: bool = False, jobs: int = 4, csrc: Path | None = None,
          out: Path | None = None) -> Path:
    """Compiles every source to an object (in parallel) and links the library
    (by default the package's sources into aladyn_amd/libaladyn_gpu.so)."""
    cc = hipcc()
    src_dir = Path(csrc) if csrc else CSRC
    lib = Path(out) if out else LIB
    objdir = lib.parent / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        objs.append(obj)
        cmd = [cc, *flags(), "-c", str(src_dir / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join(f"--- {s}\n{t}" for s, t in failed)
        raise RuntimeError(f"HIP compilation failed:\n{msg}")
    tmp = lib.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
