"""ctypes binding of the C-ABI declared in include/aladyn_gpu.h.

The shared library is built in-tree (aladyn_amd/libaladyn_gpu.so) by
aladyn_amd/build.py.  There is no fallback: if the library is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# ALADYN_LIB names another build of the same library (a variant under
# comparison); it must exist like the default one
LIB_PATH = Path(os.environ.get("ALADYN_LIB") or Path(__file__).resolve().parent / "libaladyn_gpu.so")

PREC_F64 = 64
PREC_F32 = 32
MAX_SPECIES = 4

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ESTATE", 5: "ERANGE"}


class AladynConfig(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("ibx", C.c_int32), ("iby", C.c_int32), ("ibz", C.c_int32),
        ("der_ord", C.c_int32),
        ("precision", C.c_int32),
        ("nsp", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dt", C.c_double),
        ("cfl", C.c_double),
        ("ompe", C.c_double),
        ("lorentz_fact", C.c_double * MAX_SPECIES),
        ("tile", C.c_int32 * 3),
        ("block", C.c_int32),
        ("sort_every", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


# every exported symbol with its argument types (all return int unless noted)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.c_int64
SIGNATURES = {
    "aladyn_create": [C.POINTER(AladynConfig), C.POINTER(_P)],
    "aladyn_destroy": [_P],
    "aladyn_last_error": [_P],
    "aladyn_put_species": [_P, C.c_int32, _D, _I64, _I64],
    "aladyn_species_count": [_P, C.c_int32, C.POINTER(_I64)],
    "aladyn_get_species": [_P, C.c_int32, _D, _I64, C.POINTER(_I64)],
    "aladyn_put_fields": [_P, _D],
    "aladyn_get_fields": [_P, _D],
    "aladyn_get_current": [_P, _D],
    "aladyn_lpf2_evolve": [_P, C.c_int32, C.c_int32],
    "aladyn_set_graphs": [_P, C.c_int32],
    "aladyn_set_dt": [_P, C.c_double],
    "aladyn_pfields_prepare": [_P],
    "aladyn_curr_clean": [_P],
    "aladyn_particle_stage": [_P, C.c_int32, C.c_int32],
    "aladyn_curr_mpi_collect": [_P],
    "aladyn_advance_lpf_fields": [_P],
    "aladyn_lp_window_xshift": [_P, C.c_int32, C.c_int32, C.c_double, C.POINTER(C.c_int32)],
    "aladyn_inject_species": [_P, C.c_int32, _D, _I64, _I64],
    "aladyn_box": [_P, _D],
    "aladyn_set_box_origin": [_P, C.c_double],
    "aladyn_set_injection": [_P, C.c_int32, _D, _I64, _D, C.c_int32, _D, C.c_int32, _D, _D, _I64,
                             C.c_double, C.c_double],
    "aladyn_injection_state": [_P, C.c_int32, C.POINTER(C.c_int64)],
    "aladyn_put_envelope": [_P, C.c_double, _D],
    "aladyn_get_envelope": [_P, _D],
    "aladyn_env_lpf2_evolve": [_P, C.c_int32],
    "aladyn_energy_momenta": [_P, C.c_int32, _D, _D],
    "aladyn_field_energy": [_P, _D],
    "aladyn_grid_charge": [_P, C.c_int32, _D],
    "aladyn_grid_charge_noalloc": [_P, C.c_int32, _D],
    "aladyn_beam_fields": [_P, C.c_int32, C.c_double, C.c_double],
    "aladyn_continuity_residual": [_P, _D],
    "aladyn_set_diagnostics": [_P, C.c_int32],
    "aladyn_synchronize": [_P],
    "aladyn_kernel_timing": [_P, C.c_int32, _D, _D, C.POINTER(_I64)],
    "aladyn_env_kernel_timing": [_P, C.c_int32, _D, _D, C.POINTER(_I64)],
    "aladyn_batch_create": [C.POINTER(_P), C.c_int32, C.c_int32, C.POINTER(_P)],
    "aladyn_batch_lpf2_evolve": [_P, C.c_int32, C.c_int32],
    "aladyn_batch_synchronize": [_P],
    "aladyn_batch_destroy": [_P],
    "aladyn_batch_last_error": [_P],
}

_lib = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Loads the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is missing: build the HIP library first (python -m aladyn_amd.build)")
    lib = C.CDLL(str(p))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_char_p if name.endswith("last_error") else C.c_int
    if path is None:
        _lib = lib
    return lib


class AladynError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{ERRORS.get(code, code)}] {message}")
        self.code = code
