"""The C-ABI library loads and exports every symbol include/aladyn_gpu.h
declares (CPU only: no entry point that touches a device is called)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "aladyn_gpu.h"


def declared_symbols():
    text = HEADER.read_text()
    names = re.findall(r"^(?:int|const char \*)\s*(aladyn_\w+)\s*\(", text, flags=re.M)
    assert names, "no declarations found"
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_symbols()
    for required in ("aladyn_create", "aladyn_lpf2_evolve", "aladyn_particle_stage",
                     "aladyn_advance_lpf_fields", "aladyn_put_species", "aladyn_get_species"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from aladyn_amd._lib import LIB_PATH, SIGNATURES

    if not LIB_PATH.exists():
        pytest.fail(f"{LIB_PATH} not built (run python -m aladyn_amd.build)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert name in SIGNATURES, f"{name} missing from the ctypes binding"


def test_config_struct_matches_header_size():
    """sizeof(aladyn_config) in C equals the ctypes mirror."""
    from aladyn_amd._lib import AladynConfig

    # 10 int32, 6+4 doubles (dx,dy,dz,dt,cfl,ompe + lorentz_fact[4]),
    # tile[3], block, sort_every, reserved[5]
    expected = 10 * 4 + 10 * 8 + 10 * 4
    assert ctypes.sizeof(AladynConfig) == expected
