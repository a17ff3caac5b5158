"""MI355X (gfx950) implementation of ALaDyn's laser-plasma PIC step.

The product is the HIP library behind the C-ABI of include/aladyn_gpu.h
(aladyn_amd/csrc); this package holds its build script, the ctypes mirror of
the reference's LP-step interface (LPStep) and the seeded input generators.
"""
from ._lib import PREC_F32, PREC_F64, AladynConfig, AladynError, load  # noqa: F401
from .config import LPConfig  # noqa: F401
from .lp_step import LPStep  # noqa: F401

__all__ = ["LPStep", "LPConfig", "AladynError", "PREC_F32", "PREC_F64", "load"]
