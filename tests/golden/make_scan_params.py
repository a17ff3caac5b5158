"""Writes scan_2d_params.json: the parameter table of the batched 2-D scan
(64 runs, seed 7) drawn by aladyn_amd.inputs.scan_2d_params.  The fixture
pins the draw: tests/test_scan.py checks the function still returns it."""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from aladyn_amd.inputs import scan_2d_params  # noqa: E402

if __name__ == "__main__":
    doc = {"generator": "numpy.random.Generator(PCG64(7)), draws per run in the order "
                        "nx, ny, log n/nc, a0, ppc[2]",
           "runs": scan_2d_params(7, 64)}
    (HERE / "scan_2d_params.json").write_text(json.dumps(doc, indent=1) + "\n")
