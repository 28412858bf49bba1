"""Shared test set-up: the `gpu` marker and the import paths of the package and the oracle."""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "paulistrings-rs_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # The benchmark input is generated with torch's device generator.  torch's HIP runtime has to come up before the
    # library's in one process (brought up after it, torch reports no device), so a run that may select GPU tests
    # starts torch's device first; the library is loaded lazily, on its first call.
    if "not gpu" not in (config.getoption("markexpr", "") or ""):
        try:
            import torch

            if torch.cuda.is_available():
                torch.zeros(1, device="cuda")
        except ImportError:
            pass
