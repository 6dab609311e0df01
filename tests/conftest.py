import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "meep-torch_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU-only check")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture
def gold():
    return golden


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def record(case, **values):
    """Append one JSON line of measured errors to $MEEP_PARITY_REPORT (if
    set), so that errors a passing test computes -- e.g. against an FP64
    run -- are kept, not only printed."""
    import json
    print(f"[{case}] " + ", ".join(f"{k} {v:.3e}" if isinstance(v, float) else f"{k} {v}"
                                   for k, v in values.items()))
    path = os.environ.get("MEEP_PARITY_REPORT")
    if path:
        with open(path, "a") as fh:
            fh.write(json.dumps({"case": case, **values}) + "\n")
