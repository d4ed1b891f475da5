import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected on a host without a GPU")
    return torch.device("cuda:0")


def pytest_terminal_summary(terminalreporter):
    """Largest error against the FP64 oracle seen by each parity check."""
    try:
        import _parity  # the same module object the test files import (tests/ is on sys.path)
    except Exception:
        return
    if not _parity.RECORD:
        return
    tr = terminalreporter
    tr.write_sep("-", "max error vs FP64 oracle (per check)")
    for name in sorted(_parity.RECORD):
        v, b = _parity.RECORD[name]
        tr.write_line(f"{name:60s} {v:10.3e}  (bound {b:.1e})")
    path = os.environ.get("WGBM_PARITY_REPORT")
    if path:
        import json
        with open(path, "w") as f:
            json.dump({k: {"max_error": v, "bound": b} for k, (v, b) in sorted(_parity.RECORD.items())}, f, indent=1)
