import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "libnegf.jl_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: production-size case (seconds to minutes)")


def relfro(a, b):
    """Relative Frobenius error ||a - b|| / ||b|| (Julia's norm(X, 2) on a
    matrix is the Frobenius norm, test/common_to_test.jl)."""
    import numpy as np
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def max_block_err(got: dict, ref: dict):
    return max(relfro(got[k], ref[k]) for k in ref)
