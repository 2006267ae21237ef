import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"ref_{name}.npz")))


GOLDEN_CASES = ["g3", "gni3", "gnj3", "g2", "g4same", "gedge4"]


def uniform_cube(n, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    x = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    q = (2 * rng.random(n) - 1).astype(np.float32)
    return x, q


def rel_l2(a, b):
    """||a - b|| / ||b||; complex inputs keep both parts."""
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = np.asarray(a, dtype=np.complex128)
        b = np.asarray(b, dtype=np.complex128)
    else:
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
