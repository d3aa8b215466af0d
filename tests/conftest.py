import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "esrrf_golden.npz")
GOLDEN_CASES = ("m5_r15_s1", "m4_r20_s2", "m2_r10_s3_noiw", "m3_r25_s17", "m8_r12_s2", "m1_r15_s1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libnpx_hip.so")


def rel_err(x, ref):
    """(max, mean) of |x - ref| / max|ref| — the parity metric (1e-4 bar on the normalised output)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.abs(ref).max()
    if scale == 0:
        scale = 1.0
    e = np.abs(x - ref) / scale
    return float(e.max()), float(e.mean())


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import ref32
    ref32.build()
    return ref32
