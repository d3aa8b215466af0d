"""Host-only checks of the serving session's job layout and slicing (npx_session_slice_plan): no device calls.

A slice that runs as frame chunks writes one partial-moment slot per chunk; the job's pool block holds cmax slots, so
no slice may run more chunks than that, for any job the served stream can submit."""
import itertools

import numpy as np
import pytest

from nanopyx_amd import _hip

SIZES = (64, 128, 256, 384, 512, 1024)
MAGS = (2, 3, 4, 5, 8)


def _check(N, H, W, M, R, u16, host, budget):
    p = _hip.Session.slice_plan(N, H, W, M, R, u16, host, budget)
    assert 1 <= p["max_chunks"] <= p["cmax"], (N, H, W, M, R, u16, host, budget, p)
    assert p["off_boundary"] == 0, (N, H, W, M, R, u16, host, budget, p)
    assert p["slices"] >= 1
    return p


def test_host_stack_slices_never_exceed_partial_slots():
    """The shapes where the upload piece cuts the frames per chunk below N / cmax (large frames from host memory)
    used to run slices of more chunks than the job's partial slots: 1000 x 512 x 512 at M2 float32 ran 16 chunks
    against 2 slots."""
    p = _check(1000, 512, 512, 2, 1.5, False, True, 1 << 29)
    assert p["pieces"] > 1 and p["cmax"] > 1
    _check(200, 512, 512, 3, 1.5, False, True, 1 << 29)
    _check(300, 384, 384, 2, 1.0, False, True, 1 << 29)


@pytest.mark.parametrize("budget", [1 << 29, 1 << 26, 1 << 31])
def test_slicing_over_served_shape_grid(budget):
    """Every (H, W, M) of the served stream's distribution at a spread of frame counts, radii, dtypes and sources."""
    rng = np.random.default_rng(5)
    for H, W, M in itertools.product(SIZES, SIZES, MAGS):
        N = int(np.exp(rng.uniform(np.log(50), np.log(1000))))
        R = float(rng.uniform(1.0, 2.5))
        u16 = bool(rng.integers(2))
        for host in (False, True):
            _check(N, H, W, M, R, u16, host, budget)


@pytest.mark.parametrize("N", [1, 7, 8, 9, 63, 64, 65, 999, 1000])
def test_slicing_frame_count_edges(N):
    for host in (False, True):
        _check(N, 512, 512, 2, 1.5, False, host, 1 << 29)
        _check(N, 64, 64, 8, 2.5, True, host, 1 << 29)
