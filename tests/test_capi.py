"""The C ABI library loads without a GPU, exports every symbol include/npx_esrrf.h declares, and its host-only
plan reproduces the reference's parameter derivation; argument errors come back as status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from nanopyx_amd import _hip
from oracle import fp64


def _declared():
    src = open(os.path.join(ROOT, "include", "npx_esrrf.h")).read()
    return sorted(set(re.findall(r"\b(npx_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    lib = _hip.lib()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n


def test_version_and_error_paths():
    lib = _hip.lib()
    assert lib.npx_version() >= 100
    rc = lib.npx_esrrf_frames(None, 0, 1, 8, 8, 4, 1.5, 1.0, 1, None, None)
    assert rc == -1 and b"NULL" in lib.npx_last_error()
    rc = lib.npx_esrrf_temporal(ctypes.c_void_p(16), 0, 1, 8, 8, 4, 1.5, 1.0, 1, 0, None, None, None, None, 0, None)
    assert rc == -1
    with pytest.raises(_hip.NpxError, match="radius"):
        _hip.plan_info(8, 8, 4, 5.0)  # window wider than the kernels are built for
    # the lattice entry points of the use_fht branch refuse bad arguments before touching a device
    p = ctypes.c_void_p(16)
    rc = lib.npx_rgc_lattice(None, p, p, 1, 8, 8, 2, 1.5, 1.0, 1, 0, p, None)
    assert rc == -1 and b"NULL" in lib.npx_last_error()
    rc = lib.npx_rgc_lattice(p, p, p, 1, 8, 8, 1, 1.5, 1.0, 1, 0, p, None)
    assert rc == -3 and b"M >= 2" in lib.npx_last_error()
    rc = lib.npx_fht_interpolate_lattice(p, 1, 8, 8, 2, 1, 0, 16, -1, p, p, ctypes.c_size_t(1 << 20), None)
    assert rc == -1 and b"stride" in lib.npx_last_error()
    # the per-frame tuner checks its stack and output before timing anything
    rc = lib.npx_autotune_frames(None, 0, 4, 8, 8, 2, 1.5, 1.0, 1, p, None, None)
    assert rc == -1 and b"NULL" in lib.npx_last_error()
    rc = lib.npx_autotune_frames(p, 0, 4, 8, 8, 2, 1.5, 1.0, 1, None, None, None)
    assert rc == -1 and b"output" in lib.npx_last_error()
    rc = lib.npx_autotune_frames(p, 0, 4, 8, 8, 0, 1.5, 1.0, 1, p, None, None)
    assert rc == -1 and b"magnification" in lib.npx_last_error()


@pytest.mark.parametrize("R,M", [(1.0, 2), (1.5, 5), (2.0, 4), (2.5, 8), (1.2345, 3)])
def test_plan_constants_match_reference_derivation(R, M):
    d = _hip.plan_info(64, 48, M, R)
    p = fp64.ref_params(R, M)
    assert d["K"] == -p["start"] == p["end"] - 1
    assert d["margin"] == p["margin"]
    assert np.float32(d["tSS"]) == np.float32(p["tSS"]) and np.float32(d["tSO"]) == np.float32(p["tSO"])
    assert (d["HM"], d["WM"]) == (64 * M, 48 * M)


def test_plan_ring_tables_against_fp64_geometry():
    H = W = 16
    for M, R in [(4, 2.0), (5, 1.5), (3, 2.5)]:
        t = _hip.plan_tables(H, W, M, R)
        K = t["info"]["K"]
        p = fp64.ref_params(R, M)
        for ph in range(M * M):
            py, px = divmod(ph, M)
            ey, ex = (2 * py + 1) // M, (2 * px + 1) // M
            for j in range(-K, K + 1):
                dy = (ey + j) / 2 - (py + 0.5) / M
                assert abs(t["dy"][ph, j + K] - dy) <= 1e-7
                for i in range(-K, K + 1):
                    dx = (ex + i) / 2 - (px + 0.5) / M
                    assert abs(t["dx"][ph, i + K] - dx) <= 1e-7
                    w, tt, dd = t["w"][ph, i + K, j + K], t["t"][ph, i + K, j + K], t["d"][ph, i + K, j + K]
                    d = np.hypot(dx, dy)
                    inring = d != 0 and d <= p["tSO"]
                    assert (w > 0) == inring and (tt > 0) == inring
                    assert abs(dd - d) <= 1e-6
                    if inring:
                        dw = (d * np.exp(-d * d / p["tSS"])) ** 4
                        assert abs(w - dw) <= 2e-5 * dw  # w is built from the float32 distance, as the reference does
                        assert abs(tt - dw / d) <= 2e-5 * dw / d


def test_temporal_static_argument_errors():
    """npx_esrrf_temporal_static checks its arguments and its scratch limit on the host, before any device call."""
    lib = _hip.lib()
    p = ctypes.c_void_p(16)
    rc = lib.npx_esrrf_temporal_static(None, 0, 4, 8, 8, 2, 1.5, 1.0, 1, 1, p, None, None, None)
    assert rc == -1 and b"NULL" in lib.npx_last_error()
    rc = lib.npx_esrrf_temporal_static(p, 0, 4, 8, 8, 2, 1.5, 1.0, 1, 0, p, None, None, None)
    assert rc == -1 and b"mode mask" in lib.npx_last_error()
    rc = lib.npx_esrrf_temporal_static(p, 0, 4, 8, 8, 2, 1.5, 1.0, 1, 4, p, None, None, None)  # TAC2 output NULL
    assert rc == -1
    rc = lib.npx_esrrf_temporal_static(p, 0, 4, 8, 8, 2, 5.0, 1.0, 1, 1, p, None, None, None)  # radius too wide
    assert rc == -3
    # tables of a 30000 x 30000 image at M 64 (taps of 1.92 M output rows and as many columns, ~79 MB) exceed the
    # 64 MiB scratch
    rc = lib.npx_esrrf_temporal_static(p, 0, 1, 30000, 30000, 64, 1.0, 1.0, 1, 1, p, None, None, None)
    assert rc == -1 and b"static scratch" in lib.npx_last_error()


def test_lattice_taps_are_catmull_rom_at_half_integers():
    t = _hip.plan_tables(9, 9, 3, 1.5)
    for a in range(1, 18):
        w = t["lrw"][a]
        if a % 2 == 0:       # integer coordinate a/2: [-1, 9, 9, -1]/16 minus taps outside the image
            full = np.array([-1, 9, 9, -1]) / 16
            base = a // 2 - 2
            exp = np.array([full[u] if 0 <= base + u < 9 else 0 for u in range(4)])
            assert t["lrb"][a] == base and np.allclose(w, exp)
        else:                # n + 1/2: the pixel itself
            assert t["lrb"][a] == (a - 1) // 2 - 1 and np.allclose(w, [0, 1, 0, 0])


def test_tune_table_host_only():
    """npx_tune_set / npx_tune_get and the launch shape npx_plan_info reports (no GPU calls)."""
    from nanopyx_amd import _hip
    H, W, M, R = 48, 80, 4, 2.0
    _hip.tune_set(H, W, M, R, 0, 0)
    mw, tb, tuned = _hip.tune_get(H, W, M, R)
    assert not tuned and mw == 16 and tb == 2048
    assert _hip.plan_info(H, W, M, R)["nwaves"] == 16
    _hip.tune_set(H, W, M, R, 4, 512)
    assert _hip.tune_get(H, W, M, R) == (4, 512, True)
    d = _hip.plan_info(H, W, M, R)
    assert d["nwaves"] == 4 and d["ngroups"] == 4
    assert _hip.tune_get(H, W, M, 2.5)[2] is False  # keyed by radius too
    _hip.tune_set(H, W, M, R, 0, 0)
    assert _hip.tune_get(H, W, M, R)[2] is False
    with pytest.raises(ValueError):
        _hip.tune_set(H, W, M, R, 17, 512)


@pytest.mark.parametrize("M,R,q_frames", [(4, 2.0, 16), (2, 1.0, 8), (3, 1.5, 8), (5, 1.5, 8), (8, 1.0, 8)])
def test_frames_rows_per_lane_by_block_shape(M, R, q_frames):
    """The per-frame and arrays kernels take 16 rows per lane only when one block of 16 waves covers every phase of a
    region (M = 4 at the default shape); smaller or several blocks per region take 8 (their smaller LDS tile lets more
    blocks share a CU). The temporal kernel always takes 8."""
    d = _hip.plan_info(40, 72, M, R)
    assert d["Q_frames"] == q_frames and d["Q_temporal"] == 8
    assert (d["nwaves"] == 16 and d["ngroups"] == 1) == (q_frames == 16)


def test_frames_launch_shape_apart_from_temporal():
    """The per-frame kernels take their own stored launch shape where there is one, else the temporal kernel's; the
    plan's per-frame rows per lane follow it."""
    H, W, M, R = 40, 72, 4, 2.0
    try:
        assert _hip.tune_frames_get(H, W, M, R)[2] is False
        _hip.tune_set(H, W, M, R, 8, 512)
        assert _hip.tune_frames_get(H, W, M, R) == (8, 512, False)  # the temporal kernel's shape
        assert _hip.plan_info(H, W, M, R)["Q_frames"] == 8          # 2 blocks of 8 waves per region
        _hip.tune_frames_set(H, W, M, R, 16, 1024)
        assert _hip.tune_frames_get(H, W, M, R) == (16, 1024, True)
        assert _hip.tune_get(H, W, M, R)[:2] == (8, 512)
        assert _hip.plan_info(H, W, M, R)["Q_frames"] == 16         # one block of 16 waves covers the 16 phases
        with pytest.raises(ValueError):
            _hip.tune_frames_set(H, W, M, R, 17, 512)
    finally:
        _hip.tune_set(H, W, M, R, 0, 0)
        _hip.tune_frames_set(H, W, M, R, 0, 0)


def test_packaged_store_entries_load():
    """The packaged launch-shape store (bench.py loads it) is well-formed: every key is device|H|W|M|R with an optional
    |frames for the per-frame kernels, every shape one the autotuner can produce, and each entry is accepted by the
    library's table for its kind (read back with the matching getter)."""
    import json
    from nanopyx_amd.tuning import PACKAGED
    with open(PACKAGED) as f:
        data = json.load(f)
    assert data
    for key, v in data.items():
        parts = key.split("|")
        assert len(parts) in (5, 6) and parts[0] == "gfx950/256", key
        frames = len(parts) == 6
        assert not frames or parts[5] == "frames", key
        H, W, M, R = int(parts[1]), int(parts[2]), int(parts[3]), float(parts[4])
        mw, tb = int(v["max_waves"]), int(v["target_blocks"])
        assert mw in (4, 8, 16) and tb in (256, 512, 1024, 2048, 4096, 8192), key
        setter, getter = (_hip.tune_frames_set, _hip.tune_frames_get) if frames else (_hip.tune_set, _hip.tune_get)
        try:
            setter(H, W, M, R, mw, tb)
            assert getter(H, W, M, R) == (mw, tb, True), key
        finally:
            setter(H, W, M, R, 0, 0)


def test_temporal_chunks_keep_six_frames():
    """The temporal kernel splits the frames into chunks for about target_blocks blocks, each of at least 6 frames:
    the workspace (5 partial maps per chunk) gives the chunk count without a GPU."""
    H, W, M, R = 64, 64, 5, 1.5
    P = H * M * W * M
    _hip.tune_set(H, W, M, R, 16, 2048)
    try:
        # 1 x 8 regions x 2 phase groups = 16 blocks per chunk, 2048 / 16 = 128 chunks wanted, at most N // 6 (then
        # rounded to whole frames per chunk: 100 frames in 16 chunks of 7 are 15 chunks)
        for N, want in ((100, 15), (16, 2), (11, 0), (1000, 125), (2000, 125)):  # 1000: 8 frames each; 2000: 16
            ws = _hip.temporal_workspace_bytes(N, H, W, M, R)
            nchunk = ws // (5 * P * 4)
            assert nchunk == want, (N, nchunk)
            if nchunk:
                assert -(-N // nchunk) >= 6 or nchunk == 1
    finally:
        _hip.tune_set(H, W, M, R, 0, 0)
