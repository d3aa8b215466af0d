"""Temporal AVG / VAR / TAC2 at full stack length against the reference's own reduction.

The fused kernel folds every frame into one-pass moments in float32 (Welford-type mean and M2, a re-centred lag-1
co-moment, and Chan-type merges of frame chunks and served slices). The reference reduces the (N, HM, WM) per-frame
maps with NumPy's float32 mean, two-pass var and calculate_tac2
(src/nanopyx/core/transform/sr_temporal_correlations.py:64-70,205-221). Here the per-frame maps (npx_esrrf_frames,
pinned per frame to the oracle) are reduced over every frame by the oracle's NumPy-exact reduction (ora_temporal)
and compared with the fused maps, so only the drift of the streamed moments over hundreds of frames is measured:
500 frames at 256 x 256 M5 (three modes), 1000 frames at 512 x 512 M4 (a strip of the output), and 1000 frames
through every way the moments are streamed (one pass, frame chunks, served slices, the static-scratch entry point).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    assert _torch.cuda.is_available(), "GPU tests need a ROCm device"
    return _torch


def _report(tag, x, ref):
    mx, mean = rel_err(x, ref)
    print(f"{tag}: max {mx:.3e} mean {mean:.3e}", flush=True)
    return mx


def test_plain_256_all_500_frames(torch, oracle_built):
    """500 x 256 x 256 float32, M5 R1.5 S2: fused AVG / VAR / TAC2 (one pass and frame chunks) against NumPy's
    reduction of all 500 per-frame maps."""
    from nanopyx_amd import _hip, synthetic
    sc = synthetic.CONFIGS["plain_256"]
    x = torch.from_numpy(synthetic.generate(sc)).cuda()
    print("plain_256", tuple(x.shape), flush=True)
    modes = ("AVG", "VAR", "TAC2")
    chunked = _hip.esrrf_temporal(x, 5, 1.5, 2, True, modes)
    direct = _hip.esrrf_temporal(x, 5, 1.5, 2, True, modes, workspace=False)
    frames = _hip.esrrf_frames(x, 5, 1.5, 2, True).cpu().numpy()
    print("frames", frames.shape, flush=True)
    for m in modes:
        ref = oracle_built.temporal(frames, m)
        assert _report(f"plain_256[:500] {m} one pass vs NumPy reduction", direct[m].cpu().numpy(), ref) <= TOL
        assert _report(f"plain_256[:500] {m} chunked vs NumPy reduction", chunked[m].cpu().numpy(), ref) <= TOL


def test_large_u16_all_1000_frames_strip(torch, oracle_built):
    """1000 x 512 x 512 uint16, M4 R2 S2: fused AVG and TAC2 of the whole image against NumPy's reduction of all 1000
    per-frame maps on a strip of 256 output rows (64 input rows) across the full width."""
    from nanopyx_amd import _hip, synthetic
    sc = synthetic.CONFIGS["large_512_u16"]
    x = torch.from_numpy(synthetic.generate(sc)).cuda()
    print("large_512_u16", tuple(x.shape), flush=True)
    M, r0, r1 = 4, 896, 1152
    fused = _hip.esrrf_temporal(x, M, 2.0, 2.0, True, ("AVG", "TAC2"))
    strip = np.empty((x.shape[0], r1 - r0, x.shape[2] * M), np.float32)
    for f0 in range(0, x.shape[0], 100):
        fr = _hip.esrrf_frames(x[f0:f0 + 100], M, 2.0, 2.0, True)
        strip[f0:f0 + 100] = fr[:, r0:r1].cpu().numpy()
        del fr
    print("strip", strip.shape, flush=True)
    for m in ("AVG", "TAC2"):
        ref = oracle_built.temporal(strip, m)
        assert _report(f"large_512_u16[:1000] {m} rows {r0}-{r1} vs NumPy reduction",
                       fused[m][r0:r1].cpu().numpy(), ref) <= TOL


def test_long_stack_every_streaming_path(torch, oracle_built):
    """1000 x 32 x 48 float32, M2 R1.5 S2: the oracle's per-frame maps reduced by NumPy against the fused moments
    streamed in one pass, in frame chunks, in served slices resumed from the resident state (host and device
    stacks), and through the static-scratch entry point."""
    from nanopyx_amd import _hip, synthetic
    N, H, W, M, R, S = 1000, 32, 48, 2, 1.5, 2
    sc = synthetic.Scene(N, H, W, 40, 1.5, 600, 0.2, 100, 0.0, 2.0, 0.0, "float32", 21)
    xh = synthetic.generate(sc)
    x = torch.from_numpy(xh).cuda()
    maps = oracle_built.esrrf(xh, M, R, S)
    modes = ("AVG", "VAR", "TAC2")
    ref = {m: oracle_built.temporal(maps, m) for m in modes}
    got = {"one pass": _hip.esrrf_temporal(x, M, R, S, True, modes, workspace=False),
           "chunks": _hip.esrrf_temporal(x, M, R, S, True, modes)}
    nb = _hip.temporal_workspace_bytes(N, H, W, M, R)
    print(f"chunked workspace {nb} bytes ({nb // (5 * H * M * W * M * 4)} chunks)", flush=True)
    outs = {m: torch.empty((H * M, W * M), device="cuda") for m in modes}
    _hip._check(_hip.lib().npx_esrrf_temporal_static(x.data_ptr(), 0, N, H, W, M, float(R), float(S), 1, 7,
                                                      outs["AVG"].data_ptr(), outs["VAR"].data_ptr(),
                                                      outs["TAC2"].data_ptr(),
                                                      int(torch.cuda.current_stream().cuda_stream)))
    got["static scratch"] = outs
    hstack = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    hstack.copy_(x)
    torch.cuda.synchronize()
    for label, src in (("served slices, device stack", x), ("served slices, host stack", hstack)):
        budget = H * M * W * M * 100  # about 100 frames per step: slices of whole frame chunks
        print(label, _hip.Session.slice_plan(N, H, W, M, R, False, src is hstack, budget), flush=True)
        sess = _hip.Session(1 << 30, 4, budget)
        o = {m: torch.empty((H * M, W * M), device="cuda") for m in modes}
        jid = sess.submit(src, M, R, S, True, modes, o)
        for _ in range(100000):
            sess.step()
            if sess.poll(jid) == _hip.Session.DONE:
                break
        print(label, sess.stats(), flush=True)
        sess.close()
        got[label] = o
    torch.cuda.synchronize()
    for label, o in got.items():
        for m in modes:
            assert _report(f"long[{N}] {m} {label} vs NumPy reduction", o[m].cpu().numpy(), ref[m]) <= TOL
