"""The serving session (npx_session_*): every job's maps are the same alone and in a mix (continuous batching with
frame slices resumed from resident moments), equal the plain fused call, match the oracle on small jobs, and jobs
wait for pool room instead of failing."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4
# Served jobs run their frames in slices and chunks of their own, so their moments are merged at other frames than a
# plain call's: the merged float32 moments then differ by rounding alone, bounded as for the temporal kernel's chunked
# against single-pass comparison (TAC2, a difference of near-equal products, moves most). Where the frame chunks are
# the same the tests ask for equality bit for bit.
CHUNK_TOL = 1e-5


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    assert _torch.cuda.is_available(), "GPU tests need a ROCm device"
    return _torch


def _report(tag, x, ref):
    mx, mean = rel_err(x, ref)
    print(f"{tag}: max {mx:.3e} mean {mean:.3e}", flush=True)
    return mx


# (N, H, W, M, R, S, mode, dtype, host)
JOBS = [
    (40, 24, 70, 2, 1.0, 1, "AVG", "float32", True),
    (33, 17, 20, 3, 2.5, 2, "VAR", "uint16", False),
    (57, 40, 64, 5, 1.7, 3, "TAC2", "float32", True),
    (9, 12, 9, 8, 1.2, 2, "VAR", "float32", False),
    (64, 32, 96, 4, 2.0, 2, "TAC2", "uint16", True),
    (21, 70, 30, 2, 2.3, 1, "AVG", "uint16", False),
]


def _requests(torch):
    from nanopyx_amd import serving, synthetic
    reqs = []
    for i, (N, H, W, M, R, S, mode, dtype, host) in enumerate(JOBS):
        sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0,
                             100.0 if dtype == "uint16" else 0.0, dtype, 40 + i)
        x = synthetic.generate_device(sc)
        if host:
            h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
            h.copy_(x)
            x = h
        reqs.append(serving.Request(x, M, R, S, mode, True))
    torch.cuda.synchronize()
    return reqs


def _run(torch, reqs, idx, px_frames_per_step, pool_bytes=8 << 30, max_active=32):
    """Submits the requests `idx` at once and steps the session until all are done; returns their maps."""
    from nanopyx_amd import _hip
    sess = _hip.Session(pool_bytes, max_active, px_frames_per_step)
    outs, ids = {}, {}
    for i in idx:
        r = reqs[i]
        N, H, W = r.stack.shape
        outs[i] = {r.mode: torch.full((H * r.magnification, W * r.magnification), -7.0, device="cuda")}
        ids[i] = sess.submit(r.stack, r.magnification, r.radius, r.sensitivity, r.iw, (r.mode,), outs[i])
    saw_queue = False
    for _ in range(100000):
        st = sess.stats()
        saw_queue = saw_queue or st["queued"] > 0 and st["active"] > 0
        sess.step()
        if all(sess.poll(j) == _hip.Session.DONE for j in ids.values()):
            break
    else:
        raise AssertionError("session did not finish")
    torch.cuda.synchronize()
    stats = sess.stats()
    sess.close()
    return {i: outs[i][reqs[i].mode].cpu().numpy() for i in idx}, stats, saw_queue


def test_alone_equals_mixed(torch):
    reqs = _requests(torch)
    idx = list(range(len(reqs)))
    mixed, st, _ = _run(torch, reqs, idx, px_frames_per_step=1 << 16)  # small steps: several slices per job
    print("mixed session stats", st, flush=True)
    assert st["slices"] > len(reqs) and st["launches"] < st["slices"]  # jobs shared launches, sliced
    for i in idx:
        alone, _, _ = _run(torch, reqs, [i], px_frames_per_step=1 << 40)  # one slice
        assert np.array_equal(alone[i], mixed[i]), f"job {i} differs alone vs mixed"
        print(f"job {i}: alone == mixed", flush=True)


def test_served_equals_plain_call(torch):
    from nanopyx_amd import _hip
    reqs = _requests(torch)
    served, _, _ = _run(torch, reqs, list(range(len(reqs))), px_frames_per_step=1 << 18)
    for i, r in enumerate(reqs):
        plain = _hip.esrrf_temporal(r.stack.cuda(), r.magnification, r.radius, r.sensitivity, True, (r.mode,),
                                    workspace=False)[r.mode].cpu().numpy()
        assert _report(f"job {i} served vs plain", served[i], plain) <= CHUNK_TOL


def test_served_matches_oracle(torch, oracle_built):
    reqs = _requests(torch)
    idx = [1, 3, 5]
    served, _, _ = _run(torch, reqs, idx, px_frames_per_step=1 << 15)
    for i in idx:
        r = reqs[i]
        x = r.stack.cpu().numpy().astype(np.float32)
        ref = oracle_built.temporal(oracle_built.esrrf(x, r.magnification, r.radius, r.sensitivity, True), r.mode)
        assert _report(f"job {i} ({r.mode}) served vs oracle", served[i], ref) <= TOL


def test_admission_waits_for_pool(torch):
    """A pool that holds about one job at a time: the rest queue, then run as room frees up."""
    reqs = _requests(torch)
    idx = [0, 2, 4]
    from nanopyx_amd import _hip
    need = []
    for i in idx:
        N, H, W, M, R, S, mode, dtype, host = JOBS[i]
        need.append(_hip.Session.job_bytes(N, H, W, M, R, dtype == "uint16", host))
    pool = max(need) + 4096
    assert sum(need) > pool
    full, _, _ = _run(torch, reqs, idx, px_frames_per_step=1 << 40)
    tight, st, saw_queue = _run(torch, reqs, idx, px_frames_per_step=1 << 40, pool_bytes=pool)
    assert saw_queue
    for i in idx:
        assert np.array_equal(full[i], tight[i])


def test_session_errors(torch):
    from nanopyx_amd import _hip
    sess = _hip.Session(1 << 20, 4, 1 << 20)
    x = torch.zeros((10, 256, 256), device="cuda")
    out = {"AVG": torch.empty((1024, 1024), device="cuda")}
    with pytest.raises(ValueError):
        sess.submit(x, 4, 1.5, 1, True, ("AVG",), out)  # state alone exceeds the 1 MiB pool
    with pytest.raises(ValueError):
        sess.submit(x, 4, 1.5, 1, True, ("MEDIAN",), out)
    sess.close()
    with pytest.raises(ValueError):
        _hip.Session(0, 4, 1)


def test_serve_loop_small(torch):
    """The request loop end to end on a short seeded stream (every request completes, latencies recorded)."""
    from nanopyx_amd import serving
    reqs = serving.make_requests(6, seed=11, host=True, max_frames=60)
    summary, lat = serving.serve(reqs, rate_hz=200.0, seed=11)
    print("serve summary", summary, flush=True)
    assert summary["requests"] == 6 and np.all(lat > 0)
    for r in reqs:
        plain = __import__("nanopyx_amd")._hip.esrrf_temporal(r.stack.cuda(), r.magnification, r.radius,
                                                                r.sensitivity, True, (r.mode,), workspace=False)
        assert rel_err(r.outputs[r.mode].cpu().numpy(), plain[r.mode].cpu().numpy())[0] <= CHUNK_TOL


@pytest.mark.parametrize("N,H,W,M,R,mode,steps", [
    (300, 256, 256, 2, 1.5, "TAC2", 1 << 40),   # 78 MB host stack: three upload pieces, chunked slices
    (150, 200, 120, 3, 2.0, "VAR", 1 << 18),    # 14 MB in one piece, many small slices
    (260, 256, 300, 1, 1.0, "AVG", 1 << 20),    # 80 MB, several pieces and slices
])
def test_host_stack_in_pieces(torch, N, H, W, M, R, mode, steps):
    """A host stack is uploaded in pieces while its first slices run: the maps equal those of the same job with the
    stack already on the device, bit for bit, and the plain fused call (<= CHUNK_TOL)."""
    from nanopyx_amd import _hip, serving, synthetic
    print(f"host pieces N{N} {H}x{W} M{M} R{R} {mode}", flush=True)
    sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 77)
    x = synthetic.generate_device(sc)
    h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    h.copy_(x)
    torch.cuda.synchronize()
    reqs = [serving.Request(h, M, R, 1, mode, True), serving.Request(x, M, R, 1, mode, True)]
    got, st, _ = _run(torch, reqs, [0, 1], px_frames_per_step=steps)
    print("session stats", st, flush=True)
    assert np.array_equal(got[0], got[1]), "host-source maps differ from device-source maps"
    plain = _hip.esrrf_temporal(x, M, R, 1, True, (mode,), workspace=False)[mode].cpu().numpy()
    assert _report(f"host pieces {mode} vs plain", got[0], plain) <= CHUNK_TOL


@pytest.mark.parametrize("N,H,W,M,mode", [
    (300, 384, 384, 2, "TAC2"),   # 177 MB host stack in 6 pieces; 4 partial slots of 56 frames
    (200, 512, 512, 3, "AVG"),    # 210 MB; the whole stack arrives before the first slice
])
def test_host_stack_stepped_after_upload(torch, N, H, W, M, mode):
    """A host stack whose upload has fully landed before the session's next step: the step may take every frame at
    once, and its slice must still run no more frame chunks than the job has partial slots (a longer slice would
    write partial moments past its slots, into the job's own stack). Maps equal the device-source job and
    the plain fused call (<= CHUNK_TOL); bit for bit when both have the same frame chunks."""
    from nanopyx_amd import _hip, synthetic
    print(f"host stack stepped after upload N{N} {H}x{W} M{M} {mode}", flush=True)
    plan = _hip.Session.slice_plan(N, H, W, M, 1.5, False, True)
    assert plan["max_chunks"] <= plan["cmax"]
    sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 78)
    x = synthetic.generate_device(sc)
    h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    h.copy_(x)
    torch.cuda.synchronize()
    maps = []
    for src in (h, x):
        sess = _hip.Session(2 << 30, 4, 1 << 29)
        out = {mode: torch.full((H * M, W * M), -7.0, device="cuda")}
        jid = sess.submit(src, M, 1.5, 1, True, (mode,), out)
        sess.step()                 # admits the job and queues its upload
        torch.cuda.synchronize()    # every piece has landed before the next step
        for _ in range(10000):
            sess.step()
            if sess.poll(jid) == _hip.Session.DONE:
                break
        else:
            raise AssertionError("session did not finish")
        torch.cuda.synchronize()
        print("session stats", sess.stats(), flush=True)
        sess.close()
        maps.append(out[mode].cpu().numpy())
    # the upload pieces may shorten a host job's chunks (different merge points, so different float32 rounding of the
    # merged moments); with the same chunks the maps are the same bit for bit
    dev_plan = _hip.Session.slice_plan(N, H, W, M, 1.5, False, False)
    if dev_plan["cfpc"] == plan["cfpc"]:
        assert np.array_equal(maps[0], maps[1]), "host-source maps differ from device-source maps"
    assert _report(f"host after upload {mode} vs device source", maps[0], maps[1]) <= CHUNK_TOL
    plain = _hip.esrrf_temporal(x, M, 1.5, 1, True, (mode,), workspace=False)[mode].cpu().numpy()
    assert _report(f"host after upload {mode} vs plain", maps[0], plain) <= CHUNK_TOL


def test_fed_job_errors(torch):
    from nanopyx_amd import _hip
    x = torch.zeros((16, 8, 8), device="cuda")
    sess = _hip.Session(64 << 20, 4, 1 << 20)
    out = {"AVG": torch.empty((16, 16), device="cuda")}
    with pytest.raises(ValueError):
        sess.submit_fed(x, 2, 1.5, 1, True, ("AVG",), out, 0)
    jid = sess.submit_fed(x, 2, 1.5, 1, True, ("AVG",), out, 8)
    sess.feed(jid, 8)
    with pytest.raises(ValueError):
        sess.feed(jid, 8)          # counts must grow
    with pytest.raises(ValueError):
        sess.feed(jid, 17)         # beyond the stack
    plain = sess.submit(x, 2, 1.5, 1, True, ("AVG",), {"AVG": torch.empty((16, 16), device="cuda")})
    with pytest.raises(ValueError):
        sess.feed(plain, 4)        # not a fed job
    sess.feed(jid, 16)
    for _ in range(10000):
        sess.step()
        if sess.poll(jid) == _hip.Session.DONE and sess.poll(plain) == _hip.Session.DONE:
            break
    torch.cuda.synchronize()
    sess.close()


@pytest.mark.parametrize("N,H,W,M,R,S,fpt,mode,dtype,mpc", [
    (300, 256, 256, 2, 1.5, 2, 100, "AVG", "float32", True),    # 79 MB: three timepoints, 128-frame pieces
    (260, 384, 384, 3, 2.0, 1, 0, "TAC2", "uint16", False),     # 77 MB uint16, one timepoint
    (210, 256, 512, 4, 1.0, 3, 70, "VAR", "float32", True),     # 110 MB, pieces straddle timepoints
])
def test_workflow_streams_host_stack(torch, N, H, W, M, R, S, fpt, mode, dtype, mpc):
    """eSRRF() on a host stack of 64 MiB or more streams it (pinned pieces, fed session jobs, one per timepoint): the
    result equals eSRRF() on the same stack already on the device (<= CHUNK_TOL: the frame chunks may differ)."""
    import nanopyx_amd
    from nanopyx_amd import synthetic
    print(f"streamed workflow N{N} {H}x{W} M{M} fpt {fpt} {mode} {dtype} mpc {mpc}", flush=True)
    sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0,
                         100.0 if dtype == "uint16" else 0.0, dtype, 79)
    x = synthetic.generate(sc)
    assert x.nbytes >= 64 << 20
    got = nanopyx_amd.eSRRF(x, magnification=M, radius=R, sensitivity=S, frames_per_timepoint=fpt,
                            temporal_correlation=mode, macro_pixel_correction=mpc)
    ref = nanopyx_amd.eSRRF(torch.from_numpy(x).cuda(), magnification=M, radius=R, sensitivity=S,
                            frames_per_timepoint=fpt, temporal_correlation=mode,
                            macro_pixel_correction=mpc).cpu().numpy()
    assert got.shape == ref.shape and got.dtype == np.float32
    assert _report(f"streamed workflow {mode} vs device stack", got, ref) <= CHUNK_TOL


def test_workflow_stream_error_leaves_it_usable(torch):
    """A streamed eSRRF() call that fails on the host (an unknown temporal mode, refused when its first job is
    queued, with the reference's ValueError) drops the shared session; the next streamed call makes a new one and
    equals eSRRF() on the same stack on the device."""
    import nanopyx_amd
    from nanopyx_amd import synthetic
    sc = synthetic.Scene(130, 256, 512, 2600, 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 81)
    x = synthetic.generate(sc)
    assert x.nbytes >= 64 << 20
    with pytest.raises(ValueError, match="Type of correlation"):
        nanopyx_amd.eSRRF(x, magnification=2, radius=1.5, temporal_correlation="XYZ")
    got = nanopyx_amd.eSRRF(x, magnification=2, radius=1.5, temporal_correlation="AVG")
    ref = nanopyx_amd.eSRRF(torch.from_numpy(x).cuda(), magnification=2, radius=1.5,
                            temporal_correlation="AVG").cpu().numpy()
    assert _report("streamed workflow after a failed call vs device stack", got, ref) <= CHUNK_TOL


@pytest.mark.parametrize("N,H,W,M,R,S,mode", [
    (2, 48, 1024, 5, 1.7, 2, "TAC2"),
    (2, 48, 1024, 3, 2.3, 1, "AVG"),
    (2, 24, 1024, 8, 1.2, 3, "VAR"),
])
def test_wide_shapes_vs_oracle(torch, oracle_built, N, H, W, M, R, S, mode):
    """The served stream's widest inputs (1024 columns) at magnifications that are not powers of two: the reference
    rounds each pixel centre to float32 ((x M + 0.5) / M per position) where the kernels use the exact phase offset,
    a difference that grows with the coordinate. Per-frame maps, the plain fused call and the served job against the
    oracle's restatement of the reference chain."""
    from nanopyx_amd import _hip, synthetic
    print(f"wide N{N} {H}x{W} M{M} R{R} S{S} {mode}", flush=True)
    sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 80 + M)
    xh = synthetic.generate(sc)
    x = torch.from_numpy(xh).cuda()
    ref_maps = oracle_built.esrrf(xh, M, R, S)
    frames = _hip.esrrf_frames(x, M, R, S, True).cpu().numpy()
    assert _report(f"wide M{M} frames vs reference restatement", frames, ref_maps) <= TOL
    # the right-most quarter of the columns, where the coordinates are largest
    q = frames.shape[2] * 3 // 4
    assert _report(f"wide M{M} frames, last quarter of columns", frames[..., q:], ref_maps[..., q:]) <= TOL
    ref = oracle_built.temporal(ref_maps, mode)
    plain = _hip.esrrf_temporal(x, M, R, S, True, (mode,))[mode].cpu().numpy()
    assert _report(f"wide M{M} {mode} plain vs reference restatement", plain, ref) <= TOL
    sess = _hip.Session(1 << 30, 4, 1 << 22)
    out = {mode: torch.empty((H * M, W * M), device="cuda")}
    jid = sess.submit(x, M, R, S, True, (mode,), out)
    for _ in range(100000):
        sess.step()
        if sess.poll(jid) == _hip.Session.DONE:
            break
    torch.cuda.synchronize()
    sess.close()
    assert _report(f"wide M{M} {mode} served vs reference restatement", out[mode].cpu().numpy(), ref) <= TOL
