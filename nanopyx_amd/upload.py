"""Streamed upload of host (NumPy) stacks for the workflows: the stack goes to the device in pieces through two
pinned staging buffers, while the eSRRF jobs of a serving session (npx_session_submit_fed) already run on the
pieces that have arrived.

Per piece: host threads copy the frames into a free staging buffer (NumPy releases the GIL for large copies), one
asynchronous host-to-device copy on a copy stream moves it into the device stack, and the session is told how many
frames each job now has (npx_session_feed, ordered after that copy). The host copy of the next piece, the transfer of
this one and the kernels of the previous ones overlap.
"""
from __future__ import annotations

import concurrent.futures
import threading

import numpy as np

from . import _hip

PIECE_BYTES = 32 << 20          # staging piece (whole frames)
STREAM_MIN_BYTES = 64 << 20     # host stacks at least this large are streamed
_COPY_THREADS = 8

_lock = threading.Lock()
_state = {}


def _resources(torch, nbytes):
    """Cached pinned staging buffers (two of at least nbytes), copy stream, host copy pool (caller holds _lock)."""
    st = _state
    if st.get("bytes", 0) < nbytes:
        st["bufs"] = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        st["bytes"] = nbytes
        st["events"] = [None, None]
    if "stream" not in st:
        st["stream"] = torch.cuda.Stream()
        st["pool"] = concurrent.futures.ThreadPoolExecutor(_COPY_THREADS)
    return st


def _host_copy(pool, dst, src):
    """dst[...] = src over frame ranges on the copy threads."""
    n = src.shape[0]
    k = max(1, min(_COPY_THREADS, n))
    cuts = np.linspace(0, n, k + 1).astype(int)
    futs = [pool.submit(np.copyto, dst[a:b], src[a:b]) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
    for f in futs:
        f.result()


def esrrf_temporal_streamed(x: np.ndarray, magnification, radius, sensitivity, iw, mode, frames_per_timepoint):
    """eSRRF + one temporal correlation of every timepoint of a host stack (float32 or uint16 NumPy array), the
    stack streamed to the device: returns a (T, H*M, W*M) device tensor of the timepoint maps."""
    torch = _hip._torch()
    N, H, W = x.shape
    M = int(magnification)
    fpt = int(frames_per_timepoint)
    T = N // fpt + (1 if N % fpt else 0)
    tdtype = torch.uint16 if x.dtype == np.uint16 else torch.float32
    fb = H * W * x.itemsize
    pf = max(8, PIECE_BYTES // fb)
    d = torch.empty((N, H, W), dtype=tdtype, device="cuda")
    out = torch.empty((T, H * M, W * M), dtype=torch.float32, device="cuda")
    spans = [(t * fpt, min(N, (t + 1) * fpt)) for t in range(T)]
    need = sum(_hip.Session.job_bytes(b - a, H, W, M, radius, tdtype == torch.uint16, False) for a, b in spans)
    d8 = d.view(torch.uint8).view(N, H * W * x.itemsize)  # byte views: one copy path for float32 and uint16
    with _lock:  # one streamed call at a time shares the staging buffers and the session
        res = _resources(torch, pf * fb)
        if _state.get("pool_bytes", 0) < need:
            if "session" in _state:
                _state.pop("session").close()
            _state["session"] = _hip.Session(need + (16 << 20), 64)
            _state["pool_bytes"] = need
        sess = _state["session"]
        cs = res["stream"]
        try:
            jobs = [sess.submit_fed(d[a:b], M, radius, sensitivity, iw, (mode,), {mode: out[t]}, pf)
                    for t, (a, b) in enumerate(spans)]
            fed = [0] * T
            cs.wait_stream(torch.cuda.current_stream())  # the device stack's memory is free of earlier work
            for p, a in enumerate(range(0, N, pf)):
                b = min(N, a + pf)
                k = p % 2
                if res["events"][k] is not None:
                    res["events"][k].synchronize()  # the buffer's previous transfer has been read
                staged = res["bufs"][k][: (b - a) * fb]
                _host_copy(res["pool"], staged.numpy().view(x.dtype).reshape(b - a, H, W), x[a:b])
                with torch.cuda.stream(cs):
                    d8[a:b].copy_(staged.view(b - a, fb), non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(cs)
                    res["events"][k] = ev
                for t, (ta, tb) in enumerate(spans):
                    have = min(b, tb) - ta
                    if have > fed[t]:
                        sess.feed(jobs[t], have, cs)
                        fed[t] = have
                sess.step()
            for _ in range(10_000_000):
                sess.step()
                if all(sess.poll(j) == _hip.Session.DONE for j in jobs):
                    break
            else:
                raise _hip.NpxError("streamed eSRRF jobs did not finish")
            for j in jobs:
                sess.forget(j)
        except BaseException:
            # a failed call leaves no job in the shared session: the session is destroyed (which waits for the
            # device, so no slice still reads this call's stack) and the next call makes a new one
            _state.pop("pool_bytes", None)
            try:
                _state.pop("session").close()
            finally:
                cs.synchronize()
            raise
        torch.cuda.current_stream().wait_stream(cs)
    return out
