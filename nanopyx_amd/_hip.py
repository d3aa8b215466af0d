"""ctypes binding of libnpx_hip.so (include/npx_esrrf.h) with PyTorch-ROCm tensors as device memory.

There is no CPU path behind these functions: if the library or a ROCm GPU is missing they raise.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import threading

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libnpx_hip.so")
_lib = None
_lock = threading.Lock()

NPX_AVG, NPX_VAR, NPX_TAC2 = 1, 2, 4
MODE_BITS = {"AVG": NPX_AVG, "VAR": NPX_VAR, "TAC2": NPX_TAC2}
DTYPE_F32, DTYPE_U16 = 0, 1

_c_int, _c_double, _c_void_p, _c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class NpxError(RuntimeError):
    """A libnpx_hip call failed (HIP error, unsupported parameters, or bad workspace)."""


class NpxUnsupported(NpxError):
    """Parameters valid for the reference but outside what these kernels build (NPX_ERR_UNSUPPORTED)."""


def library_path() -> str:
    return _LIB_PATH


def lib():
    """Loads libnpx_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise NpxError(f"{_LIB_PATH} is missing: build it with `python -m nanopyx_amd._build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
        # torch first: its bundled HIP runtime (SONAME libamdhip64.so.7) then satisfies this library's dependency,
        # so one runtime serves both; loaded the other way round, torch would bring a second runtime instance
        import torch  # noqa: F401
        L = ctypes.CDLL(_LIB_PATH)
        L.npx_version.restype = _c_int
        L.npx_last_error.restype = ctypes.c_char_p
        L.npx_device_info.argtypes = [ctypes.c_char_p, _c_int]
        L.npx_esrrf_frames.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_double,
                                       _c_int, _c_void_p, _c_void_p]
        L.npx_esrrf_temporal_workspace.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_double]
        L.npx_esrrf_temporal_workspace.restype = _c_size_t
        L.npx_esrrf_temporal.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_double,
                                         _c_int, ctypes.c_uint, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                         _c_size_t, _c_void_p]
        L.npx_esrrf_temporal_static.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double,
                                                _c_double, _c_int, ctypes.c_uint, _c_void_p, _c_void_p, _c_void_p,
                                                _c_void_p]
        L.npx_rgc.argtypes = [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_double,
                              _c_double, _c_int, _c_int, _c_void_p, _c_void_p]
        L.npx_rgc_lattice.argtypes = L.npx_rgc.argtypes
        L.npx_temporal.argtypes = [_c_void_p, _c_int, ctypes.c_longlong, ctypes.c_uint, _c_void_p, _c_void_p,
                                   _c_void_p, _c_void_p]
        L.npx_macro_pixel.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p]
        L.npx_plan_info.argtypes = [_c_int, _c_int, _c_int, _c_double, _ip, _fp]
        L.npx_plan_tables.argtypes = [_c_int, _c_int, _c_int, _c_double, _fp, _fp, _ip, _fp, _ip, _fp, _ip, _fp,
                                      _ip, _fp]
        L.npx_set_launch.argtypes = [_c_int, _c_int]
        L.npx_timing.argtypes = [_c_int]
        L.npx_timing_read.argtypes = [_ip, ctypes.POINTER(_c_double), ctypes.POINTER(_c_double)]
        L.npx_esrrf_sweep_workspace.argtypes = [_c_int, _c_int, _c_int, _c_int, _ip, _c_int,
                                                ctypes.POINTER(_c_double), _c_int]
        L.npx_esrrf_sweep_workspace.restype = _c_size_t
        L.npx_esrrf_sweep.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _ip, _c_int,
                                      ctypes.POINTER(_c_double), _c_int, ctypes.POINTER(_c_double), _c_int,
                                      ctypes.POINTER(_c_void_p), _c_void_p, _c_size_t, _c_void_p]
        L.npx_capture_begin.argtypes = [ctypes.POINTER(_c_void_p)]
        L.npx_capture_end.argtypes = [_c_void_p, ctypes.POINTER(_c_void_p)]
        L.npx_graph_launch.argtypes = [_c_void_p, _c_void_p]
        L.npx_graph_destroy.argtypes = [_c_void_p]
        L.npx_session_create.argtypes = [_c_size_t, _c_int, ctypes.c_longlong, ctypes.POINTER(_c_void_p)]
        L.npx_session_submit.argtypes = [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                         _c_double, _c_double, _c_int, ctypes.c_uint, _c_void_p, _c_void_p,
                                         _c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.npx_session_submit_fed.argtypes = [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int,
                                             _c_double, _c_double, _c_int, ctypes.c_uint, _c_void_p, _c_void_p,
                                             _c_void_p, _c_int, ctypes.POINTER(ctypes.c_longlong)]
        L.npx_session_feed.argtypes = [_c_void_p, ctypes.c_longlong, _c_int, _c_void_p]
        L.npx_session_job_bytes.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_double, _c_int, _c_int]
        L.npx_session_job_bytes.restype = _c_size_t
        L.npx_session_slice_plan.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_double, _c_int, _c_int,
                                             ctypes.c_longlong, _ip]
        L.npx_session_step.argtypes = [_c_void_p, _c_void_p, _ip]
        L.npx_session_poll.argtypes = [_c_void_p, ctypes.c_longlong]
        L.npx_session_forget.argtypes = [_c_void_p, ctypes.c_longlong]
        L.npx_session_stats.argtypes = [_c_void_p, ctypes.POINTER(ctypes.c_longlong)]
        L.npx_session_destroy.argtypes = [_c_void_p]
        L.npx_autotune_workspace.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_double]
        L.npx_autotune_workspace.restype = _c_size_t
        L.npx_autotune.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_double, _c_int,
                                   ctypes.c_uint, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_size_t, _c_void_p,
                                   ctypes.POINTER(_c_double)]
        L.npx_tune_set.argtypes = [_c_int, _c_int, _c_int, _c_double, _c_int, _c_int]
        L.npx_tune_get.argtypes = [_c_int, _c_int, _c_int, _c_double, _ip]
        L.npx_autotune_frames.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _c_double,
                                          _c_int, _c_void_p, _c_void_p, ctypes.POINTER(_c_double)]
        L.npx_tune_frames_set.argtypes = L.npx_tune_set.argtypes
        L.npx_tune_frames_get.argtypes = L.npx_tune_get.argtypes
        L.npx_radiality.argtypes = [_c_void_p, _c_int, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_double, _c_int,
                                    _c_int, _c_int, _c_void_p, _c_void_p]
        L.npx_radiality_info.argtypes = [_c_int, _c_int, _c_int, _c_double, _c_int, _ip, _fp]
        L.npx_radiality_tables.argtypes = [_c_int, _c_int, _c_int, _c_double, _c_int, _fp, _fp, _ip, _ip]
        L.npx_esrrf3d_chunk_bytes.argtypes = [_c_size_t]
        L.npx_esrrf3d_chunk_bytes.restype = _c_size_t
        L.npx_esrrf3d_workspace.argtypes = [_c_int] * 7
        L.npx_esrrf3d_workspace.restype = _c_size_t
        L.npx_esrrf3d.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double,
                                  _c_double, _c_double, _c_int, _c_int, _c_void_p, _c_void_p, _c_size_t, _c_void_p]
        L.npx_shift_magnify_shape.argtypes = [_c_int, _c_int, ctypes.c_float, ctypes.c_float, _ip, _ip]
        L.npx_shift_magnify.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, _c_void_p, _c_void_p]
        L.npx_roberts.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p]
        L.npx_fht_workspace.argtypes = [_c_int] * 7
        L.npx_fht_workspace.restype = _c_size_t
        L.npx_fht_interpolate.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                          _c_void_p, _c_void_p, _c_size_t, _c_void_p]
        L.npx_fht_interpolate_lattice.argtypes = [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                                  _c_int, _c_void_p, _c_void_p, _c_size_t, _c_void_p]
        L.npx_release.restype = None
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().npx_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        if rc == -3:
            raise NpxUnsupported(f"libnpx_hip error {rc}: {msg}")
        raise NpxError(f"libnpx_hip error {rc}: {msg}")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise NpxError("the HIP_gfx950 run type needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


_CAPTURING = set()  # raw handles of the streams Graph is capturing on


def _stream(torch, stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _on(torch, stream, *tensors) -> int:
    """Raw handle of the stream a library call launches on. A caller-given stream other than the current one first
    waits for the work queued on the current stream (the input's dtype cast or copy made by these wrappers, the
    allocation of outputs and workspaces), and every device tensor the call reads or writes is recorded on it, so the
    caching allocator does not hand their memory to other work before the call has finished. A stream under graph
    capture is taken as it is: inputs are prepared before the capture."""
    cur = torch.cuda.current_stream()
    if stream is None:
        return int(cur.cuda_stream)
    h = int(stream.cuda_stream)
    if h != int(cur.cuda_stream) and h not in _CAPTURING:
        stream.wait_stream(cur)
        for t in tensors:
            if t is not None and isinstance(t, torch.Tensor) and t.is_cuda:
                t.record_stream(stream)
    return h


def _dev_stack(x):
    """(N, H, W) float32 or uint16 device tensor from a NumPy array or tensor (uint16 stays uint16: the kernel
    converts it exactly as ``astype(float32)`` would; other dtypes are cast to float32 on the host first)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype not in (torch.float32, torch.uint16):
            t = t.to(torch.float32)
        if not t.is_cuda:
            t = t.cuda()
    else:
        a = np.asarray(x)
        if a.dtype != np.uint16 and a.dtype != np.float32:
            a = a.astype(np.float32, copy=False)
        t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if t.dim() == 2:
        t = t.unsqueeze(0)
    if t.dim() != 3:
        raise ValueError("Image must be 2D and 3D (sequence of 2D images)")
    return t.contiguous()


def _dtype_code(t) -> int:
    import torch
    return DTYPE_U16 if t.dtype == torch.uint16 else DTYPE_F32


K_FUSED_MAX = 6  # widest ring half-width int(2 R) the fused kernels are built for (R < 3.5)
_COMPOSED_CHUNK_BYTES = 2 << 30  # device bytes of one frame chunk of the composed chain


def ring_half_width(radius) -> int:
    """Half-width int(2 R) of the RGC window in lattice points, formed in float32 as the reference forms it
    (_c_sr_radial_gradient_convergence.c:27-28)."""
    return int(np.float32(2.0) * np.float32(radius))


def _wide(radius) -> bool:
    return ring_half_width(radius) > K_FUSED_MAX


def _composed_frames(x, M, radius, sensitivity, iw, stream):
    """Per-frame eSRRF maps through the reference's chain of separate steps (_le_esrrf.pyx:190-194), for radii whose
    ring window the fused kernels are not built for: Catmull-Rom xM of the frames (npx_shift_magnify), Roberts cross
    (npx_roberts), Catmull-Rom x2M of both gradients, then the RGC on those arrays weighted by the magnified frames
    (npx_rgc, which runs its per-pixel kernel for these radii). Frame chunks bound the x2M gradient arrays. The steps
    allocate their outputs, so the chain is not capturable into a graph."""
    torch = _torch()
    if stream is not None and int(stream.cuda_stream) in _CAPTURING:
        raise NpxUnsupported(f"radius {radius}: ring windows wider than {2 * K_FUSED_MAX + 1} lattice points run as "
                             "separate allocating steps, which a graph capture cannot hold")
    N, H, W = x.shape
    HM, WM = H * M, W * M
    per_frame = 4 * HM * WM * 10 + 8 * H * W  # magnified frame, two x2M gradients, map; Roberts pair
    chunk = max(1, min(N, _COMPOSED_CHUNK_BYTES // per_frame))
    if stream is not None:
        stream.wait_stream(torch.cuda.current_stream())
        x.record_stream(stream)
    ctx = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
    with ctx:
        out = torch.empty((N, HM, WM), dtype=torch.float32, device=x.device)
        for f0 in range(0, N, chunk):
            part = x[f0:f0 + chunk]
            mag = shift_magnify(part, 0, 0, M, M)
            gc, gr = roberts(part)
            gci = shift_magnify(gc, 0, 0, 2 * M, 2 * M)
            gri = shift_magnify(gr, 0, 0, 2 * M, 2 * M)
            del gc, gr
            out[f0:f0 + chunk] = rgc(gci, gri, mag, M, radius, sensitivity, iw, False)
            del gci, gri, mag
    return out


def esrrf_frames(stack, magnification, radius, sensitivity, iw, stream=None):
    """Per-frame eSRRF maps as a device tensor (N, H*M, W*M). Radii beyond the fused kernels (int(2 R) > 6) run the
    reference's chain of separate steps on the GPU."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    M = int(magnification)
    if _wide(radius):
        return _composed_frames(x, M, radius, sensitivity, iw, stream)
    out = torch.empty((N, H * M, W * M), dtype=torch.float32, device=x.device)
    _check(lib().npx_esrrf_frames(x.data_ptr(), _dtype_code(x), N, H, W, M, float(radius), float(sensitivity),
                                  int(bool(iw)), out.data_ptr(), _on(torch, stream, x, out)))
    return out


def temporal_workspace_bytes(N, H, W, M, R) -> int:
    return int(lib().npx_esrrf_temporal_workspace(N, H, W, int(M), float(R)))


def esrrf_temporal(stack, magnification, radius, sensitivity, iw, modes=("AVG",), stream=None, workspace=True):
    """Fused eSRRF + temporal correlation; returns {mode: device tensor (H*M, W*M)}. Radii beyond the fused kernels
    (int(2 R) > 6) take the per-frame maps of the reference's chain of separate steps, then the temporal correlation of
    that stack (npx_temporal), as the reference does."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    M = int(magnification)
    if _wide(radius):
        for m in modes:
            if m not in MODE_BITS:
                raise ValueError("Type of correlation must be AVG, VAR or TAC2")
        maps = _composed_frames(x, M, radius, sensitivity, iw, stream)
        return temporal(maps, modes, stream=stream)
    mask = 0
    outs = {}
    for m in modes:
        if m not in MODE_BITS:
            raise ValueError("Type of correlation must be AVG, VAR or TAC2")
        mask |= MODE_BITS[m]
        outs[m] = torch.empty((H * M, W * M), dtype=torch.float32, device=x.device)
    nbytes = temporal_workspace_bytes(N, H, W, M, radius) if workspace else 0
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=x.device) if nbytes else None
    ptr = lambda k: outs[k].data_ptr() if k in outs else None
    _check(lib().npx_esrrf_temporal(x.data_ptr(), _dtype_code(x), N, H, W, M, float(radius), float(sensitivity),
                                    int(bool(iw)), mask, ptr("AVG"), ptr("VAR"), ptr("TAC2"),
                                    work.data_ptr() if work is not None else None, nbytes,
                                    _on(torch, stream, x, work, *outs.values())))
    return outs


def _sweep_lists(magnifications, radii, sensitivities):
    mags = (ctypes.c_int * len(magnifications))(*[int(m) for m in magnifications])
    rads = (ctypes.c_double * len(radii))(*[float(r) for r in radii])
    sens = (ctypes.c_double * len(sensitivities))(*[float(s) for s in sensitivities])
    return mags, rads, sens


def sweep_workspace_bytes(N, H, W, magnifications, radii, sensitivities) -> int:
    mags, rads, _ = _sweep_lists(magnifications, radii, sensitivities)
    return int(lib().npx_esrrf_sweep_workspace(N, H, W, len(mags), mags, len(rads), rads, len(sensitivities)))


class SweepOutputs:
    """Output maps of a parameter sweep: ``maps[(M, R, S)]`` is the (H*M, W*M) AVG device tensor."""

    def __init__(self, torch, x, magnifications, radii, sensitivities, workspace):
        N, H, W = x.shape
        # what the pointer array and workspace were sized for: a later call reusing them must match
        self.shape = (int(N), int(H), int(W))
        self.lists = ([int(m) for m in magnifications], [float(r) for r in radii], [float(s) for s in sensitivities])
        self.key = [(int(m), float(r), float(s)) for m in magnifications for r in radii for s in sensitivities]
        self.maps = {}
        for (m, r, s) in self.key:
            self.maps[(m, r, s)] = torch.empty((H * m, W * m), dtype=torch.float32, device=x.device)
        # radii the batched kernel covers go through npx_esrrf_sweep (its pointer array and workspace); wider ones
        # through the composed chain, one setting at a time
        self.narrow = [float(r) for r in radii if not _wide(r)]
        self.wide = [float(r) for r in radii if _wide(r)]
        nkey = [(int(m), r, float(s)) for m in magnifications for r in self.narrow for s in sensitivities]
        self.ptrs = (_c_void_p * len(nkey))(*[self.maps[k].data_ptr() for k in nkey])
        nbytes = sweep_workspace_bytes(N, H, W, magnifications, self.narrow, sensitivities) if (
            workspace and self.narrow) else 0
        self.nbytes = nbytes
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=x.device) if nbytes else None


def esrrf_sweep(stack, magnifications, radii, sensitivities, iw=True, stream=None, workspace=True, outputs=None):
    """eSRRF temporal AVG for every (magnification, radius, sensitivity) in one batched launch.

    Returns a SweepOutputs whose ``maps[(M, R, S)]`` are device tensors. Passing the SweepOutputs of an earlier call
    with the same stack shape and lists reuses its buffers. Radii beyond the fused kernels (int(2 R) > 6) are left
    out of the launch and computed per setting through the reference's chain of separate steps."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    if not magnifications or not radii or not sensitivities:
        raise ValueError("a sweep needs at least one magnification, radius and sensitivity")
    lists = ([int(m) for m in magnifications], [float(r) for r in radii], [float(s) for s in sensitivities])
    narrow = [r for r in lists[1] if not _wide(r)]
    if outputs is not None:
        if outputs.shape != (N, H, W) or outputs.lists != lists:
            raise ValueError(f"outputs were made for a {outputs.shape} stack and settings {outputs.lists}, not "
                             f"{(N, H, W)} and {lists}")
        if len(outputs.ptrs) != len(lists[0]) * len(narrow) * len(lists[2]):
            raise ValueError("outputs hold a different number of maps than the settings need")
        if outputs.work is not None and narrow and outputs.nbytes < sweep_workspace_bytes(N, H, W, magnifications,
                                                                                           narrow, sensitivities):
            raise ValueError("outputs' workspace is smaller than this sweep needs")
    out = outputs if outputs is not None else SweepOutputs(torch, x, magnifications, radii, sensitivities, workspace)
    if narrow:
        mags, rads, sens = _sweep_lists(magnifications, narrow, sensitivities)
        _check(lib().npx_esrrf_sweep(x.data_ptr(), _dtype_code(x), N, H, W, len(mags), mags, len(rads), rads,
                                     len(sens), sens, int(bool(iw)), out.ptrs,
                                     out.work.data_ptr() if out.work is not None else None, out.nbytes,
                                     _on(torch, stream, x, out.work, *out.maps.values())))
    for m in lists[0]:
        for r in out.wide:
            for s in lists[2]:
                avg = esrrf_temporal(x, m, r, s, iw, ("AVG",), stream=stream)["AVG"]
                ctx = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
                with ctx:
                    out.maps[(m, r, s)].copy_(avg)
    return out


class Graph:
    """An instantiated HIP graph of npx_* launches (npx_capture_begin / npx_capture_end).

    ``with Graph.capture() as g: fn(stream=g.stream)`` records the calls made with ``stream=g.stream``;
    ``g.launch()`` replays them on the current stream."""

    def __init__(self):
        self._exec = None
        self.stream = None

    @classmethod
    def capture(cls):
        return cls()

    def __enter__(self):
        torch = _torch()
        s = _c_void_p()
        _check(lib().npx_capture_begin(ctypes.byref(s)))
        self._raw = s.value
        _CAPTURING.add(int(self._raw))
        self.stream = torch.cuda.ExternalStream(self._raw)
        return self

    def __exit__(self, exc_type, exc, tb):
        g = _c_void_p()
        _CAPTURING.discard(int(self._raw))
        rc = lib().npx_capture_end(self._raw, ctypes.byref(g))
        self.stream = None
        if exc_type is None:
            _check(rc)
            self._exec = g.value
        elif rc == 0:
            lib().npx_graph_destroy(g.value)
        return False

    def launch(self, stream=None):
        torch = _torch()
        _check(lib().npx_graph_launch(self._exec, _stream(torch, stream)))

    def close(self):
        if self._exec is not None:
            lib().npx_graph_destroy(self._exec)
            self._exec = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Session:
    """A serving session of the library (npx_session_*): continuous batching of eSRRF + temporal jobs."""

    QUEUED, RUNNING, DONE = 0, 1, 2

    def __init__(self, pool_bytes: int, max_active: int = 16, px_frames_per_step: int = 1 << 29):
        _torch()
        h = _c_void_p()
        _check(lib().npx_session_create(int(pool_bytes), int(max_active), int(px_frames_per_step), ctypes.byref(h)))
        self._h = h.value

    def submit(self, stack, magnification, radius, sensitivity, iw, modes, outputs):
        """Queues a job. stack: (N, H, W) float32 / uint16 tensor on the device, or on the host (pinned memory
        preferred; copied into the session pool on admission). outputs: {mode: device tensor (H*M, W*M)}.
        The caller keeps stack and outputs alive until the job is done."""
        torch = _torch()
        if stack.dim() != 3:
            raise ValueError("Image must be 2D and 3D (sequence of 2D images)")
        if stack.dtype not in (torch.float32, torch.uint16) or not stack.is_contiguous():
            raise ValueError("served stacks must be contiguous float32 or uint16 tensors")
        N, H, W = stack.shape
        mask = 0
        for m in modes:
            if m not in MODE_BITS:
                raise ValueError("Type of correlation must be AVG, VAR or TAC2")
            mask |= MODE_BITS[m]
        ptr = lambda k: outputs[k].data_ptr() if k in outputs else None
        jid = ctypes.c_longlong(0)
        _check(lib().npx_session_submit(self._h, stack.data_ptr(), int(not stack.is_cuda), _dtype_code(stack), N, H,
                                        W, int(magnification), float(radius), float(sensitivity), int(bool(iw)), mask,
                                        ptr("AVG"), ptr("VAR"), ptr("TAC2"), ctypes.byref(jid)))
        return jid.value

    def submit_fed(self, stack, magnification, radius, sensitivity, iw, modes, outputs, piece_frames):
        """Queues a job whose device stack (N, H, W) the caller fills itself, piece by piece; announce each piece
        with feed(). piece_frames: frames the caller copies at a time (the job's frame chunks are no longer)."""
        torch = _torch()
        if stack.dim() != 3 or not stack.is_cuda or not stack.is_contiguous():
            raise ValueError("a fed job's stack must be a contiguous (N, H, W) device tensor")
        if stack.dtype not in (torch.float32, torch.uint16):
            raise ValueError("served stacks must be float32 or uint16")
        N, H, W = stack.shape
        mask = 0
        for m in modes:
            if m not in MODE_BITS:
                raise ValueError("Type of correlation must be AVG, VAR or TAC2")
            mask |= MODE_BITS[m]
        ptr = lambda k: outputs[k].data_ptr() if k in outputs else None
        jid = ctypes.c_longlong(0)
        _check(lib().npx_session_submit_fed(self._h, stack.data_ptr(), _dtype_code(stack), N, H, W,
                                            int(magnification), float(radius), float(sensitivity), int(bool(iw)), mask,
                                            ptr("AVG"), ptr("VAR"), ptr("TAC2"), int(piece_frames), ctypes.byref(jid)))
        return jid.value

    def feed(self, job_id, frames, stream=None):
        """Frames [0, frames) of a fed job's stack are on the device once the work queued so far on `stream` (the
        current stream by default) has completed."""
        torch = _torch()
        _check(lib().npx_session_feed(self._h, int(job_id), int(frames), _stream(torch, stream)))

    @staticmethod
    def job_bytes(N, H, W, magnification, radius, dtype_u16=False, host=True) -> int:
        """Pool bytes a job needs (tables, running moments, chunk partials, stack when submitted from the host)."""
        return int(lib().npx_session_job_bytes(int(N), int(H), int(W), int(magnification), float(radius),
                                               DTYPE_U16 if dtype_u16 else DTYPE_F32, int(bool(host))))

    @staticmethod
    def slice_plan(N, H, W, magnification, radius, dtype_u16=False, host=True, px_frames_per_step=1 << 29):
        """Host-only replay of how a lone job is laid out and sliced (npx_session_slice_plan): partial slots, frames
        per chunk, most chunks in one slice, slices, slices off a chunk boundary, upload pieces."""
        info = (_c_int * 6)()
        _check(lib().npx_session_slice_plan(int(N), int(H), int(W), int(magnification), float(radius),
                                            DTYPE_U16 if dtype_u16 else DTYPE_F32, int(bool(host)),
                                            int(px_frames_per_step), info))
        return dict(zip(("cmax", "cfpc", "max_chunks", "slices", "off_boundary", "pieces"), list(info)))

    def step(self, stream=None) -> int:
        torch = _torch()
        n = ctypes.c_int(0)
        _check(lib().npx_session_step(self._h, _stream(torch, stream), ctypes.byref(n)))
        return n.value

    def poll(self, job_id) -> int:
        rc = lib().npx_session_poll(self._h, int(job_id))
        if rc < 0:
            _check(rc)
        return rc

    def forget(self, job_id) -> None:
        _check(lib().npx_session_forget(self._h, int(job_id)))

    def stats(self):
        a = (ctypes.c_longlong * 6)()
        _check(lib().npx_session_stats(self._h, a))
        return dict(zip(("queued", "active", "free_bytes", "launches", "slices", "steps"), list(a)))

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(lib().npx_session_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rgc(gx, gy, img, magnification, radius, sensitivity, iw, pad_edges, stream=None):
    torch = _torch()
    gx, gy, img = (_dev_stack(a) for a in (gx, gy, img))
    gx, gy, img = (a.to(torch.float32) for a in (gx, gy, img))
    N, HM, WM = img.shape
    if tuple(gx.shape) != (N, 2 * HM, 2 * WM) or tuple(gy.shape) != (N, 2 * HM, 2 * WM):
        raise ValueError("gradient arrays must be (N, 2*rowsM, 2*colsM) for an (N, rowsM, colsM) image")
    out = torch.empty_like(img)
    _check(lib().npx_rgc(gx.data_ptr(), gy.data_ptr(), img.data_ptr(), N, HM, WM, int(magnification), float(radius),
                         float(sensitivity), int(bool(iw)), int(bool(pad_edges)), out.data_ptr(),
                         _on(torch, stream, gx, gy, img, out)))
    return out


def lattice_shape(n, stride):
    """Kept positions of an axis of n points for a lattice stride (>= 2): the multiples of it and the last point."""
    return -(-n // stride) + (1 if (n - 1) % stride else 0)


def rgc_lattice(gx, gy, img, magnification, radius, sensitivity, iw, pad_edges, stream=None):
    """rgc with the gradients given only at the lattice it reads, compact (N, 2H+1, 2W+1) (npx_rgc_lattice);
    raises NpxUnsupported for shapes or radii only the per-pixel kernel covers."""
    torch = _torch()
    gx, gy, img = (_dev_stack(a).to(torch.float32).contiguous() for a in (gx, gy, img))
    N, HM, WM = img.shape
    M = int(magnification)
    if M < 2 or HM % M or WM % M:
        raise NpxUnsupported("compact lattice needs an image of H x W pixels magnified M >= 2 times")
    shape = (N, 2 * (HM // M) + 1, 2 * (WM // M) + 1)
    if tuple(gx.shape) != shape or tuple(gy.shape) != shape:
        raise ValueError(f"compact gradient lattices must be {shape}")
    out = torch.empty_like(img)
    _check(lib().npx_rgc_lattice(gx.data_ptr(), gy.data_ptr(), img.data_ptr(), N, HM, WM, M, float(radius),
                                 float(sensitivity), int(bool(iw)), int(bool(pad_edges)), out.data_ptr(),
                                 _on(torch, stream, gx, gy, img, out)))
    return out


def temporal(stack, modes, stream=None):
    """AVG / VAR / TAC2 over axis 0 of an (N, ...) stack; returns {mode: device tensor}."""
    torch = _torch()
    x = stack if isinstance(stack, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(stack, np.float32))
    x = x.to(device="cuda", dtype=torch.float32).contiguous()
    N = x.shape[0]
    P = int(np.prod(x.shape[1:]))
    outs, mask = {}, 0
    for m in modes:
        if m not in MODE_BITS:
            raise ValueError("Type of correlation must be AVG, VAR or TAC2")
        mask |= MODE_BITS[m]
        outs[m] = torch.empty(tuple(x.shape[1:]), dtype=torch.float32, device=x.device)
    ptr = lambda k: outs[k].data_ptr() if k in outs else None
    _check(lib().npx_temporal(x.data_ptr(), N, P, mask, ptr("AVG"), ptr("VAR"), ptr("TAC2"),
                              _on(torch, stream, x, *outs.values())))
    return outs


def macro_pixel_(img, magnification, stream=None):
    """In-place macro-pixel correction of a (S, HM, WM) float32 device tensor."""
    torch = _torch()
    S, HM, WM = img.shape
    M = int(magnification)
    cmap = torch.empty(S * M * M, dtype=torch.float32, device=img.device)
    _check(lib().npx_macro_pixel(img.data_ptr(), S, HM, WM, M, cmap.data_ptr(), _on(torch, stream, img, cmap)))
    return img


def radiality(stack, interp, magnification, ringRadius, border, positivity, iw, stream=None, out=None):
    """SRRF radiality maps (N, H*M, W*M) as a device tensor. `interp` is the caller's (N, H*M, W*M) intensity image
    for the weighting, or None to weight by the xM Catmull-Rom magnification computed in the kernel."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    M = int(magnification)
    ip = 0
    if interp is not None:
        interp = _dev_stack(interp).to(torch.float32).contiguous()
        if tuple(interp.shape) != (N, H * M, W * M):
            raise ValueError("image_interp must be (N, H*magnification, W*magnification) for an (N, H, W) image")
        ip = interp.data_ptr()
    if out is None:
        out = torch.empty((N, H * M, W * M), dtype=torch.float32, device=x.device)
    _check(lib().npx_radiality(x.data_ptr(), _dtype_code(x), ip, N, H, W, M, float(ringRadius), int(border),
                               int(bool(positivity)), int(bool(iw)), out.data_ptr(),
                               _on(torch, stream, x, interp, out)))
    return out


def shift_magnify(stack, shift_row, shift_col, mag_row, mag_col, stream=None):
    """Catmull-Rom shift and magnify of an (N, H, W) stack: (N, int(H*mag_row), int(W*mag_col)) device tensor."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    hm, wm = ctypes.c_int(), ctypes.c_int()
    _check(lib().npx_shift_magnify_shape(H, W, float(mag_row), float(mag_col), ctypes.byref(hm), ctypes.byref(wm)))
    out = torch.empty((N, hm.value, wm.value), dtype=torch.float32, device=x.device)
    _check(lib().npx_shift_magnify(x.data_ptr(), _dtype_code(x), N, H, W, float(shift_row), float(shift_col),
                                   float(mag_row), float(mag_col), out.data_ptr(), _on(torch, stream, x, out)))
    return out


def roberts(stack, stream=None):
    """Roberts cross gradients (gc, gr) of an (N, H, W) stack as device tensors."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    gc = torch.empty((N, H, W), dtype=torch.float32, device=x.device)
    gr = torch.empty_like(gc)
    _check(lib().npx_roberts(x.data_ptr(), _dtype_code(x), N, H, W, gc.data_ptr(), gr.data_ptr(),
                             _on(torch, stream, x, gc, gr)))
    return gc, gr


def fht_interpolate(stack, magnification, mirror=True, tile_size=0, overlap=16, lattice=0, stream=None):
    """Fourier interpolation of an (N, h, w) stack by an integer factor: (N, h*M, w*M) float32 device tensor. With
    lattice >= 2 only the rows and columns that are multiples of it and the last ones, compact:
    (N, lattice_shape(h*M, lattice), lattice_shape(w*M, lattice)) (npx_fht_interpolate_lattice)."""
    torch = _torch()
    x = _dev_stack(stack).to(torch.float32).contiguous()
    N, h, w = x.shape
    M = int(magnification)
    need = int(lib().npx_fht_workspace(N, h, w, M, int(bool(mirror)), int(tile_size), int(overlap)))
    if need == 0:
        raise ValueError("bad Fourier interpolation shape")
    work = torch.empty(need, dtype=torch.uint8, device=x.device)
    if lattice and lattice >= 2:
        shape = (N, lattice_shape(h * M, lattice), lattice_shape(w * M, lattice))
    else:
        shape = (N, h * M, w * M)
    out = torch.empty(shape, dtype=torch.float32, device=x.device)
    if lattice:
        _check(lib().npx_fht_interpolate_lattice(x.data_ptr(), N, h, w, M, int(bool(mirror)), int(tile_size),
                                                 int(overlap), int(lattice), out.data_ptr(), work.data_ptr(), need,
                                                 _on(torch, stream, x, out, work)))
    else:
        _check(lib().npx_fht_interpolate(x.data_ptr(), N, h, w, M, int(bool(mirror)), int(tile_size), int(overlap),
                                         out.data_ptr(), work.data_ptr(), need, _on(torch, stream, x, out, work)))
    return out


ESRRF3D_MODES = {"average": 0, "std": 1}


def esrrf3d(stack, magnification_xy, magnification_z, radius, radius_z, sensitivity, mode, iw, stream=None):
    """eSRRF 3D of an (F, S, R, C) stack (float32 or uint16): the (S*Mz, R*Mxy, C*Mxy) running mean ("average") or
    standard deviation ("std") over frames, as a device tensor."""
    torch = _torch()
    if isinstance(stack, torch.Tensor):
        x = stack if stack.dtype in (torch.float32, torch.uint16) else stack.to(torch.float32)
        x = x.cuda() if not x.is_cuda else x
    else:
        a = np.asarray(stack)
        if a.dtype != np.uint16 and a.dtype != np.float32:
            a = a.astype(np.float32)
        x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if x.dim() == 3:
        x = x.unsqueeze(0)
    if x.dim() != 4:
        raise ValueError("image must either be 3D or 4D")
    x = x.contiguous()
    if mode not in ESRRF3D_MODES:
        raise ValueError("mode must be 'average' or 'std'")
    F, S, R, C = x.shape
    Mxy, Mz = int(magnification_xy), int(magnification_z)
    m = ESRRF3D_MODES[mode]
    need = int(lib().npx_esrrf3d_workspace(F, S, R, C, Mxy, Mz, m))
    if need == 0:
        raise ValueError("bad eSRRF 3D shape")
    work = torch.empty(need, dtype=torch.uint8, device=x.device)
    out = torch.empty((S * Mz, R * Mxy, C * Mxy), dtype=torch.float32, device=x.device)
    _check(lib().npx_esrrf3d(x.data_ptr(), _dtype_code(x), F, S, R, C, Mxy, Mz, float(radius), float(radius_z),
                             float(sensitivity), m, int(bool(iw)), out.data_ptr(), work.data_ptr(), need,
                             _on(torch, stream, x, out, work)))
    return out


def esrrf3d_chunk_bytes(nbytes=0) -> int:
    """Sets the workspace bytes per frame chunk of later eSRRF 3D calls (0 keeps it); returns the previous value."""
    return int(lib().npx_esrrf3d_chunk_bytes(int(nbytes)))


def radiality_info(H, W, M, ringRadius, border):
    """Host-only view of a radiality plan: ring coordinates and launch geometry."""
    info = (ctypes.c_int * 12)()
    ring = (ctypes.c_float * 25)()
    _check(lib().npx_radiality_info(int(H), int(W), int(M), float(ringRadius), int(border), info, ring))
    keys = ("HM", "WM", "j_lo", "j_hi", "i_lo", "i_hi", "TH", "TW", "TJ", "TI", "bands", "tiles")
    d = dict(zip(keys, list(info)))
    d["R"] = ring[0]
    d["xRing"] = np.asarray(ring[1:13], np.float32)
    d["yRing"] = np.asarray(ring[13:25], np.float32)
    return d


def radiality_tables(H, W, M, ringRadius, border):
    """Host-only copy of a radiality plan's tables: (row (H*M, 80), col (W*M, 80), tile_r0, tile_c0)."""
    d = radiality_info(H, W, M, ringRadius, border)
    row = np.zeros((d["HM"], 80), np.float32)
    col = np.zeros((d["WM"], 80), np.float32)
    r0 = np.zeros(d["bands"], np.int32)
    c0 = np.zeros(d["tiles"], np.int32)
    _check(lib().npx_radiality_tables(int(H), int(W), int(M), float(ringRadius), int(border),
                                      row.ctypes.data_as(_fp), col.ctypes.data_as(_fp), r0.ctypes.data_as(_ip),
                                      c0.ctypes.data_as(_ip)))
    return row, col, r0, c0


def plan_info(H, W, M, R):
    """Host-only view of the kernel geometry and reference-derived constants."""
    info = (ctypes.c_int * 16)()
    finfo = (ctypes.c_float * 4)()
    _check(lib().npx_plan_info(H, W, M, float(R), info, finfo))
    keys = ("K", "margin", "HM", "WM", "nphase", "Q_frames", "Q_temporal", "ring_stride", "HG", "HI", "LP", "nwaves",
            "ngroups", "prefetch", "_r0", "_r1")
    d = dict(zip(keys, list(info)))
    d.update(fwhm=finfo[0], sigma=finfo[1], tSS=finfo[2], tSO=finfo[3])
    return d


def plan_tables(H, W, M, R):
    """Host copies of the plan tables: the exact per-phase ring geometry (ring, wsum; the kernels form each block's
    from its float32 pixel centres, equal at power-of-two magnifications) and the lattice and intensity taps."""
    d = plan_info(H, W, M, R)
    K, P, NW = d["K"], d["nphase"], 2 * d["K"] + 1
    ring = np.zeros((P, d["ring_stride"]), np.float32)
    wsum = np.zeros(P, np.float32)
    arr = {}
    for name, n in (("lr", 2 * H + 1), ("lc", 2 * W + 1), ("cr", d["HM"]), ("cc", d["WM"])):
        arr[name + "b"] = np.zeros(n, np.int32)
        arr[name + "w"] = np.zeros((n, 4), np.float32)
    f = lambda a: a.ctypes.data_as(_fp)
    i = lambda a: a.ctypes.data_as(_ip)
    _check(lib().npx_plan_tables(H, W, M, float(R), f(ring), f(wsum), i(arr["lrb"]), f(arr["lrw"]), i(arr["lcb"]),
                                 f(arr["lcw"]), i(arr["crb"]), f(arr["crw"]), i(arr["ccb"]), f(arr["ccw"])))
    # per-phase ring block: dy[j], dx[i], (d, t)[i][j], w[i][j]
    views = dict(dy=ring[:, :NW], dx=ring[:, NW:2 * NW],
                 d=ring[:, 2 * NW:2 * NW + 2 * NW * NW].reshape(P, NW, NW, 2)[..., 0],
                 t=ring[:, 2 * NW:2 * NW + 2 * NW * NW].reshape(P, NW, NW, 2)[..., 1],
                 w=ring[:, 2 * NW + 2 * NW * NW:2 * NW + 3 * NW * NW].reshape(P, NW, NW),
                 prefix=ring[:, 2 * NW + 3 * NW * NW + 1:2 * NW + 3 * NW * NW + 1 + (NW + 1) ** 2].reshape(P, NW + 1,
                                                                                                     NW + 1))
    return dict(info=d, ring=ring, wsum=wsum, **views, **arr)


def autotune(stack, magnification, radius, sensitivity, iw, modes=("AVG",), stream=None):
    """Times the candidate launch shapes on this stack (npx_autotune) and keeps the fastest for its (H, W, M, R).
    Returns (max_waves_per_block, target_blocks, ms per call)."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    M = int(magnification)
    mask, outs = 0, {}
    for m in modes:
        if m not in MODE_BITS:
            raise ValueError("Type of correlation must be AVG, VAR or TAC2")
        mask |= MODE_BITS[m]
        outs[m] = torch.empty((H * M, W * M), dtype=torch.float32, device=x.device)
    nbytes = int(lib().npx_autotune_workspace(N, H, W, M, float(radius)))
    work = torch.empty(nbytes, dtype=torch.uint8, device=x.device) if nbytes else None
    ptr = lambda k: outs[k].data_ptr() if k in outs else None
    best = ctypes.c_double(0)
    _check(lib().npx_autotune(x.data_ptr(), _dtype_code(x), N, H, W, M, float(radius), float(sensitivity),
                              int(bool(iw)), mask, ptr("AVG"), ptr("VAR"), ptr("TAC2"),
                              work.data_ptr() if work is not None else None, nbytes,
                              _on(torch, stream, x, work, *outs.values()),
                              ctypes.byref(best)))
    mw, tb, _ = tune_get(H, W, M, radius)
    return mw, tb, best.value


def autotune_frames(stack, magnification, radius, sensitivity, iw, stream=None):
    """Times the candidate launch shapes of the per-frame-map kernel on this stack (npx_autotune_frames) and keeps
    the fastest for its (H, W, M, R). Returns (max_waves_per_block, target_blocks, ms per call)."""
    torch = _torch()
    x = _dev_stack(stack)
    N, H, W = x.shape
    M = int(magnification)
    out = torch.empty((N, H * M, W * M), dtype=torch.float32, device=x.device)
    best = ctypes.c_double(0)
    _check(lib().npx_autotune_frames(x.data_ptr(), _dtype_code(x), N, H, W, M, float(radius), float(sensitivity),
                                     int(bool(iw)), out.data_ptr(), _on(torch, stream, x, out), ctypes.byref(best)))
    mw, tb, _ = tune_frames_get(H, W, M, radius)
    return mw, tb, best.value


def tune_frames_set(H, W, M, R, max_waves_per_block: int, target_blocks: int) -> None:
    """Stores a launch shape of the per-frame-map kernels for (H, W, M, R) (0, 0 forgets it)."""
    _check(lib().npx_tune_frames_set(int(H), int(W), int(M), float(R), int(max_waves_per_block),
                                     int(target_blocks)))


def tune_frames_get(H, W, M, R):
    """(max_waves_per_block, target_blocks, own?) that per-frame-map calls with these parameters use; own is False
    when they use the temporal kernel's shape."""
    cfg = (ctypes.c_int * 2)()
    rc = lib().npx_tune_frames_get(int(H), int(W), int(M), float(R), cfg)
    if rc < 0:
        _check(rc)
    return cfg[0], cfg[1], bool(rc)


def tune_set(H, W, M, R, max_waves_per_block: int, target_blocks: int) -> None:
    """Stores a launch shape for (H, W, M, R) (0, 0 forgets it)."""
    _check(lib().npx_tune_set(int(H), int(W), int(M), float(R), int(max_waves_per_block), int(target_blocks)))


def tune_get(H, W, M, R):
    """(max_waves_per_block, target_blocks, tuned?) that calls with these parameters use."""
    cfg = (ctypes.c_int * 2)()
    rc = lib().npx_tune_get(int(H), int(W), int(M), float(R), cfg)
    if rc < 0:
        _check(rc)
    return cfg[0], cfg[1], bool(rc)


def set_launch(max_waves_per_block: int = 0, target_blocks: int = 0) -> None:
    """Launch shape of the fused kernels for subsequent calls (0 = library defaults)."""
    _check(lib().npx_set_launch(int(max_waves_per_block), int(target_blocks)))


def timing(on: bool) -> None:
    """Record HIP events around every fused-kernel launch from now on (True) or stop (False)."""
    _check(lib().npx_timing(int(bool(on))))


def timing_read():
    """(calls, total fused-kernel ms, total merge ms) of the launches recorded since the last read."""
    n = ctypes.c_int(0)
    k = ctypes.c_double(0)
    m = ctypes.c_double(0)
    _check(lib().npx_timing_read(ctypes.byref(n), ctypes.byref(k), ctypes.byref(m)))
    return n.value, k.value, m.value


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    _check(lib().npx_device_info(buf, 256))
    return buf.value.decode()
