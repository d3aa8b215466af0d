"""npx_esrrf_temporal_static (no device allocation: tables and chunk partials in the library's static scratch) against
npx_esrrf_temporal with a caller workspace, through the C ABI: two calls in a row on one GPU, the second with a
smaller input (the scratch keeps the first call's contents), and a call whose chunk partials do not fit the scratch."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    assert _torch.cuda.is_available(), "GPU tests need a ROCm device"
    return _torch


def _stack(torch, N, H, W, dtype, seed):
    from nanopyx_amd import synthetic
    sc = synthetic.Scene(N, H, W, max(1, int(0.02 * H * W)), 1.5, 500, 0.3, 100, 0.0, 2.0,
                         100.0 if dtype == "uint16" else 0.0, dtype, seed)
    return synthetic.generate_device(sc)


def _both(torch, x, M, R, S, modes):
    """(static, workspace) maps of one stack through the C ABI on the current stream."""
    from nanopyx_amd import _hip
    L = _hip.lib()
    N, H, W = x.shape
    mask = sum(_hip.MODE_BITS[m] for m in modes)
    st = int(torch.cuda.current_stream().cuda_stream)
    dt = _hip._dtype_code(x)
    res = []
    for static in (True, False):
        outs = {m: torch.full((H * M, W * M), -7.0, device="cuda") for m in ("AVG", "VAR", "TAC2")}
        ptr = lambda m: outs[m].data_ptr() if m in modes else None
        if static:
            rc = L.npx_esrrf_temporal_static(x.data_ptr(), dt, N, H, W, M, float(R), float(S), 1, mask, ptr("AVG"),
                                             ptr("VAR"), ptr("TAC2"), st)
        else:
            nb = int(L.npx_esrrf_temporal_workspace(N, H, W, M, float(R)))
            work = torch.empty(max(nb, 1), dtype=torch.uint8, device="cuda")
            rc = L.npx_esrrf_temporal(x.data_ptr(), dt, N, H, W, M, float(R), float(S), 1, mask, ptr("AVG"),
                                      ptr("VAR"), ptr("TAC2"), work.data_ptr(), nb, st)
        _hip._check(rc)
        torch.cuda.synchronize()
        res.append({m: outs[m].cpu().numpy() for m in modes})
    return res


def _chunks(N, H, W, M, R):
    from nanopyx_amd import _hip
    nb = int(_hip.lib().npx_esrrf_temporal_workspace(N, H, W, M, float(R)))
    return nb // (5 * H * M * W * M * 4)


@pytest.mark.parametrize("first,second", [
    # (N, H, W, M, R, S, dtype, modes): the first call's chunk partials fit the scratch, the second is smaller
    ((96, 72, 100, 3, 1.5, 2, "float32", ("AVG", "VAR", "TAC2")), (20, 24, 40, 2, 1.0, 1, "uint16", ("AVG", "TAC2"))),
    ((64, 48, 64, 5, 2.0, 3, "uint16", ("TAC2",)), (12, 16, 20, 4, 2.5, 2, "float32", ("AVG", "VAR"))),
])
def test_static_equals_workspace_two_calls(torch, first, second):
    for k, (N, H, W, M, R, S, dtype, modes) in enumerate((first, second)):
        print(f"no-alloc call {k}: N{N} {H}x{W} M{M} R{R} S{S} {dtype} {modes}, workspace chunks "
              f"{_chunks(N, H, W, M, R)}", flush=True)
        x = _stack(torch, N, H, W, dtype, 90 + k)
        st, ws = _both(torch, x, M, R, S, modes)
        for m in modes:
            mx, mean = rel_err(st[m], ws[m])
            print(f"  {m}: max {mx:.3e} mean {mean:.3e}", flush=True)
            assert np.array_equal(st[m], ws[m]), m  # same tables, same chunks: same maps


def test_static_chunks_beyond_scratch(torch):
    """A stack whose chunk partials (5 maps of 1280 x 1280 per chunk) would not all fit the 64 MiB scratch: the call
    runs fewer chunks and its maps equal the workspace call's to float32 merge rounding (<= 1e-6)."""
    N, H, W, M, R = 200, 256, 256, 5, 1.5
    print(f"no-alloc beyond scratch: N{N} {H}x{W} M{M}, workspace chunks {_chunks(N, H, W, M, R)}", flush=True)
    x = _stack(torch, N, H, W, "float32", 99)
    st, ws = _both(torch, x, M, R, 2, ("AVG", "VAR", "TAC2"))
    for m in ("AVG", "VAR", "TAC2"):
        mx, mean = rel_err(st[m], ws[m])
        print(f"  {m}: max {mx:.3e} mean {mean:.3e}", flush=True)
        assert mx <= 1e-6, m
