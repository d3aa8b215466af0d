"""eSRRF 3D on the GPU (npx_esrrf3d) against the oracle: magnifications, average and std, sensitivity, weighting,
uint16, frame chunking and the eSRRF3D() workflow."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4
CASES = [  # (Mxy, Mz, radius, sensitivity, mode, iw)
    (2, 2, 1.5, 1, "average", True),
    (1, 1, 1.5, 2, "average", True),
    (1, 2, 1.2, 1, "std", False),
    (2, 1, 2.0, 3, "std", True),
    (3, 2, 1.5, 1, "average", True),
    (2, 3, 1.0, 1.5, "std", True),
]


@pytest.fixture(scope="module")
def torch():
    import torch as _torch
    assert _torch.cuda.is_available(), "GPU tests need a ROCm device"
    return _torch


def _stack(shape, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.poisson(50, size=shape) + rng.random(shape) * 20).astype(np.float32)


def _report(tag, x, ref):
    mx, mean = rel_err(x, ref)
    print(f"{tag}: max {mx:.3e} mean {mean:.3e}", flush=True)
    return mx


@pytest.mark.parametrize("case", CASES)
def test_esrrf3d_matches_oracle(torch, oracle_built, case):
    from nanopyx_amd import _hip
    Mxy, Mz, R, S, mode, iw = case
    print(f"esrrf3d case {case}", flush=True)
    x = _stack((4, 10, 18, 22), seed=Mxy * 10 + Mz)
    ref = oracle_built.esrrf3d(x, Mxy, Mz, R, 2.8, 4.0, S, mode, iw)
    got = _hip.esrrf3d(x, Mxy, Mz, R, R * 2.8 / 4.0, S, mode, iw).cpu().numpy()
    assert got.shape == ref.shape
    assert _report(f"esrrf3d {case}", got, ref) <= TOL


def test_esrrf3d_uint16_operator_and_chunks(torch, oracle_built):
    from nanopyx_amd import _hip
    from nanopyx_amd.core.transform import eSRRF3D_ST
    rng = np.random.default_rng(7)
    x16 = (rng.poisson(200, size=(5, 8, 16, 16)) + 100).astype(np.uint16)
    ref = oracle_built.esrrf3d(x16.astype(np.float32), 2, 2, 1.5, 2.8, 4.0, 1, "std", True)
    op = eSRRF3D_ST(verbose=False)
    got = op.run(x16, magnification_xy=2, magnification_z=2, mode="std", run_type="HIP_gfx950")
    assert isinstance(got, np.ndarray) and got.dtype == np.float32
    assert _report("eSRRF3D_ST.run u16 std", got, ref) <= TOL
    # frame chunks of one or two frames carry the running statistics through the caller's buffers bit for bit
    for mode in ("average", "std"):
        whole = _hip.esrrf3d(x16, 2, 2, 1.5, 1.05, 1, mode, True).cpu().numpy()
        prev = _hip.esrrf3d_chunk_bytes(1)          # one frame per chunk
        try:
            chunked = _hip.esrrf3d(x16, 2, 2, 1.5, 1.05, 1, mode, True).cpu().numpy()
        finally:
            _hip.esrrf3d_chunk_bytes(prev)
        assert np.array_equal(whole, chunked), mode


def test_esrrf3d_workflow(torch, oracle_built):
    from nanopyx_amd import eSRRF3D
    x = _stack((6, 8, 16, 18), seed=9)
    got = eSRRF3D(x, frames_per_timepoint=4)
    assert got.shape == (2, 16, 32, 36) and got.dtype == np.float32
    for t, part in enumerate((x[:4], x[4:])):
        ref = oracle_built.macro_pixel(oracle_built.esrrf3d(part, 2, 2), 2)
        assert _report(f"eSRRF3D timepoint {t}", got[t], ref) <= TOL
    raw = eSRRF3D(x, macro_pixel_correction=False, mode="std")
    assert _report("eSRRF3D std", raw[0], oracle_built.esrrf3d(x, 2, 2, mode="std")) <= TOL


def test_esrrf3d_degenerate_gradients(torch, oracle_built):
    """A flat volume (G = 0: the unit gradient is stored as 0) and a ramp along z against the oracle."""
    from nanopyx_amd import _hip
    z = np.arange(8, dtype=np.float32)[:, None, None]
    vols = {"flat": np.full((8, 12, 14), 40.0, np.float32),
            "ramp_z": np.broadcast_to(3.0 * z + 10.0, (8, 12, 14)).astype(np.float32)}
    for name, v in vols.items():
        print(f"esrrf3d degenerate {name}", flush=True)
        x = np.stack([v, v * 1.2 + 1.0]).astype(np.float32)
        ref = oracle_built.esrrf3d(x, 2, 2, 1.5, 2.8, 4.0, 1, "average", True)
        got = _hip.esrrf3d(x, 2, 2, 1.5, 1.5 * 2.8 / 4.0, 1, "average", True).cpu().numpy()
        assert np.isfinite(got).all()
        if name == "flat":
            assert not np.abs(ref).any() and not np.abs(got).any()
        assert _report(f"esrrf3d degenerate {name}", got, ref) <= TOL
