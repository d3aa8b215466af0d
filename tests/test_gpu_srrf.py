"""SRRF radiality on the GPU (npx_radiality) against the oracle: the Radiality operator on caller arrays, the fused
SRRF() workflow (in-kernel magnification + macro-pixel correction), uint16 input, frame chunking, errors."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4
CASES = [  # (N, H, W, M, ringRadius, border, positivity, iw)
    (3, 20, 24, 2, 0.5, 0, True, True),
    (2, 18, 70, 5, 0.5, 0, True, True),
    (2, 16, 30, 3, 1.7, 1, False, False),
    (2, 21, 17, 4, 2.0, 0, True, True),
    (2, 12, 12, 1, 1.0, 2, True, False),
    (1, 40, 33, 8, 1.25, 0, False, True),
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
def test_radiality_matches_oracle(torch, oracle_built, case):
    from nanopyx_amd import _hip
    N, H, W, M, R, b, pos, iw = case
    print(f"radiality case {case}", flush=True)
    x = _stack((N, H, W), seed=M)
    interp = oracle_built.shift_magnify(x, M, M)
    ref = oracle_built.radiality(x, interp, M, R, b, pos, iw)
    got = _hip.radiality(x, interp if iw else None, M, R, b, pos, iw).cpu().numpy()
    assert got.shape == ref.shape
    assert _report(f"radiality {case}", got, ref) <= TOL
    # in-kernel magnification instead of the caller's image
    fused = _hip.radiality(x, None, M, R, b, pos, iw).cpu().numpy()
    assert _report(f"radiality fused {case}", fused, ref) <= TOL


def test_radiality_uint16_and_tensor(torch, oracle_built):
    from nanopyx_amd import _hip
    from nanopyx_amd.core.transform import Radiality
    rng = np.random.default_rng(5)
    x16 = (rng.poisson(300, size=(3, 24, 28)) + 100).astype(np.uint16)
    x = x16.astype(np.float32)
    interp = oracle_built.shift_magnify(x, 4, 4)
    ref = oracle_built.radiality(x, interp, 4, 0.5)
    got = _hip.radiality(torch.from_numpy(x16.astype(np.int32)).to(torch.uint16).cuda(), None, 4, 0.5, 0, True,
                         True)
    assert _report("radiality u16", got.cpu().numpy(), ref) <= TOL
    op = Radiality(verbose=False)
    out = op.run(x, interp, magnification=4, ringRadius=0.5, run_type="HIP_gfx950")
    assert isinstance(out, np.ndarray) and out.dtype == np.float32
    assert _report("Radiality.run", out, ref) <= TOL
    out_t = op.run(torch.from_numpy(x).cuda(), torch.from_numpy(interp).cuda(), magnification=4, ringRadius=0.5)
    assert out_t.is_cuda and np.array_equal(out_t.cpu().numpy(), out)


@pytest.mark.parametrize("mpc", [True, False])
def test_srrf_workflow_matches_oracle(torch, oracle_built, mpc):
    from nanopyx_amd import SRRF
    x = _stack((4, 26, 31), seed=9)
    ref = oracle_built.srrf_workflow(x, 5, 0.5, 0, True, True, mpc)
    got = SRRF(x, magnification=5, macro_pixel_correction=mpc)
    assert got.shape == ref.shape and got.dtype == np.float32
    assert _report(f"SRRF mpc={mpc}", got, ref) <= TOL


def test_radiality_chunking_and_size(torch, oracle_built):
    """A 64-frame 128 x 160 stack at M = 5 (frames split over block chunks) equals per-frame launches, and its
    first and last frames match the oracle."""
    from nanopyx_amd import _hip
    x = torch.from_numpy(_stack((64, 128, 160), seed=11)).cuda()
    full = _hip.radiality(x, None, 5, 0.5, 0, True, True)
    for f in (0, 31, 63):
        one = _hip.radiality(x[f:f + 1], None, 5, 0.5, 0, True, True)
        assert torch.equal(one[0], full[f]), f
    xs = x[[0, 63]].cpu().numpy()
    ref = oracle_built.radiality(xs, oracle_built.shift_magnify(xs, 5, 5), 5, 0.5)
    assert _report("radiality 128x160 frames 0, 63", full[[0, 63]].cpu().numpy(), ref) <= TOL


def test_radiality_errors(torch):
    from nanopyx_amd import _hip
    x = _stack((2, 10, 12))
    with pytest.raises(ValueError):
        _hip.radiality(x, np.zeros((2, 20, 20), np.float32), 2, 0.5, 0, True, True)   # wrong interp shape
    with pytest.raises(ValueError):
        _hip.radiality(x, None, 2, -0.5, 0, True, True)
    with pytest.raises(ValueError):
        _hip.radiality(x, None, 2, 0.5, -1, True, True)


@pytest.mark.parametrize("shift,mag", [((0.0, 0.0), (5.0, 5.0)), ((0.3, -1.2), (2.0, 3.0)), ((-0.7, 0.25), (2.5, 1.5)),
                                       ((0.0, 0.0), (1.0, 1.0))])
def test_cr_shift_magnify_matches_oracle(torch, oracle_built, shift, mag):
    from nanopyx_amd.core.transform import CRShiftAndMagnify
    x = _stack((3, 17, 23), seed=12)
    ref = oracle_built.shift_magnify(x, mag[0], mag[1], shift[0], shift[1])
    got = CRShiftAndMagnify(verbose=False).run(x, shift[0], shift[1], mag[0], mag[1], run_type="HIP_gfx950")
    assert got.shape == ref.shape and got.dtype == np.float32
    assert _report(f"CRShiftAndMagnify shift {shift} mag {mag}", got, ref) <= 1e-5


def test_reference_srrf_call_sequence(torch, oracle_built):
    """The reference's own SRRF test sequence: CRShiftAndMagnify x5, then Radiality on the stack and its
    magnification (uint16 camera counts)."""
    from nanopyx_amd.core.transform import CRShiftAndMagnify, Radiality
    rng = np.random.default_rng(13)
    x = (rng.poisson(120, size=(10, 20, 20)) + 100).astype(np.uint16)
    interp = CRShiftAndMagnify(verbose=False).run(x, 0, 0, 5, 5)
    rad = Radiality(verbose=False).run(x, interp)
    xf = x.astype(np.float32)
    ref = oracle_built.radiality(xf, oracle_built.shift_magnify(xf, 5, 5), 5, 0.5)
    assert _report("CRShiftAndMagnify + Radiality", rad, ref) <= TOL


def test_release_and_rebuild_caches(torch):
    """npx_release frees every cached plan (eSRRF, radiality, eSRRF 3D, Fourier interpolation, shift-magnify); the
    next calls rebuild them and give the same results."""
    from nanopyx_amd import _hip
    x = _stack((2, 20, 24), seed=14)
    x3 = _stack((2, 6, 16, 16), seed=15)

    def run_all():
        torch.cuda.synchronize()
        return [_hip.radiality(x, None, 2, 0.5, 0, True, True).cpu(), _hip.esrrf3d(x3, 2, 2, 1.5, 1.05, 1, "average",
                                                                                 True).cpu(),
                _hip.fht_interpolate(x, 2).cpu(), _hip.shift_magnify(x, 0.2, 0.1, 2, 3).cpu(),
                _hip.esrrf_frames(x, 2, 1.5, 1, True).cpu()]

    a = run_all()
    torch.cuda.synchronize()
    _hip.lib().npx_release()
    b = run_all()
    for u, v in zip(a, b):
        assert torch.equal(u, v)


def test_radiality_degenerate_gradients(torch, oracle_built):
    """Flat image (G = 0: Dk is 0 there, not 0 * inf), ramps and a step edge against the oracle. On a linear ramp the
    opposite ring samples cancel, so both sides are rounding noise around 0 (about 1e-7 of the intensity): those are
    held to an absolute bound of 1e-5 of the image's largest value instead of the relative metric."""
    from nanopyx_amd import _hip
    y, x = np.mgrid[0:20, 0:40].astype(np.float32)
    images = {"flat": np.full((20, 40), 50.0, np.float32), "ramp_x": 2.0 * x + 1.0, "ramp_diag": x - y + 60.0,
              "step": np.where(y < 10, 5.0, 90.0).astype(np.float32)}
    for name, img in images.items():
        print(f"radiality degenerate {name}", flush=True)
        st = np.stack([img, img + 7.0]).astype(np.float32)
        ref = oracle_built.radiality(st, oracle_built.shift_magnify(st, 5, 5), 5, 0.5)
        got = _hip.radiality(st, None, 5, 0.5, 0, True, True).cpu().numpy()
        assert np.isfinite(got).all()
        if name == "flat":
            assert not np.abs(ref).any() and not np.abs(got).any()
        if name.startswith("ramp"):
            diff = float(np.abs(got - ref).max())
            print(f"radiality degenerate {name}: max |diff| {diff:.3e}, max |ref| {np.abs(ref).max():.3e}", flush=True)
            assert diff <= 1e-5 * float(img.max())
        else:
            assert _report(f"radiality degenerate {name}", got, ref) <= TOL
