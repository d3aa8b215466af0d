"""Radii whose ring window is wider than the fused kernels' (int(2 R) > 6, i.e. R >= 3.5): the reference accepts any
radius (_c_sr_radial_gradient_convergence.c:27-28), and this build runs them through the reference's chain of
separate steps on the GPU (_le_esrrf.pyx:190-194: Catmull-Rom xM, Roberts cross, Catmull-Rom x2M of both gradients,
the per-pixel RGC on those arrays). Per-frame maps, the temporal correlations, a sweep mixing narrow and wide radii,
the eSRRF() workflow and a caller-given stream, against the oracle's restatement of that chain."""
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


def _stack(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.poisson(60, size=shape) + rng.random(shape) * 30).astype(np.float32)


def test_half_width_rule():
    from nanopyx_amd import _hip
    assert [_hip.ring_half_width(r) for r in (3.4, 3.49, 3.5, 3.6, 4.3, 5.1)] == [6, 6, 7, 7, 8, 10]


@pytest.mark.parametrize("shape,M,R,S,iw", [
    ((3, 24, 30), 2, 3.6, 1.0, True),    # K 7
    ((2, 20, 22), 3, 4.3, 2.0, False),   # K 8
    ((2, 16, 20), 4, 5.1, 1.5, True),    # K 10, general power
])
def test_frames_vs_oracle(torch, oracle_built, shape, M, R, S, iw):
    from nanopyx_amd.core.transform._le_esrrf import eSRRF
    x = _stack(shape, seed=shape[1] + M)
    print(f"wide radius frames {shape} M{M} R{R} S{S} iw={iw}", flush=True)
    out = eSRRF(verbose=False).run(x, magnification=M, radius=R, sensitivity=S, doIntensityWeighting=iw)
    ref = oracle_built.esrrf(x, M, R, S, iw)
    assert out.shape == ref.shape
    assert _report(f"frames M{M} R{R}", out, ref) <= TOL


def test_temporal_vs_oracle(torch, oracle_built):
    from nanopyx_amd import _hip
    x = _stack((12, 20, 24), seed=7)
    M, R, S = 2, 3.6, 2.0
    got = _hip.esrrf_temporal(x, M, R, S, True, ("AVG", "VAR", "TAC2"))
    maps = oracle_built.esrrf(x, M, R, S, True)
    for m in ("AVG", "VAR", "TAC2"):
        assert _report(f"temporal {m} R{R}", got[m].cpu().numpy(), oracle_built.temporal(maps, m)) <= TOL


def test_sweep_mixed_radii(torch, oracle_built):
    """Narrow radii through the batched launch, the wide one through the composed chain, in one call."""
    from nanopyx_amd import _hip
    x = _stack((6, 16, 24), seed=11)
    mags, radii, sens = [2, 3], [1.5, 3.6], [1.0, 2.0]
    out = _hip.esrrf_sweep(x, mags, radii, sens)
    for M in mags:
        for S in sens:
            wide = out.maps[(M, 3.6, S)].cpu().numpy()
            ref = oracle_built.temporal(oracle_built.esrrf(x, M, 3.6, S, True), "AVG")
            assert _report(f"sweep wide M{M} S{S}", wide, ref) <= TOL
            one = _hip.esrrf_temporal(x, M, 1.5, S, True, ("AVG",), workspace=False)["AVG"].cpu().numpy()
            assert _report(f"sweep narrow M{M} S{S}", out.maps[(M, 1.5, S)].cpu().numpy(), one) <= 1e-6


def test_workflow_vs_oracle(torch, oracle_built):
    from nanopyx_amd.methods.esrrf.eSRRF_workflow import eSRRF as esrrf_workflow
    x = _stack((8, 20, 20), seed=13)
    out = esrrf_workflow(x, magnification=2, radius=3.6, sensitivity=1, frames_per_timepoint=4,
                         temporal_correlation="AVG", macro_pixel_correction=True)
    ref = oracle_built.esrrf_workflow(x, magnification=2, radius=3.6, sensitivity=1, frames_per_timepoint=4,
                                      temporal_correlation="AVG", macro_pixel_correction=True)
    assert out.shape == ref.shape
    assert _report("workflow R3.6", out, ref) <= TOL


def test_explicit_stream(torch):
    """The composed chain on a caller-given stream gives the maps of the same call on the current stream."""
    from nanopyx_amd import _hip
    x = _stack((4, 16, 20), seed=17)
    ref = _hip.esrrf_temporal(x, 2, 3.6, 1.0, True, ("AVG", "TAC2"))
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    got = _hip.esrrf_temporal(x, 2, 3.6, 1.0, True, ("AVG", "TAC2"), stream=s)
    torch.cuda.current_stream().wait_stream(s)
    for k in ("AVG", "TAC2"):
        assert torch.equal(got[k], ref[k]), k
