"""Parity of the batched parameter sweep (npx_esrrf_sweep): every (magnification, radius, sensitivity) AVG map
of one launch against the oracle's single-parameter eSRRF + temporal mean, against the single-job fused kernel,
chunked against unchunked, and the hipGraph replay against the eager call."""
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


def _stack(shape, seed=3):
    rng = np.random.default_rng(seed)
    return (rng.poisson(50, size=shape) + rng.random(shape) * 20).astype(np.float32)


@pytest.mark.parametrize("shape,mags,radii,sens", [
    ((6, 20, 70), [2, 3], [1.0, 1.5, 2.0, 2.5], [1.0, 2.0, 2.5]),   # C4's radii, K 2..5 on one lattice
    ((5, 9, 13), [4], [0.4, 3.4], [1.0, 3.0]),                        # K 0 beside K 6
    ((4, 12, 16), [1, 5], [1.0, 1.2, 1.4, 1.6, 1.8], [1, 2, 3, 4, 5]),  # > 4 radii and > 4 sensitivities: job split
])
def test_sweep_matches_oracle(torch, oracle_built, shape, mags, radii, sens):
    from nanopyx_amd import _hip
    x = _stack(shape)
    print(f"sweep shape {shape} mags {mags} radii {radii} sens {sens}", flush=True)
    out = _hip.esrrf_sweep(x, mags, radii, sens, iw=True)
    torch.cuda.synchronize()
    for M in mags:
        for R in radii:
            for S in sens:
                maps = oracle_built.esrrf(x, M, R, S, True)
                ref = oracle_built.temporal(maps, "AVG")
                got = out.maps[(M, float(R), float(S))].cpu().numpy()
                assert _report(f"M{M} R{R} S{S}", got, ref) <= TOL


def test_sweep_equals_single_launches(torch):
    """Each sweep map equals the single-job fused kernel's AVG for that setting (same arithmetic)."""
    from nanopyx_amd import _hip
    x = _stack((40, 48, 80), seed=5)
    mags, radii, sens = [2, 4], [1.0, 2.0], [1.0, 2.0, 3.0]
    out = _hip.esrrf_sweep(x, mags, radii, sens, workspace=False)
    for M in mags:
        for R in radii:
            for S in sens:
                one = _hip.esrrf_temporal(x, M, R, S, True, ("AVG",), workspace=False)["AVG"]
                assert _report(f"sweep vs single M{M} R{R} S{S}", out.maps[(M, R, S)].cpu().numpy(),
                               one.cpu().numpy()) <= 1e-6


def test_sweep_every_half_width_body(torch):
    """M4 and M8 (phases filling whole blocks: one job per radius) at half-widths K 0..6, 16 jobs in two launches:
    each job runs the batched body built for its K; every map against the single-job kernel for its setting (the
    sweep's mean takes the weighted value unrounded, so within 1e-6 rather than bit for bit)."""
    from nanopyx_amd import _hip
    x = _stack((12, 24, 40), seed=9)
    mags, radii, sens = [4, 8], [0.4, 0.6, 1.0, 1.5, 2.0, 2.5, 3.0, 3.4], [1.0, 2.0]
    print("half-widths", [_hip.plan_info(24, 40, 4, R)["K"] for R in radii], flush=True)
    out = _hip.esrrf_sweep(x, mags, radii, sens, workspace=False)
    for M in mags:
        for R in radii:
            for S in sens:
                one = _hip.esrrf_temporal(x, M, R, S, True, ("AVG",), workspace=False)["AVG"]
                assert _report(f"sweep vs single M{M} R{R} S{S}", out.maps[(M, R, S)].cpu().numpy(),
                               one.cpu().numpy()) <= 1e-6


def test_sweep_chunked_equals_unchunked(torch):
    from nanopyx_amd import _hip
    x = _stack((200, 32, 64), seed=6)
    mags, radii, sens = [2, 4], [1.5, 2.5], [1.0, 2.0]
    assert _hip.sweep_workspace_bytes(200, 32, 64, mags, radii, sens) > 0
    a = _hip.esrrf_sweep(x, mags, radii, sens, workspace=True)
    b = _hip.esrrf_sweep(x, mags, radii, sens, workspace=False)
    for k in a.key:
        assert _report(f"chunked vs single chunk {k}", a.maps[k].cpu().numpy(), b.maps[k].cpu().numpy()) <= 1e-5


def test_sweep_c4_full_size(torch):
    """The 32-setting sweep at 200 x 256 x 256 against single fused launches of a few of its settings."""
    from nanopyx_amd import _hip, synthetic
    sc = synthetic.CONFIGS["sweep_256"]
    x = torch.from_numpy(synthetic.generate(sc)).cuda()
    e = sc.esrrf
    mags, radii, sens = e["magnification"], e["radius"], e["sensitivity"]
    out = _hip.esrrf_sweep(x, mags, radii, sens)
    for (M, R, S) in [(2, 1.0, 1.0), (4, 2.5, 4.0), (4, 1.5, 2.0), (2, 2.0, 3.0)]:
        one = _hip.esrrf_temporal(x, M, R, S, True, ("AVG",))["AVG"]
        got = out.maps[(M, float(R), float(S))]
        assert got.shape == (256 * M, 256 * M)
        assert _report(f"C4 M{M} R{R} S{S} vs single", got.cpu().numpy(), one.cpu().numpy()) <= 1e-5


def test_sweep_graph_replays(torch):
    """A captured sweep replayed three times, with eager work between replays, equals the eager sweep."""
    from nanopyx_amd import _hip
    x = _stack((64, 24, 70), seed=7)
    xd = torch.from_numpy(x).cuda()
    mags, radii, sens = [2, 3], [1.0, 2.0], [1.0, 2.0]
    ref = _hip.esrrf_sweep(xd, mags, radii, sens)
    torch.cuda.synchronize()
    ref = {k: v.cpu().numpy() for k, v in ref.maps.items()}
    outs = _hip.SweepOutputs(torch, xd, mags, radii, sens, workspace=True)
    with _hip.Graph.capture() as g:
        _hip.esrrf_sweep(xd, mags, radii, sens, stream=g.stream, outputs=outs)
    for rep in range(3):
        for v in outs.maps.values():
            v.fill_(-1.0)
        g.launch()
        torch.cuda.synchronize()
        for k, v in outs.maps.items():
            assert np.array_equal(v.cpu().numpy(), ref[k]), (rep, k)
        _hip.esrrf_sweep(xd, mags, radii, sens)  # eager work between replays
        torch.cuda.synchronize()
    g.close()


def test_sweep_errors(torch):
    from nanopyx_amd import _hip
    x = _stack((2, 8, 8))
    with pytest.raises(ValueError):
        _hip.esrrf_sweep(x, [], [1.0], [1.0])
    with pytest.raises(ValueError):
        _hip.esrrf_sweep(x, [2], [1.0], [])


def test_sweep_outputs_reuse_checked(torch):
    """Outputs (map pointers, workspace) made for one stack shape and setting list are refused for another before
    anything launches."""
    from nanopyx_amd import _hip
    x = torch.from_numpy(_stack((12, 16, 24))).cuda()
    outs = _hip.SweepOutputs(torch, x, [2], [1.0, 1.5], [1.0, 2.0], workspace=True)
    _hip.esrrf_sweep(x, [2], [1.0, 1.5], [1.0, 2.0], outputs=outs)
    with pytest.raises(ValueError, match="outputs were made"):
        _hip.esrrf_sweep(torch.from_numpy(_stack((24, 16, 24))).cuda(), [2], [1.0, 1.5], [1.0, 2.0], outputs=outs)
    with pytest.raises(ValueError, match="outputs were made"):
        _hip.esrrf_sweep(x, [2], [1.0, 1.5, 2.0], [1.0, 2.0], outputs=outs)
    with pytest.raises(ValueError, match="outputs were made"):
        _hip.esrrf_sweep(x, [2, 4], [1.0, 1.5], [1.0, 2.0], outputs=outs)
    torch.cuda.synchronize()


def test_parameter_sweep_class(torch, oracle_built):
    from nanopyx_amd.core.analysis import ParameterSweep
    x = _stack((5, 16, 20), seed=8)
    ps = ParameterSweep()
    rec = ps.reconstructions(x, 2, [1, 2], [1.0, 1.5])
    assert rec[2].shape == (2, 2, 32, 40)
    for si, S in enumerate([1, 2]):
        for ri, R in enumerate([1.0, 1.5]):
            ref = oracle_built.temporal(oracle_built.esrrf(x, 2, R, S, True), "AVG")
            assert _report(f"ParameterSweep[{si},{ri}]", rec[2][si, ri], ref) <= TOL
    tac = ps.reconstructions(x, 2, [2], [1.5], temporal_correlation="TAC2")[2][0, 0]
    ref = oracle_built.temporal(oracle_built.esrrf(x, 2, 1.5, 2, True), "TAC2")
    assert _report("ParameterSweep TAC2", tac, ref) <= TOL


def test_parameter_sweep_qnr(torch, oracle_built):
    """QnR of the GPU sweep equals the same scoring (error map RSP, FIRE of odd/even frames, logistic harmonic mean)
    applied to the oracle's reconstructions, and run_esrrf_parameter_sweep picks its argmax."""
    from nanopyx_amd.core.analysis import ParameterSweep
    from nanopyx_amd import run_esrrf_parameter_sweep
    x = _stack((10, 24, 28), seed=9)
    sens, radii = [1, 2], [1.0, 1.5, 2.0]
    qnr = ParameterSweep().run(x, 2, sens, radii)
    assert qnr.shape == (2, 3) and np.isfinite(qnr).all()
    ps = ParameterSweep()
    rsp = np.zeros((2, 3))
    frc = np.zeros((2, 3))
    for si, S in enumerate(sens):
        for ri, R in enumerate(radii):
            full = oracle_built.temporal(oracle_built.esrrf(x, 2, R, S, True), "AVG")
            odd = oracle_built.temporal(oracle_built.esrrf(np.ascontiguousarray(x[1::2]), 2, R, S, True), "AVG")
            even = oracle_built.temporal(oracle_built.esrrf(np.ascontiguousarray(x[::2]), 2, R, S, True), "AVG")
            rsp[si, ri] = ps.calculate_rsp(x, full)
            frc[si, ri] = ps.calculate_frc(odd, even)
    expect = ps.calculate_qnr_score(rsp, frc)
    print("QnR gpu", np.round(qnr, 5).tolist(), "oracle", np.round(expect, 5).tolist())
    assert np.allclose(qnr, expect, rtol=1e-3, atol=1e-4)
    q2 = run_esrrf_parameter_sweep(x, 2, sens, radii, return_qnr=True)
    assert np.allclose(q2, qnr, rtol=1e-6, atol=1e-7)
    best = run_esrrf_parameter_sweep(x, 2, sens, radii)
    si, ri = np.unravel_index(np.argmax(qnr), qnr.shape)
    assert best == (sens[si], radii[ri])



def test_parameter_sweep_decorr_and_slices(torch, oracle_built):
    """use_decorr=True scores the all-frames reconstruction by its decorrelation resolution; with n_frames only the
    last (sensitivity, radius) pair keeps a non-zero sliced RSP, as in the reference. An emitter scene (noise-only
    reconstructions have no finite decorrelation resolution)."""
    from nanopyx_amd import synthetic
    from nanopyx_amd.core.analysis import DecorrAnalysis, ParameterSweep
    sc = synthetic.Scene(frames=12, height=48, width=56, emitters=60, sigma=1.5, photons=500, p_on=0.3,
                         background=100, seed=14)
    x = synthetic.generate(sc)
    sens, radii = [1, 1.5], [1.0, 1.5, 2.0]
    ps = ParameterSweep()
    qnr = ps.run(x, 2, sens, radii, use_decorr=True)
    rsp = np.zeros((2, 3))
    res = np.zeros((2, 3))
    frc = np.zeros((2, 3))
    for si, S in enumerate(sens):
        for ri, R in enumerate(radii):
            full = oracle_built.temporal(oracle_built.esrrf(x, 2, R, S, True), "AVG")
            rsp[si, ri] = ps.calculate_rsp(x, full)
            d = DecorrAnalysis()
            d.run_analysis(full)
            res[si, ri] = d.resolution
            odd = oracle_built.temporal(oracle_built.esrrf(np.ascontiguousarray(x[1::2]), 2, R, S, True), "AVG")
            even = oracle_built.temporal(oracle_built.esrrf(np.ascontiguousarray(x[::2]), 2, R, S, True), "AVG")
            frc[si, ri] = ps.calculate_frc(odd, even)
    assert np.isfinite(res).all()
    expect = ps.calculate_qnr_score(rsp, res)
    print("decorr QnR gpu", np.round(qnr, 5).tolist(), "oracle", np.round(expect, 5).tolist(), flush=True)
    assert np.allclose(qnr, expect, rtol=1e-3, atol=1e-4)
    sliced = ps.run(x, 2, sens, radii, n_frames=4)
    assert np.count_nonzero(sliced) == 1 and sliced[1, 2] > 0
    per = [ps.calculate_rsp(x[i:i + 4], oracle_built.temporal(oracle_built.esrrf(
        np.ascontiguousarray(x[i:i + 4]), 2, 2.0, 1.5, True), "AVG")) for i in (0, 4, 8)]
    r = np.zeros((2, 3))
    r[1, 2] = np.mean(per)
    assert np.allclose(sliced, ps.calculate_qnr_score(r, frc), rtol=1e-3, atol=1e-4)


def test_parameter_sweep_reference_call(torch):
    """The reference's own sweep test call: float64 noise, two-frame slices, decorrelation scoring."""
    from nanopyx_amd import run_esrrf_parameter_sweep
    img = np.random.default_rng(11).random((10, 100, 100))
    s, r = run_esrrf_parameter_sweep(img, n_frames=2, use_decorr=True, plot_sweep=False)
    assert s in (1, 2) and r in (1, 1.5)
