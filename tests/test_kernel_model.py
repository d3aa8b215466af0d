"""CPU checks of the fused kernel's launch geometry and arithmetic before it runs on a GPU: the block-by-block
model in kernel_model.py repeats the kernel's index math on the library's own tables (asserting every index
stays in its buffer) and must reproduce the oracle."""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, rel_err
from kernel_model import model_frames, model_multi


@pytest.mark.parametrize("name", ["m5_r15_s1", "m2_r10_s3_noiw", "m3_r25_s17", "m8_r12_s2", "m1_r15_s1"])
def test_model_matches_golden(golden, name):
    M, R, S, iw = golden[f"{name}/params"]
    x = golden[f"{name}/input"][:1]
    out = model_frames(x, int(M), float(R), float(S), bool(iw))
    assert rel_err(out, golden[f"{name}/rgc_maps"][:1])[0] <= 2e-6


@pytest.mark.parametrize("shape,M,R,S", [
    ((1, 3, 5), 4, 2.0, 2.0),        # image smaller than the ring window
    ((1, 9, 65), 2, 3.4, 1.0),       # widest window (K = 6), two column regions
    ((1, 17, 3), 5, 0.4, 1.0),       # K = 0
    ((1, 1, 1), 3, 1.5, 1.0),        # one pixel
    ((1, 11, 7), 7, 1.7, 2.5),       # two phase groups (49 phases), non-integer sensitivity
])
@pytest.mark.parametrize("Q", [None, 8])
def test_model_edge_shapes(oracle_built, shape, M, R, S, Q):
    """Q=None: the per-frame kernel's rows per lane; Q=8: the temporal kernel's."""
    rng = np.random.default_rng(1)
    x = (rng.random(shape) * 100 + 1).astype(np.float32)
    out = model_frames(x, M, R, S, Q=Q)
    ref = oracle_built.esrrf(x, M, R, S)
    assert rel_err(out, ref)[0] <= 1e-5


@pytest.mark.parametrize("shape,M,radii,KL", [
    ((1, 10, 70), 2, [1.0, 1.5, 2.0, 2.5], None),  # sweep radii: K 2..5 on a lattice of K 5
    ((1, 9, 13), 3, [1.0, 2.0], 6),               # lattice wider than every ring (batched launch of K 6)
    ((1, 5, 4), 4, [0.4, 3.4], None),             # K 0 beside K 6
])
def test_model_shared_lattice(oracle_built, shape, M, radii, KL):
    """Batched/sweep blocks: one lattice of half-width KL, each wave reads its own ring at offset KL - K; every
    (radius, sensitivity) must match the oracle's single-parameter run."""
    rng = np.random.default_rng(2)
    x = (rng.random(shape) * 100 + 1).astype(np.float32)
    sens = [1.0, 2.0, 2.5]
    outs = model_multi(x, M, radii, sens, Q=8, KL=KL)
    for R in radii:
        for S in sens:
            ref = oracle_built.esrrf(x, M, R, S)
            assert rel_err(outs[(float(R), float(S))], ref)[0] <= 1e-5, (R, S)


def _group_rows(nphase, M, nwaves, nrad, group):
    rows = set()
    for w in range(nwaves):
        sw = group * nwaves + w
        rw = sw // nphase
        if rw >= nrad:
            break
        rows.add((sw - rw * nphase) // M)
    return rows


@pytest.mark.parametrize("M", list(range(1, 21)) + [32, 64])
def test_phase_row_sets_fit(M):
    """The fused block shares the intensity row pass between waves of one phase row and keeps at most 4 such rows:
    every launch shape the library makes (single job: nwaves = ceil(M^2 / ceil(M^2 / max_waves)) for each
    autotuner candidate; batched: 16 slots over up to 4 radii) must have at most 4 distinct phase rows per group."""
    nphase = M * M
    shapes = []
    for mw in (1, 2, 4, 8, 12, 16):
        ngroups = -(-nphase // mw)
        shapes.append((-(-nphase // ngroups), 1, ngroups))
    for nrad in (1, 2, 3, 4):
        shapes.append((16, nrad, -(-(nphase * nrad) // 16)))
    for nwaves, nrad, ngroups in shapes:
        for g in range(ngroups):
            assert len(_group_rows(nphase, M, nwaves, nrad, g)) <= 4, (M, nwaves, nrad, g)


@pytest.mark.parametrize("N,H,W,M,R,S,iw,pad", [
    (1, 20, 35, 2, 1.0, 3.0, False, False),
    (2, 10, 9, 5, 1.5, 1.0, True, False),
    (1, 16, 34, 4, 2.0, 2.0, True, True),
    (1, 10, 5, 7, 2.5, 1.7, True, True),      # 49 phases
])
def test_model_arrays_mode(oracle_built, N, H, W, M, R, S, iw, pad):
    """The kernel's arrays mode (npx_rgc on caller gradients) block by block: the lattice gathered from the x2M
    arrays and the caller's weight image reproduce the oracle's restatement of the reference loop."""
    from kernel_model import model_multi
    rng = np.random.default_rng(H * W + M)
    HM, WM = H * M, W * M
    gx = rng.normal(0, 10, (N, 2 * HM, 2 * WM)).astype(np.float32)
    gy = rng.normal(0, 10, (N, 2 * HM, 2 * WM)).astype(np.float32)
    img = (rng.random((N, HM, WM)) * 100 + 1).astype(np.float32)
    out = model_multi(np.zeros((N, H, W), np.float32), M, [R], [S], iw=iw, arrays=(gx, gy, img, pad))
    ref = oracle_built.rgc(gx, gy, img, M, R, S, iw, pad)
    assert rel_err(out[(float(R), float(S))], ref)[0] <= 1e-5


@pytest.mark.parametrize("M,R", [(2, 1.0), (4, 2.0), (8, 0.6), (3, 1.5), (5, 0.5), (7, 2.5)])
def test_block_ring_against_plan_tables(M, R):
    """The geometry a block forms from its rounded pixel centres against the library's host tables of the exact
    geometry: identical at power-of-two magnifications (every centre is exact in float32), at any magnification for
    the first block (centres below 64 carry rounding errors under 2^-18); far along a row at other magnifications the
    offsets move by the rounding of the centre, which the reference's maps carry too."""
    from nanopyx_amd import _hip
    from kernel_model import block_ring
    H = W = 1200
    t = _hip.plan_tables(H, W, M, R)
    d = t["info"]
    pow2 = M & (M - 1) == 0
    for xr, yr in ((63, 7), (1023, 15), (1100, 600)):
        br = block_ring(M, d["K"], xr, yr, d["tSS"], d["tSO"])
        exact = pow2 or (xr, yr) == (63, 7)
        tol = 2.0 ** -18 if exact else 1e-3
        for k in ("dy", "dx", "d", "t"):
            assert np.allclose(br[k], t[k], rtol=0, atol=tol * max(1.0, float(np.abs(t[k]).max()))), (k, xr, yr)
        # w ~ d^4 near the centre: a position error e moves w by about 4 e / d relative (d >= 0.1 here)
        assert np.allclose(br["wsum"], t["wsum"], rtol=tol * 100, atol=0)
        assert np.allclose(br["prefix"], t["prefix"], rtol=0, atol=tol * 100 * float(t["prefix"].max()))
        if pow2:
            for k in ("dy", "dx", "d", "t", "wsum", "prefix"):
                assert np.array_equal(br[k], t[k]), (k, xr, yr)
        elif xr > 64:
            assert not np.array_equal(br["dx"], t["dx"])


@pytest.mark.parametrize("M,R,W", [(7, 0.54, 300), (5, 0.5, 330), (3, 1.0, 520)])
def test_model_follows_reference_pixel_centres(oracle_built, M, R, W):
    """Rows wider than 256 at magnifications that are not powers of two: the reference rounds each pixel centre to
    float32, which moves its maps 5e-5 - 1.1e-4 from the exact geometry here; the blocks' rounded-centre geometry
    follows the reference (measured 5.6e-6 - 1.5e-5)."""
    from oracle import fp64
    rng = np.random.default_rng(3)
    x = (rng.random((1, 3, W)) * 100 + 1).astype(np.float32)
    out = model_frames(x, M, R, 2.0)
    ref = oracle_built.esrrf(x, M, R, 2.0)
    d = rel_err(fp64.esrrf_frames(x, M, R, 2.0, True), ref)[0]
    e = rel_err(out, ref)[0]
    assert e <= 2.5e-5 and e < d / 3, (e, d)
