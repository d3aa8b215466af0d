"""The oracle is pinned before it is trusted: the FP32 restatement against the reference's own compiled C
kernels (bit-exact), both against the golden vectors, the FP64 statement against the FP32 one, NumPy's own
temporal functions, closed forms of Catmull-Rom and Roberts, and the invariants of the method."""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, rel_err
from oracle import fp64


def _case(golden, name):
    M, R, S, iw = golden[f"{name}/params"]
    return golden[f"{name}/input"], int(M), float(R), float(S), bool(iw)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_restatement_matches_golden_bit_exact(golden, oracle_built, name):
    ref32 = oracle_built
    x, M, R, S, iw = _case(golden, name)
    xf = x.astype(np.float32)
    gc, gr = ref32.roberts(xf)
    assert np.array_equal(gc, golden[f"{name}/roberts_gc"])
    assert np.array_equal(gr, golden[f"{name}/roberts_gr"])
    assert np.array_equal(ref32.shift_magnify(xf, float(M), float(M)), golden[f"{name}/magnified"])
    maps = ref32.esrrf(xf, M, R, S, iw)
    # same kernels, same flags -> identical bits; a different FMA contraction on another host is a 1-ulp matter
    assert rel_err(maps, golden[f"{name}/rgc_maps"])[0] <= 1e-6
    for mode in ("AVG", "VAR", "TAC2"):
        assert rel_err(ref32.temporal(golden[f"{name}/rgc_maps"], mode), golden[f"{name}/{mode}"])[0] <= 1e-6


def test_restatement_bit_exact_vs_reference_kernels(oracle_built):
    ref32 = oracle_built
    if not ref32.reference_available():
        pytest.skip("oracle/_ref/libnpx_ref.so not built (reference sources absent)")
    rng = np.random.default_rng(7)
    for M, R, S in [(5, 1.5, 1.0), (4, 2.0, 2.0), (3, 2.5, 1.3), (2, 1.0, 4.0)]:
        x = (rng.random((2, 11, 13)) * 300).astype(np.float32)
        a = ref32.esrrf(x, M, R, S)
        ref32.use_reference_kernels(True)
        try:
            b = ref32.esrrf(x, M, R, S)
        finally:
            ref32.use_reference_kernels(False)
        assert np.array_equal(a, b)


def test_numpy_temporal_emulation_bit_exact(oracle_built):
    ref32 = oracle_built
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((37, 9, 11)) * 50 + 200).astype(np.float32)
    assert np.array_equal(ref32.temporal(x, "AVG"), np.mean(x, axis=0))
    assert np.array_equal(ref32.temporal(x, "VAR"), np.var(x, axis=0))
    c = x - np.mean(x, axis=0)
    assert np.array_equal(ref32.temporal(x, "TAC2"), np.mean(c[:-1] * c[1:], axis=0))


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fp64_oracle_matches_golden(golden, name):
    x, M, R, S, iw = _case(golden, name)
    maps = fp64.esrrf_frames(x, M, R, S, iw)
    assert rel_err(maps, golden[f"{name}/rgc_maps"])[0] <= 2e-5
    for mode in ("AVG", "VAR", "TAC2"):
        ref = golden[f"{name}/{mode}"]
        assert rel_err(fp64.temporal(maps, mode), ref)[0] <= 1e-4


def test_fp64_rgc_from_arrays_matches_golden(golden):
    name = "m2_r10_s3_noiw"
    gx, gy, mag = (golden[f"{name}/{k}"] for k in ("grad_col_interp", "grad_row_interp", "magnified"))
    out = fp64.rgc_from_arrays(gx, gy, mag, 2, 1.0, 3.0, False)
    assert rel_err(out, golden[f"{name}/rgc_maps"])[0] <= 2e-5
    out_pad = fp64.rgc_from_arrays(gx, gy, mag, 2, 1.0, 3.0, True, pad_edges=True)
    assert rel_err(out_pad, golden[f"{name}/rgc_pad_edges"])[0] <= 2e-5


def test_catmull_rom_closed_forms():
    # weights at integer coordinates are [-1, 9, 9, -1]/16 and the exact pixel value at n + 1/2
    assert np.allclose(fp64.cubic(np.array([1.5, 0.5, -0.5, -1.5])), [-1 / 16, 9 / 16, 9 / 16, -1 / 16])
    assert np.allclose(fp64.cubic(np.array([1.0, 0.0, -1.0, -2.0])), [0, 1, 0, 0])
    rng = np.random.default_rng(0)
    img = rng.random((7, 9))
    half = fp64.cr_sample(img, np.arange(7) + 0.5, np.arange(9) + 0.5)
    assert np.allclose(half, img, atol=1e-14)
    # linear ramps are reproduced away from the zero-padded border
    ramp = np.add.outer(np.arange(12) * 0.5, np.arange(12) * 2.0) + 3
    up = fp64.magnify(ramp, 4)
    r = np.arange(48) / 4
    exp = np.add.outer((r - 0.5) * 0.5, (r - 0.5) * 2.0) + 3
    assert np.allclose(up[10:38, 10:38], exp[10:38, 10:38], atol=1e-12)


def test_roberts_closed_form():
    img = np.arange(20, dtype=np.float64).reshape(4, 5) ** 2
    gc, gr = fp64.roberts(img)
    for r in range(4):
        for c in range(5):
            r0, c0 = max(r - 1, 0), max(c - 1, 0)
            assert gc[r, c] == img[r0, c] - img[r, c0] + img[r, c] - img[r0, c0]
            assert gr[r, c] == -img[r0, c] + img[r, c0] + img[r, c] - img[r0, c0]


def test_fp64_brute_force_single_pixel():
    """One pixel of the vectorised RGC against a literal loop over the reference's window."""
    rng = np.random.default_rng(5)
    H, W, M, R = 6, 7, 3, 1.5
    img = rng.random((H, W)) * 100
    p = fp64.ref_params(R, M)
    gc, gr = fp64.roberts(img)
    n_a, n_b = 2 * H + 8, 2 * W + 8
    GX, GY = fp64._lattice_gradients(gc, gr, H, W, M, n_a, n_b)
    out = fp64.rgc_frame_from_lattice(GX, GY, H * M, W * M, M, R, 1.0)
    rM, cM = 9, 11
    yc, xc = (rM + 0.5) / M, (cM + 0.5) / M
    num = den = 0.0
    for j in range(p["start"], p["end"]):
        vy = (int(2 * yc) + j) / 2
        if not (0 < vy <= H * M - 1):
            continue
        for i in range(p["start"], p["end"]):
            vx = (int(2 * xc) + i) / 2
            if not (0 < vx <= W * M - 1):
                continue
            dx, dy = vx - xc, vy - yc
            d = np.hypot(dx, dy)
            if d == 0 or d > p["tSO"]:
                continue
            gx = fp64.cr_sample(gc, [min(vy * 2 * M, 2 * H * M - 1) / (2 * M)], [min(vx * 2 * M, 2 * W * M - 1) / (2 * M)])[0, 0]
            gy = fp64.cr_sample(gr, [min(vy * 2 * M, 2 * H * M - 1) / (2 * M)], [min(vx * 2 * M, 2 * W * M - 1) / (2 * M)])[0, 0]
            dw = (d * np.exp(-d * d / p["tSS"])) ** 4
            den += dw
            if gx * dx + gy * dy < 0:
                num += (1 - abs(gy * dx - gx * dy) / np.hypot(gx, gy) / d) * dw
    assert abs(out[rM, cM] - num / den) < 1e-12


def test_fp64_invariants():
    rng = np.random.default_rng(11)
    x = rng.random((2, 10, 12)) * 100 + 10
    # constant image -> all zeros
    assert np.all(fp64.esrrf_frames(np.full((1, 8, 9), 42.0), 4, 1.5, 1) == 0)
    # transpose equivariance and scale equivariance (alpha > 0)
    a = fp64.esrrf_frames(x, 3, 1.5, 2)
    b = fp64.esrrf_frames(np.transpose(x, (0, 2, 1)), 3, 1.5, 2)
    assert np.allclose(np.transpose(b, (0, 2, 1)), a, rtol=1e-10, atol=1e-10)
    c = fp64.esrrf_frames(3.0 * x, 3, 1.5, 2)
    assert np.allclose(c, 3.0 * a, rtol=1e-10, atol=1e-10)
    # zero margin
    m = fp64.ref_params(1.5, 3)["margin"]
    assert np.all(a[:, :m] == 0) and np.all(a[:, :, :m] == 0) and np.all(a[:, -m:] == 0)
    # AVG / VAR are permutation invariant, TAC2 is not
    st = rng.random((6, 5, 5))
    perm = st[[3, 1, 5, 0, 2, 4]]
    assert np.allclose(fp64.temporal(st, "AVG"), fp64.temporal(perm, "AVG"))
    assert np.allclose(fp64.temporal(st, "VAR"), fp64.temporal(perm, "VAR"))
    assert not np.allclose(fp64.temporal(st, "TAC2"), fp64.temporal(perm, "TAC2"))


def test_macro_pixel_closed_form(oracle_built):
    ref32 = oracle_built
    M = 3
    phase = np.arange(1, M * M + 1, dtype=np.float64).reshape(M, M)
    img = np.tile(phase, (4, 5))[None] * 2.0
    out = fp64.macro_pixel(img, M)
    assert np.allclose(out, 1.0)
    assert np.allclose(ref32.macro_pixel(img.astype(np.float32), M), 1.0)
    # a zero phase stays untouched
    img2 = img.copy()
    img2[0, ::M, ::M] = 0
    assert np.all(fp64.macro_pixel(img2, M)[0, ::M, ::M] == 0)


@pytest.mark.parametrize("name,M", [("mpc5", 5), ("mpc4", 4), ("mpc2", 2)])
def test_macro_pixel_matches_reference_vectors(oracle_built, name, M):
    """The restatement of macro_pixel_corrector against the reference's own compiled function
    (core/transform/mpcorrector.pyx:7-73, built by tests/golden/make_cython_golden.py): bit for bit, including phases
    whose mean is 0 (left undivided)."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cython_golden.npz"))
    got = oracle_built.macro_pixel(g[f"{name}_in"], M)
    assert np.array_equal(got.reshape(g[f"{name}_out"].shape), g[f"{name}_out"])


def test_temporal_unknown_mode_raises():
    with pytest.raises(ValueError):
        fp64.temporal(np.zeros((2, 2, 2)), "XYZ")
