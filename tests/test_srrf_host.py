"""SRRF radiality on the host: the oracle against the reference's own C kernel, the gfx950 plan's ring and tables
against the reference's arithmetic (a NumPy evaluation of the kernel's table walk against the oracle), and closed
forms. No GPU needed."""
import numpy as np
import pytest

from oracle import ref32

CASES = [  # (N, H, W, M, ringRadius, border, positivity, iw)
    (2, 20, 24, 2, 0.5, 0, True, True),
    (2, 18, 22, 5, 0.5, 0, True, True),
    (1, 16, 30, 3, 1.7, 1, False, False),
    (2, 21, 17, 4, 2.0, 0, True, True),
    (1, 12, 12, 1, 1.0, 2, True, False),
]


def _stack(shape, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.poisson(50, size=shape) + rng.random(shape) * 20).astype(np.float32)


@pytest.fixture(scope="module")
def oracle():
    ref32.build()
    return ref32


@pytest.mark.parametrize("case", CASES)
def test_radiality_restatement_bit_exact_vs_reference(oracle, case):
    if not oracle.reference_srrf_available():
        pytest.skip("oracle/_ref/libnpx_ref_srrf.so is not built (reference sources absent)")
    N, H, W, M, R, b, pos, iw = case
    x = _stack((N, H, W), seed=M)
    interp = oracle.shift_magnify(x, M, M)
    mine = oracle.radiality(x, interp, M, R, b, pos, iw)
    oracle.use_reference_radiality(True)
    oracle.use_reference_kernels(True)
    try:
        ref = oracle.radiality(x, interp, M, R, b, pos, iw)
    finally:
        oracle.use_reference_radiality(False)
        oracle.use_reference_kernels(False)
    assert np.array_equal(mine, ref)


def test_radiality_closed_forms(oracle):
    # a constant image has zero gradients: every Dk is 0
    x = np.full((1, 12, 14), 7.0, np.float32)
    assert not oracle.radiality(x, oracle.shift_magnify(x, 3, 3), 3, 0.5).any()
    # a symmetric spot: gradients point at its centre, so the centre pixel's ring sees +Dk everywhere
    yy, xx = np.mgrid[0:15, 0:15]
    spot = np.exp(-((yy - 7) ** 2 + (xx - 7) ** 2) / (2 * 1.5 ** 2)).astype(np.float32)[None] * 100 + 10
    rad = oracle.radiality(spot, oracle.shift_magnify(spot, 2, 2), 2, 0.5, 0, True, False)
    # output pixels 14 and 15 have their centres at input 7.25 and 7.75, either side of the spot centre 7.5
    c = rad[0, 14:16, 14:16]
    assert c.min() > 0.3 and c.max() == rad[0].max() and np.ptp(c) < 1e-5
    # computed range: rows and columns [(1 + border) M, (H - 1 - border) M), zero elsewhere
    x = _stack((1, 10, 12), seed=3)
    rad = oracle.radiality(x, oracle.shift_magnify(x, 2, 2), 2, 0.5, 1, False, False)
    inside = np.zeros_like(rad, bool)
    inside[:, 4:16, 4:20] = True
    assert not rad[~inside].any() and np.count_nonzero(rad[inside]) > 0.9 * inside.sum()


def test_radiality_transpose(oracle):
    """The 12-sample ring is symmetric under x <-> y (30 degree steps), so transposing the input transposes the
    output up to the float32 rounding of the ring coordinates."""
    x = _stack((1, 16, 16), seed=4)
    a = oracle.radiality(x, oracle.shift_magnify(x, 3, 3), 3, 0.6)
    xt = np.ascontiguousarray(x.transpose(0, 2, 1))
    b = oracle.radiality(xt, oracle.shift_magnify(xt, 3, 3), 3, 0.6)
    assert np.abs(a - b.transpose(0, 2, 1)).max() <= 1e-4 * np.abs(a).max()


def test_radiality_plan_ring_matches_oracle(oracle):
    from nanopyx_amd import _hip
    for M, R in [(2, 0.5), (5, 0.5), (3, 1.7), (8, 1.25)]:
        d = _hip.radiality_info(20, 24, M, R, 0)
        x, y, r = oracle.radiality_ring(R, M)
        assert np.array_equal(d["xRing"], x) and np.array_equal(d["yRing"], y) and d["R"] == np.float32(r)


def _emulate(x, interp, M, R, border, pos, iw):
    """NumPy statement of the gfx950 kernel's table walk: gradient taps (bases + weights from the plan's tables),
    Dk, the sign rule and the positivity clamp, evaluated in float64."""
    from nanopyx_amd import _hip
    N, H, W = x.shape
    d = _hip.radiality_info(H, W, M, R, border)
    row, col, r0, c0 = _hip.radiality_tables(H, W, M, R, border)
    HM, WM = H * M, W * M
    rb = row[:, :12].view(np.int32) + r0[np.arange(HM) // d["TJ"]][:, None]
    cb = col[:, :12].view(np.int32) + c0[np.arange(WM) // d["TI"]][:, None]
    wy, wx = row[:, 12:60].reshape(HM, 12, 4), col[:, 12:60].reshape(WM, 12, 4)
    dy, dx = row[:, 60:72], col[:, 60:72]
    out = np.zeros((N, HM, WM), np.float32)
    j0, j1, i0, i1 = d["j_lo"], d["j_hi"], d["i_lo"], d["i_hi"]
    for f in range(N):
        img = x[f].astype(np.float64)
        gx = np.zeros((H, W))
        gy = np.zeros((H, W))
        gx[1:-1, 1:-1] = img[1:-1, 2:] - img[1:-1, :-2]
        gy[1:-1, 1:-1] = img[2:, 1:-1] - img[:-2, 1:-1]
        pad = 8
        gxp, gyp = np.pad(gx, pad), np.pad(gy, pad)
        J, I = np.meshgrid(np.arange(j0, j1), np.arange(i0, i1), indexing="ij")
        div = np.zeros(J.shape)
        for k in range(12):
            qx = np.zeros(J.shape)
            qy = np.zeros(J.shape)
            for u in range(4):
                for v in range(4):
                    rr = rb[J, k] + v + pad
                    cc = cb[I, k] + u + pad
                    w = wy[J, k, v] * wx[I, k, u]
                    qx += w * gxp[rr, cc]
                    qy += w * gyp[rr, cc]
            g = np.sqrt(qx * qx + qy * qy)
            cr = np.abs(qy * dx[I, k] - qx * dy[J, k])       # offsets are stored over the ring radius R
            with np.errstate(divide="ignore", invalid="ignore"):
                dk = np.where(g > 0, (1 - cr / g) ** 2, 0.0)
            dot = qx * dx[I, k] + qy * dy[J, k]
            div += np.where(dot < 0, -dk, dk)
        div /= 12
        v = np.maximum(div, 0) if pos else div
        if iw:
            v = v * interp[f, j0:j1, i0:i1]
        out[f, j0:j1, i0:i1] = v
    return out


@pytest.mark.parametrize("case", CASES)
def test_radiality_tables_reproduce_oracle(oracle, case):
    N, H, W, M, R, b, pos, iw = case
    x = _stack((N, H, W), seed=M + 1)
    interp = oracle.shift_magnify(x, M, M)
    ref = oracle.radiality(x, interp, M, R, b, pos, iw)
    emu = _emulate(x, interp, M, R, b, pos, iw)
    assert np.abs(emu - ref).max() <= 1e-5 * np.abs(ref).max()


def test_radiality_plan_shapes():
    """Plans build (every tap inside its tile) over small, ragged and large-ring shapes; bad ones are refused."""
    from nanopyx_amd import _hip
    for H, W, M, R, b in [(3, 3, 1, 0.5, 0), (5, 200, 2, 0.5, 0), (200, 7, 7, 3.0, 1), (33, 65, 8, 2.5, 0),
                          (64, 64, 4, 0.0, 0), (9, 11, 10, 1.0, 3)]:
        d = _hip.radiality_info(H, W, M, R, b)
        assert d["HM"] == H * M and d["j_lo"] == (1 + b) * M and d["i_hi"] == (W - 1 - b) * M
        assert d["bands"] == -(-H * M // d["TJ"]) and d["tiles"] == -(-W * M // d["TI"])
    with pytest.raises(ValueError):
        _hip.radiality_info(0, 5, 2, 0.5, 0)
    with pytest.raises(ValueError):
        _hip.radiality_info(5, 5, 2, -1.0, 0)
