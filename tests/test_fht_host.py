"""The oracle's restatement of the use_fht Fourier interpolation (oracle/fht.py), pinned by closed forms and by the
reference's own interpolation.py (tests/golden/fht_golden.npz)."""
import os

import numpy as np
import pytest

from oracle import fht


@pytest.mark.parametrize("k,M", [(1, 2), (3, 3), (5, 4), (7, 5)])
def test_fht_band_limited_cosine(k, M):
    """A 32 x 32 image (a power of two, so mirror padding adds nothing) of 2 + cos(2 pi k x / 32) is interpolated to
    2 + cos(2 pi k (x / M) / 32) on the M-times finer grid; the rescale keeps [1, 3], which both grids reach."""
    n = 32
    x = np.arange(n)
    img = (2 + np.cos(2 * np.pi * k * x / n))[None, :].repeat(n, 0).astype(np.float32)
    out = fht.fht_2d(img, M)
    xf = np.arange(n * M) / M
    expect = (2 + np.cos(2 * np.pi * k * xf / n))[None, :].repeat(n * M, 0)
    assert out.shape == (n * M, n * M) and np.abs(out - expect).max() < 2e-5


def test_fht_constant_and_mirror_layout():
    # a constant image: only the DC term, scaled by 1 / M^2 by the larger inverse transform, and no rescale
    assert np.allclose(fht.fht_2d(np.full((20, 24), 4.0, np.float32), 2), 1.0)
    # mirror padding of a 3 x 5 image to 8 x 8: the image in the middle, flipped copies around it
    small = np.arange(15, dtype=np.float32).reshape(3, 5)
    pad = fht._mirror_pad(small)
    full = np.zeros((9, 15), np.float32)
    for th in range(3):
        for tw in range(3):
            t = small[::-1] if th % 2 == 0 else small
            t = t[:, ::-1] if tw % 2 == 0 else t
            full[th * 3:(th + 1) * 3, tw * 5:(tw + 1) * 5] = t
    assert np.array_equal(pad, full[0:8, 3:11])


def test_fht_tiles():
    assert fht.tile_starts(512, 256, 16) == [0, 224, 448]
    assert fht.tile_starts(300, 256, 16) == [0, 224]
    assert fht.tile_starts(700, 256, 16) == [0, 224, 448, 672]
    assert fht.tile_starts(100, 64, 8) == [0, 48, 96]
    # a constant frame larger than a tile: every tile gives c / M^2, and the Hann blend of equal tiles is that value
    out = fht.fht_space_interpolation(np.full((300, 270), 9.0, np.float32), 3)
    assert out.shape == (900, 810) and np.allclose(out, 1.0, atol=1e-6)
    # frames no larger than the tile take the whole-frame path; mirror padding can be switched off there only
    rng = np.random.default_rng(0)
    x = rng.random((2, 40, 50)).astype(np.float32)
    assert np.array_equal(fht.fht_space_interpolation(x, 2)[1], fht.fht_2d(x[1], 2, True))
    assert np.array_equal(fht.fht_space_interpolation(x, 2, doMirrorPadding=False)[0], fht.fht_2d(x[0], 2, False))


def _rows_inverse_model(X, Ww, M, k_term=2.0):
    """The row inverse as the fht_rows kernel (nanopyx_amd/csrc/npx_fht.hip) computes it, vectorised: rows of
    X[k], k <= K = Ww/2 (zero beyond, up to N/2, N = Ww M) -> irfft's unnormalised values. Phase r of n = M m + r is
    the length-Ww inverse DFT over m of G_r (G_r[0] = Re X[0], G_r[k] = Y_r[k], G_r[Ww-k] = conj(Y_r[k]) for
    0 < k < K, G_r[K] = 2 Re Y_r[K], Y_r[k] = X[k] e^{2 pi i k r / N}); phases 2p and 2p+1 share one complex FFT
    (G_2p + i G_2p+1), run as mixed radix-4/2 Stockham stages in natural order; twiddles e^{2 pi i j / N} are formed
    from a 64-entry fine and a coarse table as the kernel's LDS tables hold them."""
    rows = X.shape[0]
    N, K, P = Ww * M, Ww // 2, (M + 1) // 2
    j = np.arange(K + 1)[None, :] * np.arange(M)[:, None]                       # [r, k]: k r
    tw = np.exp(2j * np.pi * 64 * (j >> 6) / N) * np.exp(2j * np.pi * (j & 63) / N)
    Y = X[:, None, :] * tw[None]                                               # [row, r, k]
    G = np.zeros((rows, M, Ww), np.complex128)
    G[:, :, 0] = X[:, None, 0].real
    G[:, :, 1:K] = Y[:, :, 1:K]
    G[:, :, Ww - np.arange(1, K)] = np.conj(Y[:, :, 1:K])
    G[:, :, K] = k_term * Y[:, :, K].real
    Z = np.zeros((rows, P, Ww), np.complex128)
    Z += G[:, 0::2]
    Z[:, :M // 2] += 1j * G[:, 1::2]
    Ns = 1
    while Ns < Ww:                                                             # Stockham: inputs j + r Ww/R
        R = 4 if 4 * Ns <= Ww else 2
        nb = Ww // R
        jj = np.arange(nb)
        k = jj % Ns
        v = [Z[..., jj + r * nb] * np.exp(2j * np.pi * r * k / (Ns * R)) for r in range(R)]
        if R == 4:
            a0, a1, b0, b1 = v[0] + v[2], v[0] - v[2], v[1] + v[3], 1j * (v[1] - v[3])
            out = [a0 + b0, a1 + b1, a0 - b0, a1 - b1]
        else:
            out = [v[0] + v[1], v[0] - v[1]]
        Zn = np.empty_like(Z)
        idx = (jj // Ns) * Ns * R + k
        for r in range(R):
            Zn[..., idx + r * Ns] = out[r]
        Z, Ns = Zn, Ns * R
    x = np.empty((rows, N))
    for r in range(M):
        x[:, r::M] = Z[:, r // 2].imag if r % 2 else Z[:, r // 2].real
    return x


@pytest.mark.parametrize("Ww,M", [(2, 2), (4, 3), (8, 5), (16, 10), (32, 4), (64, 7), (256, 10), (256, 5)])
def test_rows_inverse_scheme(Ww, M):
    """The fused row inverse's polyphase pairing and Stockham index scheme against numpy's irfft of the zero-extended
    half spectrum (the path it replaces), for radix-4-only, radix-2-only and mixed lengths, odd and even M."""
    rng = np.random.default_rng(Ww * 100 + M)
    K, N = Ww // 2, Ww * M
    X = rng.standard_normal((3, K + 1)) + 1j * rng.standard_normal((3, K + 1))
    full = np.zeros((3, N // 2 + 1), np.complex128)
    full[:, :K + 1] = X
    ref = np.fft.irfft(full, n=N, axis=1) * N
    got = _rows_inverse_model(X, Ww, M)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # a mistake planted in the model is seen: the K term taken once, as a Nyquist term would be
    assert np.abs(_rows_inverse_model(X, Ww, M, k_term=1.0) - ref).max() > 1e-6 * np.abs(ref).max()


FHT_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fht_golden.npz")
FHT_CASES = ("whole_mirror_m4", "whole_nopad_m3", "tiled_m3_t32", "tiled_m2_t16")


@pytest.mark.parametrize("name", FHT_CASES)
def test_oracle_fht_against_reference_vectors(name):
    """oracle/fht.py against the reference's own interpolation.py on non-constant frames (whole frames with and
    without mirror padding, overlapping ragged tiles blended with the Hann window), written by
    tests/golden/make_fht_golden.py."""
    from oracle import fht
    g = np.load(FHT_GOLDEN)
    x, ref = g[f"{name}/input"], g[f"{name}/output"]
    M, mirror, tile, overlap = (int(v) for v in g[f"{name}/params"])
    got = fht.fht_space_interpolation(x, M, bool(mirror), tile, overlap)
    assert got.shape == ref.shape
    err = np.abs(got.astype(np.float64) - ref) / np.abs(ref).max()
    print(f"fht {name}: max {err.max():.3e}")
    assert err.max() <= 1e-6
