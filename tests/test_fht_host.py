"""The oracle's restatement of the use_fht Fourier interpolation (oracle/fht.py), pinned by closed forms."""
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
