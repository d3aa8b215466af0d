"""Sweep scoring on the host (nanopyx_amd.core.analysis / core.transform): Fourier interpolation, Pearson, error
map, FIRE number and QnR, pinned by closed forms, invariants and scalar restatements of the reference's loops, and
FIRE, decorrelation and Pearson against the reference's own compiled analyses (tests/golden/cython_golden.npz)."""
import math

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from nanopyx_amd.core.analysis.frc import FIRECalculator, make_even_square
from nanopyx_amd.core.analysis.pearson_correlation import pearson_correlation
from nanopyx_amd.core.transform.error_map import ErrorMap
from nanopyx_amd.core.transform.fht_interpolation import fht_space_interpolation


@pytest.mark.parametrize("k,f", [(1, 2), (3, 4), (5, 2)])
def test_fht_band_limited_cosine(k, f):
    """Fourier interpolation of a band-limited cosine on a power-of-two grid is the cosine on the fine grid (the
    range rescale keeps [1, 3], which the samples reach)."""
    N = 32
    x = np.arange(N)
    img = (2 + np.cos(2 * np.pi * k * x / N))[None, :].repeat(N, 0).astype(np.float32)
    out = fht_space_interpolation(img, f, doMirrorPadding=False)
    xf = np.arange(N * f) / f
    expect = (2 + np.cos(2 * np.pi * k * xf / N))[None, :].repeat(N * f, 0)
    assert out.shape == (N * f, N * f)
    assert np.abs(out - expect).max() < 2e-5


def test_fht_constant_and_shapes():
    # a constant image keeps its shape; its level is divided by factor^2 (no range to rescale to), as in the reference
    out = fht_space_interpolation(np.full((5, 7), 3.0, np.float32), 2)
    assert out.shape == (10, 14) and np.allclose(out, 3.0 / 4)
    rng = np.random.default_rng(1)
    st = rng.random((3, 9, 13)).astype(np.float32)
    o3 = fht_space_interpolation(st, 3)
    assert o3.shape == (3, 27, 39)
    assert np.array_equal(o3[1], fht_space_interpolation(st[1], 3))
    # without mirror padding nothing is cropped, so the rescaled output spans the input's range exactly
    sq = rng.random((16, 16)).astype(np.float32)
    o = fht_space_interpolation(sq, 2, doMirrorPadding=False)
    assert np.isclose(o.min(), sq.min(), atol=1e-6) and np.isclose(o.max(), sq.max(), atol=1e-6)
    # mirror padding of a 3x5 image to an 8x8 square: tiles alternate between the image and its flips
    from nanopyx_amd.core.transform.fht_interpolation import _mirror_padding_even_square
    small = np.arange(15, dtype=np.float32).reshape(3, 5)
    pad = _mirror_padding_even_square(small)
    full = np.zeros((9, 15), np.float32)
    for th in range(3):
        for tw in range(3):
            t = small[::-1] if th % 2 == 0 else small      # middle tile (1, 1) is the image itself
            t = t[:, ::-1] if tw % 2 == 0 else t
            full[th * 3:(th + 1) * 3, tw * 5:(tw + 1) * 5] = t
    assert np.array_equal(pad, full[0:8, 3:11])


def test_pearson():
    rng = np.random.default_rng(2)
    a = rng.random((20, 30))
    b = 0.3 * a + rng.random((20, 30))
    assert abs(pearson_correlation(a, b) - np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 1e-6
    assert pearson_correlation(a, -2 * a + 1) == pytest.approx(-1, abs=1e-6)
    assert pearson_correlation(a, np.ones_like(a)) == 0


def test_error_map_recovers_blur_and_scale():
    """Reference = block mean of the super-resolved truth scaled by alpha, offset by beta and blurred by sigma; the
    error map finds sigma, alpha, beta and an RSP near 1."""
    rng = np.random.default_rng(3)
    M = 2
    gt = gaussian_filter(rng.random((128, 128)), 4).astype(np.float32)
    gt = (gt - gt.min()) / (gt.max() - gt.min())
    alpha, beta, sigma = 3.0, 0.5, 2.5
    ref = gaussian_filter(gt * alpha + beta, sigma)
    ref = ref.reshape(64, M, 64, M).mean(axis=(1, 3)).astype(np.float32)
    em = ErrorMap()
    em.optimise(ref, gt)
    assert em.getRSP() > 0.98
    assert abs(em.get_sigma() - sigma) < 0.5
    assert abs(em._alpha - alpha) < 0.3 and abs(em._beta - beta) < 0.3
    em2 = ErrorMap()
    em2.optimise(ref, gt, fixedSigma=1.25)
    assert em2.get_sigma() == pytest.approx(1.25 * M)


def _ring_scalar(images, centre, size):
    """The reference's _calculate_frc_value loop (frc.pyx:121-158), scalar float32, for a few radii."""
    f32 = np.float32
    max_r = centre - 1
    out = {}
    for r in range(1, min(max_r, 6)):
        step = f32(1.0 / r)
        angles = np.arange(0, float(f32(math.pi * 2)), float(step), dtype=np.float32)
        s = [f32(0)] * 3
        for ang in angles:
            c = f32(math.cos(float(ang)))
            sn = f32(math.sqrt(float(f32(f32(1) - f32(c * c)))))
            if float(ang) > math.pi:
                sn = -sn
            x = f32(f32(max_r + 1) + f32(f32(r) * c))
            y = f32(f32(max_r + 1) + f32(f32(r) * sn))
            xb, yb = int(x), int(y)
            xf, yf = max(f32(x - f32(xb)), f32(0)), max(f32(y - f32(yb)), f32(0))
            ll = yb * size + xb
            for i, im in enumerate(images):
                up = f32(im[ll + size] + f32(xf * f32(im[ll + size + 1] - im[ll + size])))
                lo = f32(im[ll] + f32(xf * f32(im[ll + 1] - im[ll])))
                s[i] = f32(s[i] + f32(lo + f32(yf * f32(up - lo))))
        out[r] = f32(float(s[0]) / math.sqrt(float(f32(s[1] * s[2]))))
    return out


def test_frc_ring_values_match_scalar_loop():
    rng = np.random.default_rng(4)
    size = 24
    images = tuple(rng.random(size * size).astype(np.float32) + 0.1 for _ in range(3))
    fc = FIRECalculator()
    res = fc._frc_value(size // 2, size, images)
    for r, v in _ring_scalar(images, size // 2, size).items():
        assert res[r, 1] == v, r


def test_fire_number_behaviour():
    rng = np.random.default_rng(5)
    img = gaussian_filter(rng.random((96, 96)), 2).astype(np.float32)
    assert FIRECalculator().calculate_fire_number(img, img) == 0  # FRC = 1 everywhere: no crossing
    fires = []
    for noise in (0.002, 0.02, 0.2):
        a = img + noise * rng.standard_normal(img.shape)
        b = img + noise * rng.standard_normal(img.shape)
        fires.append(FIRECalculator().calculate_fire_number(a, b))
    assert all(f > 0 for f in fires) and fires[0] < fires[1] < fires[2]  # more noise: coarser resolution
    fc = FIRECalculator()
    fc.calculate_fire_number(img[:95, :90], img[:95, :90])
    assert fc.field_of_view == 90 and make_even_square(img[:95, :90]).shape == (90, 90)


def test_qnr_score():
    from nanopyx_amd.core.analysis.parameter_sweep import ParameterSweep
    ps = ParameterSweep()
    frc = np.array([[2.0, 3.0], [4.0, 6.0]])
    n = ps.logistic_image_conversion(frc)
    assert n.min() == pytest.approx(0.075) and n.max() == pytest.approx(0.925)
    rsp = np.array([[0.9, 0.5], [0.7, 0.8]])
    q = ps.calculate_qnr_score(rsp, frc)
    assert np.allclose(q, 2 * rsp * n / (rsp + n))


# ---- decorrelation analysis (nanopyx_amd.core.analysis.decorr) ------------------------------------------------------

def _band_limited(n, k0, noise, seed):
    """White noise with an ideal low-pass at normalised radius k0 (1 = Nyquist), plus white noise of std `noise`."""
    rng = np.random.default_rng(seed)
    F = np.fft.fftshift(np.fft.fft2(rng.standard_normal((n, n))))
    c = np.arange(n) - n / 2
    F[np.sqrt(c[:, None] ** 2 + c[None, :] ** 2) * 2 / n > k0] = 0
    img = np.real(np.fft.ifft2(np.fft.ifftshift(F)))
    return (img / img.std() + noise * rng.standard_normal((n, n)) + 10).astype(np.float32)


@pytest.mark.parametrize("k0", [0.2, 0.3, 0.5, 0.7])
def test_decorr_finds_band_limit(k0):
    """The decorrelation curve of a band-limited image peaks at its cut-off frequency: kc0 (from the unfiltered
    spectrum) lands within one bin spacing-and-a-half of k0, and the resolution is 2 * pixel_size / kc_max."""
    from nanopyx_amd.core.analysis.decorr import DecorrAnalysis
    d = DecorrAnalysis(pixel_size=3)
    d.run_analysis(_band_limited(256, k0, 0.02, seed=int(k0 * 10)))
    assert abs(float(d.kc0) - k0) <= 1.5 / 49
    assert float(d.resolution) == pytest.approx(6 / float(d.kc_max))


def test_decorr_resolution_follows_blur():
    from nanopyx_amd import calculate_decorr_analysis
    rng = np.random.default_rng(6)
    base = rng.random((256, 256))
    res = [calculate_decorr_analysis(gaussian_filter(base, s).astype(np.float32)) for s in (1.0, 2.0, 4.0)]
    assert res[0] < res[1] < res[2]
    assert 1.5 < res[1] / res[0] < 2.5 and 1.5 < res[2] / res[1] < 2.5   # resolution ~ proportional to the blur


def _ring_loop(fr, fi, nr_, ni_, crmin, crmax, n_r, height, width):
    """The reference's ring loop as written (decorr.pyx:169-203): x outer, y inner, early return, float32 bins."""
    f32 = np.float32
    out = np.zeros(2 * n_r, np.float32)
    ox = int(f32(width) * (f32(1) - f32(crmax)) / f32(2))
    oy = int(f32(height) * (f32(1) - f32(crmax)) / f32(2))
    w = int(f32(width) * f32(crmax))
    h = int(f32(height) * f32(crmax))
    for x in range(ox, ox + w):
        for y in range(oy, oy + h):
            dist = f32((x - width / 2) ** 2 + (y - height / 2) ** 2)
            dist = f32(math.sqrt(float(f32(f32(4) * dist / f32(width * width)))))
            if x * height + y > width * height / 2 + height / 2:
                return out
            if 0 <= dist <= f32(crmax):
                v = f32(f32(dist - f32(crmin)) / f32(f32(crmax) - f32(crmin)))
                v = f32(f32(v * f32(n_r - 1)) + f32(0))
                if v < 0:
                    v = f32(0)
                d = int(math.floor(float(v) + 0.5))
                if d + n_r < 2 * n_r:
                    out[d] = f32(out[d] + f32(f32(fr[y, x] * nr_[y, x]) + f32(fi[y, x] * ni_[y, x])))
                    out[d + n_r] = f32(out[d + n_r] + f32(f32(nr_[y, x] * nr_[y, x]) + f32(ni_[y, x] * ni_[y, x])))
    return out


@pytest.mark.parametrize("crmin,crmax", [(0.0, 1.0), (0.2, 0.75), (0.44, 0.79)])
def test_decorr_ring_sums_match_loop(crmin, crmax):
    from nanopyx_amd.core.analysis.decorr import DecorrAnalysis, _normalize_fft
    rng = np.random.default_rng(7)
    n = 32
    fr = rng.standard_normal((n, n)).astype(np.float32)
    fi = rng.standard_normal((n, n)).astype(np.float32)
    nf = _normalize_fft(fr, fi)
    d = DecorrAnalysis(n_r=12)
    d.img_ref = np.zeros((n, n), np.float32)
    got = d._get_corr_coef_ring(fr, fi, nf[0], nf[1], np.float32(crmin), np.float32(crmax))
    ref = _ring_loop(fr, fi, nf[0], nf[1], crmin, crmax, 12, n, n)
    assert np.array_equal(got, ref)


def test_decorr_pieces():
    from nanopyx_amd.core.analysis.decorr import (DecorrAnalysis, _apodize_edges, _get_circular_mask,
                                                   _normalize_fft)
    w = 10
    m = _get_circular_mask(w, 1)
    brute = np.array([[1.0 if (y - w / 2) ** 2 + (x - w / 2) ** 2 < w * w / 4 else 0.0 for x in range(w)]
                      for y in range(w)], np.float32)
    assert np.array_equal(m, brute)
    nf = _normalize_fft(np.array([[3.0, 0.0]], np.float32), np.array([[4.0, 0.0]], np.float32))
    assert np.allclose(nf[:, 0, 0], [0.6, 0.8]) and np.all(nf[:, 0, 1] == 0)
    # apodisation: the centre is untouched, the outermost ring of pixels (distance 0 from the edge) takes the mean
    # of the pixels outside the inscribed circle
    rng = np.random.default_rng(8)
    img = rng.random((64, 64)).astype(np.float32)
    ap = _apodize_edges(img)
    assert np.array_equal(ap[22:42, 22:42], img[22:42, 22:42])
    c = np.arange(64) - 32
    outside = (c[:, None] ** 2 + c[None, :] ** 2).astype(np.float32) > np.float32(12) ** 2
    assert np.allclose(ap[0, :], img[outside].mean(), atol=1e-6)
    # padding: power-of-two square filled with the first pixel, image at rows [0, h), centred along x only
    pad = DecorrAnalysis._get_preprocessed_image(img[:40, :50])
    assert pad.shape == (64, 64) and np.array_equal(pad[:40, 7:57], img[:40, :50])
    assert np.all(pad[40:] == img[0, 0]) and np.all(pad[:, :7] == img[0, 0])
    # highest local maximum; a maximum on a rise at the curve's end is skipped
    d = DecorrAnalysis(n_r=11)
    peak = np.array([0, .2, .5, .9, .6, .4, .3, .3, .2, .2, .1], np.float32)
    assert np.allclose(d._get_d_corr_max(peak, 0, 1), [0.3, 0.9])
    rising = np.linspace(0, 1, 11).astype(np.float32)
    assert np.allclose(d._get_d_corr_max(rising, 0, 1), [0, 0])
    bump_then_rise = np.array([0, .4, .7, .5, .45, .5, .6, .7, .8, .9, 1.0], np.float32)
    assert np.allclose(d._get_d_corr_max(bump_then_rise, 0.2, 0.7), [0.2 + 0.5 * 2 / 10, 0.7])


def _scoring_golden():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cython_golden.npz")
    return np.load(path)


@pytest.mark.parametrize("n", [64, 96])
def test_fire_number_matches_reference_vectors(n):
    """FIRE numbers of two noisy realisations of one emitter field, against the reference's own FIRECalculator
    (core/analysis/frc.pyx:17-294, compiled from the reference checkout by tests/golden/make_cython_golden.py)."""
    g = _scoring_golden()
    got = FIRECalculator(pixel_size=1, units="pixel").calculate_fire_number(g[f"pair{n}_a"], g[f"pair{n}_b"])
    ref = float(g[f"fire{n}"])
    assert abs(got - ref) <= 1e-5 * abs(ref), (got, ref)


@pytest.mark.parametrize("name,kw", [("pattern80", {}), ("blur72", {}), ("pair64_a", {}), ("pair96_a", {}),
                                     ("pair96_b", dict(rmin=0.05, rmax=0.8, n_r=40, n_g=8, pixel_size=2.5,
                                                       units="nm"))])
def test_decorr_resolution_matches_reference_vectors(name, kw):
    """Decorrelation resolutions of a band-limited pattern, a blurred noise field and emitter images (default and
    non-default analysis parameters), against the reference's own DecorrAnalysis (core/analysis/decorr.pyx:22-413,
    compiled from the reference checkout by tests/golden/make_cython_golden.py)."""
    from nanopyx_amd.core.analysis.decorr import DecorrAnalysis
    g = _scoring_golden()
    d = DecorrAnalysis(**kw)
    d.run_analysis(g[name])
    ref = float(g[f"decorr_{name}" + ("_params" if kw else "")])
    assert abs(d.resolution - ref) <= 1e-5 * abs(ref), (d.resolution, ref)


def test_scoring_golden_sources_recorded():
    """The fixture names the reference files it was built from (paths and SHA-256)."""
    s = str(_scoring_golden()["sources_sha256"])
    assert "src/nanopyx/core/analysis/frc.pyx" in s and "src/nanopyx/core/analysis/decorr.pyx" in s


@pytest.mark.parametrize("a,b,key", [("pair64_a", "pair64_b", "pearson_pair64"),
                                     ("pattern80", "blur72", "pearson_pattern_blur")])
def test_pearson_matches_reference_vectors(a, b, key):
    """Pearson correlation (the error map's RSP) against the reference's own pearson_correlation
    (core/analysis/pearson_correlation.pyx:36-60, float32 sums): within 1e-5 (the reference accumulates in float32)."""
    g = _scoring_golden()
    x, y = g[a], g[b]
    n = min(x.shape[0], y.shape[0])
    got = pearson_correlation(x[:n, :n], y[:n, :n])
    assert abs(got - float(g[key])) <= 1e-5, (got, float(g[key]))
