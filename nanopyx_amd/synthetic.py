"""Seeded synthetic blinking-emitter stacks for the eSRRF workloads.

Emitters sit at uniform random sub-pixel positions in [0, W) x [0, H) (pixel c covers [c, c+1)); each is
an isotropic Gaussian PSF integrated over the pixel area (erf) and scaled so its photons sum to the
per-emitter count; per frame each emitter is on with independent Bernoulli(p_on). The expected image is
background + the on-emitters, then Poisson shot noise, then additive Gaussian read noise; camera stacks
add an offset, round and clip to uint16. One ``numpy.random.Generator(PCG64(seed))`` stream is drawn in
frame order, so a stack is reproducible from its parameters and seed alone.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from math import sqrt

import numpy as np
from scipy.special import erf


@dataclass
class Scene:
    frames: int
    height: int
    width: int
    emitters: int
    sigma: float
    photons: float
    p_on: float
    background: float = 100.0
    bg_gradient: float = 0.0      # background * (1 - g/2 + g * x / (W - 1)): mean preserved, +-g/2 across x
    read_noise: float = 0.0
    offset: float = 0.0
    dtype: str = "float32"        # "float32" or "uint16"
    seed: int = 0
    esrrf: dict = field(default_factory=dict)


# Workloads of the eSRRF benchmarks (frames x H x W, scene, eSRRF settings).
CONFIGS = {
    "cpu_small": Scene(100, 64, 64, 200, 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 1,
                       dict(magnification=5, radius=1.5, sensitivity=1, doIntensityWeighting=True,
                            temporal=("AVG",))),
    "plain_256": Scene(500, 256, 256, 1500, 1.5, 1000, 0.1, 100, 0.0, 0.0, 0.0, "float32", 2,
                       dict(magnification=5, radius=1.5, sensitivity=2, doIntensityWeighting=True,
                            temporal=("AVG", "VAR", "TAC2"))),
    "large_512_u16": Scene(1000, 512, 512, 6000, 1.8, 800, 0.05, 200, 0.5, 3.0, 100.0, "uint16", 3,
                           dict(magnification=4, radius=2.0, sensitivity=2, doIntensityWeighting=True,
                                temporal=("AVG", "TAC2"))),
    "sweep_256": Scene(200, 256, 256, 3200, 1.5, 500, 0.3, 100, 0.0, 2.0, 0.0, "float32", 4,
                       dict(magnification=(2, 4), radius=(1.0, 1.5, 2.0, 2.5), sensitivity=(1, 2, 3, 4),
                            doIntensityWeighting=True, temporal=("AVG",))),
    # SRRF radiality at the reference's SRRF() defaults on the plain_256 scene (per-frame maps, 1280 x 1280)
    "srrf_256": Scene(100, 256, 256, 1500, 1.5, 1000, 0.1, 100, 0.0, 0.0, 0.0, "float32", 6,
                      dict(magnification=5, ringRadius=0.5, border=0, radialityPositivityConstraint=True,
                           doIntensityWeighting=True)),
}


def _pixel_profile(center: np.ndarray, sigma: float, n: int, half: int):
    """Pixel-integrated 1D Gaussian profiles: (len(center), 2*half+1) weights and their first pixel index."""
    start = np.floor(center).astype(np.int64) - half
    edges = start[:, None] + np.arange(2 * half + 2)[None, :]
    z = (edges - center[:, None]) / (sqrt(2.0) * sigma)
    cdf = 0.5 * (1.0 + erf(z))
    prof = np.diff(cdf, axis=1)
    return prof, start


def _dense_profiles(center, sigma, n):
    half = int(np.ceil(5 * sigma)) + 1
    prof, start = _pixel_profile(center, sigma, n, half)
    A = np.zeros((center.size, n), dtype=np.float64)
    for k in range(prof.shape[1]):
        idx = start + k
        ok = (idx >= 0) & (idx < n)
        A[np.nonzero(ok)[0], idx[ok]] = prof[ok, k]
    return A


def generate(scene: Scene, frames: int | None = None) -> np.ndarray:
    """Render ``scene`` (optionally only its first ``frames`` frames) as an (N, H, W) stack."""
    N = scene.frames if frames is None else min(frames, scene.frames)
    H, W = scene.height, scene.width
    rng = np.random.Generator(np.random.PCG64(scene.seed))
    ys = rng.uniform(0, H, scene.emitters)
    xs = rng.uniform(0, W, scene.emitters)
    Ay = _dense_profiles(ys, scene.sigma, H)          # (E, H)
    Ax = _dense_profiles(xs, scene.sigma, W)          # (E, W)
    bg_row = np.full(W, scene.background, dtype=np.float64)
    if scene.bg_gradient:
        g = scene.bg_gradient
        bg_row = scene.background * (1 - g / 2 + g * np.arange(W) / max(W - 1, 1))
    out = np.empty((N, H, W), dtype=np.uint16 if scene.dtype == "uint16" else np.float32)
    for f in range(N):
        on = rng.random(scene.emitters) < scene.p_on
        idx = np.nonzero(on)[0]
        expected = (Ay[idx].T * scene.photons) @ Ax[idx] + bg_row[None, :]
        img = rng.poisson(expected).astype(np.float64)
        if scene.read_noise:
            img += rng.normal(0.0, scene.read_noise, size=img.shape)
        if scene.dtype == "uint16":
            out[f] = np.clip(np.rint(img + scene.offset), 0, 65535).astype(np.uint16)
        else:
            out[f] = (img + scene.offset).astype(np.float32)
    return out


def generate_3d(frames, slices, rows, cols, emitters, sigma_xy=1.5, sigma_z=1.2, photons=800.0, p_on=0.2,
                background=100.0, seed=7):
    """(F, S, R, C) float32 volumes of blinking emitters: uniform 3D positions, a separable Gaussian PSF (sigma_xy,
    sigma_z in voxels), per-frame Bernoulli(p_on) emission, background and Poisson noise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    z, y, x = (rng.uniform(0, n, emitters) for n in (slices, rows, cols))
    gz = np.exp(-((np.arange(slices)[None, :] + 0.5 - z[:, None]) ** 2) / (2 * sigma_z ** 2))
    gy = np.exp(-((np.arange(rows)[None, :] + 0.5 - y[:, None]) ** 2) / (2 * sigma_xy ** 2))
    gx = np.exp(-((np.arange(cols)[None, :] + 0.5 - x[:, None]) ** 2) / (2 * sigma_xy ** 2))
    norm = photons / (gz.sum(1) * gy.sum(1) * gx.sum(1))
    out = np.empty((frames, slices, rows, cols), np.float32)
    for f in range(frames):
        on = np.nonzero(rng.random(emitters) < p_on)[0]
        vol = np.einsum("e,ez,ey,ex->zyx", norm[on], gz[on], gy[on], gx[on], optimize=True) + background
        out[f] = rng.poisson(vol).astype(np.float32)
    return out


def request_scene(rng: np.random.Generator, seed: int) -> Scene:
    """One served-stream job: frames log-uniform on 50-1000, H and W from {64,...,1024}, density 0.02/px^2."""
    sizes = (64, 128, 256, 384, 512, 1024)
    N = int(round(np.exp(rng.uniform(np.log(50), np.log(1000)))))
    H = int(rng.choice(sizes))
    W = int(rng.choice(sizes))
    M = int(rng.choice((2, 3, 4, 5, 8)))
    R = float(rng.uniform(1.0, 2.5))
    S = int(rng.choice((1, 2, 3)))
    mode = str(rng.choice(("AVG", "VAR", "TAC2")))
    dtype = str(rng.choice(("float32", "uint16")))
    return Scene(N, H, W, max(1, int(round(0.02 * H * W))), 1.5, 500, 0.3, 100, 0.0, 2.0,
                 100.0 if dtype == "uint16" else 0.0, dtype, seed,
                 dict(magnification=M, radius=R, sensitivity=S, doIntensityWeighting=True, temporal=(mode,)))


def generate_device(scene: Scene, frames: int | None = None, device="cuda"):
    """The recipe of ``generate`` rendered on the GPU with torch's generator seeded by ``scene.seed`` (same scene
    model and statistics, not the same draws as ``generate``). Returns an (N, H, W) float32 or uint16 device tensor;
    used for the served-stream workload, whose many large stacks would take minutes to render on the host."""
    import torch
    N = scene.frames if frames is None else min(frames, scene.frames)
    H, W, E = scene.height, scene.width, scene.emitters
    g = torch.Generator(device=device)
    g.manual_seed(int(scene.seed))
    ys = torch.rand(E, generator=g, device=device, dtype=torch.float64) * H
    xs = torch.rand(E, generator=g, device=device, dtype=torch.float64) * W

    def profiles(center, n):
        edges = torch.arange(n + 1, device=device, dtype=torch.float64)
        cdf = 0.5 * (1.0 + torch.special.erf((edges[None, :] - center[:, None]) / (sqrt(2.0) * scene.sigma)))
        return (cdf[:, 1:] - cdf[:, :-1]).to(torch.float32)  # (E, n)

    Ay, Ax = profiles(ys, H), profiles(xs, W)
    bg_row = torch.full((W,), float(scene.background), device=device, dtype=torch.float32)
    if scene.bg_gradient:
        gr = scene.bg_gradient
        bg_row = scene.background * (1 - gr / 2 + gr * torch.arange(W, device=device, dtype=torch.float32) /
                                     max(W - 1, 1))
    u16 = scene.dtype == "uint16"
    out = torch.empty((N, H, W), dtype=torch.float32, device=device)
    AyT = Ay.t().contiguous()                                   # (H, E)
    fc = max(1, (1 << 27) // max(1, H * E))
    for f0 in range(0, N, fc):
        f1 = min(N, f0 + fc)
        on = (torch.rand((f1 - f0, E), generator=g, device=device) < scene.p_on).to(torch.float32) * scene.photons
        expected = torch.matmul(AyT[None] * on[:, None, :], Ax) + bg_row[None, None, :]
        img = torch.poisson(expected, generator=g)
        if scene.read_noise:
            img = img + scene.read_noise * torch.randn(img.shape, generator=g, device=device)
        out[f0:f1] = img + scene.offset
    if u16:
        out = torch.clamp(torch.round(out), 0, 65535)
        try:
            return out.to(torch.uint16)
        except (RuntimeError, TypeError):
            return torch.from_numpy(out.cpu().numpy().astype(np.uint16)).to(device)
    return out
