"""Benchmark of the fused eSRRF path on one MI355X.

A step is one fused eSRRF + temporal-correlation pass (libnpx_hip npx_esrrf_temporal) over the whole synthetic
stack of the workload, resident in HBM before timing starts. Default workload: 1000 frames x 512x512 uint16
camera counts, magnification 4, radius 2.0, sensitivity 2, intensity weighting, AVG + TAC2 -> 2048x2048 maps.

Prints one JSON line: throughput in output Mpx*frames/s (N * H*M * W*M / 1e6 / seconds), the FP32 roofline of
the fused kernel (algorithmic flops per launch / its HIP-event duration) and the reference CPU path timed on a
bounded frame sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "eSRRF output Mpx·frames/s vs plain GPU & CPU; sweep speedup; served p50/p99"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# the fused kernels are bound by FP32 vector ALU issue (a stencil: no MFMA on this path)
BOUND = "fp32-valu"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ring_samples_mean(H, W, M, R):
    """Mean in-disc ring-sample count over the M*M output phases (interior pixels)."""
    from nanopyx_amd import _hip
    t = _hip.plan_tables(H, W, M, R)
    return float((t["w"] > 0).reshape(M * M, -1).sum(1).mean())


def traffic_from_profile(workload):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", f"pmc_esrrf_fused_{workload}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def sweep_bench(reps):
    """The 32-setting parameter sweep (200 x 256 x 256 float32, radius {1, 1.5, 2, 2.5} x sensitivity {1, 2, 3, 4} x
    magnification {2, 4}, AVG) as one batched launch, as its captured hipGraph, and as 32 sequential single-setting
    fused launches; times from the host clock around `reps` synchronised repetitions of each."""
    import torch
    from nanopyx_amd import _hip, synthetic
    sc = synthetic.CONFIGS["sweep_256"]
    e = sc.esrrf
    mags, radii, sens = list(e["magnification"]), list(e["radius"]), list(e["sensitivity"])
    x = torch.from_numpy(synthetic.generate(sc)).cuda()
    N, H, W = x.shape

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    outs = _hip.SweepOutputs(torch, x, mags, radii, sens, workspace=True)
    t_sweep = timed(lambda: _hip.esrrf_sweep(x, mags, radii, sens, outputs=outs))
    _hip.timing(True)
    _hip.esrrf_sweep(x, mags, radii, sens, outputs=outs)
    torch.cuda.synchronize()
    _hip.timing(False)
    _, k_ms, m_ms = _hip.timing_read()
    with _hip.Graph.capture() as g:
        _hip.esrrf_sweep(x, mags, radii, sens, stream=g.stream, outputs=outs)
    t_graph = timed(g.launch)
    g.close()
    settings = [(m, r, s) for m in mags for r in radii for s in sens]
    t_seq = timed(lambda: [_hip.esrrf_temporal(x, m, r, s, True, ("AVG",)) for (m, r, s) in settings])
    # deduplicated algorithmic work: ring sums once per (R, M), the final power/weight/mean per setting
    flops = 0.0
    for m in mags:
        U = N * H * m * W * m
        for r in radii:
            flops += U * (9.0 * ring_samples_mean(H, W, m, r) + 32.0)
        flops += U * 8.0 * len(radii) * len(sens)
    k_s = k_ms / 1e3
    return {"workload": f"sweep_256: {N} frames x {H}x{W} float32, M{mags} x R{radii} x S{sens}, AVG "
                        f"({len(settings)} settings)",
            "batched_ms": round(t_sweep * 1e3, 3), "graph_ms": round(t_graph * 1e3, 3),
            "sequential_32_ms": round(t_seq * 1e3, 3),
            "speedup_vs_sequential": round(t_seq / min(t_sweep, t_graph), 2),
            "kernel_ms": round(k_ms, 3), "merge_ms": round(m_ms, 3),
            "flops_dedup": flops, "tflops_kernel": round(flops / k_s / 1e12, 2),
            "frac_fp32_peak": round(flops / k_s / 1e12 / FP32_PEAK_TFLOPS, 4)}


# per output pixel-frame, reference-equivalent: 12 ring samples x 95 (two direct 16-tap Catmull-Rom gradient
# interpolations of 40 flops each + the convergence term 15), + 43 per pixel (see srrf_bench)
SRRF_REF_FLOPS_PER_PX = 12 * 95 + 43


def srrf_flops_per_px(H, W, M, ringRadius, border=0):
    """FP32 flops the separable SRRF kernel executes per output pixel-frame: the row pass interpolates every
    gradient-tile row (TH per band of TJ output rows) at each of the 12 sample x positions (4 taps x 2 gradient
    components), the column pass combines 4 of those rows per sample (4 x 2) and forms the convergence term (15), and
    each pixel adds 43 (ring mean and clamp, 16-tap intensity interpolation, weight)."""
    from nanopyx_amd import _hip
    d = _hip.radiality_info(H, W, M, ringRadius, border)
    return 12 * d["TH"] * 4 * 2 * 2 / d["TJ"] + 12 * (4 * 2 * 2 + 15) + 43


def srrf_bench(reps, cpu_frames=2, workload="srrf_256"):
    """SRRF radiality (npx_radiality: gradients, 12-sample Catmull-Rom ring, in-kernel xM intensity weighting) over
    the srrf_256 stack, per-frame maps out. Kernel time from HIP events on the launch stream over `reps` launches.
    `frac` counts the flops the separable kernel executes (srrf_flops_per_px); `frac_reference_equivalent` counts
    the reference's direct evaluation: per ring sample two 16-tap Catmull-Rom interpolations (40 flops each) and the
    convergence term (|G|^2, cross and dot products, Dk, accumulate: 15), 12 x 95; per pixel the ring mean and clamp
    (2), the 16-tap intensity interpolation (40) and the weight (1). The CPU baseline is the reference's own compiled
    radiality and Catmull-Rom kernels (oracle/_ref) driven like its threaded SRRF() on `cpu_frames` frames."""
    import torch
    from nanopyx_amd import _hip, synthetic
    sc = synthetic.CONFIGS[workload]
    e = sc.esrrf
    M, R = e["magnification"], e["ringRadius"]
    x_host = synthetic.generate(sc)
    x = torch.from_numpy(x_host).cuda()
    N, H, W = x.shape
    out = torch.empty((N, H * M, W * M), dtype=torch.float32, device="cuda")

    def launch():
        _hip.radiality(x, None, M, R, e["border"], e["radialityPositivityConstraint"], e["doIntensityWeighting"],
                       out=out)

    for _ in range(2):
        launch()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        launch()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    U = N * H * M * W * M
    SRRF_FLOPS_PER_PX = srrf_flops_per_px(H, W, M, R, e["border"])
    flops = U * float(SRRF_FLOPS_PER_PX)
    achieved = flops / (ms / 1e3) / 1e12
    ref_achieved = U * float(SRRF_REF_FLOPS_PER_PX) / (ms / 1e3) / 1e12
    res = {"workload": f"{workload}: {N} frames x {H}x{W} float32, SRRF M{M} ringRadius {R}, positivity, "
                       f"intensity-weighted, per-frame maps {H * M}x{W * M}",
           "kernel_ms": round(ms, 4), "value": round(U / 1e6 / (ms / 1e3), 1), "unit": "Mpx·frames/s",
           "roofline": {"bound": BOUND, "compute": "FP32 vector ALU", "achieved": round(achieved, 2),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                        "flops_per_launch": flops, "flops_per_px": SRRF_FLOPS_PER_PX,
                        "flops_per_px_reference_equivalent": SRRF_REF_FLOPS_PER_PX,
                        "frac_reference_equivalent": round(ref_achieved / FP32_PEAK_TFLOPS, 4),
                        "algorithmic_bytes": N * H * W * 4 + U * 4,
                        "hbm_gbs_algorithmic": round((N * H * W * 4 + U * 4) / (ms / 1e3) / 1e9, 1)}}
    if cpu_frames:
        from oracle import ref32
        ref32.build()
        kind = "port"
        if ref32.reference_srrf_available() and ref32.reference_available():
            ref32.use_reference_radiality(True)
            ref32.use_reference_kernels(True)  # the Catmull-Rom magnification of the intensity
            kind = "reference"
        xs = x_host[:cpu_frames]
        t0 = time.perf_counter()
        ref = ref32.srrf_workflow(xs, M, R, e["border"], e["radialityPositivityConstraint"],
                                  e["doIntensityWeighting"], macro_pixel_correction=False)
        dt = time.perf_counter() - t0
        ref32.use_reference_radiality(False)
        ref32.use_reference_kernels(False)
        Us = cpu_frames * H * M * W * M
        res["cpu_baseline"] = {"value": round(Us / 1e6 / dt, 2), "unit": "Mpx·frames/s",
                               "cores": ref32.num_threads(), "kind": kind,
                               "sample": f"first {cpu_frames} frames (CR xM + radiality)"}
        got = out[:cpu_frames].cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        res["parity_max_rel_err"] = float(err)
    return res


def operators_bench(reps, cpu_seconds=6.0, frames=100, rgc_frames=20):
    """The two operators of the path measured on their own, on the first frames of the plain_256 scene (256 x 256
    float32, M5 R1.5 S2, intensity weighting):
    - eSRRF_ST.run, per-frame maps (npx_esrrf_frames) of `frames` frames;
    - RadialGradientConvergence.run on caller arrays (npx_rgc) of `rgc_frames` frames: the x2M Catmull-Rom of both
      Roberts gradients and the xM Catmull-Rom of the image, made on the GPU by the library's separate steps
      (npx_shift_magnify, npx_roberts) as the reference's eSRRF_ST composes them, then the RGC on those arrays.
    Kernel time from HIP events on the launch stream over `reps` calls; FP32 roofline on F = U (9 S + 40) as for the
    fused kernel; the CPU baseline is the reference's own C kernels (oracle/_ref) on a bounded frame sample of the
    same inputs (the chain for eSRRF_ST, _c_calculate_rgc for the RGC), and parity is taken on that sample."""
    import torch
    from nanopyx_amd import _hip, synthetic
    from oracle import ref32
    sc = synthetic.CONFIGS["plain_256"]
    e = sc.esrrf
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    x_host = synthetic.generate(sc, frames=frames)
    x = torch.from_numpy(x_host).cuda()
    N, H, W = x.shape
    s_bar = ring_samples_mean(H, W, M, R)
    ref32.build()
    kind = "port"
    if ref32.reference_available():
        ref32.use_reference_kernels(True)
        kind = "reference"

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            out = fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, out

    def record(ms, n, in_bytes, what):
        U = n * H * M * W * M
        flops = U * (9.0 * s_bar + 40.0)
        achieved = flops / (ms / 1e3) / 1e12
        alg = in_bytes + U * 4
        return {"workload": what, "kernel_ms": round(ms, 4), "value": round(U / 1e6 / (ms / 1e3), 1),
                "unit": "Mpx·frames/s",
                "roofline": {"bound": BOUND, "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                             "flops_per_launch": flops, "ring_samples_mean": s_bar, "algorithmic_bytes": alg,
                             "hbm_gbs_algorithmic": round(alg / (ms / 1e3) / 1e9, 1)}}

    def cpu_record(fn, n_total, what):
        done, kept, t0 = 0, [], time.perf_counter()
        while done < n_total:
            kept.append(fn(done, min(n_total, done + 2)))
            done = min(n_total, done + 2)
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        dt = time.perf_counter() - t0
        return {"value": round(done * H * M * W * M / 1e6 / dt, 3), "unit": "Mpx·frames/s",
                "cores": ref32.num_threads(), "kind": kind,
                "sample": f"first {done} of {n_total} frames, {what}, in {dt:.1f} s"}, np.concatenate(kept), done

    def parity(got, ref):
        g, r = got.astype(np.float64), ref.astype(np.float64)
        err = np.abs(g - r) / max(np.abs(r).max(), 1e-30)
        return {"max": float(err.max()), "mean": float(err.mean())}

    try:
        ms, maps = timed(lambda: _hip.esrrf_frames(x, M, R, S, iw))
        fr = record(ms, N, N * H * W * 4, f"eSRRF_ST.run per-frame maps: {N} frames x {H}x{W} float32 (plain_256), "
                                          f"M{M} R{R} S{S} intensity-weighted -> {N} x {H * M}x{W * M}")
        cpu, ref, n = cpu_record(lambda a, b: ref32.esrrf(x_host[a:b], M, R, S, iw), N, "reference CPU chain")
        fr["cpu_baseline"] = cpu
        fr["parity"] = dict(frames=n, against="reference CPU chain (oracle/_ref kernels)",
                            **parity(maps[:n].cpu().numpy(), ref))
        del maps
        # RadialGradientConvergence on caller arrays
        xs = x[:rgc_frames]
        img_i = _hip.shift_magnify(xs, 0.0, 0.0, M, M)
        gc, gr = _hip.roberts(xs)
        gci = _hip.shift_magnify(gc, 0.0, 0.0, 2 * M, 2 * M)
        gri = _hip.shift_magnify(gr, 0.0, 0.0, 2 * M, 2 * M)
        del gc, gr
        nr = img_i.shape[0]
        ms, out = timed(lambda: _hip.rgc(gci, gri, img_i, M, R, S, iw, False))
        # algorithmic input bytes: the lattice positions the ring reads (every M-th row and column of the x2M
        # arrays, 8 B for the gradient pair) and the xM image
        lat = nr * (2 * H + 1) * (2 * W + 1) * 8 + nr * H * M * W * M * 4
        rg = record(ms, nr, lat, f"RadialGradientConvergence.run on caller arrays: {nr} frames, gradients "
                                 f"{2 * H * M}x{2 * W * M} and image {H * M}x{W * M} float32 (from plain_256), "
                                 f"M{M} R{R} S{S} intensity-weighted")
        g_h, r_h, i_h = gci.cpu().numpy(), gri.cpu().numpy(), img_i.cpu().numpy()
        cpu, ref, n = cpu_record(lambda a, b: ref32.rgc(g_h[a:b], r_h[a:b], i_h[a:b], M, R, S, iw, False), nr,
                                 "reference _c_calculate_rgc")
        rg["cpu_baseline"] = cpu
        rg["parity"] = dict(frames=n, against="reference _c_calculate_rgc (oracle/_ref) on the same arrays",
                            **parity(out[:n].cpu().numpy(), ref))
        del out, gci, gri, img_i, g_h, r_h, i_h
    finally:
        ref32.use_reference_kernels(False)
    del x
    torch.cuda.empty_cache()
    return {"esrrf_frames": fr, "rgc_arrays": rg}


def esrrf3d_samples_mean(Mxy, Mz, radius, radius_z):
    """Mean count of window samples that pass the distance conditions, over the Mz x Mxy x Mxy output phases (interior
    voxels)."""
    sigma, sigma_z = np.float32(radius / 2.355), np.float32(radius_z / 2.355)
    tSO, tSO_z = 2 * sigma + 1, 2 * sigma_z + 1
    j = np.arange(-int(radius), int(np.float32(radius) + 1) + 1)
    k = np.arange(-int(radius_z), int(np.float32(radius_z) + 1) + 1)
    J, I, K = np.meshgrid(j, j, k, indexing="ij")
    dxy = np.sqrt(I ** 2 + J ** 2)
    d = np.sqrt(I ** 2 + J ** 2 + K ** 2)
    return float(((d != 0) & (dxy <= tSO) & (K <= tSO_z)).sum())


def esrrf3d_bench(reps, cpu_frames=1):
    """eSRRF 3D (npx_esrrf3d: gradients, xMxy Catmull-Rom of intensity and gradients, per-voxel convergence folded
    into the running mean) of 10 frames x 32 slices x 128 x 128 at Mxy 2, Mz 2, radius 1.5 (radius_z 1.05),
    sensitivity 1, intensity weighting, 'average'. Times from HIP events on the launch stream; algorithmic flops per
    output voxel-frame: 43 per window sample (distances 7, weight 6, G.d 5, cross product and norms 20, Dk and
    accumulate 5) x the mean sample count, + 6 per voxel, + 160 / Mz for the xy magnification of 4 fields."""
    import torch
    from nanopyx_amd import _hip, synthetic
    F, S, R, C, Mxy, Mz, rad = 10, 32, 128, 128, 2, 2, 1.5
    rad_z = rad * 2.8 / 4.0
    x_host = synthetic.generate_3d(F, S, R, C, emitters=400)
    x = torch.from_numpy(x_host).cuda()

    def launch():
        return _hip.esrrf3d(x, Mxy, Mz, rad, rad_z, 1.0, "average", True)

    for _ in range(2):
        launch()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        out = launch()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    U = F * S * Mz * R * Mxy * C * Mxy
    s_bar = esrrf3d_samples_mean(Mxy, Mz, rad, rad_z)
    flops = U * (43.0 * s_bar + 6.0 + 160.0 / Mz)
    achieved = flops / (ms / 1e3) / 1e12
    res = {"workload": f"esrrf3d: {F} frames x {S}x{R}x{C} float32, Mxy {Mxy} Mz {Mz}, radius {rad} "
                       f"(radius_z {rad_z:.3f}), sensitivity 1, average -> {S * Mz}x{R * Mxy}x{C * Mxy}",
           "ms_per_call": round(ms, 4), "value": round(U / 1e6 / (ms / 1e3), 1), "unit": "Mvoxel·frames/s",
           "roofline": {"bound": BOUND, "compute": "FP32 vector ALU", "achieved": round(achieved, 2),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                        "flops_per_call": flops, "samples_mean": s_bar}}
    if cpu_frames:
        from oracle import ref32
        ref32.build()
        kind = "port"
        if ref32.reference_available():
            ref32.use_reference_3d(True)
            ref32.use_reference_kernels(True)  # the xy Catmull-Rom magnification
            kind = "reference"
        t0 = time.perf_counter()
        ref = ref32.esrrf3d(x_host[:cpu_frames], Mxy, Mz, rad, 2.8, 4.0, 1, "average", True)
        dt = time.perf_counter() - t0
        ref32.use_reference_3d(False)
        ref32.use_reference_kernels(False)
        Us = cpu_frames * S * Mz * R * Mxy * C * Mxy
        res["cpu_baseline"] = {"value": round(Us / 1e6 / dt, 3), "unit": "Mvoxel·frames/s",
                               "cores": ref32.num_threads(), "kind": kind,
                               "sample": f"first {cpu_frames} frame(s) of the same stack"}
        got = _hip.esrrf3d(x[:cpu_frames], Mxy, Mz, rad, rad_z, 1.0, "average", True).cpu().numpy()
        res["parity_max_rel_err"] = float(np.abs(got - ref).max() / np.abs(ref).max())
    return res


def fht_bench(reps, cpu_frames=1, frames=20):
    """eSRRF with use_fht=True (Fourier interpolation branch) on the first `frames` frames of the srrf_256 scene
    (256 x 256 float32), magnification 5, radius 1.5, sensitivity 1, intensity weighting, per-frame maps: Fourier x5
    of the image, Roberts cross, Fourier x10 of both gradients (2560 x 2560 outputs per frame), RGC on those arrays.
    HIP events on the launch stream around the whole chain. HBM bytes of the Fourier steps as they run, per enlarged
    S x S grid of a 256 x 256 working tile: the Ww/2+1 = 129 non-zero spectrum columns (16 B per entry, S entries
    each) written by the placement, read and written by the column FFT and read by the fused row inverse (4 passes),
    the float32 grid written by the row inverse, read and written by the rescale and crop (12 B per point); the
    padding and forward FFT (about 1.6 MB per grid) are left out. The dominant kernel, the row inverse, computes
    ceil(M/2) float64 FFTs of length 256 per row in LDS and is not bound by these bytes."""
    import torch
    from nanopyx_amd import _hip, synthetic
    from nanopyx_amd.core.transform._le_esrrf import _esrrf_fht
    sc = synthetic.CONFIGS["srrf_256"]
    x_host = synthetic.generate(sc, frames=frames)
    x = torch.from_numpy(x_host).cuda()
    N, H, W = x.shape
    M, R = 5, 1.5

    def launch():
        return _esrrf_fht(x, M, R, 1.0, True, False)

    launch()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        out = launch()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    U = N * H * M * W * M
    hbm_min = float(N * sum(12 * S * S + 4 * 16 * 129 * S for S in (256 * M, 256 * 2 * M, 256 * 2 * M)))
    res = {"workload": f"fht_256: {N} frames x {H}x{W} float32, eSRRF use_fht M{M} R{R} S1, per-frame maps "
                       f"{H * M}x{W * M}",
           "ms_per_call": round(ms, 3), "value": round(U / 1e6 / (ms / 1e3), 1), "unit": "Mpx·frames/s",
           "roofline": {"bound": "hbm", "achieved": round(hbm_min / (ms / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(hbm_min / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": None, "bytes_min_per_call": hbm_min}}
    if cpu_frames:
        from oracle import fht, ref32
        ref32.build()
        t0 = time.perf_counter()
        ref = fht.esrrf_fht(x_host[:cpu_frames], M, R, 1, True, False)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(cpu_frames * H * M * W * M / 1e6 / dt, 3), "unit": "Mpx·frames/s",
                               "cores": ref32.num_threads(), "kind": "port",
                               "sample": f"first {cpu_frames} frame(s): NumPy Fourier interpolation + oracle RGC"}
        got = out[:cpu_frames].cpu().numpy()
        res["parity_max_rel_err"] = float(np.abs(got - ref).max() / np.abs(ref).max())
    return res


def serving_bench(n, rate_hz, seed=5):
    """The served stream: n seeded requests (frames log-uniform 50-1000, H and W from {64..1024}, M from
    {2, 3, 4, 5, 8}, R uniform on [1, 2.5], S from {1, 2, 3}, one of AVG / VAR / TAC2, float32 or uint16) arriving as
    a Poisson process at rate_hz (and at 2x and 3x that rate for the rate curve, latency split by job size), stacks
    in pinned host memory; served by the continuous-batching session and, on the same arrival trace, by one plain
    fused call per request."""
    import torch
    from nanopyx_amd import serving
    t0 = time.perf_counter()
    reqs = serving.make_requests(n, seed=seed, host=True)
    log(f"serving: {n} requests rendered in {time.perf_counter() - t0:.1f} s")
    t_arr = serving.arrivals(n, rate_hz, seed)
    warm = serving.make_requests(2, seed=seed + 1, host=True, max_frames=16)
    serving.serve(warm, 1000.0, seed)  # first-use costs (code objects, pool) outside the measured stream
    served, lat_s = serving.serve(reqs, rate_hz, seed, t_arr=t_arr)
    log(f"served: {served}")
    plain, lat_p = serving.serve_plain(reqs, rate_hz, seed, t_arr=t_arr)
    log(f"plain: {plain}")
    # the same requests at 1x, 2x and 3x the rate, latency split by job size (terciles of output pixel-frames)
    curve = [{"rate_hz": rate_hz, "served": served, "plain_per_request": plain,
              "served_by_size": serving.by_size(reqs, lat_s), "plain_by_size": serving.by_size(reqs, lat_p)}]
    for k in (2, 3):
        ta = serving.arrivals(n, k * rate_hz, seed)
        sv, ls = serving.serve(reqs, k * rate_hz, seed, t_arr=ta)
        pl, lp = serving.serve_plain(reqs, k * rate_hz, seed, t_arr=ta)
        curve.append({"rate_hz": k * rate_hz, "served": sv, "plain_per_request": pl,
                      "served_by_size": serving.by_size(reqs, ls), "plain_by_size": serving.by_size(reqs, lp)})
        log(f"rate {k * rate_hz}: served {sv} plain {pl}")
    for c in curve:
        c["p99_speedup"] = round(c["plain_per_request"]["p99_ms"] / c["served"]["p99_ms"], 2)
        c["p50_speedup"] = round(c["plain_per_request"]["p50_ms"] / c["served"]["p50_ms"], 2)
        c["throughput_speedup"] = round(c["served"]["throughput_mpx_frames_s"] /
                                        c["plain_per_request"]["throughput_mpx_frames_s"], 2)
    # saturated: every request arrives at t = 0 (throughput of the packed launches vs one call per request)
    t0s = np.zeros(n)
    served_sat, _ = serving.serve(reqs, rate_hz, seed, t_arr=t0s)
    plain_sat, _ = serving.serve_plain(reqs, rate_hz, seed, t_arr=t0s)
    log(f"saturated: served {served_sat} plain {plain_sat}")
    px = sum(r.out_px_frames for r in reqs)
    stack_gb = sum(r.stack.numel() * r.stack.element_size() for r in reqs) / 1e9
    del reqs
    torch.cuda.empty_cache()
    return {"workload": f"{n} requests, Poisson {rate_hz}/s, seed {seed}: {px / 1e9:.1f} G output px-frames, "
                        f"{stack_gb:.1f} GB of stacks in pinned host memory",
            "rate_hz": rate_hz, "served": served, "plain_per_request": plain,
            "saturated_served": served_sat, "saturated_plain_per_request": plain_sat, "rate_curve": curve,
            "p99_speedup": round(plain["p99_ms"] / served["p99_ms"], 2),
            "throughput_speedup": round(served["throughput_mpx_frames_s"] / plain["throughput_mpx_frames_s"], 2)}


def cpu_baseline(x, e, seconds):
    """The reference CPU chain (CR xM, Roberts, CR x2M twice, RGC, temporal) on the first frames of the same
    stack, chunk after chunk until `seconds` of wall clock have passed. Returns the baseline record and the
    reference's per-frame maps of the frames it processed (for the parity check)."""
    from oracle import ref32
    ref32.build()
    kind = "port"
    if ref32.reference_available():
        ref32.use_reference_kernels(True)
        kind = "reference"
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    N, H, W = x.shape
    chunk = 4
    done = 0
    kept = []
    t0 = time.perf_counter()
    while done < N:
        part = x[done:done + chunk].astype(np.float32)
        maps = ref32.esrrf(part, M, R, S, iw)
        for mode in e["temporal"]:
            ref32.temporal(maps, mode)
        kept.append(maps)
        done += part.shape[0]
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    ref32.use_reference_kernels(False)
    value = done * H * M * W * M / 1e6 / dt
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(value, 3), "unit": "Mpx·frames/s", "cores": ref32.num_threads(), "kind": kind,
            "sample": f"first {done} of {N} frames ({H}x{W}, M{M}, R{R}, S{S}, {'+'.join(e['temporal'])}), "
                      f"reference CPU chain in {dt:.1f} s; {cpu}, nproc {os.cpu_count()}, "
                      f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}"}, np.concatenate(kept)


def parity_on_sample(x_dev, ref_maps, e):
    """Max and mean of |gpu - ref| / max|ref| for each temporal map of the frames the CPU baseline processed: the
    fused GPU path on those frames against NumPy's temporal step over the reference's per-frame maps."""
    from oracle import ref32
    from nanopyx_amd import _hip
    n = ref_maps.shape[0]
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    gpu = _hip.esrrf_temporal(x_dev[:n], M, R, S, iw, tuple(e["temporal"]))
    out = {"frames": int(n), "against": "reference CPU chain (oracle/_ref kernels) + NumPy float32 temporal"}
    for mode in e["temporal"]:
        ref = ref32.temporal(ref_maps, mode).astype(np.float64)
        g = gpu[mode].cpu().numpy().astype(np.float64)
        err = np.abs(g - ref) / max(np.abs(ref).max(), 1e-30)
        out[mode] = {"max": float(err.max()), "mean": float(err.mean())}
    return out


def config_leg(workload, reps, cpu_seconds, full_length_parity=False):
    """A plain-eSRRF workload of the benchmark set other than the headline, measured as the headline is: the fused
    eSRRF + temporal call over the whole stack resident in HBM (host wall clock per call, and the HIP-event time of
    the fused kernel and of the moment merge on the launch stream, over `reps` calls), the fused kernel's FP32
    roofline, the reference CPU chain (the reference's own C kernels in oracle/_ref) on a bounded frame sample of the
    same stack, and parity of every temporal map against NumPy's reduction of the reference's per-frame maps on the
    frames the CPU processed; with full_length_parity also against NumPy's reduction of the GPU per-frame maps of
    every frame of the stack (the drift of the streamed moments at full length)."""
    import torch
    from nanopyx_amd import _hip, synthetic
    from oracle import ref32
    sc = synthetic.CONFIGS[workload]
    e = sc.esrrf
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    modes = tuple(e["temporal"])
    x_host = synthetic.generate(sc)
    x = torch.from_numpy(x_host).cuda()
    N, H, W = x.shape
    for _ in range(2):
        _hip.esrrf_temporal(x, M, R, S, iw, modes)
    torch.cuda.synchronize()
    _hip.timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        _hip.esrrf_temporal(x, M, R, S, iw, modes)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    _hip.timing(False)
    n, k_ms, m_ms = _hip.timing_read()
    k_ms, m_ms = k_ms / max(n, 1), m_ms / max(n, 1)
    U = N * H * M * W * M
    s_bar = ring_samples_mean(H, W, M, R)
    flops = U * (9.0 * s_bar + 40.0)
    achieved = flops / (k_ms / 1e3) / 1e12
    res = {"workload": f"{workload}: {N} frames x {H}x{W} {sc.dtype}, M{M} R{R} S{S} intensity-weighted, "
                       f"{'+'.join(modes)} -> {H * M}x{W * M}",
           "ms_per_call": round(dt * 1e3, 4), "kernel_ms": round(k_ms, 4), "merge_ms": round(m_ms, 4),
           "value": round(U / 1e6 / dt, 1), "value_kernels": round(U / 1e6 / ((k_ms + m_ms) / 1e3), 1),
           "unit": "Mpx·frames/s",
           "roofline": {"bound": BOUND, "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "flops_per_launch": flops,
                        "ring_samples_mean": s_bar}}
    cpu, ref_maps = cpu_baseline(x_host, e, cpu_seconds)
    res["cpu_baseline"] = cpu
    res["parity"] = parity_on_sample(x, ref_maps, e)
    del ref_maps
    if full_length_parity:
        ref32.build()
        fused = _hip.esrrf_temporal(x, M, R, S, iw, modes)
        frames = _hip.esrrf_frames(x, M, R, S, iw).cpu().numpy()
        full = {"frames": int(N), "against": "NumPy float32 reduction of the GPU per-frame maps of every frame"}
        for mode in modes:
            ref = ref32.temporal(frames, mode).astype(np.float64)
            g = fused[mode].cpu().numpy().astype(np.float64)
            err = np.abs(g - ref) / max(np.abs(ref).max(), 1e-30)
            full[mode] = {"max": float(err.max()), "mean": float(err.mean())}
        res["parity_full_length"] = full
        del frames
    del x
    torch.cuda.empty_cache()
    return res


def end_to_end(x_host, e, reps=2):
    """Host stack -> device copy -> fused eSRRF + temporal -> maps back on the host, wall clock per call."""
    import torch
    from nanopyx_amd import _hip
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    pinned = torch.from_numpy(x_host).pin_memory()
    best = None
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xd = pinned.to("cuda", non_blocking=True)
        outs = _hip.esrrf_temporal(xd, M, R, S, iw, tuple(e["temporal"]))
        host = {k: v.cpu() for k, v in outs.items()}
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None or dt < best else best
        del xd, outs, host
    N, H, W = x_host.shape
    return {"ms": round(best * 1e3, 3), "value": round(N * H * M * W * M / 1e6 / best, 1),
            "unit": "Mpx·frames/s", "note": "pinned host stack copied in, maps copied out, best of "
                                            f"{reps} after one warm-up"}


def workflow_numpy(x_host, e, reps=2):
    """eSRRF() (the workflow function) called with the NumPy stack, first temporal mode, no macro-pixel correction:
    the stack is streamed (pinned pieces, one fed session job); wall clock per call including the map's copy back
    into a NumPy array, best of `reps` after one warm-up."""
    import torch
    import nanopyx_amd
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    mode = e["temporal"][0]
    best = None
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = nanopyx_amd.eSRRF(x_host, magnification=M, radius=R, sensitivity=S, temporal_correlation=mode,
                                doIntensityWeighting=iw, macro_pixel_correction=False)
        dt = time.perf_counter() - t0
        best = dt if best is None or dt < best else best
        del out
    N, H, W = x_host.shape
    return {"ms": round(best * 1e3, 3), "value": round(N * H * M * W * M / 1e6 / best, 1),
            "unit": "Mpx·frames/s", "mode": mode,
            "note": f"eSRRF() on the NumPy stack ({x_host.dtype}, pageable host memory): streamed through pinned pieces "
                    f"into a fed session job, {mode} map copied back, best of {reps} after one warm-up"}


def end_to_end_streamed(x_host, e, reps=2):
    """Host stack -> maps on the host through a serving session: the session uploads the pinned stack in pieces of
    about 32 MB and runs each slice of frames as soon as its piece has arrived, so the copy overlaps the kernels;
    wall clock per call (submit, step until done, maps copied out)."""
    import torch
    from nanopyx_amd import _hip
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    modes = tuple(e["temporal"])
    pinned = torch.from_numpy(x_host).pin_memory()
    N, H, W = x_host.shape
    need = _hip.Session.job_bytes(N, H, W, M, R, x_host.dtype == np.uint16, True)
    best = None
    for _ in range(reps + 1):
        outs = {m: torch.empty((H * M, W * M), device="cuda") for m in modes}
        sess = _hip.Session(need + (64 << 20), 1, 1 << 29)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        jid = sess.submit(pinned, M, R, S, iw, modes, outs)
        while sess.poll(jid) != _hip.Session.DONE:
            sess.step()
        host = {k: v.cpu() for k, v in outs.items()}
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = sess.stats()
        sess.close()
        best = dt if best is None or dt < best else best
        del outs, host
    return {"ms": round(best * 1e3, 3), "value": round(N * H * M * W * M / 1e6 / best, 1),
            "unit": "Mpx·frames/s", "slices": st["slices"], "launches": st["launches"],
            "note": f"pinned host stack uploaded in pieces overlapping the kernels (serving session), maps copied "
                    f"out, best of {reps} after one warm-up"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large_512_u16")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="re-run the launch-shape autotuner on the workload")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-srrf", action="store_true")
    ap.add_argument("--srrf-only", action="store_true", help="run only the SRRF radiality leg and print its JSON")
    ap.add_argument("--no-esrrf3d", action="store_true")
    ap.add_argument("--no-fht", action="store_true")
    ap.add_argument("--no-operators", action="store_true", help="skip the eSRRF_ST.run and RGC-on-arrays leg")
    ap.add_argument("--operators-only", action="store_true",
                    help="run only the eSRRF_ST.run and RGC-on-arrays leg and print its JSON")
    ap.add_argument("--no-configs", action="store_true", help="skip the cpu_small and plain_256 legs")
    ap.add_argument("--configs-only", action="store_true", help="run only the cpu_small and plain_256 legs")
    ap.add_argument("--serving-only", action="store_true", help="run only the served-stream leg and print its JSON")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one plain 1 GiB device-to-device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--fht-only", action="store_true", help="run only the use_fht leg and print its JSON")
    ap.add_argument("--esrrf3d-only", action="store_true", help="run only the eSRRF 3D leg and print its JSON")
    ap.add_argument("--sweep-only", action="store_true", help="run only the parameter-sweep leg and print its JSON")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the timed fused launches of the workload (no other legs, no CPU baseline): for "
                         "profiling the fused kernel")
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--serving-rate", type=float, default=50.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the eSRRF path is built and measured on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is a design note only)", flush=True)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = 1 << 30
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)                        # warm-up copy
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 2}), flush=True)
        return
    if args.srrf_only:
        print(json.dumps(srrf_bench(max(5, args.steps), cpu_frames=0 if args.no_cpu else 2)), flush=True)
        return
    if args.operators_only:
        print(json.dumps(operators_bench(max(5, args.steps), cpu_seconds=0.0 if args.no_cpu else 6.0)), flush=True)
        return
    if args.fht_only:
        print(json.dumps(fht_bench(max(3, args.steps // 2), cpu_frames=0 if args.no_cpu else 1)), flush=True)
        return
    if args.serving_only:
        print(json.dumps(serving_bench(args.serving_requests, args.serving_rate)), flush=True)
        return
    if args.configs_only:
        print(json.dumps({"cpu_small": config_leg("cpu_small", max(5, args.steps), 10.0),
                          "plain_256": config_leg("plain_256", max(5, args.steps), 10.0, True)}), flush=True)
        return
    if args.esrrf3d_only:
        print(json.dumps(esrrf3d_bench(max(3, args.steps), cpu_frames=0 if args.no_cpu else 1)), flush=True)
        return
    if args.sweep_only:
        print(json.dumps(sweep_bench(max(3, args.steps // 2))), flush=True)
        return
    if args.kernel_only:
        for k in ("no_configs", "no_sweep", "no_srrf", "no_esrrf3d", "no_fht", "no_serving", "no_cpu", "no_operators"):
            setattr(args, k, True)

    import torch
    from nanopyx_amd import _hip, synthetic

    sc = synthetic.CONFIGS[args.workload]
    e = sc.esrrf
    M, R, S, iw = e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"]
    modes = tuple(e["temporal"])
    log(f"generating {args.workload}: {sc.frames}x{sc.height}x{sc.width} {sc.dtype}")
    t = time.perf_counter()
    x_host = synthetic.generate(sc)
    log(f"generated in {time.perf_counter() - t:.1f} s")
    N, H, W = x_host.shape
    HM, WM = H * M, W * M
    x = torch.from_numpy(x_host).cuda()
    torch.cuda.synchronize()
    log(f"device: {_hip.device_info()}")

    from nanopyx_amd import autotune
    from nanopyx_amd.tuning import PACKAGED
    tune_info = {"store": os.path.relpath(PACKAGED, ROOT), "loaded": autotune.load(PACKAGED) if os.path.exists(
        PACKAGED) else 0}
    if args.autotune:
        t0 = time.perf_counter()
        tune_info["live"] = autotune.tune(x, M, R, S, iw, modes, persist=False)
        tune_info["tune_seconds"] = round(time.perf_counter() - t0, 2)
    mw, tb, tuned = _hip.tune_get(H, W, M, R)
    tune_info.update(max_waves=mw, target_blocks=tb, tuned=tuned)
    log(f"launch shape: {tune_info}")

    def step():
        return _hip.esrrf_temporal(x, M, R, S, iw, modes)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _hip.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _hip.timing(False)
    ncalls, kern_ms, merge_ms = _hip.timing_read()
    log(f"{args.steps} steps in {dt:.3f} s; fused kernel {kern_ms / max(ncalls, 1):.3f} ms/launch, "
        f"merge {merge_ms / max(ncalls, 1):.3f} ms/launch")

    U = N * HM * WM
    value = U / 1e6 / (dt / args.steps)
    s_bar = ring_samples_mean(H, W, M, R)
    flops = U * (9.0 * s_bar + 40.0)
    k_s = kern_ms / max(ncalls, 1) / 1e3
    achieved = flops / k_s / 1e12
    in_bytes = N * H * W * (2 if sc.dtype == "uint16" else 4)
    alg_bytes = in_bytes + len(modes) * HM * WM * 4
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpx·frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {N} frames x {H}x{W} {sc.dtype}, M{M} R{R} S{S} "
                               f"intensity-weighted, {'+'.join(modes)} -> {HM}x{WM}",
                   "frames": N, "height": H, "width": W, "magnification": M, "radius": R, "sensitivity": S,
                   "temporal": list(modes), "input_dtype": sc.dtype},
        "roofline": {"bound": BOUND, "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic_from_profile(args.workload),
                     "compute": "FP32 vector ALU (no MFMA on this stencil path)",
                     "flops_per_launch": flops, "ring_samples_mean": s_bar,
                     "kernel_ms": round(k_s * 1e3, 4), "merge_ms": round(merge_ms / max(ncalls, 1), 4),
                     "algorithmic_bytes": alg_bytes,
                     "hbm_gbs_algorithmic": round(alg_bytes / k_s / 1e9, 1)},
        "launch_shape": tune_info,
        "cpu_baseline": None,
        "configs": None,
        "sweep": None,
        "serving": None,
        "srrf": None,
        "esrrf3d": None,
        "use_fht": None,
        "operators": None,
    }
    if not args.no_configs:
        log("cpu_small and plain_256 workloads ...")
        result["configs"] = {"cpu_small": config_leg("cpu_small", max(5, args.steps), 0.0 if args.no_cpu else 10.0),
                             "plain_256": config_leg("plain_256", max(5, args.steps), 0.0 if args.no_cpu else 10.0,
                                                     not args.no_cpu)}
        log(f"configs: {result['configs']}")
    if not args.no_sweep:
        log("parameter sweep ...")
        result["sweep"] = sweep_bench(max(3, args.steps // 2))
        log(f"sweep: {result['sweep']}")
    if not args.no_srrf:
        log("SRRF radiality ...")
        result["srrf"] = srrf_bench(max(5, args.steps), cpu_frames=0 if args.no_cpu else 2)
        log(f"srrf: {result['srrf']}")
    if not args.no_esrrf3d:
        log("eSRRF 3D ...")
        result["esrrf3d"] = esrrf3d_bench(max(3, args.steps), cpu_frames=0 if args.no_cpu else 1)
        log(f"esrrf3d: {result['esrrf3d']}")
    if not args.no_fht:
        log("eSRRF use_fht ...")
        result["use_fht"] = fht_bench(max(3, args.steps // 2), cpu_frames=0 if args.no_cpu else 1)
        log(f"use_fht: {result['use_fht']}")
    if not args.no_operators:
        log("eSRRF_ST.run and RGC on arrays ...")
        result["operators"] = operators_bench(max(5, args.steps), cpu_seconds=0.0 if args.no_cpu else 6.0)
        log(f"operators: {result['operators']}")
    if not args.no_serving:
        log("served stream ...")
        result["serving"] = serving_bench(args.serving_requests, args.serving_rate)
    if not args.kernel_only:
        log("end to end (host stack in, maps out) ...")
        result["end_to_end_with_pcie"] = end_to_end(x_host, e)
        result["end_to_end_streamed"] = end_to_end_streamed(x_host, e)
        result["workflow_numpy_stack"] = workflow_numpy(x_host, e)
        log(f"end to end: {result['end_to_end_with_pcie']} {result['end_to_end_streamed']} "
            f"{result['workflow_numpy_stack']}")
    if not args.no_cpu:
        log("timing the reference CPU path ...")
        result["cpu_baseline"], ref_maps = cpu_baseline(x_host, e, args.cpu_seconds)
        result["parity"] = parity_on_sample(x, ref_maps, e)
        log(f"parity: {result['parity']}")
        del ref_maps
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
