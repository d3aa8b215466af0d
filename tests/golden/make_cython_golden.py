"""Writes tests/golden/cython_golden.npz: outputs of the reference's own Cython functions that no other fixture pins —
the sweep scoring's FRC/FIRE, decorrelation and Pearson correlation, and the macro-pixel corrector — on fixed inputs.

The reference's Cython modules core/analysis/frc.pyx, core/analysis/decorr.pyx, core/analysis/pearson_correlation.pyx
(the error map's RSP) and core/transform/mpcorrector.pyx, and the modules they import or
cimport (core/analysis/ccm_helper_functions, core/utils/math, core/utils/array, core/utils/timeit,
core/generate/mask, core/transform/padding, core/transform/edges, core/transform/normalize), are compiled from where
they lie under the reference checkout into a temporary directory outside this repository, with the reference's Linux
wheel flags (-O3 -fopenmp). The directory holds only those modules, in their package folders and without any
__init__ file: Python imports the folders as namespace packages, so none of the reference's package initialisation
(its installed-metadata lookup, the Liquid Engine) runs and no stand-in for any reference module is written. The
script imports the analyses and the corrector from there and records inputs and outputs; nothing of the reference
is copied into the repository. Run it in the build container (it needs the reference checkout, Cython, gcc):

    python tests/golden/make_cython_golden.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODULES = [  # (package folder under nanopyx/core, module, compiled)
    ("analysis", "frc", True), ("analysis", "decorr", True), ("analysis", "ccm_helper_functions", True),
    ("analysis", "pearson_correlation", True), ("transform", "mpcorrector", True),
    ("utils", "math", True), ("utils", "array", True), ("utils", "timeit", False),
    ("generate", "mask", True), ("transform", "padding", False), ("transform", "edges", True),
    ("transform", "normalize", True),
]


def build(ref_root: str, tmp: str) -> list[str]:
    """Places the modules' sources in tmp/nanopyx/core/<folder>/ (symbolic links to the reference's files), builds the
    Cython ones in place; returns the source paths with their SHA-256."""
    core = os.path.join(ref_root, "src", "nanopyx", "core")
    exts, digests = [], []
    for folder, mod, compiled in MODULES:
        dst = os.path.join(tmp, "nanopyx", "core", folder)
        os.makedirs(dst, exist_ok=True)
        for ext in ((".pyx", ".pxd") if compiled else (".py",)):
            src = os.path.join(core, folder, mod + ext)
            if os.path.exists(src):
                os.symlink(src, os.path.join(dst, mod + ext))
                with open(src, "rb") as fh:
                    digest = hashlib.sha256(fh.read()).hexdigest()
                digests.append(f"{digest}  src/nanopyx/core/{folder}/{mod}{ext}")
        if compiled:
            exts.append(f"nanopyx/core/{folder}/{mod}.pyx")
    setup = os.path.join(tmp, "setup_golden.py")
    with open(setup, "w") as f:
        f.write(
            "import numpy\n"
            "from setuptools import setup, Extension\n"
            "from Cython.Build import cythonize\n"
            f"mods = {exts!r}\n"
            "ext = [Extension(m[:-4].replace('/', '.'), [m], include_dirs=[numpy.get_include()],\n"
            "                 extra_compile_args=['-O3', '-fopenmp'], extra_link_args=['-fopenmp']) for m in mods]\n"
            "setup(ext_modules=cythonize(ext, language_level=3, quiet=True), script_args=['build_ext', '--inplace'])\n")
    r = subprocess.run([sys.executable, setup], cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise SystemExit("building the reference's Cython modules failed:\n" + r.stdout[-4000:])
    return digests


def images():
    """Fixed test images: two noisy realisations of one blurred emitter field (FRC pairs), a band-limited pattern and a
    blurred noise field (decorrelation), with seeds and sizes recorded."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(20261204)
    out = {}
    for n in (64, 96):
        truth = np.zeros((n, n), np.float64)
        ys, xs = rng.integers(4, n - 4, size=(2, n // 2))
        truth[ys, xs] = rng.uniform(500, 1500, size=ys.size)
        truth = gaussian_filter(truth, 1.6) + 20.0
        a = rng.poisson(truth).astype(np.float32)
        b = rng.poisson(truth).astype(np.float32)
        out[f"pair{n}_a"], out[f"pair{n}_b"] = a, b
    yy, xx = np.mgrid[0:80, 0:80]
    out["pattern80"] = (100 + 40 * np.cos(2 * np.pi * xx / 9.0) * np.cos(2 * np.pi * yy / 13.0) +
                        rng.normal(0, 4, (80, 80))).astype(np.float32)
    out["blur72"] = gaussian_filter(rng.normal(0, 1, (72, 72)), 2.0).astype(np.float32) * 100 + 300
    return out


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tmp = tempfile.mkdtemp(prefix="npx_scoring_ref_")
    try:
        digests = build(ref_root, tmp)
        sys.path.insert(0, tmp)
        from nanopyx.core.analysis.frc import FIRECalculator
        from nanopyx.core.analysis.decorr import DecorrAnalysis
        from nanopyx.core.analysis.pearson_correlation import pearson_correlation
        from nanopyx.core.transform.mpcorrector import macro_pixel_corrector
        imgs = images()
        rec = dict(imgs)
        for n in (64, 96):
            a, b = imgs[f"pair{n}_a"], imgs[f"pair{n}_b"]
            fc = FIRECalculator(pixel_size=1, units="pixel")
            rec[f"fire{n}"] = np.float64(fc.calculate_fire_number(a, b))
        for name in ("pattern80", "blur72", "pair64_a", "pair96_a"):
            d = DecorrAnalysis(pixel_size=1, units="pixel")
            d.run_analysis(imgs[name])
            rec[f"decorr_{name}"] = np.float64(d.resolution)
        # non-default analysis parameters: radius range, sampling, pixel size
        d = DecorrAnalysis(rmin=0.05, rmax=0.8, n_r=40, n_g=8, pixel_size=2.5, units="nm")
        d.run_analysis(imgs["pair96_b"])
        rec["decorr_pair96_b_params"] = np.float64(d.resolution)
        rec["pearson_pair64"] = np.float64(pearson_correlation(imgs["pair64_a"], imgs["pair64_b"]))
        rec["pearson_pattern_blur"] = np.float64(pearson_correlation(imgs["pattern80"][:72, :72], imgs["blur72"]))
        # macro-pixel corrector on eSRRF-like maps: a positive phase pattern times a smooth field (3 slices at M5),
        # one slice at M4 with zero phase entries (the division skips them), and one slice at M2
        rng = np.random.default_rng(7)
        yy, xx = np.mgrid[0:40, 0:35]
        field = 1.0 + 0.5 * np.sin(xx / 6.0) * np.cos(yy / 5.0)
        ph = 1.0 + 0.3 * rng.random((5, 5))
        mpc5 = np.stack([field * np.tile(ph, (8, 7)) * (s + 1) + rng.random((40, 35)) * 0.01 for s in range(3)])
        mpc4 = (rng.random((1, 24, 32)) * np.tile(rng.random((4, 4)) > 0.2, (6, 8))).astype(np.float32)
        mpc2 = (rng.random((1, 18, 22)) + 0.5).astype(np.float32)
        for name, img, M in (("mpc5", mpc5, 5), ("mpc4", mpc4, 4), ("mpc2", mpc2, 2)):
            rec[f"{name}_in"] = img.astype(np.float32)
            rec[f"{name}_out"] = np.asarray(macro_pixel_corrector(img.astype(np.float32), M), np.float32)
        rec["sources_sha256"] = np.array("\n".join(digests))
        path = os.path.join(HERE, "cython_golden.npz")
        np.savez_compressed(path, **rec)
        print(f"wrote {path}: " + ", ".join(f"{k}={float(v):.6g}" for k, v in rec.items() if np.ndim(v) == 0
                                              and k != "sources_sha256"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
