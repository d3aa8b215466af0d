"""Build of libnpx_hip.so (gfx950) in-tree: nanopyx_amd/lib/libnpx_hip.so.

Device code is compiled with packed FP32 disabled (``-packed-fp32-ops``): every FP32 operation stays a
single-lane VALU op. The host-side plan (npx_plan.cpp) is built with ``-ffp-contract=off`` so its float32
roundings follow the reference's arithmetic operation by operation.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "lib")
BUILD_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(OUT_DIR, "libnpx_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({' '.join(cmd[:3])} ...):\n{r.stdout}")
    if verbose and out:
        print(out, flush=True)


def build(verbose: bool = False, extra: list[str] | None = None) -> str:
    """Compile every source of the library; always rebuilds. Returns the library path."""
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(BUILD_DIR, exist_ok=True)
    extra = list(extra or [])
    inc = ["-I", CSRC, "-I", os.path.join(os.path.dirname(PKG), "include")]
    plan_o = os.path.join(BUILD_DIR, "npx_plan.o")
    kern_o = os.path.join(BUILD_DIR, "npx_kernels.o")
    batch_o = os.path.join(BUILD_DIR, "npx_batched.o")
    capi_o = os.path.join(BUILD_DIR, "npx_capi.o")
    sess_o = os.path.join(BUILD_DIR, "npx_session.o")
    srrf_o = os.path.join(BUILD_DIR, "npx_srrf.o")
    e3d_o = os.path.join(BUILD_DIR, "npx_esrrf3d.o")
    fht_o = os.path.join(BUILD_DIR, "npx_fht.o")
    interp_o = os.path.join(BUILD_DIR, "npx_interp.o")
    hip = [HIPCC, "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", *DEVICE_FLAGS, *inc]
    steps = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", *inc, "-c",
         os.path.join(CSRC, "npx_plan.cpp"), "-o", plan_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_kernels.hip"), "-o", kern_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_batched.hip"), "-o", batch_o],
        [*hip, "-x", "hip", "-c", os.path.join(CSRC, "npx_capi.cpp"), "-o", capi_o],
        [*hip, "-x", "hip", "-c", os.path.join(CSRC, "npx_session.cpp"), "-o", sess_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_srrf.hip"), "-o", srrf_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_esrrf3d.hip"), "-o", e3d_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_fht.hip"), "-o", fht_o],
        [*hip, *extra, "-c", os.path.join(CSRC, "npx_interp.hip"), "-o", interp_o],
    ]
    # the translation units are independent: compile them concurrently
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(steps)) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd in steps]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", kern_o, batch_o, capi_o, sess_o, srrf_o, e3d_o, fht_o,
          interp_o, plan_o, "-L", os.path.join(ROCM, "lib"), "-lhipfft", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o",
          tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
