"""Launch-shape autotuning with a persistent store.

``tune(stack, M, R, S, iw, modes)`` runs the library's autotuner (npx_autotune: every candidate shape timed on the
given stack) and records the winner for the stack's (H, W, M, R) in a JSON store; ``tune_frames(stack, M, R, S, iw)``
does the same for the per-frame-map kernels (npx_autotune_frames; entries keyed with a trailing ``|frames``), whose
best shape can differ; ``load()`` hands the stored shapes back to the library in a new process. The store is ``$NANOPYX_AMD_TUNE_FILE`` or
``~/.cache/nanopyx_amd/launch_tuning.json``; entries are keyed by the device (architecture and CU count) as well,
since the best shape is a property of the GPU. It plays the part of the Liquid Engine's per-run-type benchmark store (the reference times run
types and lets its Agent pick the fastest; here the run type is fixed and the launch shape is what gets timed).
"""
from __future__ import annotations

import json
import os

from . import _hip


def store_path() -> str:
    return os.environ.get("NANOPYX_AMD_TUNE_FILE",
                          os.path.join(os.path.expanduser("~"), ".cache", "nanopyx_amd", "launch_tuning.json"))


def device_id() -> str:
    """Architecture and CU count of the current device, e.g. "gfx950/256" (the product name HIP reports varies)."""
    name, arch, cus = _hip.device_info().split(";")[:3]
    return f"{arch.split(':')[0]}/{cus}"


def _key(device, H, W, M, R, frames=False):
    return f"{device}|{int(H)}|{int(W)}|{int(M)}|{float(R)!r}" + ("|frames" if frames else "")


def _read(path):
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _record(stack, magnification, radius, mw, tb, ms, frames, persist):
    N, H, W = (stack.shape if len(stack.shape) == 3 else (1,) + tuple(stack.shape))
    entry = {"max_waves": mw, "target_blocks": tb, "ms": round(ms, 4), "frames": int(N)}
    if persist:
        path = store_path()
        data = _read(path)
        data[_key(device_id(), H, W, magnification, radius, frames)] = entry
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    return entry


def tune(stack, magnification, radius, sensitivity=1, iw=True, modes=("AVG",), persist=True):
    """Autotunes the launch shape for this stack's (H, W, M, R); returns {"max_waves", "target_blocks", "ms"}."""
    mw, tb, ms = _hip.autotune(stack, magnification, radius, sensitivity, iw, modes)
    return _record(stack, magnification, radius, mw, tb, ms, False, persist)


def tune_frames(stack, magnification, radius, sensitivity=1, iw=True, persist=True):
    """Autotunes the per-frame-map kernels' launch shape for this stack's (H, W, M, R) (eSRRF_ST.run and the RGC on
    arrays of that size); returns {"max_waves", "target_blocks", "ms"}."""
    mw, tb, ms = _hip.autotune_frames(stack, magnification, radius, sensitivity, iw)
    return _record(stack, magnification, radius, mw, tb, ms, True, persist)


def load(path=None) -> int:
    """Hands every stored shape of this device to the library; returns how many were loaded."""
    data = _read(path or store_path())
    dev = device_id()
    n = 0
    for k, v in data.items():
        parts = k.split("|")
        frames = len(parts) == 6 and parts[5] == "frames"
        if len(parts) not in (5, 6) or (len(parts) == 6 and not frames) or parts[0] != dev:
            continue
        setter = _hip.tune_frames_set if frames else _hip.tune_set
        setter(int(parts[1]), int(parts[2]), int(parts[3]), float(parts[4]), int(v["max_waves"]),
               int(v["target_blocks"]))
        n += 1
    return n
