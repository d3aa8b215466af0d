"""Autotunes the launch shape of the fused kernel for the benchmark workloads on this GPU and writes the packaged
store nanopyx_amd/tuning/mi355x.json (or the path given as the first argument).

Workloads: the synthetic stacks of nanopyx_amd.synthetic.CONFIGS with single eSRRF settings (frames x H x W,
magnification, radius, temporal modes as listed there). Each is rendered, uploaded and tuned with
nanopyx_amd.autotune.tune (every candidate shape timed by npx_autotune on that stack); the per-frame-map kernels are
tuned with nanopyx_amd.autotune.tune_frames on the first 100 frames of plain_256 (the operators the benchmark times
on their own). ``--frames-only`` runs only that step.
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main(path=None, frames_only=False):
    import torch
    from nanopyx_amd import autotune, synthetic
    from nanopyx_amd.tuning import PACKAGED
    path = path or PACKAGED
    os.environ["NANOPYX_AMD_TUNE_FILE"] = path
    for name in () if frames_only else ("cpu_small", "plain_256", "large_512_u16"):
        sc = synthetic.CONFIGS[name]
        e = sc.esrrf
        x = torch.from_numpy(synthetic.generate(sc)).cuda()
        entry = autotune.tune(x, e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"],
                              tuple(e["temporal"]))
        print(name, json.dumps(entry), flush=True)
        del x
    sc = synthetic.CONFIGS["plain_256"]
    e = sc.esrrf
    x = torch.from_numpy(synthetic.generate(sc, frames=100)).cuda()
    entry = autotune.tune_frames(x, e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"])
    print("plain_256[:100] per-frame maps", json.dumps(entry), flush=True)
    print("wrote", path, flush=True)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "--frames-only"]
    main(args[0] if args else None, "--frames-only" in sys.argv[1:])
