"""Autotunes the launch shape of the fused kernel for the benchmark workloads on this GPU and writes the packaged
store nanopyx_amd/tuning/mi355x.json (or the path given as the first argument).

Workloads: the synthetic stacks of nanopyx_amd.synthetic.CONFIGS with single eSRRF settings (frames x H x W,
magnification, radius, temporal modes as listed there). Each is rendered, uploaded and tuned with
nanopyx_amd.autotune.tune (every candidate shape timed by npx_autotune on that stack).
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main(path=None):
    import torch
    from nanopyx_amd import autotune, synthetic
    from nanopyx_amd.tuning import PACKAGED
    path = path or PACKAGED
    os.environ["NANOPYX_AMD_TUNE_FILE"] = path
    for name in ("cpu_small", "plain_256", "large_512_u16"):
        sc = synthetic.CONFIGS[name]
        e = sc.esrrf
        x = torch.from_numpy(synthetic.generate(sc)).cuda()
        entry = autotune.tune(x, e["magnification"], e["radius"], e["sensitivity"], e["doIntensityWeighting"],
                              tuple(e["temporal"]))
        print(name, json.dumps(entry), flush=True)
        del x
    print("wrote", path, flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
