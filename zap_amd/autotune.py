"""Launch-shape autotuner for the grid-stride kernels (zaptex_set_launch), measured on the real workloads.

For each tunable kernel in turn (coordinate descent, in the order of their share of the time) it tries every
workgroup size (64, 128, 256) and workgroup cap, timing the whole workload with HIP events:
  c4  one graph launch of the 512-texture batch (zaptex_batch_run; the plan is rebuilt per shape)
  c2  one step of the 4096^2 chain (zaptex_mip_chain_dev + zaptex_encode_dev)
and keeps the fastest.  Results go to zap_amd/tuning/mi355x.json, which `apply` loads.

    python -m zap_amd.autotune            # both workloads, about a minute on one MI355X
"""
from __future__ import annotations

import json
import os
import time

from . import binding

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning", "mi355x.json")
KERNELS = {"c4": ["zt_bc1_climb", "zt_bc7_stage6", "zt_bc14_encode", "zt_bc7_stage5", "zt_mip_batch", "zt_bc7_s1"],
           "c2": ["zt_bc7_stage6", "zt_bc7_stage5", "zt_bc7_s1"]}
WGS = (64, 128, 256)
CAPS = (256, 512, 1024, 2048, 4096, 8192, 16384, 1 << 20)


def _time(fn, reps: int) -> float:
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(lib, workload: str, step, prepare=None, reps: int = 3, log=print) -> dict:
    """Coordinate descent over KERNELS[workload]; `prepare()` runs after every shape change (e.g. replan)."""
    res = {"kernels": {}, "trials": []}
    base = _time(step, reps)
    res["default_ms"] = base
    best_ms = base
    for k in KERNELS[workload]:
        cur = lib.get_launch(k)
        best = (cur, best_ms)
        for wg in WGS:
            for cap in CAPS:
                lib.set_launch(k, wg, cap)
                if prepare:
                    prepare()
                ms = _time(step, reps)
                res["trials"].append([k, wg, cap, round(ms, 4)])
                if ms < best[1] * 0.995:
                    best = ((wg, cap), ms)
        lib.set_launch(k, *best[0])
        if prepare:
            prepare()
        best_ms = best[1]
        res["kernels"][k] = {"wg": best[0][0], "max_groups": best[0][1]}
        if log:
            log(f"  {workload} {k}: wg {best[0][0]} cap {best[0][1]} -> {best_ms:.3f} ms", flush=True)
    res["tuned_ms"] = _time(step, reps * 2)
    return res


def apply(lib, workload: str, path: str = TUNING_PATH) -> dict | None:
    """Set the launch shapes tuned for `workload`; returns them (None when no tuning file exists)."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        cfg = json.load(fh).get(workload)
    if not cfg:
        return None
    for k, v in cfg["kernels"].items():
        lib.set_launch(k, v["wg"], v["max_groups"])
    return cfg["kernels"]


DEFAULTS = {"zt_bc14_encode": (256, 8192), "zt_bc7_stage6": (256, 8192), "zt_bc7_stage5": (256, 2048), "zt_bc7_s1": (256, 1024),
            "zt_mip_batch": (256, 8192), "zt_bc1_climb": (256, 8192)}


def reset(lib) -> None:
    for k, (wg, cap) in DEFAULTS.items():
        lib.set_launch(k, wg, cap)


if __name__ == "__main__":
    import torch

    from . import batch as B
    from . import synth
    lib = binding.load()
    out = {"device": torch.cuda.get_device_name(0), "date": time.strftime("%Y-%m-%d")}
    # c2: the 4096^2 chain step
    reset(lib)
    c2 = torch.from_numpy(synth.c2_bc7(4096, 2)).cuda()
    sizes = binding.mip_sizes(4096, 4096)
    levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
    outs = [torch.empty(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    srcs = [c2] + levels
    imgs = lib.images((srcs[i], outs[i], w, h, w * 4, "bc7") for i, (w, h) in enumerate(sizes))
    ws = torch.empty(lib.workspace_size(imgs), dtype=torch.uint8, device="cuda")

    def c2_step():
        lib.mip_chain_dev(c2, 4096, 4096, 4096 * 4, levels, srgb=0)
        lib.encode_dev(imgs, ws)
    out["c2"] = tune(lib, "c2", c2_step, reps=20)
    print(f"c2: {out['c2']['default_ms']:.4f} -> {out['c2']['tuned_ms']:.4f} ms", flush=True)
    del c2, levels, srcs, outs, imgs, ws
    # c4: the 512-texture batch
    reset(lib)
    specs, bases = B.c4_textures(512, 4)
    bt = B.Batch(lib, [(bases[i], w, h, fmt, srgb, None) for i, (w, h, fmt, srgb, kind) in enumerate(specs)])
    out["c4"] = tune(lib, "c4", bt.run, prepare=bt.replan, reps=3)
    out["c4"]["pixels"] = bt.pixels
    print(f"c4: {out['c4']['default_ms']:.3f} -> {out['c4']['tuned_ms']:.3f} ms ({bt.pixels / 1e6:.0f} Mpx)", flush=True)
    bt.close()
    os.makedirs(os.path.dirname(TUNING_PATH), exist_ok=True)
    dest = os.environ.get("ZAPTEX_TUNING_OUT", TUNING_PATH)
    with open(dest, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {dest}", flush=True)
