#!/usr/bin/env python3
"""BC7 encode throughput on one MI355X: the 4096x4096 RGBA8 texture plus its 13-level box mip chain.

One step = generate the 12 mip levels from the resident 4096^2 base (zap_mip_next, linear) and BC7-encode all
13 levels (zap_bc_encode's search: mode 6, then mode 5 and mode 1 where the reference takes them), inputs
already in HBM.  value = 22,369,621 source pixels per step / step time, in Mpix/s.  The step's output is
checked against the reference's golden hashes before timing.

Also reported, measured inside this run:
  roofline      the dominant kernel (BC7 mode-6 stage) against the VALU issue ceiling measured in this run: its
                VALU wave instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU on this build and workload) over its
                average duration from HIP events on its launch stream, set against the fastest issue rate of an
                inline-asm probe (libztprobe.so); traffic = FETCH_SIZE + WRITE_SIZE per launch, each scaled by what it
                reads for a plain copy of known size; the oracle's scalar-op rate is given beside it as work, and the
                mip chain's GB/s against the HBM peak and the plain-copy rate measured in this run
  cpu_baseline  the reference encoder (oracle/_ref, built from zap_tex.h) on this host's cores, banded into
                8-block rows like the reference viewer, over the same 13-level chain
  batch         the C4 workload: 512 seeded textures (sides 64..4096 incl. non-powers of two; 30% BC1, 15% BC3,
                15% BC5 normal maps, 40% BC7) with full mip chains, one graph launch (zaptex_batch_run) per step,
                launch shapes from the autotuner (zap_amd/tuning/mi355x.json)
  learned       the learned-proposal encode (zap_amd/models/bc7_proposal.npy, top-4 exact + exhaustive fallback at
                the model's calibrated threshold) against the exhaustive all-mode search on the same chain:
                Mpix/s of both, chain PSNR of both (tex_bench.c's RGBA definition) and the gap; the network
                kernel's TFLOP/s against the dense bf16 MFMA peak ("network")
  configs       BC1 of the 1024^2 image, BC6H of the 2048^2 HDR field, ASTC 4x4 of the 2048^2 texture: device
                Mpix/s, golden-hash check, kernel times, and the reference encoder's rate on the host cores
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_CU, N_XCD = 256, 8                         # MI355X: GRBM_GUI_ACTIVE sums the 8 XCDs' busy cycles
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0                    # dense bf16 MFMA
MIP_CHAIN_BYTES = 89_478_484                 # read the 4096^2 base once + write every generated level once


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(levels_np, threads: int, repeats: int = 5, configs: dict | None = None) -> dict:
    """Reference zap_tex.h encoder (oracle/_ref; the oracle restatement if that is absent) on host cores; with
    `configs` ({name: image} of the other workloads) also each one's rate, banded the same way, median of 3."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import loader
    kind = "reference" if loader.reference_available() else "port"
    lib = loader.reference() if kind == "reference" else loader.oracle()
    bands = []
    for lvl in levels_np:
        h = lvl.shape[0]
        for y0 in range(0, h, 32):
            bands.append(lvl[y0:min(h, y0 + 32)])
    npx = sum(l.shape[0] * l.shape[1] for l in levels_np)
    times = []
    with ThreadPoolExecutor(threads) as ex:
        for _ in range(repeats):
            t0 = time.perf_counter()
            list(ex.map(lambda b: lib.bc_encode(b, "bc7"), bands))
            times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    one = levels_np[2]  # 1024^2, single thread
    t1 = time.perf_counter()
    lib.bc_encode(one, "bc7")
    t1 = time.perf_counter() - t1
    res = {"value": round(npx / t / 1e6, 3), "unit": "Mpix/s", "cores": threads, "kind": kind,
           "sample": f"BC7 of the full 13-level 4096^2 chain ({npx} px), 8-block-row bands on {threads} threads, median of {repeats}",
           "single_thread_mpix_s": round(one.shape[0] * one.shape[1] / t1 / 1e6, 3),
           "cpu": _cpu_model(), "nproc": os.cpu_count()}
    if configs:
        import numpy as np
        fmts = {name: fmt for name, fmt, _, _ in CONFIG_WORK}
        res["configs_mpix_s"] = {}
        for name, img in configs.items():
            fmt = fmts[name]
            enc = {"bc6h": lib.bc6h_encode, "astc": lib.astc_encode}.get(fmt, lambda b, f=fmt: lib.bc_encode(b, f))
            bands = [np.ascontiguousarray(img[y0:y0 + 32]) for y0 in range(0, img.shape[0], 32)]
            ts = []
            with ThreadPoolExecutor(threads) as ex:
                for _ in range(3):
                    t0 = time.perf_counter()
                    list(ex.map(enc, bands))
                    ts.append(time.perf_counter() - t0)
            res["configs_mpix_s"][name] = round(img.shape[0] * img.shape[1] / statistics.median(ts) / 1e6, 2)
    return res


VALU_PROBE_KINDS = ("v_add_u32", "v_dot4_u32_u8", "v_fma_f32", "v_max_i32", "v_lshl_add_u32", "v_perm_b32", "v_mad_u64_u32", "v_mul_lo_u32",
                    "v_mad_u32_u24")
COPY_PROBE_BYTES = 256 << 20


def _probe_lib():
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "zap_amd", "_lib", "libztprobe.so"))
    lib.ztprobe_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.ztprobe_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    return lib


def copy_probe(steps: int = 10) -> dict:
    """A plain 256 MiB device copy (zt_copy_probe): its rate, 2 x bytes / time, is the practical HBM bar."""
    import torch
    lib = _probe_lib()
    a = torch.empty(COPY_PROBE_BYTES, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    st = torch.cuda.current_stream().cuda_stream
    assert lib.ztprobe_copy(a.data_ptr(), b.data_ptr(), COPY_PROBE_BYTES, st) == 0
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        lib.ztprobe_copy(a.data_ptr(), b.data_ptr(), COPY_PROBE_BYTES, st)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"bytes": COPY_PROBE_BYTES, "ms": round(ms, 4), "gb_s": round(2 * COPY_PROBE_BYTES / (ms / 1e3) / 1e9, 1)}


def valu_probe(groups: int = 256 * 8, iters: int = 4096) -> dict:
    """The chip's VALU issue ceiling, measured: independent inline-asm chains of one instruction (32 wave64
    instructions per wave and round), 8 workgroups of 4 waves per CU, timed with HIP events on the launch stream.
    Returns wave instructions per second per kind and the fastest."""
    import torch
    lib = _probe_lib()
    out = torch.empty(groups * 256, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    insts = groups * 4 * iters * 32
    res = {"groups": groups, "waves": groups * 4, "iters": iters, "wave_insts_per_launch": insts}
    for kind, name in enumerate(VALU_PROBE_KINDS):
        assert lib.ztprobe_valu(kind, groups, iters, out.data_ptr(), stream.cuda_stream) == 0
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            lib.ztprobe_valu(kind, groups, iters, out.data_ptr(), stream.cuda_stream)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        res[name] = {"ms": round(ms, 4), "wave_insts_per_s": round(insts / (ms / 1e3) / 1e9, 2)}   # G wave instructions / s
        log(f"valu probe {name}: {ms:.3f} ms, {res[name]['wave_insts_per_s']} G wave-insts/s")
    res["ceiling_g_wave_insts_per_s"] = max(res[n]["wave_insts_per_s"] for n in VALU_PROBE_KINDS)
    return res


COUNTER_PASSES = (("SQ_INSTS_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE", "SQ_BUSY_CYCLES"), ("FETCH_SIZE",), ("WRITE_SIZE",))


def counters_child(c2_path: str) -> int:
    """The program the counter passes profile: one plain copy of COPY_PROBE_BYTES (calibrates the byte counters),
    then two steps of the workload (mip chain + BC7 encode of the 4096^2 chain) with the bench's launch shapes."""
    import numpy as np
    import torch

    import zap_amd
    from zap_amd import autotune, binding
    lib = zap_amd.load()
    autotune.apply(lib, "c2")
    plib = _probe_lib()
    a = torch.empty(COPY_PROBE_BYTES, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    plib.ztprobe_copy(a.data_ptr(), b.data_ptr(), COPY_PROBE_BYTES, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    del a, b
    c2 = np.load(c2_path)
    H, W = c2.shape[:2]
    sizes = binding.mip_sizes(W, H)
    base = torch.from_numpy(c2).cuda()
    levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
    outs = [torch.empty(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    imgs = lib.images(((base if i == 0 else levels[i - 1]), outs[i], w, h, w * 4, "bc7") for i, (w, h) in enumerate(sizes))
    ws = torch.empty(lib.workspace_size(imgs), dtype=torch.uint8, device="cuda")
    for _ in range(2):
        lib.mip_chain_dev(base, W, H, W * 4, levels, srgb=0)
        lib.encode_dev(imgs, ws)
    torch.cuda.synchronize()
    log("counters child: done")
    return 0


def rocprof_counters(c2_path: str, timeout: int = 180) -> dict:
    """Per-dispatch means of the COUNTER_PASSES counters for every kernel of counters_child, one rocprofv3 --pmc
    run per pass (a pass holds what one run can collect); {"error": ...} when the profiler is missing or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return {"error": "rocprofv3 not found"}
    out: dict = {}
    with tempfile.TemporaryDirectory() as tmp:
        for k, counters in enumerate(COUNTER_PASSES):
            d = os.path.join(tmp, f"p{k}")
            cmd = [exe, "--pmc", *counters, "-d", d, "-o", "c", "--output-format", "csv", "--",
                   sys.executable, os.path.abspath(__file__), "--counters-child", c2_path]
            log(f"counters: pass {k + 1}/{len(COUNTER_PASSES)} ({' '.join(counters)})")
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
            except subprocess.TimeoutExpired:
                return {"error": f"rocprofv3 pass {counters} timed out"}
            files = [os.path.join(root, f) for root, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
            if r.returncode != 0 and not files:
                return {"error": f"rocprofv3 pass {counters} failed (rc {r.returncode}): {r.stderr[-300:]}"}
            per = {}
            for row in csv.DictReader(open(files[0])):
                name = row["Kernel_Name"].split("(")[0].replace("void ", "")
                key = (name, row["Dispatch_Id"])
                per.setdefault(key, {})[row["Counter_Name"]] = float(row["Counter_Value"])
                per[key]["_ns"] = int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
            for (name, _), vals in per.items():
                agg = out.setdefault(name, {})
                for c, v in vals.items():
                    agg.setdefault(c, []).append(v)
    return {name: {c: sum(v) / len(v) for c, v in cs.items()} for name, cs in out.items()}


def learned_leg(lib, c2, steps: int = 5) -> dict:
    """Learned proposal vs exhaustive all-mode search on the 4096^2 chain (mip chain regenerated on the GPU), with
    the shipped fallback (the network's top fallback_k) and with the exhaustive search as the fallback
    (fallback_k = 0).  The plain search's encode-only rate on the same chain is set beside it: the learned encode
    trades speed against that heuristic for near-exhaustive quality."""
    from zap_amd import learned
    state, meta = learned.load(lib)
    chain = learned.Chain(lib, c2)
    fk = int(meta.get("fallback_k", 0))
    r = learned.evaluate(lib, chain, state, meta["topk"], [meta["threshold"]], steps, fallback_k=fk)
    lr, full, plain = r["learned"][0], r["full"], r["plain"]
    out = {"threshold": meta["threshold"], "topk": meta["topk"], "fallback_k": fk, "features": meta["features"], "train_blocks": meta["blocks"],
           "mpix_s": round(lr["mpix_s"], 1), "exhaustive_mpix_s": round(full["mpix_s"], 1),
           "speedup_vs_exhaustive": round(lr["mpix_s"] / full["mpix_s"], 3), "psnr_db": round(lr["psnr"], 4),
           "exhaustive_psnr_db": round(full["psnr"], 4), "gap_db": round(lr["gap_db"], 4), "fallback": round(lr["fallback"], 4),
           "plain_psnr_db": round(plain["psnr"], 4), "ms": round(lr["ms"], 3), "exhaustive_ms": round(full["ms"], 3),
           "plain_encode_mpix_s": round(plain["mpix_s"], 1), "speed_vs_plain": round(lr["mpix_s"] / plain["mpix_s"], 4)}
    if fk:
        r0 = learned.evaluate(lib, chain, state, meta["topk"], [meta["threshold"]], steps, fallback_k=0)["learned"][0]
        out["exhaustive_fallback"] = {"fallback_k": 0, "mpix_s": round(r0["mpix_s"], 1), "ms": round(r0["ms"], 3),
                                      "speedup_vs_exhaustive": round(r0["mpix_s"] / full["mpix_s"], 3), "psnr_db": round(r0["psnr"], 4),
                                      "gap_db": round(r0["gap_db"], 4), "fallback": round(r0["fallback"], 4)}
    # the network kernel against the bf16 MFMA peak: 2 (64*128 + 128*128 + 128*288) FLOP per block on the rows it
    # keeps, over its mean duration from the library's HIP events on its launch stream
    import torch
    imgs = chain.images()
    lws = lib.bc7_learned_workspace(imgs, meta["topk"], fk)
    name = "zt_mlp_propose4" if meta["topk"] <= 4 else "zt_mlp_propose16"
    lib.profile(True)
    lib.profile_reset()
    for _ in range(3):
        lib.bc7_learned_dev(imgs, state, meta["topk"], float(meta["threshold"]), lws, fk)
    torch.cuda.synchronize()
    t, n = lib.profile_get(name)
    lib.profile(False)
    if n:
        flop = 2 * (64 * 128 + 128 * 128 + 128 * 288) * chain.blocks
        tf = flop / (t / n / 1e3) / 1e12
        out["network"] = {"kernel": name, "bound": "mfma", "achieved": round(tf, 1), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(tf / PEAK_BF16_TFLOPS, 4), "flop_per_launch": flop, "ms_per_launch": round(t / n, 5)}
    log(f"learned: {out['mpix_s']} Mpix/s ({out['speedup_vs_exhaustive']}x exhaustive {out['exhaustive_mpix_s']}), "
        f"gap {out['gap_db']} dB, fallback {out['fallback']}")
    return out


def batch_leg(lib, steps: int = 20) -> dict:
    """The 512-texture batch with full chains, one graph launch per step, autotuned launch shapes."""
    import torch

    from zap_amd import autotune
    from zap_amd import batch as B
    autotune.reset(lib)
    shapes = autotune.apply(lib, "c4")
    t0 = time.perf_counter()
    specs, bases = B.c4_textures(512, 4)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    bt = B.Batch(lib, [(bases[i], w, h, fmt, srgb, None) for i, (w, h, fmt, srgb, kind) in enumerate(specs)])
    bt.run()
    torch.cuda.synchronize()
    # every texture's levels and blocks against the reference's hashes, before timing
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "c4_batch.json")))["textures"]
    got = bt.digests()
    bad = [i for i, (g, d) in enumerate(zip(gold, got)) if (g["mips"], g["blocks"]) != d]
    exact = not bad and len(gold) == len(got)
    log(f"batch: {len(got) - len(bad)}/{len(gold)} textures byte-identical to the reference")
    for _ in range(3):   # warm-up launches, then the timed ones back to back
        bt.run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        bt.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    fmts = {f: sum(1 for s in specs if s[2] == f) for f in ("bc1", "bc3", "bc5", "bc7")}
    out = {"textures": len(specs), "formats": fmts, "pixels": bt.pixels, "blocks": bt.blocks, "ms": round(ms, 3),
           "mpix_s": round(bt.pixels / (ms / 1e3) / 1e6, 1), "byte_exact_vs_reference": exact, "textures_differing": bad[:16],
           "graph_launches_per_step": 1, "launch_shapes": shapes,
           "generate_s": round(gen_s, 2)}
    bt.close()
    autotune.reset(lib)
    log(f"batch: {len(specs)} textures, {bt.pixels / 1e6:.0f} Mpx incl. chains, {ms:.2f} ms -> {out['mpix_s']} Mpix/s")
    return out


CONFIG_WORK = [  # (name, format, golden key, kernels)
    ("c1_bc1_1024", "bc1", "c1_1024_bc1", ["zt_bc14_encode", "zt_bc1_climb", "zt_bc1_finish"]),
    ("c3_bc6h_2048", "bc6h", "c3_bc6h_2048", ["zt_bc6h_encode"]),
    ("c3_astc_2048", "astc", "c3_astc_2048", ["zt_astc_classify", "zt_astc_single", "zt_astc_pick1", "zt_astc_search",
                                              "zt_astc_pair", "zt_astc_pick2"]),
]


def config_images() -> dict:
    from zap_amd import synth
    return {"c1_bc1_1024": synth.c1_bc1(1024, 1), "c3_bc6h_2048": synth.c3_hdr(2048, 5), "c3_astc_2048": synth.c3_astc(2048, 3)}


def configs_leg(lib, golden: dict, images: dict, steps: int) -> dict:
    """The other single-texture workloads, on the device entry point with inputs resident in HBM: BC1 of the
    1024^2 gradient+noise+rectangles image, BC6H of the 2048^2 HDR field, ASTC 4x4 of the 2048^2 noise texture.
    Each is checked against its golden hash; per-kernel durations come from the library's HIP events."""
    import numpy as np
    import torch

    from zap_amd import binding
    out = {}
    for name, fmt, gkey, kernels in CONFIG_WORK:
        img = images[name]
        h, w = img.shape[:2]
        src = torch.from_numpy(np.ascontiguousarray(img)).cuda()
        dst = torch.empty(binding.bc_size(w, h, fmt), dtype=torch.uint8, device="cuda")
        imgs = lib.images([(src, dst, w, h, w * img.itemsize * 4, fmt)])
        ws = torch.empty(max(1, lib.workspace_size(imgs)), dtype=torch.uint8, device="cuda")
        lib.encode_dev(imgs, ws)
        torch.cuda.synchronize()
        exact = hashlib.sha256(dst.cpu().numpy().tobytes()).hexdigest() == golden["outputs"][gkey]
        for _ in range(3):
            lib.encode_dev(imgs, ws)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            lib.encode_dev(imgs, ws)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        lib.profile(True)
        lib.profile_reset()
        for _ in range(5):
            lib.encode_dev(imgs, ws)
        torch.cuda.synchronize()
        kern = {}
        for k in kernels:
            t, n = lib.profile_get(k)
            if n:
                kern[k] = round(t / 5, 5)   # ms per encode
        lib.profile(False)
        out[name] = {"pixels": w * h, "blocks": ((w + 3) // 4) * ((h + 3) // 4), "ms": round(ms, 4),
                     "mpix_s": round(w * h / (ms / 1e3) / 1e6, 1), "byte_exact_vs_reference": exact, "kernel_ms_per_encode": kern}
        log(f"{name}: {out[name]['mpix_s']} Mpix/s ({ms:.3f} ms), exact {exact}, kernels {kern}")
        del src, dst, imgs, ws
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-texture batch leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the BC1 1024^2 / BC6H 2048^2 / ASTC 2048^2 legs")
    ap.add_argument("--no-host", action="store_true", help="skip the host-entry (PCIe-inclusive) leg")
    ap.add_argument("--batch-only", action="store_true", help="only build and time the 512-texture batch (for profiling)")
    ap.add_argument("--configs-only", action="store_true",
                    help="only the BC1 1024^2 / BC6H 2048^2 / ASTC 2048^2 legs, --steps encodes each (for profiling)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--counters-child", metavar="C2_NPY", help=argparse.SUPPRESS)
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 counter passes (traffic, VALU instructions)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU issue probe (zap_amd/_lib/libztprobe.so) and print its rates")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates the profiler's HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps(copy_probe(3)), flush=True)
        return 0
    if args.counters_child:
        return counters_child(args.counters_child)
    if args.valu_probe:
        print(json.dumps({"valu": valu_probe(), "copy": copy_probe()}), flush=True)
        return 0
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs the hot path on a single MI355X (no multi-GPU path); use --gpus 1")
        return 2

    import numpy as np
    import torch

    import zap_amd
    from zap_amd import binding, synth

    lib = zap_amd.load()
    torch.cuda.set_device(0)
    from zap_amd import autotune
    if args.batch_only:
        print(json.dumps(batch_leg(lib, args.steps)), flush=True)
        return 0
    if args.configs_only:
        golden = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
        res = configs_leg(lib, golden, config_images(), args.steps)
        print(json.dumps(res), flush=True)
        return 0 if all(r["byte_exact_vs_reference"] for r in res.values()) else 1
    shapes_c2 = autotune.apply(lib, "c2")
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))
    ops = json.load(open(os.path.join(ROOT, "tests", "golden", "ops.json")))["c2_bc7_chain"]

    log("generating the 4096^2 texture ...")
    c2 = synth.c2_bc7(4096, 2)
    W = H = 4096
    sizes = binding.mip_sizes(W, H)
    npx = sum(w * h for w, h in sizes)
    base = torch.from_numpy(c2).cuda()
    levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
    src = [base] + levels
    outs = [torch.empty(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    imgs = lib.images((src[i], outs[i], w, h, w * 4, "bc7") for i, (w, h) in enumerate(sizes))
    ws = torch.empty(lib.workspace_size(imgs), dtype=torch.uint8, device="cuda")

    def step():
        lib.mip_chain_dev(base, W, H, W * 4, levels, srgb=0)
        lib.encode_dev(imgs, ws)

    step()
    torch.cuda.synchronize()
    got = [hashlib.sha256(o.cpu().numpy().tobytes()).hexdigest() for o in outs]
    exact = got == golden["outputs"]["c2_bc7_chain"]
    log(f"output byte-identical to the reference on all 13 levels: {exact}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    value = npx / (ms / 1e3) / 1e6
    log(f"{args.steps} steps: {ms:.3f} ms/step (wall {wall / args.steps * 1e3:.3f}) -> {value:.1f} Mpix/s")

    # per-kernel durations: HIP events recorded by the library on the launch stream
    lib.profile(True)
    lib.profile_reset()
    for _ in range(args.profile_steps):
        step()
    torch.cuda.synchronize()
    kern = {}
    for k in ["zt_mip_tiles", "zt_mip_tail", "zt_mip_level", "zt_zero_u32", "zt_bc7_stage6", "zt_bc7_stage5",
              "zt_bc7_s1_rank", "zt_bc7_s1_fit", "zt_bc7_s1_pick", "zt_bc7_s1_fitpick", "zt_bc7_stage1"]:
        t, n = lib.profile_get(k)
        if n:
            kern[k] = {"launches_per_step": n // args.profile_steps, "ms_per_launch": round(t / n, 5),
                       "ms_per_step": round(t / args.profile_steps, 5)}
    lib.profile(False)

    def group_ms(names):
        return sum(kern[k]["ms_per_step"] for k in names if k in kern)

    stages = {"stage6": group_ms(["zt_bc7_stage6"]), "stage5": group_ms(["zt_bc7_stage5"]),
              "stage1": group_ms(["zt_bc7_s1_rank", "zt_bc7_s1_fit", "zt_bc7_s1_pick", "zt_bc7_s1_fitpick", "zt_bc7_stage1"]),
              "mips": group_ms(["zt_mip_tiles", "zt_mip_tail", "zt_mip_level"])}
    s6 = kern["zt_bc7_stage6"]["ms_per_launch"]
    # the mip chain alone, back to back without per-kernel events (which add a few microseconds to every launch):
    # the mips' HBM figure is its whole-chain rate, tail launch and launch gaps included
    for _ in range(20):
        lib.mip_chain_dev(base, W, H, W * 4, levels, srgb=0)
    torch.cuda.synchronize()
    m0, m1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    m0.record()
    for _ in range(200):
        lib.mip_chain_dev(base, W, H, W * 4, levels, srgb=0)
    m1.record()
    torch.cuda.synchronize()
    chain_ms = m0.elapsed_time(m1) / 200
    mip_gbs = MIP_CHAIN_BYTES / (chain_ms / 1e3) / 1e9
    stage_tops = {k: round(ops[f"ops_{k}"] / (v / 1e3) / 1e12, 3) for k, v in stages.items() if k != "mips" and v > 0}
    log(json.dumps(kern))
    enc_ms = stages["stage6"] + stages["stage5"] + stages["stage1"] + group_ms(["zt_zero_u32"])
    enc_tops = ops["ops_total"] / (enc_ms / 1e3) / 1e12

    # ceilings measured on this box now: the fastest VALU issue rate of the probe, the plain-copy bandwidth
    probe = valu_probe()
    copy = copy_probe()
    ceiling = probe["ceiling_g_wave_insts_per_s"]
    # stage 6's VALU instructions and HBM bytes per launch: rocprofv3 counter passes over this build and workload
    if args.no_counters:
        counters = {"error": "skipped (--no-counters)"}
    else:
        import tempfile
        fd, c2_path = tempfile.mkstemp(suffix=".npy")
        os.close(fd)
        np.save(c2_path, c2)
        try:
            counters = rocprof_counters(c2_path)
        finally:
            os.unlink(c2_path)
    # the mode-6 stage runs as zt_bc7_stage6q (refits queued per wave, the default) or zt_bc7_stage6 (one pass)
    s6name = "zt_bc7_stage6q" if "zt_bc7_stage6q" in counters else "zt_bc7_stage6"
    s6c, cpc = counters.get(s6name), counters.get("zt_copy_probe")
    roof = {"bound": "valu", "kernel": s6name, "unit": "G VALU wave-instructions/s", "peak": ceiling,
            "peak_is": "the fastest issue rate of the VALU probe (zt_valu_probe, inline-asm chains) measured in this run",
            "ms_per_launch": s6, "achieved": None, "frac": None, "traffic": None}
    if s6c and cpc and "SQ_INSTS_VALU" in s6c:
        insts = s6c["SQ_INSTS_VALU"]
        achieved = insts / (s6 / 1e3) / 1e9
        scale_f = COPY_PROBE_BYTES / (cpc["FETCH_SIZE"] * 1024)
        scale_w = COPY_PROBE_BYTES / (cpc["WRITE_SIZE"] * 1024)
        cycles = s6c["GRBM_GUI_ACTIVE"] / N_XCD
        roof.update({"achieved": round(achieved, 2), "frac": round(achieved / ceiling, 4),
                     "traffic": int(s6c["FETCH_SIZE"] * 1024 * scale_f + s6c["WRITE_SIZE"] * 1024 * scale_w),
                     "sq_insts_valu": int(insts), "sq_waves": int(s6c["SQ_WAVES"]),
                     "fetch_kib": round(s6c["FETCH_SIZE"], 1), "write_kib": round(s6c["WRITE_SIZE"], 1),
                     "byte_scale": {"fetch": round(scale_f, 4), "write": round(scale_w, 4), "copy_bytes": COPY_PROBE_BYTES},
                     "clock_ghz_profiled": round(cycles / (s6c["_ns"] / 1e9) / 1e9, 3),
                     "issue_per_cu_clock": round(insts / N_CU / cycles, 4),
                     "counters_source": "rocprofv3 --pmc passes in this run: " + "; ".join(" ".join(p) for p in COUNTER_PASSES)})
        # the other kernels of the step from the same passes: VALU instructions per launch and calibrated HBM bytes
        roof["chain_kernels"] = {
            k: {"us": round(v["_ns"] / 1e3, 2), "sq_insts_valu": int(v.get("SQ_INSTS_VALU", 0)), "sq_waves": int(v.get("SQ_WAVES", 0)),
                "fetch_bytes": int(v.get("FETCH_SIZE", 0) * 1024 * scale_f), "write_bytes": int(v.get("WRITE_SIZE", 0) * 1024 * scale_w)}
            for k, v in sorted(counters.items()) if k.startswith("zt_") and "_ns" in v}
    else:
        roof["counters_source"] = f"unavailable: {counters.get('error', 'stage-6 or copy counters missing')}"
    roof.update({"valu_probe": probe,
                 "algorithmic": {"ops_per_launch": ops["ops_stage6"], "tops": round(ops["ops_stage6"] / (s6 / 1e3) / 1e12, 3),
                                 "is": "the oracle's scalar ops (tests/golden/ops.json) per second: a work rate, not a utilisation "
                                       "(the kernel does four multiply-adds in one v_dot4_u32_u8)",
                                 "stage_tops": stage_tops, "encoder_tops": round(enc_tops, 3), "encoder_ms": round(enc_ms, 5)},
                 "mips": {"bound": "hbm", "achieved": round(mip_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(mip_gbs / PEAK_HBM_GBS, 4), "practical_peak": copy["gb_s"],
                          "frac_of_practical": round(mip_gbs / copy["gb_s"], 4), "bytes_per_chain": MIP_CHAIN_BYTES,
                          "ms": round(chain_ms, 5), "ms_is": "whole 12-level chain (tile pass + tail), 200 chains back to back",
                          "kernel_event_ms": round(stages["mips"], 5)}})
    log(f"roofline: stage 6 {roof['achieved']} of {ceiling} G wave-insts/s (frac {roof['frac']}), traffic {roof['traffic']}; "
        f"mips {mip_gbs:.0f} GB/s of {copy['gb_s']} copy")

    # end to end through the drop-in's host entry (zap_bc_encode: host buffers in, PCIe both ways), per level
    levels_np = [c2] + [l.cpu().numpy() for l in levels]
    e2e_mpix = None
    if not args.no_host:
        lib.bc_encode(levels_np[0], "bc7")
        t0 = time.perf_counter()
        for lvl in levels_np:
            lib.bc_encode(lvl, "bc7")
        e2e_mpix = round(npx / (time.perf_counter() - t0) / 1e6, 1)
        log(f"host entry (H2D + encode + D2H per level): {e2e_mpix} Mpix/s")

    autotune.reset(lib)
    learned_res = None if args.no_learned else learned_leg(lib, c2)
    batch_res = None if args.no_batch else batch_leg(lib)
    cfg_imgs = None if args.no_configs else config_images()
    configs_res = None if args.no_configs else configs_leg(lib, golden, cfg_imgs, 20)
    cpu = None if args.no_cpu else cpu_baseline(levels_np, args.cpu_threads, configs=cfg_imgs)
    if cpu and configs_res:
        for name, v in cpu.get("configs_mpix_s", {}).items():
            configs_res[name]["cpu_reference_mpix_s"] = v
    line = {
        "metric": "BC7 Mpix/s on 1 MI355X: plain vs learned-proposal at \u22640.05 dB PSNR gap",
        "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/int32/f32 (bit-exact)", "data": "synthetic (seeded value noise + rectangles + independent alpha, 25% opaque blocks)",
        "config": {"workload": "BC7 plain search of a 4096x4096 RGBA8 texture and its 13-level box mip chain "
                               "(mip generation + encode per step; 1,398,103 blocks, 22,369,621 px)",
                   "value_is": "the plain search (the reference's mode 6/5/1 search, byte-exact); learned-proposal figures in 'learned'",
                   "global_batch": 1, "seq_len": None, "parallelism": "single GPU", "bc7_m1k": lib.get_option(1),
                   "byte_exact_vs_reference": exact},
        "roofline": roof,
        "stages_ms": {k: round(v, 5) for k, v in stages.items()},
        "host_entry_mpix_s": e2e_mpix,
        "launch_shapes": shapes_c2,
        "kernels": kern,
        "learned": learned_res,
        "batch": batch_res,
        "configs": configs_res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0 if exact else 1


if __name__ == "__main__":
    sys.exit(main())
