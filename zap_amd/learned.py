"""Learned BC7 proposal: training data, training loop, threshold calibration, evaluation and the model file, all
on the GPU through libzaptex.

Training blocks come from seeded textures of the benchmark's family (value noise + hard-edged rectangles +
independent alpha, 25% opaque blocks) and their box mip chains, at seeds disjoint from the benchmark
texture's.  The family's noise scales with the texture (finest lattice cell = side / 512), so the training set
mixes sizes (4096^2, 2048^2, 1024^2 and all their mip levels) to cover smooth and busy blocks alike.  Labels are the candidates the exhaustive all-mode search picks (zaptex_bc7_full_dev).  The network
trains with the library's own MFMA forward/backward kernels and Adam (zaptex_mlp_train_step).

Quality is the reference's PSNR (tex_bench.c:17-22: 10 log10(255^2 n / SSE) over R, G, B, A for BC7), summed
over every level of a texture's chain; the learned encode is judged against the exhaustive search on the same
chain.  The risk threshold is calibrated on held-out textures of three sizes (their own seeds), never on the
benchmark's.
"""
from __future__ import annotations

import json
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import binding, synth

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "bc7_proposal.npy")
TRAIN_SEED0 = 1000                       # training textures use seeds TRAIN_SEED0, TRAIN_SEED0 + 1, ...
TRAIN_MIX = ((4096, 16), (2048, 64), (1024, 128))  # (side, count): about 56M blocks with the mip chains
CALIB = ((4096, 500), (1024, 501), (512, 502))   # held-out (side, seed) textures for the threshold


def _texture(args):
    size, seed = args
    return synth.c2_bc7(size, seed, rects=max(16, size // 16))


class Chain:
    """A texture and its box mip chain on the GPU, with one BC7 output buffer per level."""

    def __init__(self, lib, img: np.ndarray, srgb: int = 0):
        import torch
        self.lib = lib
        h, w = img.shape[:2]
        self.sizes = binding.mip_sizes(w, h)
        base = torch.from_numpy(np.ascontiguousarray(img)).cuda()
        levels = [torch.empty((hh, ww, 4), dtype=torch.uint8, device="cuda") for (ww, hh) in self.sizes[1:]]
        lib.mip_chain_dev(base, w, h, w * 4, levels, srgb=srgb)
        self.srcs = [base] + levels
        self.outs = [torch.empty(binding.bc_size(ww, hh, "bc7"), dtype=torch.uint8, device="cuda") for (ww, hh) in self.sizes]
        self.pixels = sum(ww * hh for ww, hh in self.sizes)
        self.blocks = sum(((ww + 3) // 4) * ((hh + 3) // 4) for ww, hh in self.sizes)

    def images(self, lo: int = 0, hi: int | None = None):
        hi = len(self.sizes) if hi is None else hi
        return self.lib.images((self.srcs[i], self.outs[i], self.sizes[i][0], self.sizes[i][1], self.sizes[i][0] * 4, "bc7")
                               for i in range(lo, hi))

    def psnr(self) -> float:
        """The reference's BC7 PSNR (RGBA) of the current outputs over the whole chain, decoded on the GPU."""
        import torch
        se = 0
        for (w, h), src, out in zip(self.sizes, self.srcs, self.outs):
            dec = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
            self.lib.bc_decode_dev(out, w, h, "bc7", dec)
            d = dec.to(torch.int32) - src.to(torch.int32)
            se += int((d * d).sum(dtype=torch.int64).item())
        return 99.0 if se == 0 else 10 * math.log10(255.0 * 255.0 * 4 * self.pixels / se)


def training_set(lib, mix=TRAIN_MIX, seed0: int = TRAIN_SEED0, threads: int = 16, log=print):
    """(blocks uint8 [N, 64], labels int32 [N], deltas int16 [N, 512] holding uint16 SSE deltas), all on the GPU,
    from the seeded textures of `mix` ((side, count) pairs, seeds counting up from seed0) and their mip chains;
    textures are generated on host threads while the GPU labels the previous ones."""
    import torch
    blocks, labels, deltas = [], [], []
    specs = [side for side, count in mix for _ in range(count)]
    textures = len(specs)
    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        for t, img in enumerate(ex.map(_texture, [(side, seed0 + t) for t, side in enumerate(specs)])):
            ch = Chain(lib, img, srgb=t & 1)
            for i0 in range(0, len(ch.sizes), 16):
                i1 = min(len(ch.sizes), i0 + 16)
                nb = sum(((ch.sizes[i][0] + 3) // 4) * ((ch.sizes[i][1] + 3) // 4) for i in range(i0, i1))
                imgs = ch.images(i0, i1)
                blk = torch.empty((nb, 64), dtype=torch.uint8, device="cuda")
                lab = torch.empty(nb, dtype=torch.int32, device="cuda")
                dl = torch.empty((nb, 512), dtype=torch.int16, device="cuda")
                lib.bc7_gather_dev(imgs, blk)
                lib.bc7_full_dev(imgs, lab, dl)
                blocks.append(blk)
                labels.append(lab)
                deltas.append(dl)
            torch.cuda.synchronize()
            if log and (t + 1) % max(1, textures // 8) == 0:
                log(f"  labelled {t + 1}/{textures} textures ({sum(b.shape[0] for b in blocks)} blocks, {time.time() - t0:.1f}s)",
                    flush=True)
    return torch.cat(blocks), torch.cat(labels), torch.cat(deltas)


def train(lib, blocks, labels, deltas=None, temperature: float = 0.0, epochs: int = 4, batch: int = 8192, lr: float = 2e-3,
          seed: int = 1, risk_k: int = 4, risk_weight: float = 0.2, features: int = 4, log=print):
    """Adam on shuffled minibatches with a cosine learning rate; soft targets softmax(-delta / temperature) when
    temperature > 0, else the one-hot labels; the risk slot learns the SSE the top-risk_k give away.  Returns
    (state tensor, history)."""
    import torch
    n = blocks.shape[0]
    st = lib.mlp_state(seed=seed, features=features)
    ws = lib.mlp_workspace(batch)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    step, hist = 0, []
    total_steps = max(1, epochs * (n // batch))
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        stats = torch.zeros(2, dtype=torch.float32, device="cuda")   # this step's (mean loss, top-1 hits)
        total = torch.zeros(2, dtype=torch.float64, device="cuda")
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            xb, yb = blocks[idx].contiguous(), labels[idx].contiguous()
            db = deltas[idx].contiguous() if deltas is not None and (temperature > 0 or risk_k > 0) else None
            step += 1
            rate = lr * 0.5 * (1 + math.cos(math.pi * step / total_steps))
            lib.mlp_train_step(st, xb, yb, batch, float(rate), step, ws, stats, deltas=db, temperature=temperature,
                               risk_k=risk_k if db is not None else 0, risk_weight=risk_weight)
            total += stats
            nb += 1
        torch.cuda.synchronize()
        loss, acc = float(total[0].item()) / max(1, nb), float(total[1].item()) / max(1, nb * batch)
        hist.append({"epoch": ep + 1, "loss": round(loss, 4), "top1": round(acc, 4)})
        if log:
            log(f"  epoch {ep + 1}: loss {loss:.4f} top-1 {acc:.4f}", flush=True)
    return st, hist


def _time(fn, steps: int) -> float:
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def evaluate(lib, chain: Chain, state, topk: int, thresholds, steps: int = 5, fallback_k: int = 0) -> dict:
    """Encode time (ms, inputs resident) and chain PSNR for the plain search, the exhaustive search and the learned
    proposal at each threshold; 'fallback' is the fraction of blocks the learned encode sent to its fallback (the
    exhaustive search with fallback_k = 0, else the network's top fallback_k)."""
    import torch
    imgs = chain.images()
    ws = torch.empty(lib.workspace_size(imgs), dtype=torch.uint8, device="cuda")
    res = {"blocks": chain.blocks, "pixels": chain.pixels}
    res["plain"] = {"ms": _time(lambda: lib.encode_dev(imgs, ws), steps)}
    res["plain"]["psnr"] = chain.psnr()
    fws = lib.bc7_full_workspace(imgs)
    res["full"] = {"ms": _time(lambda: lib.bc7_full_dev(imgs, ws=fws), steps)}
    res["full"]["psnr"] = chain.psnr()
    lws = lib.bc7_learned_workspace(imgs, topk, fallback_k)
    res["learned"] = []
    for thr in thresholds:
        ms = _time(lambda: lib.bc7_learned_dev(imgs, state, topk, float(thr), lws, fallback_k), steps)
        fb = int(lws[:256].view(torch.int32)[8].item())
        p = chain.psnr()
        res["learned"].append({"threshold": float(thr), "ms": ms, "psnr": p, "fallback": fb / chain.blocks,
                               "gap_db": res["full"]["psnr"] - p})
    for r in [res["plain"], res["full"]] + res["learned"]:
        r["mpix_s"] = chain.pixels / (r["ms"] / 1e3) / 1e6
    return res


def calibrate(lib, state, topk: int, textures=CALIB, max_gap_db: float = 0.04, fallback_k: int = 0, log=print) -> dict:
    """Fastest threshold (predicted SSE a block may give away before it falls back) whose chain-PSNR gap to the
    exhaustive search stays within max_gap_db on every held-out texture; speed is pooled over them.
    The margin under 0.05 dB leaves room for texture-to-texture variation."""
    grid = [1e30, 64, 32, 16, 12, 8, 6, 5, 4, 3.5, 3, 2.5, 2, 1.75, 1.5, 1.25, 1.125, 1, 0.9, 0.8, 0.75, 0.625, 0.5, 0.25, -1]
    per = []
    for side, seed in textures:
        per.append(evaluate(lib, Chain(lib, _texture((side, seed))), state, topk, grid, fallback_k=fallback_k))
    px = sum(r["pixels"] for r in per)
    full_ms = sum(r["full"]["ms"] for r in per)
    rows = []
    for i, thr in enumerate(grid):
        ms = sum(r["learned"][i]["ms"] for r in per)
        rows.append({"threshold": thr, "mpix_s": px / (ms / 1e3) / 1e6, "max_gap_db": max(r["learned"][i]["gap_db"] for r in per),
                     "gaps_db": [round(r["learned"][i]["gap_db"], 4) for r in per],
                     "fallback": sum(r["learned"][i]["fallback"] * r["blocks"] for r in per) / sum(r["blocks"] for r in per)})
    ok = [r for r in rows if r["max_gap_db"] <= max_gap_db]
    best = max(ok, key=lambda r: r["mpix_s"])
    full = px / (full_ms / 1e3) / 1e6
    if log:
        for r in rows:
            log(f"  thr {r['threshold']:8.4g}: {r['mpix_s']:9.1f} Mpix/s  gaps {r['gaps_db']} dB  fallback {r['fallback']:.3f}", flush=True)
        log(f"  exhaustive {full:.1f} Mpix/s -> threshold {best['threshold']}", flush=True)
    return {"threshold": best["threshold"], "topk": topk, "fallback_k": fallback_k, "textures": [list(t) for t in textures],
            "exhaustive_mpix_s": full, "rows": rows}


def save(lib, state, meta: dict, path: str = MODEL_PATH) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, lib.mlp_params(state).astype(np.float32), allow_pickle=False)
    with open(os.path.splitext(path)[0] + ".json", "w") as fh:
        json.dump(meta, fh, indent=1)


def load(lib, path: str = MODEL_PATH):
    """(state tensor, meta dict) of a trained model."""
    if not os.path.exists(path):
        raise binding.ZapTexError(f"no trained proposal model at {path} (train with python -m zap_amd.learned)")
    params = np.load(path, allow_pickle=False)
    with open(os.path.splitext(path)[0] + ".json") as fh:
        meta = json.load(fh)
    return lib.mlp_state(params=params, features=int(meta.get("features", 0))), meta


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="train the BC7 proposal network on the GPU")
    ap.add_argument("--mix", default=",".join(f"{s}x{c}" for s, c in TRAIN_MIX), help="training textures: side x count, ...")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--topk", type=int, default=4)
    ap.add_argument("--temperature", type=float, default=0.0, help="soft-target temperature in SSE units (0: one-hot labels)")
    ap.add_argument("--risk-weight", type=float, default=0.2)
    ap.add_argument("--features", type=int, default=4, help="input normalisation (zaptex_mlp_set_features)")
    ap.add_argument("--fallback-k", type=int, default=64,
                    help="candidates (the network's top-k) for blocks over the risk threshold; 0: the exhaustive search")
    ap.add_argument("--out", default=MODEL_PATH, help="model file (.npy; the .json beside it holds the metadata)")
    ap.add_argument("--recalibrate", action="store_true", help="keep the weights of --out, only recalibrate its threshold")
    a = ap.parse_args()
    lib = binding.load()
    if a.recalibrate:
        st, meta = load(lib, a.out)
        cal = calibrate(lib, st, int(meta["topk"]), fallback_k=a.fallback_k)
        meta.update({"threshold": cal["threshold"], "fallback_k": a.fallback_k,
                     "calibration": {k: cal[k] for k in ("textures", "exhaustive_mpix_s", "rows")}})
        save(lib, st, meta, a.out)
        print(json.dumps({"threshold": cal["threshold"], "fallback_k": a.fallback_k}), flush=True)
        raise SystemExit(0)
    t0 = time.time()
    mix = tuple(tuple(int(v) for v in m.split("x")) for m in a.mix.split(","))
    xb, yb, db = training_set(lib, mix)
    t_label = time.time() - t0
    print(f"training set: {xb.shape[0]} blocks labelled in {t_label:.1f}s", flush=True)
    t1 = time.time()
    st, hist = train(lib, xb, yb, db, a.temperature, a.epochs, a.batch, a.lr, risk_k=a.topk, risk_weight=a.risk_weight,
                     features=a.features)
    t_train = time.time() - t1
    print(f"trained in {t_train:.1f}s", flush=True)
    n_blocks = int(xb.shape[0])
    del xb, yb, db
    cal = calibrate(lib, st, a.topk, fallback_k=a.fallback_k)
    save(lib, st, {"blocks": n_blocks, "mix": [list(m) for m in mix], "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                   "temperature": a.temperature, "risk_weight": a.risk_weight, "features": a.features,
                   "label_seconds": round(t_label, 1), "train_seconds": round(t_train, 1), "history": hist,
                   "train_seeds": f"{TRAIN_SEED0}..{TRAIN_SEED0 + sum(c for _, c in mix) - 1}",
                   "threshold": cal["threshold"], "topk": a.topk, "fallback_k": a.fallback_k,
                   "calibration": {k: cal[k] for k in ("textures", "exhaustive_mpix_s", "rows")}}, a.out)
    print(json.dumps({"last_epoch": hist[-1], "threshold": cal["threshold"]}), flush=True)
