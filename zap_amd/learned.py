"""Learned BC7 proposal: training data, training loop and model file, all on the GPU through libzaptex.

Training blocks come from seeded textures of the benchmark's family (value noise + hard-edged rectangles +
independent alpha, 25% opaque blocks) and their box mip chains, at seeds disjoint from the benchmark
texture's.  Labels are the candidates the exhaustive all-mode search picks (zaptex_bc7_full_dev).  The network
trains with the library's own MFMA forward/backward kernels and Adam (zaptex_mlp_train_step).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import binding, synth

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "bc7_proposal.npy")
META_PATH = os.path.join(HERE, "models", "bc7_proposal.json")


def training_set(lib, textures: int, size: int, seed0: int = 1000, log=print):
    """(blocks uint8 [N, 64] cuda, labels int32 [N] cuda) from `textures` seeded size^2 textures + mip chains."""
    import torch
    blocks, labels = [], []
    t0 = time.time()
    for t in range(textures):
        img = synth.c2_bc7(size, seed0 + t, rects=max(16, size // 16))
        base = torch.from_numpy(img).cuda()
        sizes = binding.mip_sizes(size, size)
        levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
        lib.mip_chain_dev(base, size, size, size * 4, levels, srgb=t & 1)
        srcs = [base] + levels
        for i0 in range(0, len(sizes), 16):
            chunk = list(range(i0, min(len(sizes), i0 + 16)))
            nb = sum(((sizes[i][0] + 3) // 4) * ((sizes[i][1] + 3) // 4) for i in chunk)
            outs = [torch.empty(binding.bc_size(sizes[i][0], sizes[i][1], "bc7"), dtype=torch.uint8, device="cuda") for i in chunk]
            imgs = lib.images((srcs[i], outs[k], sizes[i][0], sizes[i][1], sizes[i][0] * 4, "bc7") for k, i in enumerate(chunk))
            blk = torch.empty((nb, 64), dtype=torch.uint8, device="cuda")
            lab = torch.empty(nb, dtype=torch.int32, device="cuda")
            lib.bc7_gather_dev(imgs, blk)
            lib.bc7_full_dev(imgs, lab)
            blocks.append(blk)
            labels.append(lab)
        torch.cuda.synchronize()
        if log and (t + 1) % max(1, textures // 8) == 0:
            log(f"  labelled {t + 1}/{textures} textures ({sum(b.shape[0] for b in blocks)} blocks, {time.time() - t0:.1f}s)")
    return torch.cat(blocks), torch.cat(labels)


def train(lib, blocks, labels, epochs: int = 4, batch: int = 8192, lr: float = 2e-3, seed: int = 1, log=print):
    """Adam on shuffled minibatches; returns (state tensor, history)."""
    import torch
    n = blocks.shape[0]
    st = lib.mlp_state(seed=seed)
    ws = lib.mlp_workspace(batch)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    step, hist = 0, []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        stats = torch.zeros(2, dtype=torch.float32, device="cuda")
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            xb, yb = blocks[idx].contiguous(), labels[idx].contiguous()
            step += 1
            rate = lr * (0.5 * (1 + np.cos(np.pi * step / max(1, epochs * (n // batch)))))
            lib.mlp_train_step(st, xb, yb, batch, float(rate), step, ws, stats)
            nb += 1
        torch.cuda.synchronize()
        loss, acc = float(stats[0].item()) / max(1, nb), float(stats[1].item()) / max(1, nb * batch)
        hist.append({"epoch": ep + 1, "loss": round(loss, 4), "top1": round(acc, 4)})
        if log:
            log(f"  epoch {ep + 1}: loss {loss:.4f} top-1 {acc:.4f}")
    return st, hist


def save(lib, state, meta: dict, path: str = MODEL_PATH) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, lib.mlp_params(state).astype(np.float32), allow_pickle=False)
    with open(META_PATH, "w") as fh:
        json.dump(meta, fh, indent=1)


def load(lib, path: str = MODEL_PATH):
    if not os.path.exists(path):
        raise binding.ZapTexError(f"no trained proposal model at {path} (train with python -m zap_amd.learned)")
    params = np.load(path, allow_pickle=False)
    return lib.mlp_state(params=params)


if __name__ == "__main__":
    import argparse

    import torch
    ap = argparse.ArgumentParser(description="train the BC7 proposal network on the GPU")
    ap.add_argument("--textures", type=int, default=24)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    a = ap.parse_args()
    lib = binding.load()
    t0 = time.time()
    xb, yb = training_set(lib, a.textures, a.size)
    t_label = time.time() - t0
    print(f"training set: {xb.shape[0]} blocks labelled in {t_label:.1f}s", flush=True)
    t1 = time.time()
    st, hist = train(lib, xb, yb, a.epochs, a.batch, a.lr)
    t_train = time.time() - t1
    save(lib, st, {"blocks": int(xb.shape[0]), "textures": a.textures, "size": a.size, "epochs": a.epochs, "batch": a.batch,
                   "lr": a.lr, "label_seconds": round(t_label, 1), "train_seconds": round(t_train, 1), "history": hist,
                   "seeds": f"1000..{1000 + a.textures - 1}"})
    print(json.dumps(hist[-1]), flush=True)
