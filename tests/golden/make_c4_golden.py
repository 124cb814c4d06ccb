"""Writes tests/golden/c4_batch.json: per texture of the 512-texture batch workload, SHA-256 of the reference's
bytes — the base image (generated by zap_amd.batch.texture_cpu), its generated mip levels 1.. concatenated
(zap_mip_next with the texture's filter), and the blocks of every level concatenated (zap_bc_encode, rdo = 0).

Runs the compiled reference (oracle/_ref, built from zap_tex.h with -std=c11 -O2 -ffp-contract=off) in this
container only; the GPU test and bench.py compare the batch's output with these hashes.
Usage: python tests/golden/make_c4_golden.py [workers]   (about 5 minutes on 8 cores)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_batch.json")
N, SEED = 512, 4


def one(i: int) -> dict:
    import numpy as np

    from oracle import loader
    from zap_amd import batch
    w, h, fmt, srgb, kind = batch.c4_specs(N, SEED)[i]
    ref = loader.reference()
    base = batch.texture_cpu(w, h, kind, SEED * 100000 + i)
    chain = ref.mip_chain(base, srgb)
    mh, bh = hashlib.sha256(), hashlib.sha256()
    for lv in chain[1:]:
        mh.update(np.ascontiguousarray(lv).tobytes())
    for lv in chain:
        bh.update(ref.bc_encode(lv, fmt).tobytes())
    return {"w": w, "h": h, "fmt": fmt, "srgb": srgb, "levels": len(chain),
            "base": hashlib.sha256(base.tobytes()).hexdigest(), "mips": mh.hexdigest(), "blocks": bh.hexdigest()}


def main() -> None:
    from zap_amd import batch
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else min(8, os.cpu_count() or 1)
    specs = batch.c4_specs(N, SEED)
    # largest first, so the pool ends together
    order = sorted(range(N), key=lambda i: -specs[i][0] * specs[i][1])
    t0 = time.time()
    res = {}
    with ProcessPoolExecutor(workers) as ex:
        for k, (i, r) in enumerate(zip(order, ex.map(one, order))):
            res[i] = r
            if k % 64 == 63:
                print(f"{k + 1}/{N} textures, {time.time() - t0:.0f} s", flush=True)
    ref_sha = hashlib.sha256(open("/root/reference/zap_tex.h", "rb").read()).hexdigest()
    doc = {"meta": {"reference": "zap_tex.h 1.3.0 sha256 " + ref_sha, "build": "gcc 11.4 -std=c11 -O2 -ffp-contract=off",
                    "generator": "zap_amd.batch.texture_cpu, seeds 400000 + i", "specs": f"zap_amd.batch.c4_specs({N}, {SEED})"},
           "textures": [res[i] for i in range(N)]}
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=0)
    print(f"wrote {OUT} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
