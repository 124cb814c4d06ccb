"""Writes tests/golden/golden.json: SHA-256 of the reference encoder's outputs on the seeded workloads.

Runs the compiled reference (oracle/_ref, built from zap_tex.h with -std=c11 -O2 -ffp-contract=off) in this
container only; the GPU tests compare against the committed hashes and never need the reference at run time.
Usage: python tests/golden/make_golden.py   (about a minute on 8 cores)
       python tests/golden/make_golden.py --only near_singular containers rdo   (rewrites those fixtures alone)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import loader  # noqa: E402
from zap_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
SMALL_SIZES = [(1, 1), (3, 2), (5, 3), (37, 21), (67, 45), (64, 64)]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def banded_bc(lib, img, fmt, threads=8):
    """Reference encode split into 8-block-row bands (byte-identical at rdo = 0, as the viewer does)."""
    h, w = img.shape[:2]
    bh = (h + 3) // 4
    bs = loader.BLOCK_BYTES[loader.FMT[fmt]]
    out = np.zeros(((w + 3) // 4) * bh * bs, dtype=np.uint8)
    band = 8
    def run(b0):
        y0 = b0 * 4
        sub = img[y0:min(h, y0 + band * 4)]
        o = lib.bc_encode(sub, fmt)
        out[b0 * ((w + 3) // 4) * bs: b0 * ((w + 3) // 4) * bs + o.size] = o
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, range(0, bh, band)))
    return out


def banded_rows(lib, img, kind, threads=8):
    h, w = img.shape[:2]
    bh = (h + 3) // 4
    out = np.zeros(((w + 3) // 4) * bh * 16, dtype=np.uint8)
    band = 8
    def run(b0):
        y0 = b0 * 4
        sub = img[y0:min(h, y0 + band * 4)]
        o = lib.bc6h_encode(sub) if kind == "bc6h" else lib.astc_encode(sub)
        off = b0 * ((w + 3) // 4) * 16
        out[off:off + o.size] = o
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, range(0, bh, band)))
    return out


def small_image(w, h, seed):
    img = synth.c2_bc7(max(w, h, 4), seed)[:h, :w]
    return np.ascontiguousarray(img)


def near_singular(ref, n_max: int = 192) -> dict:
    """Blocks whose BC1 least-squares step passes the reference's det guard with a nearly singular system (every
    index 2 or 3: det = -3.8e-6), found with the oracle's counting build, and the reference's BC1/BC3/ASTC bytes."""
    import ctypes as C
    cnt = loader.oracle(counting=True)
    cnt.lib.zto_cast_stats.argtypes = [C.c_void_p] * 3
    m, bad, near = np.zeros(3), np.zeros(3, np.uint64), np.zeros(2, np.uint64)
    picked = []
    for blk in synth.degenerate_blocks(20000, 24):
        img = np.ascontiguousarray(blk.reshape(4, 4, 4))
        cnt.lib.zto_cast_reset()
        cnt.bc_encode(img, "bc1")
        cnt.lib.zto_cast_stats(m.ctypes.data, bad.ctypes.data, near.ctypes.data)
        if near[0]:
            picked.append(blk)
            if len(picked) == n_max:
                break
    img = synth.blocks_to_image(np.array(picked))
    return {"blocks": len(picked), "input": img.tobytes().hex(), "bc1": ref.bc_encode(img, "bc1").tobytes().hex(),
            "bc3": ref.bc_encode(img, "bc3").tobytes().hex(), "astc": ref.astc_encode(img).tobytes().hex()}


# `zap tex` runs: (name, image key, CLI arguments before "w h in out")
CONTAINER_RUNS = [
    ("c1_bc1", "c1_1024", ["-f", "bc1"]), ("c1_bc1_mips_srgb", "c1_1024", ["-f", "bc1", "-m", "-S"]),
    ("c1_bc3_mips", "c1_1024", ["-f", "bc3", "-m"]), ("c1_bc4_mips_srgb", "c1_1024", ["-f", "bc4", "-m", "-S"]),
    ("c1_bc5_mips", "c1_1024", ["-f", "bc5", "-m"]), ("c1_bc7", "c1_1024", ["-f", "bc7"]),
    ("c2_bc7_mips", "c2_4096", ["-f", "bc7", "-m"]), ("c2_bc7_mips_srgb", "c2_4096", ["-f", "bc7", "-m", "-S"]),
    ("c3_astc", "c3_astc_2048", ["-f", "astc"]), ("c3_astc_mips", "c3_astc_2048", ["-f", "astc", "-m"]),
    ("odd_bc3_mips_srgb", "small_37x21", ["-f", "bc3", "-m", "-S"]), ("odd_bc7_mips", "small_5x3", ["-f", "bc7", "-m"]),
    ("odd_bc1_mips", "small_1x1", ["-f", "bc1", "-m"]),
]


def container_images() -> dict:
    return {"c1_1024": synth.c1_bc1(1024, 1), "c2_4096": synth.c2_bc7(4096, 2), "c3_astc_2048": synth.c3_astc(2048, 3),
            "small_37x21": small_image(37, 21, 40 + 37 * 7 + 21), "small_5x3": small_image(5, 3, 40 + 5 * 7 + 3),
            "small_1x1": small_image(1, 1, 40 + 1 * 7 + 1)}


def containers() -> dict:
    """SHA-256, size and header of the files the reference CLI (oracle/_ref/zapcli, built from zap.c) writes."""
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "oracle", "_ref", "zapcli")
    imgs = container_images()
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for key, img in imgs.items():
            img.tofile(os.path.join(tmp, key + ".rgba"))
        for name, key, args in CONTAINER_RUNS:
            h, w = imgs[key].shape[:2]
            dst = os.path.join(tmp, name + (".astc" if "astc" in args else ".dds"))
            subprocess.run([cli, "tex", *args, str(w), str(h), os.path.join(tmp, key + ".rgba"), dst], check=True, capture_output=True)
            data = open(dst, "rb").read()
            hn = 16 if "astc" in args else 148 if data[84:88] == b"DX10" else 128
            out[name] = {"image": key, "args": args, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest(),
                         "header": data[:hn].hex(), "payload_sha256": hashlib.sha256(data[hn:]).hexdigest()}
    return out


RDO_LAMBDAS = (10.0, 30.0, 100.0, 300.0)   # tex_bench.c:199-225 (0 is the plain encoding)


def rdo_hashes(threads: int = 8) -> dict:
    """SHA-256 of the reference's rdo > 0 encodes (single calls: RDO's raster order cannot be banded): the C1 1024^2
    image in every format and level 2 (1024^2) of the C2 chain in BC7, at each of RDO_LAMBDAS."""
    ref = loader.reference()
    c1 = synth.c1_bc1(1024, 1)
    c2l2 = ref.mip_chain(synth.c2_bc7(4096, 2), 0)[2]
    jobs = [(f"c1_1024_{f}_rdo{int(l)}", c1, f, l) for f in ["bc1", "bc3", "bc4", "bc5", "bc7"] for l in RDO_LAMBDAS]
    jobs += [(f"c2_level2_bc7_rdo{int(l)}", c2l2, "bc7", l) for l in RDO_LAMBDAS]
    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(lambda j: sha(ref.bc_encode(j[1], j[2], rdo=j[3])), jobs))
    return {"inputs": {"c1_1024": sha(c1), "c2_level2": sha(c2l2)}, "outputs": {j[0]: o for j, o in zip(jobs, outs)}}


def main():
    if sys.argv[1:2] == ["--only"]:
        with open(OUT) as fh:
            g = json.load(fh)
        for part in sys.argv[2:]:
            g[part] = {"near_singular": lambda: near_singular(loader.reference()), "containers": containers, "rdo": rdo_hashes}[part]()
            print(f"{part}: rewritten")
        with open(OUT, "w") as fh:
            json.dump(g, fh, indent=1, sort_keys=True)
        return
    ref = loader.reference()
    ref64 = loader.reference(m1k64=True)
    t0 = time.time()
    g = {"meta": {"reference": "zap_tex.h 1.3.0 sha256 " + hashlib.sha256(open("/root/reference/zap_tex.h", "rb").read()).hexdigest(),
                  "build": "gcc 11.4 -std=c11 -O2 -ffp-contract=off", "generator": "zap_amd.synth"},
         "inputs": {}, "outputs": {}, "small": {}}
    # C1: BC1 (and BC3/BC4/BC5 on the same image) 1024^2
    c1 = synth.c1_bc1(1024, 1)
    g["inputs"]["c1_1024"] = sha(c1)
    for f in ["bc1", "bc3", "bc4", "bc5"]:
        g["outputs"][f"c1_1024_{f}"] = sha(banded_bc(ref, c1, f))
    nm = synth.normal_map(1024, 1024, 11)
    g["inputs"]["normal_1024"] = sha(nm)
    g["outputs"]["normal_1024_bc5"] = sha(banded_bc(ref, nm, "bc5"))
    # C2: BC7 4096^2 + 13-level box chain (linear; sRGB chain hashed too)
    c2 = synth.c2_bc7(4096, 2)
    g["inputs"]["c2_4096"] = sha(c2)
    for srgb in (0, 1):
        chain = ref.mip_chain(c2, srgb)
        g["outputs"][f"c2_mips_srgb{srgb}"] = [sha(l) for l in chain]
        if srgb == 0:
            g["outputs"]["c2_bc7_chain"] = [sha(banded_bc(ref, l, "bc7")) for l in chain]
            g["outputs"]["c2_bc7_1024_m1k64"] = sha(banded_bc(ref64, chain[2], "bc7"))
            g["outputs"]["c2_bc1_chain"] = [sha(banded_bc(ref, l, "bc1")) for l in chain[2:]]
    print(f"c2 done {time.time() - t0:.1f}s", flush=True)
    # C3: BC6H 2048^2 and ASTC 2048^2
    hdr = synth.c3_hdr(2048, 5)
    g["inputs"]["c3_hdr_2048"] = sha(hdr)
    g["outputs"]["c3_bc6h_2048"] = sha(banded_rows(ref, hdr, "bc6h"))
    a = synth.c3_astc(2048, 3)
    g["inputs"]["c3_astc_2048"] = sha(a)
    g["outputs"]["c3_astc_2048"] = sha(banded_rows(ref, a, "astc"))
    fl = hdr[:512, :512]
    g["outputs"]["c3_hdr_mipf_512"] = sha(ref.mip_next_f(fl))
    print(f"c3 done {time.time() - t0:.1f}s", flush=True)
    # adversarial blocks through every encoder
    blocks = synth.blocks_to_image(synth.random_blocks(8192, 7))
    g["inputs"]["blocks_8192"] = sha(blocks)
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        g["outputs"][f"blocks_8192_{f}"] = sha(ref.bc_encode(blocks, f))
    g["outputs"]["blocks_8192_bc7_m1k64"] = sha(ref64.bc_encode(blocks, "bc7"))
    g["outputs"]["blocks_8192_astc"] = sha(ref.astc_encode(blocks))
    # small full-byte fixtures (edge clamps, partial blocks, odd mips)
    for (w, h) in SMALL_SIZES:
        img = small_image(w, h, 40 + w * 7 + h)
        key = f"{w}x{h}"
        ent = {"input": img.tobytes().hex()}
        for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
            ent[f] = ref.bc_encode(img, f).tobytes().hex()
        ent["astc"] = ref.astc_encode(img).tobytes().hex()
        hf = (img.astype(np.float32) / 16.0) ** 2
        hf[..., 3] = 1.0
        ent["hdr_input"] = hf.astype(np.float32).tobytes().hex()
        ent["bc6h"] = ref.bc6h_encode(hf).tobytes().hex()
        ent["mip"] = ref.mip_next(img, 0).tobytes().hex()
        ent["mip_srgb"] = ref.mip_next(img, 1).tobytes().hex()
        g["small"][key] = ent
    g["near_singular"] = near_singular(ref)
    g["containers"] = containers()
    g["rdo"] = rdo_hashes()
    with open(OUT, "w") as fh:
        json.dump(g, fh, indent=1, sort_keys=True)
    print(f"wrote {OUT} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
