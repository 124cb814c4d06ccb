"""Batched encode of many textures with their mip chains (zaptex_batch_*), and the batch workload's textures.

`Batch` owns one texture list's device arenas (generated mip levels, output blocks), the workspace and the
plan; `run()` is one graph launch.  `c4_specs` / `c4_textures` make the mixed batch (512 seeded textures, sides
64..4096 including non-powers of two, 30% BC1 / 15% BC3 / 15% BC5 unit-vector normal maps / 40% BC7, full
chains).  The textures are generated on the GPU with the same integer arithmetic as `synth` (a 4096^2 texture
takes milliseconds instead of seconds); `texture_cpu` gives the same bytes from `synth` for small checks.
"""
from __future__ import annotations

import numpy as np

from . import binding, synth

SIDES = (64, 128, 256, 512, 1024, 2048, 4096)
FORMAT_MIX = (("bc1", 30), ("bc3", 15), ("bc5", 15), ("bc7", 40))


def mip_levels(w: int, h: int) -> int:  # zap_tex.h:1435
    n = 1
    while w > 1 or h > 1:
        w, h, n = (w // 2 if w > 1 else 1), (h // 2 if h > 1 else 1), n + 1
    return n


def layout(w: int, h: int, fmt, levels: int):
    """(level sizes, mip-arena offsets of levels 1.., out-arena offsets of levels 0..) as zaptex_batch_* lays them out."""
    al = lambda v: (v + 255) // 256 * 256
    sizes, mo, oo, m, o = [], [], [], 0, 0
    cw, ch = w, h
    for lv in range(levels):
        sizes.append((cw, ch))
        oo.append(o)
        o += al(binding.bc_size(cw, ch, fmt))
        if lv + 1 < levels:
            cw, ch = (cw // 2 if cw > 1 else 1), (ch // 2 if ch > 1 else 1)
            mo.append(m)
            m += al(cw * ch * 4)
    return sizes, mo, oo


class Batch:
    """specs: (base uint8 tensor [h, w, 4] on the GPU (or with a row stride), w, h, fmt, srgb, levels or None)."""

    def __init__(self, lib, specs):
        import torch
        self.lib = lib
        self.items = []
        for base, w, h, fmt, srgb, levels in specs:
            levels = mip_levels(w, h) if levels is None else levels
            stride = base.stride(0) * base.element_size() if base.dim() == 3 else w * 4
            mips = torch.empty(max(1, lib.batch_mip_bytes(w, h, levels)), dtype=torch.uint8, device="cuda")
            out = torch.zeros(lib.batch_out_bytes(w, h, fmt, levels), dtype=torch.uint8, device="cuda")
            self.items.append(dict(base=base, stride=stride, w=w, h=h, fmt=fmt, srgb=srgb, levels=levels, mips=mips, out=out,
                                   layout=layout(w, h, fmt, levels)))
        self.texs = lib.textures((it["base"], it["stride"], it["w"], it["h"], it["fmt"], it["srgb"], it["levels"],
                                  it["mips"] if it["levels"] > 1 else None, it["out"]) for it in self.items)
        self.ws = torch.empty(lib.batch_workspace_size(self.texs), dtype=torch.uint8, device="cuda")
        self.handle = lib.batch_create(self.texs, self.ws)
        self.pixels = sum(w * h for it in self.items for (w, h) in it["layout"][0])
        self.blocks = sum(((w + 3) // 4) * ((h + 3) // 4) for it in self.items for (w, h) in it["layout"][0])

    def run(self, stream=None) -> None:
        self.lib.batch_run(self.handle, stream)

    def replan(self) -> None:
        """Rebuild the plan (and its graph) over the same textures, e.g. after zaptex_set_launch.  The workspace is
        sized again first: its task queue is planned from the launch shapes, so new shapes can need a larger one."""
        import torch
        self.close()
        need = self.lib.batch_workspace_size(self.texs)
        if need > self.ws.numel():
            self.ws = None
            self.ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        self.handle = self.lib.batch_create(self.texs, self.ws)

    def level(self, i: int, lv: int):
        """Level lv of texture i as a [h, w, 4] tensor (level 0 is the caller's base)."""
        it = self.items[i]
        if lv == 0:
            return it["base"]
        w, h = it["layout"][0][lv]
        o = it["layout"][1][lv - 1]
        return it["mips"][o:o + w * h * 4].view(h, w, 4)

    def blocks_of(self, i: int, lv: int):
        it = self.items[i]
        w, h = it["layout"][0][lv]
        o = it["layout"][2][lv]
        return it["out"][o:o + binding.bc_size(w, h, it["fmt"])]

    def digests(self) -> list[tuple[str, str]]:
        """Per texture, SHA-256 of its generated levels 1.. concatenated and of every level's blocks concatenated."""
        import hashlib
        out = []
        for i, it in enumerate(self.items):
            mh, bh = hashlib.sha256(), hashlib.sha256()
            for lv in range(1, it["levels"]):
                mh.update(self.level(i, lv).cpu().numpy().tobytes())
            for lv in range(it["levels"]):
                bh.update(self.blocks_of(i, lv).cpu().numpy().tobytes())
            out.append((mh.hexdigest(), bh.hexdigest()))
        return out

    def close(self) -> None:
        if self.handle:
            self.lib.batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------------------- the C4 textures
def c4_specs(n: int = 512, seed: int = 4):
    """(w, h, fmt, srgb, kind) per texture: sides from SIDES (aspect <= 8) or, for one in five, any side in
    64..4096; textures 0 and 1 are 300x200 and 1000x750.  kind: 'colour' (c1 family, opaque) for BC1, 'alpha'
    (c2 family) for BC3 and BC7, 'normal' for BC5."""
    out = []
    for i in range(n):
        u = [int(synth.hash32(seed, 31, i, k)) for k in range(6)]
        pick, acc = u[0] % 100, 0
        for fmt, pct in FORMAT_MIX:
            acc += pct
            if pick < acc:
                break
        if i == 0:
            w, h = 300, 200
        elif i == 1:
            w, h = 1000, 750
        elif u[1] % 5 == 0:
            w = 64 + u[2] % (4096 - 64)
            h = max(64, (w + 7) // 8, min(4096, w * (1 + u[3] % 8) // (1 + u[4] % 8)))
        else:
            w = SIDES[u[2] % len(SIDES)]
            h = SIDES[u[3] % len(SIDES)]
            while max(w, h) > 8 * min(w, h):
                h, w = (h // 2, w) if h > w else (h, w // 2)
        kind = "normal" if fmt == "bc5" else "colour" if fmt == "bc1" else "alpha"
        srgb = 0 if kind == "normal" else u[5] & 1
        out.append((w, h, fmt, srgb, kind))
    return out


def _noise_gpu(h: int, w: int, cell: int, seed: int, chan: int, dev: str = "cuda"):
    """synth.value_noise on the GPU: lattice hashed on the host (small), interpolation in int64 on the device."""
    import torch
    x = torch.arange(w, dtype=torch.int64, device=dev)
    y = torch.arange(h, dtype=torch.int64, device=dev)
    gx, gy = x // cell, y // cell

    def smooth(t):
        return (t * t * (3 * 65536 - 2 * t)) >> 32
    tx = smooth(((x % cell) << 16) // cell)[None, :]
    ty = smooth(((y % cell) << 16) // cell)[:, None]
    lx = np.arange((w - 1) // cell + 2, dtype=np.int64)[None, :]
    ly = np.arange((h - 1) // cell + 2, dtype=np.int64)[:, None]
    lat = torch.from_numpy((synth.hash32(seed, chan, cell, lx, ly) >> np.uint32(16)).astype(np.int64)).to(dev)
    r0, r1 = lat[gy], lat[gy + 1]
    top = r0[:, gx] * (65536 - tx) + r0[:, gx + 1] * tx
    bot = r1[:, gx] * (65536 - tx) + r1[:, gx + 1] * tx
    return ((top >> 16) * (65536 - ty) + (bot >> 16) * ty) >> 16


def _fractal_gpu(h: int, w: int, cells, seed: int, chan: int, dev: str = "cuda"):
    import torch
    acc = torch.zeros((h, w), dtype=torch.int64, device=dev)
    wsum = 0
    for o, n in enumerate(cells):
        cell = max(1, max(h, w) // n)
        wt = 1 << (len(cells) - 1 - o)
        acc += _noise_gpu(h, w, cell, seed, chan * 16 + o, dev) * wt
        wsum += wt
    return acc // wsum


def _rects_gpu(img, count: int, seed: int, side, channels: int = 3) -> None:
    """synth.rectangles on a CUDA tensor: the same hashes (computed for all rectangles at once), painted in order."""
    import torch
    h, w = img.shape[:2]
    v = synth.hash32(seed, 7777, np.arange(count)[:, None], np.arange(4 + channels)[None, :]).astype(np.int64)
    colours = torch.from_numpy((v[:, 4:] & 255).astype(np.uint8)).to(img.device)
    for r in range(count):
        rw = side[0] + int(v[r, 0]) % (side[1] - side[0] + 1)
        rh = side[0] + int(v[r, 1]) % (side[1] - side[0] + 1)
        x0, y0 = int(v[r, 2]) % max(1, w), int(v[r, 3]) % max(1, h)
        img[y0:y0 + rh, x0:x0 + rw, :channels] = colours[r]


def texture_gpu(w: int, h: int, kind: str, seed: int, dev: str = "cuda"):
    """One C4 texture as a uint8 [h, w, 4] tensor on `dev` (same bytes as texture_cpu)."""
    import torch
    img = torch.empty((h, w, 4), dtype=torch.uint8, device=dev)
    if kind == "normal":
        ht = _fractal_gpu(h, w, [4, 16, 64], seed, 0, dev).to(torch.float64) / 65536.0 * 64.0
        dx = torch.roll(ht, -1, 1) - ht
        dy = torch.roll(ht, -1, 0) - ht
        ln = torch.sqrt(dx * dx + dy * dy + 1.0)
        img[..., 0] = torch.clamp((-dx / ln * 0.5 + 0.5) * 255.0, 0, 255).to(torch.uint8)
        img[..., 1] = torch.clamp((-dy / ln * 0.5 + 0.5) * 255.0, 0, 255).to(torch.uint8)
        img[..., 2] = torch.clamp((1.0 / ln * 0.5 + 0.5) * 255.0, 0, 255).to(torch.uint8)
        img[..., 3] = 255
        return img
    for c in range(3):
        img[..., c] = (_fractal_gpu(h, w, [8, 32, 128, 512], seed, c, dev) * 255 >> 16).to(torch.uint8)
    side = max(w, h)
    if kind == "colour":
        img[..., 3] = 255
        _rects_gpu(img, max(8, side // 32), seed, (4, max(8, side // 16)))
        return img
    img[..., 3] = (_fractal_gpu(h, w, [4, 16, 64, 256], seed, 3, dev) * 255 >> 16).to(torch.uint8)
    _rects_gpu(img, max(8, side // 16), seed, (4, max(8, side // 16)))
    bx = np.arange((w + 3) // 4)[None, :]
    by = np.arange((h + 3) // 4)[:, None]
    opaque = torch.from_numpy((synth.hash32(seed, 4242, bx, by) & np.uint32(3)) == 0).to(dev)
    mask = opaque.repeat_interleave(4, 0).repeat_interleave(4, 1)[:h, :w]
    img[..., 3][mask] = 255
    return img


def texture_cpu(w: int, h: int, kind: str, seed: int) -> np.ndarray:
    """texture_gpu's bytes from the host generator (numpy), for small parity checks."""
    img = np.empty((h, w, 4), dtype=np.uint8)
    if kind == "normal":
        ht = synth.fractal(h, w, [4, 16, 64], seed, 0).astype(np.float64) / 65536.0 * 64.0
        dx = np.roll(ht, -1, axis=1) - ht
        dy = np.roll(ht, -1, axis=0) - ht
        ln = np.sqrt(dx * dx + dy * dy + 1.0)
        img[..., 0] = np.clip((-dx / ln * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
        img[..., 1] = np.clip((-dy / ln * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
        img[..., 2] = np.clip((1.0 / ln * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
        img[..., 3] = 255
        return img
    for c in range(3):
        img[..., c] = (synth.fractal(h, w, [8, 32, 128, 512], seed, c) * 255 >> 16).astype(np.uint8)
    side = max(w, h)
    if kind == "colour":
        img[..., 3] = 255
        synth.rectangles(img, max(8, side // 32), seed, (4, max(8, side // 16)))
        return img
    img[..., 3] = (synth.fractal(h, w, [4, 16, 64, 256], seed, 3) * 255 >> 16).astype(np.uint8)
    synth.rectangles(img, max(8, side // 16), seed, (4, max(8, side // 16)))
    bx = np.arange((w + 3) // 4)[None, :]
    by = np.arange((h + 3) // 4)[:, None]
    opaque = (synth.hash32(seed, 4242, bx, by) & np.uint32(3)) == 0
    mask = np.repeat(np.repeat(opaque, 4, axis=0), 4, axis=1)[:h, :w]
    img[..., 3][mask] = 255
    return img


def c4_textures(n: int = 512, seed: int = 4):
    """(specs, base tensors): the C4 batch generated on the GPU; texture i uses seed * 100000 + i."""
    specs = c4_specs(n, seed)
    bases = [texture_gpu(w, h, kind, seed * 100000 + i) for i, (w, h, fmt, srgb, kind) in enumerate(specs)]
    return specs, bases
