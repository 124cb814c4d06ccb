"""Seeded synthetic textures for the encoder workloads (deterministic, integer arithmetic throughout).

Every generator is a pure function of its arguments: a counter-based 32-bit hash (no RNG state) feeds
integer value noise with fixed-point smoothstep interpolation, so the same bytes come out on any host.

Workloads (sizes are the ones the parity tests and the benchmark use):
  * ``c1_bc1(1024)``   smooth gradients + 4-octave value noise (amplitude 0..64) + 32 hard-edged rectangles, alpha 255
  * ``c2_bc7(4096)``   full-range RGB value noise + 256 rectangles; alpha is independent noise, 25% of 4x4 blocks opaque
  * ``c3_hdr(2048)``   RGB half-float fields x 2^U(-8, 8) with 5% specular spikes up to 2^12, as float32 RGBA
  * ``c3_astc(2048)``  the c2 family at 2048^2
  * ``normal_map``     unit-vector XY normals (BC5 inputs)
"""
from __future__ import annotations

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def hash32(*keys) -> np.ndarray:
    """Counter-based hash (lowbias32 chain) of integer arrays/scalars, elementwise, as uint32."""
    h = np.uint64(0x9E3779B9)
    out = None
    for k in keys:
        k = np.asarray(k, dtype=np.uint64) & _M32
        x = ((h if out is None else out) ^ k) & _M32
        x = (x ^ (x >> np.uint64(16))) * np.uint64(0x7FEB352D) & _M32
        x = (x ^ (x >> np.uint64(15))) * np.uint64(0x846CA68B) & _M32
        x = x ^ (x >> np.uint64(16))
        out = x
    return out.astype(np.uint32)


def _smooth_q16(t: np.ndarray) -> np.ndarray:
    """Fixed-point smoothstep: t in [0, 65536) (Q16) -> t^2 (3 - 2t) in Q16."""
    t = t.astype(np.int64)
    return (t * t * (3 * 65536 - 2 * t)) >> 32


def value_noise(h: int, w: int, cell: int, seed: int, chan: int) -> np.ndarray:
    """Value noise in Q16 [0, 65536): random lattice values every `cell` pixels, smoothstep-bilinear."""
    y = np.arange(h, dtype=np.int64)
    x = np.arange(w, dtype=np.int64)
    gx, gy = x // cell, y // cell
    tx = _smooth_q16(((x % cell) << 16) // cell)[None, :]
    ty = _smooth_q16(((y % cell) << 16) // cell)[:, None]
    # lattice values (0..65535) hashed once per lattice point, then gathered per pixel
    lx = np.arange(int(gx[-1]) + 2, dtype=np.int64)[None, :]
    ly = np.arange(int(gy[-1]) + 2, dtype=np.int64)[:, None]
    lat = (hash32(seed, chan, cell, lx, ly) >> np.uint32(16)).astype(np.int64)
    r0, r1 = lat[gy], lat[gy + 1]
    top = r0[:, gx] * (65536 - tx) + r0[:, gx + 1] * tx
    bot = r1[:, gx] * (65536 - tx) + r1[:, gx + 1] * tx
    return ((top >> 16) * (65536 - ty) + (bot >> 16) * ty) >> 16


def fractal(h: int, w: int, cells_per_side: list[int], seed: int, chan: int) -> np.ndarray:
    """Octave sum with weights 1, 1/2, 1/4, ... normalised; Q16 result."""
    acc = np.zeros((h, w), dtype=np.int64)
    wsum = 0
    for o, n in enumerate(cells_per_side):
        cell = max(1, max(h, w) // n)
        wt = 1 << (len(cells_per_side) - 1 - o)
        acc += value_noise(h, w, cell, seed, chan * 16 + o) * wt
        wsum += wt
    return acc // wsum


def rectangles(img: np.ndarray, count: int, seed: int, side: tuple[int, int], channels: int = 3) -> None:
    """Paint `count` axis-aligned hard-edged rectangles of uniform random colour in place."""
    h, w = img.shape[:2]
    for r in range(count):
        v = [int(hash32(seed, 7777, r, k)) for k in range(4 + channels)]
        rw = side[0] + v[0] % (side[1] - side[0] + 1)
        rh = side[0] + v[1] % (side[1] - side[0] + 1)
        x0 = v[2] % max(1, w)
        y0 = v[3] % max(1, h)
        for c in range(channels):
            img[y0:y0 + rh, x0:x0 + rw, c] = v[4 + c] & 255


def c1_bc1(size: int = 1024, seed: int = 1) -> np.ndarray:
    """Smooth gradients + 4-octave value noise (0..64) + 32 rectangles, alpha 255 (RGBA8, HxWx4)."""
    h = w = size
    y = np.arange(h, dtype=np.int64)[:, None]
    x = np.arange(w, dtype=np.int64)[None, :]
    base = [255 * x // max(1, w - 1) + 0 * y, 255 * y // max(1, h - 1) + 0 * x, 255 * (x + y) // max(1, w + h - 2)]
    img = np.zeros((h, w, 4), dtype=np.uint8)
    for c in range(3):
        n = fractal(h, w, [8, 16, 32, 64], seed, c) * 64 >> 16
        img[..., c] = np.clip(base[c] + n, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    rectangles(img, 32, seed, (8, 256))
    return img


def c2_bc7(size: int = 4096, seed: int = 2, rects: int = 256) -> np.ndarray:
    """Full-range RGB value noise + hard-edged rectangles; independent alpha noise, 25% opaque 4x4 blocks."""
    h = w = size
    img = np.zeros((h, w, 4), dtype=np.uint8)
    for c in range(3):
        img[..., c] = (fractal(h, w, [8, 32, 128, 512], seed, c) * 255 >> 16).astype(np.uint8)
    img[..., 3] = (fractal(h, w, [4, 16, 64, 256], seed, 3) * 255 >> 16).astype(np.uint8)
    rectangles(img, rects, seed, (4, max(8, size // 16)))
    by = np.arange((h + 3) // 4)[:, None]
    bx = np.arange((w + 3) // 4)[None, :]
    opaque = (hash32(seed, 4242, bx, by) & np.uint32(3)) == 0
    mask = np.repeat(np.repeat(opaque, 4, axis=0), 4, axis=1)[:h, :w]
    img[..., 3][mask] = 255
    return img


def c3_astc(size: int = 2048, seed: int = 3) -> np.ndarray:
    return c2_bc7(size, seed, rects=128)


def c3_hdr(size: int = 2048, seed: int = 5) -> np.ndarray:
    """RGB half-float values as float32 RGBA (alpha 1): smooth fields x 2^U(-8, 8), 5% spikes to 2^12."""
    h = w = size
    expo = fractal(h, w, [4, 16], seed, 9).astype(np.float64) / 65536.0 * 16.0 - 8.0
    out = np.ones((h, w, 4), dtype=np.float32)
    spike_sel = (hash32(seed, 555, np.arange(h)[:, None], np.arange(w)[None, :]) % np.uint32(100)) < 5
    spike_mag = 2.0 ** ((hash32(seed, 556, np.arange(h)[:, None], np.arange(w)[None, :]) % np.uint32(13)).astype(np.float64))
    for c in range(3):
        field = fractal(h, w, [8, 32, 128], seed, c).astype(np.float64) / 65536.0
        v = field * np.exp2(expo)
        v = np.where(spike_sel, spike_mag, v)  # spikes: 2^0 .. 2^12
        out[..., c] = v.astype(np.float16).astype(np.float32)  # exact half values
    return out


def normal_map(h: int, w: int, seed: int = 11) -> np.ndarray:
    """Unit-vector XY normals from a height field (R = X, G = Y in [0, 255], B = Z, alpha 255)."""
    ht = fractal(h, w, [4, 16, 64], seed, 0).astype(np.float64) / 65536.0 * 64.0
    dx = np.roll(ht, -1, axis=1) - ht
    dy = np.roll(ht, -1, axis=0) - ht
    l = np.sqrt(dx * dx + dy * dy + 1.0)
    img = np.empty((h, w, 4), dtype=np.uint8)
    img[..., 0] = np.clip((-dx / l * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
    img[..., 1] = np.clip((-dy / l * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
    img[..., 2] = np.clip((1.0 / l * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    return img


def random_blocks(n: int, seed: int, kind: str = "mixed") -> np.ndarray:
    """n adversarial 4x4 RGBA blocks as an (n*4)x4 RGBA8 image column-stacked into a 4-wide strip.

    kinds mix: uniform noise, solid, two-colour, alpha-only variation, gradients, near-solid with one outlier.
    """
    k = np.arange(n, dtype=np.int64)[:, None]
    i = np.arange(64, dtype=np.int64)[None, :]
    r = hash32(seed, k, i).astype(np.int64)
    sel = (hash32(seed, 99, k[:, 0]) % np.uint32(8)).astype(np.int64) if kind == "mixed" else np.zeros(n, np.int64)
    blk = (r & 255).astype(np.uint8)                                   # 0: noise
    base = (hash32(seed, 5, k, i % 4) & np.uint32(255)).astype(np.uint8)
    two = np.where(((r >> 8) & 1) == 1, base, (base ^ 0x5A).astype(np.uint8))
    px = np.arange(16)[None, :]
    grad = np.repeat(((px * 17 + (r[:, :16] & 7)) & 255).astype(np.uint8), 4, axis=1)
    out = blk.copy()
    out[sel == 1] = base[sel == 1]                                      # solid
    out[sel == 2] = two[sel == 2]                                       # two-colour
    out[sel == 3] = base[sel == 3]
    out[sel == 3, 3::4] = blk[sel == 3, 3::4]                           # alpha-only variation
    out[sel == 4] = grad[sel == 4]                                      # gradient
    out[sel == 5] = base[sel == 5]
    out[sel == 5, 20:24] = blk[sel == 5, 20:24]                         # near-solid + outlier
    opaque = (sel == 6) | (sel == 7) | (((r[:, 0] >> 12) & 1) == 1)
    out[opaque, 3::4] = 255
    return out.reshape(n, 16, 4)


def degenerate_blocks(n: int, seed: int) -> np.ndarray:
    """n near-degenerate 4x4 RGBA blocks (n, 16, 4), the inputs that drive the encoders' least-squares solves
    towards a singular system: solid with +-1 on one channel, two colours, collinear steps (small and large),
    solid with one outlier, solid with +-3 noise; families cycle with the block index."""
    k = np.arange(n, dtype=np.int64)
    r = hash32(seed, k[:, None], np.arange(80)[None, :]).astype(np.int64)
    base = r[:, 0:4] & 255
    px = np.repeat(base[:, None, :], 16, axis=1)
    fam = k % 6
    i16 = np.arange(16)
    one = r[:, 4] % 3                                                    # 0: +-1 on one channel
    noise1 = (r[:, 16:32] % 3) - 1
    b0 = np.nonzero(fam == 0)[0]
    px[b0[:, None], i16[None, :], one[b0][:, None]] = base[b0, one[b0]][:, None] + noise1[b0]
    c2 = r[:, 5:9] & 255                                                 # 1: two colours
    m2 = (r[:, 32:48] & 255) < (r[:, 9:10] & 255)
    sel = fam == 1
    px[sel] = np.where(m2[sel][:, :, None], c2[sel][:, None, :], px[sel])
    d_small = (r[:, 10:13] % 17) - 8                                     # 2: collinear steps
    s4 = r[:, 48:64] % 4
    sel = fam == 2
    px[sel, :, :3] = base[sel][:, None, :3] + s4[sel][:, :, None] * d_small[sel][:, None, :]
    d_big = (r[:, 13:16] % 81) - 40                                      # 4: two levels far apart
    s2 = (r[:, 64:80] & 1) * (1 + r[:, 14:15] % 3)
    sel = fam == 4
    px[sel, :, :3] = base[sel][:, None, :3] + s2[sel][:, :, None] * d_big[sel][:, None, :]
    sel = fam == 3                                                       # 3: one outlier
    out_i = r[:, 15] % 16
    px[sel, out_i[sel], :] = r[sel][:, 40:44] & 255
    sel = fam == 5                                                       # 5: +-3 noise
    px[sel, :, :3] += ((r[sel][:, 16:64] % 7) - 3).reshape(-1, 16, 3)
    return np.clip(px, 0, 255).astype(np.uint8)


def blocks_to_image(blocks: np.ndarray) -> np.ndarray:
    """(n,16,4) blocks -> 4 x (4n) RGBA image whose k-th 4x4 block is blocks[k]."""
    n = blocks.shape[0]
    b = blocks.reshape(n, 4, 4, 4)               # block, row, col, ch
    return np.ascontiguousarray(b.transpose(1, 0, 2, 3).reshape(4, n * 4, 4))
