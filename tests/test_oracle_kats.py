"""Pins the CPU oracle with the reference's own known-answer tests (tex_bench.c:45-179) and with closed
forms and invariants that do not depend on the reference build (mip arithmetic, palette argmins, the
BC7 decoder/encoder SSE identity, band invariance)."""
import ctypes as C

import numpy as np
import pytest

from oracle import loader
from zap_amd import synth


def psnr(a, b, chmask):
    se, n = 0.0, 0
    for c in range(4):
        if chmask >> c & 1:
            d = a[..., c].astype(np.float64) - b[..., c].astype(np.float64)
            se += float((d * d).sum()); n += d.size
    return 99.0 if se == 0 else 10 * np.log10(255.0 * 255.0 * n / se)


CHMASK = {"bc1": 7, "bc3": 15, "bc4": 1, "bc5": 3, "bc7": 15}


def test_half_conversion(oracle):  # tex_bench.c:47-48
    h2f = np.array([oracle.h2f(h) for h in range(0, 0x7C00)], dtype=np.float32)
    assert all(oracle.f2h(float(v)) == h for h, v in enumerate(h2f))
    assert oracle.f2h(-1.0) == 0 and oracle.f2h(1e9) == 0x7BFF and oracle.f2h(1.0) == 0x3C00 and oracle.f2h(1.0 + 1.0 / 4096) == 0x3C00


def test_bc6h_quality_and_tiny(oracle):  # tex_bench.c:51-72
    W, H = 37, 21
    img = np.zeros((H, W, 4), np.float32)
    for y in range(H):
        for x in range(W):
            img[y, x] = [0.05 * (1 << (x // 6)), 0.5 + y * 0.25, 10000.0 if (x == 20 and y == 10) else 0.01 * (x + 1), 1]
    bc = oracle.bc6h_encode(img)
    dec = oracle.bc6h_decode(bc, W, H)
    f = np.vectorize(oracle.h2f)(dec[..., :3]).astype(np.float64)
    a = img[..., :3].astype(np.float64)
    rel = (np.abs(a - f) / (a + 0.01)).mean()
    assert rel < 0.08 and dec[0, 0, 3] == 0x3C00
    for s, hh in [(1, 1), (5, 3)]:
        oracle.bc6h_decode(oracle.bc6h_encode(img[:hh, :s]), s, hh)


def test_bc7_mode5_chosen(oracle):  # tex_bench.c:76-84
    i = np.arange(64 * 64)
    img = np.zeros((64 * 64, 4), np.uint8)
    img[:, 0] = i % 64 * 4; img[:, 1] = i // 64 * 4; img[:, 2] = 90; img[:, 3] = ((255 - i % 64 * 4) ^ (i // 64 * 4)) & 255
    img = img.reshape(64, 64, 4)
    bc = oracle.bc_encode(img, "bc7")
    m5 = int(((bc.reshape(-1, 16)[:, 0] & 0x3F) == 0x20).sum())
    assert m5 > 0 and psnr(img, oracle.bc_decode(bc, 64, 64, "bc7"), 15) > 30


def test_mip_known_answers(oracle):  # tex_bench.c:87-100
    assert oracle.mip_levels(1, 1) == 1 and oracle.mip_levels(256, 128) == 9 and oracle.mip_levels(5, 3) == 3
    bw = np.array([0, 0, 0, 255, 255, 255, 255, 255, 0, 0, 0, 255, 255, 255, 255, 255], np.uint8).reshape(2, 2, 4)
    m = oracle.mip_next(bw, 0)
    assert m[0, 0, 0] == 128 and m[0, 0, 3] == 255
    m = oracle.mip_next(bw, 1)
    assert m[0, 0, 0] == 188 and m[0, 0, 1] == 188 and m[0, 0, 3] == 255
    for v in range(256):
        assert oracle.mip_next(np.full((2, 2, 4), v, np.uint8), 1)[0, 0, 0] == v
    m2 = oracle.mip_next(np.full((3, 5, 4), 77, np.uint8), 1)
    assert m2.shape == (1, 2, 4) and m2[0, 0, 0] == 77 and m2[0, 1, 3] == 77
    fl = np.array([1, 2, 3, 4, 3, 2, 1, 0, 5, 5, 5, 5, 7, 7, 7, 7], np.float32).reshape(2, 2, 4)
    fo = oracle.mip_next_f(fl)
    assert fo[0, 0, 0] == 4 and fo[0, 0, 3] == 4


def test_mip_linear_closed_form(oracle):
    img = synth.c2_bc7(67, 3)[:, :61]
    got = oracle.mip_next(img, 0).astype(np.int64)
    s = img.astype(np.int64)
    h, w = img.shape[:2]
    q = s[0:h // 2 * 2:2, 0:w // 2 * 2:2] + s[0:h // 2 * 2:2, 1:w // 2 * 2:2] + s[1:h // 2 * 2:2, 0:w // 2 * 2:2] + s[1:h // 2 * 2:2, 1:w // 2 * 2:2]
    assert np.array_equal(got, (q + 2) >> 2)


def test_srgb_tables_match_reference_hashes(oracle):  # SHA-256 prefixes of the reference's literal tables
    import hashlib
    lin, thr = oracle.srgb_tables()
    assert hashlib.sha256(lin.astype(np.float32).tobytes()).hexdigest()[:16] == "8f94c94b2c231174"
    assert hashlib.sha256(thr.astype(np.float32).tobytes()).hexdigest()[:16] == "6b19516f9c3dd088"


def test_ise_round_trip(oracle):  # tex_bench.c:104-117
    rng = np.random.default_rng(0)
    levels = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256]
    for ri, lv in enumerate(levels):
        for n in range(1, 17):
            v = rng.integers(0, lv, n).astype(np.int32)
            assert oracle.lib.zto_ise_roundtrip(ri, n, v.ctypes.data_as(C.c_void_p)) >= 0


def test_astc_partition_checksum(oracle):  # tex_bench.c:118-121
    h = 2166136261
    for sd in range(1024):
        for i in range(16):
            h ^= oracle.lib.zto_astc_partition(sd, i & 3, i >> 2, 2)
            h = (h * 16777619) & 0xFFFFFFFF
    assert h == 0x5D5FC459
    assert oracle.lib.zto_astc_pattern_count() == 822


def test_astc_void_extent_and_quality(oracle):  # tex_bench.c:122-141
    ve = np.array([0xFC, 0xFD, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0x00, 0x40, 0x00, 0x80, 0x00, 0xC0, 0xFF, 0xFF], np.uint8)
    assert list(oracle.astc_decode(ve, 4, 4)[0, 0]) == [0x40, 0x80, 0xC0, 0xFF]
    W, H = 64, 48
    y, x = np.mgrid[0:H, 0:W]
    for k in range(2):
        edge = (x + 2 * y) % 13 < 6
        img = np.stack([np.where(edge, 220, 30 + x), np.where(edge, 40, 200 - y), x * 3,
                        ((x * 5 + y * 3) & 255) if k else np.full_like(x, 255)], -1).astype(np.uint8)
        out = oracle.astc_decode(oracle.astc_encode(img), W, H)
        assert psnr(img, out, 15) > (26 if k else 41)
    oracle.astc_encode(img[:2, :3])


def test_bcn_quality_and_split_merge(oracle):  # tex_bench.c:144-170 (rdo 0)
    W, H = 67, 45
    y, x = np.mgrid[0:H, 0:W]
    rnd = np.asarray(synth.hash32(3, x, y) % 8, np.int64)
    img = np.stack([x * 255 // W, y * 255 // H, ((x + y) * 2 + rnd) & 255, np.where(x < 30, 255, (x * y) & 255)], -1).astype(np.uint8)
    q = {}
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        bc = oracle.bc_encode(img, f)
        out = oracle.bc_decode(bc, W, H, f)
        q[f] = psnr(img, out, CHMASK[f])
        assert q[f] > 32
        if f == "bc1":
            assert (out[..., 3] == 255).all()
        sp = np.zeros_like(bc); mg = np.zeros_like(bc)
        fi = loader.FMT[f]
        oracle.lib.zto_bc_shuffle(bc.ctypes.data_as(C.c_void_p), sp.ctypes.data_as(C.c_void_p), bc.size, fi, 1)
        oracle.lib.zto_bc_shuffle(sp.ctypes.data_as(C.c_void_p), mg.ctypes.data_as(C.c_void_p), bc.size, fi, 0)
        assert np.array_equal(bc, mg)
    assert q["bc7"] > q["bc3"]


def test_solid_1x1(oracle):  # tex_bench.c:172-177
    one = np.array([10, 200, 30, 255], np.uint8).reshape(1, 1, 4)
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        o = oracle.bc_decode(oracle.bc_encode(one, f), 1, 1, f)[0, 0]
        for c in range(4):
            if CHMASK[f] >> c & 1:
                assert abs(int(o[c]) - int(one[0, 0, c])) <= 4


@pytest.mark.parametrize("fmt", ["bc1", "bc4", "bc7"])
def test_indices_are_palette_argmins(oracle, fmt):
    """Decoding gives, per pixel, the palette entry nearest the source among the block's own palette entries."""
    blocks = synth.random_blocks(512, 21)
    img = synth.blocks_to_image(blocks)
    bc = oracle.bc_encode(img, fmt)
    dec = oracle.bc_decode(bc, img.shape[1], 4, fmt).reshape(4, -1, 4, 4).transpose(1, 0, 2, 3).reshape(-1, 16, 4).astype(np.int64)
    src = blocks.astype(np.int64)
    ch = {"bc1": [0, 1, 2], "bc4": [0], "bc7": [0, 1, 2, 3]}[fmt]
    for b in range(len(blocks)):
        pal = np.unique(dec[b][:, ch], axis=0)
        if fmt == "bc7" and (bc[16 * b] & 0x7F) != 0x40:
            continue  # modes 5/1 use per-subset / per-channel palettes
        d = ((src[b][:, None, ch] - pal[None, :, :]) ** 2).sum(-1)
        got = ((src[b][:, ch] - dec[b][:, ch]) ** 2).sum(-1)
        assert np.array_equal(got, d.min(1)), b


def test_bc7_decoded_error_equals_fit_error(oracle):
    """For modes 6, 5 and 1 the SSE each encoder returns equals the SSE of its decoded block (mode 1, which
    has no alpha, on opaque blocks: the only blocks the search gives it)."""
    blocks = synth.random_blocks(1500, 5)
    res = np.zeros(6, np.int32)
    for b in range(len(blocks)):
        px = np.ascontiguousarray(blocks[b])
        oracle.lib.zto_bc7_mode_errors(px.ctypes.data_as(C.c_void_p), b % 64, res.ctypes.data_as(C.c_void_p))
        assert res[0] == res[1] and res[2] == res[3], (b, list(res))
        px[:, 3] = 255
        oracle.lib.zto_bc7_mode_errors(px.ctypes.data_as(C.c_void_p), b % 64, res.ctypes.data_as(C.c_void_p))
        assert res[4] == res[5], (b, list(res))


def test_band_split_invariance(oracle):
    img = synth.c2_bc7(96, 8)[:, :90]
    for f in ["bc1", "bc5", "bc7"]:
        whole = oracle.bc_encode(img, f)
        parts = np.concatenate([oracle.bc_encode(img[y:y + 32], f) for y in range(0, 96, 32)])
        assert np.array_equal(whole, parts)
