"""GPU parity: libzaptex's gfx950 kernels produce the reference encoder's bytes.

Checked against the committed golden hashes (made from the compiled reference) at the workload sizes, and
against the CPU oracle on seeded, adversarial, odd-sized and strided inputs.  Every call goes through the
C ABI (host entry points, or the device entry points on torch tensors).
"""
import hashlib

import numpy as np
import pytest

from zap_amd import binding, synth

pytestmark = pytest.mark.gpu
FMTS = ["bc1", "bc3", "bc4", "bc5", "bc7"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_small_fixtures(zt, golden):
    for key, ent in golden["small"].items():
        w, h = map(int, key.split("x"))
        img = np.frombuffer(bytes.fromhex(ent["input"]), np.uint8).reshape(h, w, 4)
        for f in FMTS:
            assert zt.bc_encode(img, f).tobytes().hex() == ent[f], (key, f)
        assert zt.astc_encode(img).tobytes().hex() == ent["astc"], key
        hf = np.frombuffer(bytes.fromhex(ent["hdr_input"]), np.float32).reshape(h, w, 4)
        assert zt.bc6h_encode(hf).tobytes().hex() == ent["bc6h"], key
        assert zt.mip_next(img, 0).tobytes().hex() == ent["mip"], key
        assert zt.mip_next(img, 1).tobytes().hex() == ent["mip_srgb"], key


def test_near_singular_blocks(zt, golden):
    """Blocks whose BC1 least-squares step passes the det guard with a nearly singular system (the reference's
    largest unclamped casts), as the compiled reference encodes them."""
    ent = golden["near_singular"]
    img = np.frombuffer(bytes.fromhex(ent["input"]), np.uint8).reshape(4, 4 * ent["blocks"], 4)
    for f in ("bc1", "bc3"):
        assert zt.bc_encode(img, f).tobytes().hex() == ent[f], f
    assert zt.astc_encode(img).tobytes().hex() == ent["astc"]


def test_adversarial_blocks(zt, golden):
    blocks = synth.blocks_to_image(synth.random_blocks(8192, 7))
    for f in FMTS:
        assert sha(zt.bc_encode(blocks, f)) == golden["outputs"][f"blocks_8192_{f}"], f
    assert sha(zt.astc_encode(blocks)) == golden["outputs"]["blocks_8192_astc"]
    zt.set_m1k(64)
    try:
        assert sha(zt.bc_encode(blocks, "bc7")) == golden["outputs"]["blocks_8192_bc7_m1k64"]
    finally:
        zt.set_m1k(4)


def test_c1_bc1_and_bc345(zt, golden):
    c1 = synth.c1_bc1(1024, 1)
    for f in ["bc1", "bc3", "bc4", "bc5"]:
        assert sha(zt.bc_encode(c1, f)) == golden["outputs"][f"c1_1024_{f}"], f
    nm = synth.normal_map(1024, 1024, 11)
    assert sha(zt.bc_encode(nm, "bc5")) == golden["outputs"]["normal_1024_bc5"]


def test_c2_bc7_4096_chain_on_device(zt, golden):
    """The headline workload: 4096^2 BC7 plus its 13-level box chain, generated and encoded in HBM."""
    import torch
    c2 = synth.c2_bc7(4096, 2)
    base = torch.from_numpy(c2).cuda()
    sizes = binding.mip_sizes(4096, 4096)
    for srgb in (0, 1):
        levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
        zt.mip_chain_dev(base, 4096, 4096, 4096 * 4, levels, srgb=srgb)
        torch.cuda.synchronize()
        got = [sha(c2)] + [sha(l.cpu().numpy()) for l in levels]
        assert got == golden["outputs"][f"c2_mips_srgb{srgb}"], srgb
        if srgb == 0:
            lin_levels = [base] + levels
    outs = [torch.empty(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    imgs = zt.images((lin_levels[i], outs[i], w, h, w * 4, "bc7") for i, (w, h) in enumerate(sizes))
    ws = torch.empty(zt.workspace_size(imgs), dtype=torch.uint8, device="cuda")
    zt.encode_dev(imgs, ws)
    torch.cuda.synchronize()
    got = [sha(o.cpu().numpy()) for o in outs]
    bad = [i for i, (a, b) in enumerate(zip(got, golden["outputs"]["c2_bc7_chain"])) if a != b]
    assert not bad, bad
    # the same chain as BC1 below level 2 in one mixed call
    outs1 = [torch.empty(binding.bc_size(w, h, "bc1"), dtype=torch.uint8, device="cuda") for (w, h) in sizes[2:]]
    imgs1 = zt.images((lin_levels[i + 2], outs1[i], w, h, w * 4, "bc1") for i, (w, h) in enumerate(sizes[2:]))
    zt.encode_dev(imgs1, torch.empty(zt.workspace_size(imgs1), dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()
    assert [sha(o.cpu().numpy()) for o in outs1] == golden["outputs"]["c2_bc1_chain"]


def test_c2_m1k64_level(zt, golden):
    c2 = synth.c2_bc7(4096, 2)
    lvl = zt.mip_next(zt.mip_next(c2, 0), 0)
    zt.set_m1k(64)
    try:
        assert sha(zt.bc_encode(lvl, "bc7")) == golden["outputs"]["c2_bc7_1024_m1k64"]
    finally:
        zt.set_m1k(4)


def test_c3_bc6h_and_astc(zt, golden):
    hdr = synth.c3_hdr(2048, 5)
    assert sha(zt.bc6h_encode(hdr)) == golden["outputs"]["c3_bc6h_2048"]
    assert sha(zt.mip_next_f(np.ascontiguousarray(hdr[:512, :512]))) == golden["outputs"]["c3_hdr_mipf_512"]
    a = synth.c3_astc(2048, 3)
    assert sha(zt.astc_encode(a)) == golden["outputs"]["c3_astc_2048"]


@pytest.mark.parametrize("w,h,seed", [(1, 1, 1), (3, 2, 2), (5, 3, 3), (129, 67, 4), (300, 200, 5), (1000, 750, 6)])
def test_vs_oracle_odd_sizes(zt, oracle, w, h, seed):
    img = np.ascontiguousarray(synth.c2_bc7(max(w, h, 4), seed)[:h, :w])
    for f in FMTS:
        assert np.array_equal(zt.bc_encode(img, f), oracle.bc_encode(img, f)), f
    if w * h <= 300 * 200:
        assert np.array_equal(zt.astc_encode(img), oracle.astc_encode(img))
        hf = (img.astype(np.float32) / 8.0) ** 2.5
        assert np.array_equal(zt.bc6h_encode(hf), oracle.bc6h_encode(hf))
    for srgb in (0, 1):
        assert np.array_equal(zt.mip_next(img, srgb), oracle.mip_next(img, srgb))


def test_strided_subimage_and_bands(zt, oracle):
    big = synth.c2_bc7(256, 12)
    view = big[16:16 + 96, 8:8 + 90]          # row stride 1024 bytes, not 90*4
    assert view.strides[0] == 256 * 4
    for f in ["bc1", "bc7"]:
        whole = zt.bc_encode(view, f)
        assert np.array_equal(whole, oracle.bc_encode(np.ascontiguousarray(view), f)), f
        bands = np.concatenate([zt.bc_encode(view[y:y + 32], f) for y in range(0, 96, 32)])
        assert np.array_equal(whole, bands), f


def test_host_calls_from_threads(zt):
    """The host entry points are callable from several threads at once (the reference's callers encode bands on
    threads, viewer.cpp:461-470): each thread's output equals the same call made alone."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = [(synth.c2_bc7(128, 40 + i)[: 64 + 8 * i, : 96 - 4 * i], f) for i in range(8) for f in ("bc1", "bc3", "bc7")]
    alone = [zt.bc_encode(np.ascontiguousarray(img), f) for img, f in jobs]
    with ThreadPoolExecutor(6) as ex:
        together = list(ex.map(lambda j: zt.bc_encode(np.ascontiguousarray(j[0]), j[1]), jobs * 3))
    for k, out in enumerate(together):
        assert np.array_equal(out, alone[k % len(jobs)]), k


def test_decoders_match_oracle(zt, oracle):
    rng = np.random.default_rng(3)
    raw = rng.integers(0, 256, 16 * 300, dtype=np.uint8)
    for f in FMTS:
        n = binding.bc_size(40, 120, f)
        assert np.array_equal(zt.bc_decode(raw[:n], 40, 120, f), oracle.bc_decode(raw[:n], 40, 120, f)), f
    assert np.array_equal(zt.astc_decode(raw, 40, 120), oracle.astc_decode(raw, 40, 120))
    assert np.array_equal(zt.bc6h_decode(raw, 40, 120), oracle.bc6h_decode(raw, 40, 120))
    img = synth.c3_astc(64, 9)
    bc = zt.astc_encode(img)
    assert np.array_equal(zt.astc_decode(bc, 64, 64), oracle.astc_decode(bc, 64, 64))
    # every BC7 mode: 256 random blocks each, mode bits forced (random bytes reach modes 6 and 7 only ~1 in 100)
    blk = rng.integers(0, 256, (8, 256, 16), dtype=np.uint8)
    for m in range(8):
        blk[m, :, 0] = (blk[m, :, 0] & np.uint8(0xFF << (m + 1) & 0xFF)) | np.uint8(1 << m)
    raw7 = np.ascontiguousarray(blk.reshape(-1))
    assert np.array_equal(zt.bc_decode(raw7, 4 * 64, 4 * 32, "bc7"), oracle.bc_decode(raw7, 4 * 64, 4 * 32, "bc7"))
    # ASTC blocks in the decoder's subset: 4x4 single-plane modes (bits 1:0 != 0, 3:2 = 0, 6:5 = 2, 8:7 = 0, bit 10 = 0),
    # one or two partitions, endpoint modes 8 / 12 (some mixed, which decode to magenta), the rest random; about a
    # third leave room for their colour values (the rest decode to magenta too)
    n = 16384
    a = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    lo = a[:, :8].copy().view(np.uint64)[:, 0]
    modes = np.array([m for m in range(2048) if m & 3 and not (m >> 2) & 3 and (m >> 5) & 3 == 2 and not (m >> 7) & 3 and not (m >> 10) & 1],
                     np.uint64)
    mode = modes[rng.integers(0, len(modes), n)]
    parts = rng.integers(0, 2, n).astype(np.uint64)
    cem = np.where(rng.integers(0, 2, n) == 1, 12, 8).astype(np.uint64)
    lo = (lo & ~np.uint64(0x1FFF)) | mode | parts << np.uint64(11)              # mode, partition count
    one, two = parts == 0, parts == 1
    lo[one] = (lo[one] & ~np.uint64(0xF << 13)) | cem[one] << np.uint64(13)       # endpoint mode
    cf = (cem[two] << np.uint64(2)) | (rng.integers(0, 8, int(two.sum())) == 0).astype(np.uint64)
    lo[two] = (lo[two] & ~np.uint64(0x3F << 23)) | cf << np.uint64(23)           # seed (bits 13..22) stays random
    a[:, :8] = lo.view(np.uint8).reshape(n, 8)
    rawa = np.ascontiguousarray(a.reshape(-1))
    assert np.array_equal(zt.astc_decode(rawa, 4 * 128, 4 * 128), oracle.astc_decode(rawa, 4 * 128, 4 * 128))


def test_split_merge_match_oracle(zt, oracle):
    import ctypes as C
    rng = np.random.default_rng(4)
    for f in FMTS:
        n = binding.bc_size(64, 36, f) + 5
        raw = rng.integers(0, 256, n, dtype=np.uint8)
        exp = np.zeros_like(raw)
        oracle.lib.zto_bc_shuffle(raw.ctypes.data_as(C.c_void_p), exp.ctypes.data_as(C.c_void_p), n, binding.fmt_id(f), 1)
        sp = zt.bc_split(raw, f)
        assert np.array_equal(sp, exp), f
        assert np.array_equal(zt.bc_merge(sp, f), raw), f



@pytest.mark.parametrize("m1k", [1, 2, 3, 8])
def test_mode1_fit_and_pick_forms(zt, oracle, m1k):
    """Mode 1's fits pick their winner in the same launch when a block's 2 * M1K fits make whole lane groups
    (M1K = 1, 2, 4, 8; zt_bc7_s1_fitpick), else in a second launch (zt_bc7_s1_fit + zt_bc7_s1_pick, M1K = 3):
    both against the oracle at the same M1K, on the adversarial blocks."""
    blocks = synth.blocks_to_image(synth.random_blocks(8192, 7))
    zt.set_m1k(m1k)
    oracle.lib.zto_set_m1k(m1k)
    try:
        assert np.array_equal(zt.bc_encode(blocks, "bc7"), oracle.bc_encode(blocks, "bc7")), m1k
    finally:
        zt.set_m1k(4)
        oracle.lib.zto_set_m1k(4)


@pytest.mark.parametrize("chunks", [0, 1, 2, 5])
def test_mode6_chunks_per_wave_do_not_change_bytes(zt, golden, oracle, chunks):
    """The BC7 mode-6 stage in one pass per block (0) or with its refits queued over k 64-block chunks per wave
    (ZAPTEX_OPT_S6_CHUNKS): the reference's bytes either way, also where a wave's chunks cross from one level into
    the next (a 10-job chain in one call) and where the queues are flushed part full."""
    import torch
    OPT_S6_CHUNKS = 14
    assert zt.get_option(OPT_S6_CHUNKS) == 2
    blocks = synth.blocks_to_image(synth.random_blocks(8192, 7))
    c2 = synth.c2_bc7(512, 2)
    base = torch.from_numpy(c2).cuda()
    sizes = binding.mip_sizes(512, 512)
    levels = [torch.empty((h, w, 4), dtype=torch.uint8, device="cuda") for (w, h) in sizes[1:]]
    zt.mip_chain_dev(base, 512, 512, 512 * 4, levels, srgb=0)
    torch.cuda.synchronize()
    src = [base] + levels
    zt.set_option(OPT_S6_CHUNKS, chunks)
    try:
        assert sha(zt.bc_encode(blocks, "bc7")) == golden["outputs"]["blocks_8192_bc7"]
        outs = [torch.zeros(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda") for (w, h) in sizes]
        imgs = zt.images((src[i], outs[i], w, h, w * 4, "bc7") for i, (w, h) in enumerate(sizes))
        ws = torch.empty(zt.workspace_size(imgs), dtype=torch.uint8, device="cuda")
        zt.encode_dev(imgs, ws)
        torch.cuda.synchronize()
        got = [o.cpu().numpy() for o in outs]
    finally:
        zt.set_option(OPT_S6_CHUNKS, 2)
    for lv in range(len(sizes)):
        assert np.array_equal(got[lv], oracle.bc_encode(src[lv].cpu().numpy(), "bc7")), (chunks, lv)
