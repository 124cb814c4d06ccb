"""Batched encode (zaptex_batch_*): every texture's generated mip levels and every level's blocks equal the
reference path (the oracle's zap_mip_next chain + zap_bc_encode per level) for odd, non-power-of-two and 1x1
sizes, all five formats, linear and sRGB filters, partial chains and a strided base; graph replays give the same
bytes every time; a slice of the C4 batch equals the per-texture device entry points; other launch shapes
(as the autotuner may choose) change nothing."""
import numpy as np
import pytest

from zap_amd import batch, binding

pytestmark = pytest.mark.gpu

SMALL = [  # (w, h, fmt, srgb, levels, kind)
    (1, 1, "bc7", 0, None, "alpha"), (3, 5, "bc1", 1, None, "colour"), (7, 2, "bc4", 0, None, "alpha"),
    (64, 64, "bc5", 0, None, "normal"), (300, 200, "bc7", 1, None, "alpha"), (129, 67, "bc3", 1, None, "alpha"),
    (256, 4, "bc1", 0, None, "colour"), (5, 260, "bc7", 0, None, "alpha"), (1000, 750, "bc1", 1, 3, "colour"),
    (96, 96, "bc7", 1, 2, "alpha"), (33, 17, "bc5", 0, None, "normal"), (130, 66, "bc4", 1, None, "alpha"),
]


def _make(zt, specs, seed0=900, stride_pad=0):
    import torch
    bases, cpu = [], []
    for k, (w, h, fmt, srgb, levels, kind) in enumerate(specs):
        img = batch.texture_cpu(w, h, kind, seed0 + k)
        cpu.append(img)
        if stride_pad and k % 3 == 1:   # a strided view: rows padded by stride_pad pixels
            full = torch.zeros((h, w + stride_pad, 4), dtype=torch.uint8, device="cuda")
            full[:, :w] = torch.from_numpy(img).cuda()
            bases.append(full[:, :w])
        else:
            bases.append(torch.from_numpy(img).cuda())
    b = batch.Batch(zt, [(bases[k], w, h, fmt, srgb, levels) for k, (w, h, fmt, srgb, levels, kind) in enumerate(specs)])
    return b, cpu


def _snapshot(b):
    return [([b.level(i, lv).cpu().numpy().copy() for lv in range(1, it["levels"])],
             [b.blocks_of(i, lv).cpu().numpy().copy() for lv in range(it["levels"])]) for i, it in enumerate(b.items)]


def test_batch_matches_oracle(zt, oracle):
    import torch
    b, cpu = _make(zt, SMALL, stride_pad=3)
    b.run()
    torch.cuda.synchronize()
    for i, (w, h, fmt, srgb, levels, kind) in enumerate(SMALL):
        levels = b.items[i]["levels"]
        chain = oracle.mip_chain(cpu[i], srgb)[:levels]
        for lv in range(1, levels):
            assert np.array_equal(b.level(i, lv).cpu().numpy(), chain[lv]), (i, lv)
        for lv in range(levels):
            assert np.array_equal(b.blocks_of(i, lv).cpu().numpy(), oracle.bc_encode(chain[lv], fmt)), (i, lv, fmt)
    b.close()


def test_batch_graph_replays(zt):
    import torch
    b, _ = _make(zt, SMALL[:8])
    b.run()
    torch.cuda.synchronize()
    first = _snapshot(b)
    for rep in range(3):
        for it in b.items:           # eager work between replays: clear every output
            it["out"].zero_()
            it["mips"].zero_()
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        again = _snapshot(b)
        for (m0, o0), (m1, o1) in zip(first, again):
            assert all(np.array_equal(x, y) for x, y in zip(m0, m1)) and all(np.array_equal(x, y) for x, y in zip(o0, o1)), rep
    b.close()


def test_c4_slice_equals_single_calls(zt):
    import torch
    specs = batch.c4_specs(512, 4)[:24]
    bases = [batch.texture_gpu(w, h, kind, 4 * 100000 + i) for i, (w, h, fmt, srgb, kind) in enumerate(specs)]
    b = batch.Batch(zt, [(bases[i], w, h, fmt, srgb, None) for i, (w, h, fmt, srgb, kind) in enumerate(specs)])
    b.run()
    torch.cuda.synchronize()
    for i, (w, h, fmt, srgb, kind) in enumerate(specs):
        sizes = binding.mip_sizes(w, h)
        levels = [torch.empty((hh, ww, 4), dtype=torch.uint8, device="cuda") for (ww, hh) in sizes[1:]]
        if levels:
            zt.mip_chain_dev(bases[i], w, h, w * 4, levels, srgb=srgb)
        srcs = [bases[i]] + levels
        for l0 in range(0, len(sizes), 16):
            idx = range(l0, min(len(sizes), l0 + 16))
            outs = {l: torch.empty(binding.bc_size(*sizes[l], fmt), dtype=torch.uint8, device="cuda") for l in idx}
            imgs = zt.images((srcs[l], outs[l], sizes[l][0], sizes[l][1], sizes[l][0] * 4, fmt) for l in idx)
            ws = torch.empty(max(1, zt.workspace_size(imgs)), dtype=torch.uint8, device="cuda")
            zt.encode_dev(imgs, ws)
            torch.cuda.synchronize()
            for l in idx:
                if l:
                    assert torch.equal(b.level(i, l), srcs[l]), (i, l)
                assert torch.equal(b.blocks_of(i, l), outs[l]), (i, l, fmt, (w, h))
    b.close()


def test_c4_full_batch_matches_reference_hashes(zt):
    """The whole 512-texture workload in one graph launch: every texture's generated levels and every level's
    blocks hash to the compiled reference's bytes (tests/golden/c4_batch.json, made by make_c4_golden.py), with
    the autotuned launch shapes the benchmark uses."""
    import hashlib
    import json
    import os

    import torch

    from zap_amd import autotune
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c4_batch.json")))["textures"]
    specs, bases = batch.c4_textures(512, 4)
    assert [(w, h, fmt, srgb) for w, h, fmt, srgb, _ in specs] == [(g["w"], g["h"], g["fmt"], g["srgb"]) for g in gold]
    autotune.apply(zt, "c4")
    try:
        b = batch.Batch(zt, [(bases[i], w, h, fmt, srgb, None) for i, (w, h, fmt, srgb, kind) in enumerate(specs)])
        b.run()
        torch.cuda.synchronize()
        got = b.digests()
        b.close()
    finally:
        autotune.reset(zt)
    bad = [i for i, (g, (m, o)) in enumerate(zip(gold, got)) if (g["mips"], g["blocks"]) != (m, o)]
    # a mismatch names whether the generated base already differs
    bases_bad = [i for i in bad if hashlib.sha256(bases[i].cpu().numpy().tobytes()).hexdigest() != gold[i]["base"]]
    assert not bad, (bad[:10], "bases differ:", bases_bad[:10])


def test_launch_shapes_do_not_change_bytes(zt):
    import torch
    b, _ = _make(zt, SMALL[:10], seed0=700)
    b.run()
    torch.cuda.synchronize()
    ref = _snapshot(b)
    b.close()
    names = ["zt_bc14_encode", "zt_bc1_climb", "zt_bc7_stage6", "zt_bc7_stage5", "zt_bc7_s1", "zt_mip_batch"]
    saved = {k: zt.get_launch(k) for k in names}
    try:
        for wg, cap in ((64, 3), (128, 37)):
            for k in names:
                zt.set_launch(k, wg, cap)
            b2, _ = _make(zt, SMALL[:10], seed0=700)
            b2.run()
            torch.cuda.synchronize()
            got = _snapshot(b2)
            b2.close()
            for (m0, o0), (m1, o1) in zip(ref, got):
                assert all(np.array_equal(x, y) for x, y in zip(m0, m1)) and all(np.array_equal(x, y) for x, y in zip(o0, o1)), (wg, cap)
    finally:
        for k, (wg, cap) in saved.items():
            zt.set_launch(k, wg, cap)


def test_replan_sizes_the_workspace_again(zt):
    """A re-plan under launch shapes whose task queue is larger than the first plan's (tests/test_batch_plan.py pins
    the growth): Batch.replan allocates the larger workspace, and the bytes stay the same."""
    import torch
    names = ["zt_bc14_encode", "zt_bc1_climb", "zt_bc7_stage6", "zt_bc7_stage5", "zt_bc7_s1", "zt_mip_batch"]
    saved = {k: zt.get_launch(k) for k in names}
    try:
        for k in names:
            zt.set_launch(k, 256, 1)
        b, _ = _make(zt, SMALL[:10], seed0=700)
        first = b.ws.numel()
        b.run()
        torch.cuda.synchronize()
        ref = _snapshot(b)
        for k in names:
            zt.set_launch(k, 64, 1 << 20)
        b.replan()
        b.run()
        torch.cuda.synchronize()
        got = _snapshot(b)
        grown = b.ws.numel()
        b.close()
        assert grown >= zt.batch_workspace_size(b.texs) and grown > first
        for (m0, o0), (m1, o1) in zip(ref, got):
            assert all(np.array_equal(x, y) for x, y in zip(m0, m1)) and all(np.array_equal(x, y) for x, y in zip(o0, o1))
    finally:
        for k, (wg, cap) in saved.items():
            zt.set_launch(k, wg, cap)


@pytest.mark.skip(reason="the persistent batch kernel's first GPU run faulted the GPU; it stays off the GPU (and off by "
                         "default) until the cause is found")
@pytest.mark.parametrize("mix,chunks,groups", [(2, 1, 0), (0, 1, 1), (64, 8, 0), (3, 64, 2)])
def test_persistent_kernel_equals_stage_graph(zt, mix, chunks, groups):
    """The persistent batch kernel (one launch over the task queue) against the graph of per-stage kernels on the
    same textures: queue mix, chunks per task and resident workgroups change the schedule only.  Three replays of
    the one graph launch each give the same bytes (the kernel's last workgroup clears the queue for the next)."""
    import torch
    saved = {k: zt.get_option(k) for k in (6, 7, 8, 9)}
    try:
        zt.set_option(6, 0)
        b, _ = _make(zt, SMALL, seed0=500, stride_pad=2)
        b.run()
        torch.cuda.synchronize()
        ref = _snapshot(b)
        b.close()
        zt.set_option(6, 1)
        zt.set_option(7, mix)
        zt.set_option(8, chunks)
        zt.set_option(9, groups)
        b, _ = _make(zt, SMALL, seed0=500, stride_pad=2)
        for rep in range(3):
            for it in b.items:   # clear the outputs so every replay has to write them again
                it["out"].zero_()
                it["mips"].zero_()
            b.run()
            torch.cuda.synchronize()
            got = _snapshot(b)
            for i, ((m0, o0), (m1, o1)) in enumerate(zip(ref, got)):
                assert all(np.array_equal(x, y) for x, y in zip(m0, m1)), (rep, i, "mips")
                assert all(np.array_equal(x, y) for x, y in zip(o0, o1)), (rep, i, "blocks")
        b.close()
    finally:
        for k, v in saved.items():
            zt.set_option(k, v)
