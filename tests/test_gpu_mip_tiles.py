"""The fused mip passes with either level-1 tile side (ZAPTEX_OPT_MIP_TILE = 32: six levels a pass, the default;
64: seven levels a pass, a quarter of the workgroups) give the reference's bytes: the 4096^2 chain against the
reference's hashes in both filters, odd and thin chains against the oracle's zap_mip_next chain, and a batch of odd
textures against the oracle."""
import hashlib

import numpy as np
import pytest

from zap_amd import batch, binding, synth

pytestmark = pytest.mark.gpu
OPT_MIP_TILE = 10


@pytest.fixture(params=[32, 64])
def tile(zt, request):
    saved = zt.get_option(OPT_MIP_TILE)
    zt.set_option(OPT_MIP_TILE, request.param)
    yield request.param
    zt.set_option(OPT_MIP_TILE, saved)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _chain(zt, img, srgb):
    import torch
    h, w = img.shape[:2]
    base = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    levels = [torch.empty((hh, ww, 4), dtype=torch.uint8, device="cuda") for (ww, hh) in binding.mip_sizes(w, h)[1:]]
    zt.mip_chain_dev(base, w, h, w * 4, levels, srgb=srgb)
    torch.cuda.synchronize()
    return [l.cpu().numpy() for l in levels]


def test_tiles_c2_chain_matches_reference_hashes(zt, golden, tile):
    c2 = synth.c2_bc7(4096, 2)
    for srgb in (0, 1):
        got = [sha(c2)] + [sha(l) for l in _chain(zt, c2, srgb)]
        assert got == golden["outputs"][f"c2_mips_srgb{srgb}"], srgb


@pytest.mark.parametrize("w,h", [(300, 200), (1000, 750), (129, 67), (5, 260), (2048, 64), (4096, 33), (257, 1030)])
def test_tiles_odd_chains_match_oracle(zt, oracle, tile, w, h):
    img = np.ascontiguousarray(synth.c2_bc7(max(w, h), w + h)[:h, :w])
    for srgb in (0, 1):
        exp = oracle.mip_chain(img, srgb)[1:]
        got = _chain(zt, img, srgb)
        assert len(got) == len(exp)
        for lv, (a, b) in enumerate(zip(got, exp)):
            assert np.array_equal(a, b), (srgb, lv + 1)


def test_tiles_batch_matches_oracle(zt, oracle, tile):
    import torch
    specs = [(300, 200, "bc7", 1), (129, 67, "bc3", 0), (5, 260, "bc1", 1), (1000, 750, "bc5", 0), (64, 64, "bc4", 1), (2048, 96, "bc7", 0)]
    cpu = [batch.texture_cpu(w, h, "alpha", 300 + k) for k, (w, h, f, s) in enumerate(specs)]
    b = batch.Batch(zt, [(torch.from_numpy(cpu[k]).cuda(), w, h, f, s, None) for k, (w, h, f, s) in enumerate(specs)])
    b.run()
    torch.cuda.synchronize()
    for i, (w, h, fmt, srgb) in enumerate(specs):
        levels = b.items[i]["levels"]
        chain = oracle.mip_chain(cpu[i], srgb)[:levels]
        for lv in range(1, levels):
            assert np.array_equal(b.level(i, lv).cpu().numpy(), chain[lv]), (i, lv)
        for lv in range(levels):
            assert np.array_equal(b.blocks_of(i, lv).cpu().numpy(), oracle.bc_encode(chain[lv], fmt)), (i, lv, fmt)
    b.close()
