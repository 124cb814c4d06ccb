"""GPU exhaustive all-mode BC7 search: byte-identical to the oracle's exhaustive search (same candidates,
same tie rule), and never worse than the reference's mode-6/5/1 search on any block."""
import numpy as np
import pytest

from zap_amd import binding, synth

pytestmark = pytest.mark.gpu


def _full_gpu(zt, img):
    import torch
    h, w = img.shape[:2]
    src = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    dst = torch.zeros(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda")
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    cand = torch.full((nb,), -1, dtype=torch.int32, device="cuda")
    zt.bc7_full_dev(zt.images([(src, dst, w, h, w * 4, "bc7")]), cand)
    torch.cuda.synchronize()
    return dst.cpu().numpy(), cand.cpu().numpy()


def test_full_search_matches_oracle(zt, oracle):
    blocks = synth.random_blocks(512, 13)
    img = synth.blocks_to_image(blocks)
    got, cand = _full_gpu(zt, img)
    exp = oracle.bc7_full_encode(img)
    assert np.array_equal(got, exp)
    for b in range(0, 512, 37):
        assert oracle.bc7_full_block(blocks[b])[1] == cand[b]


def test_full_search_texture_and_odd_size(zt, oracle):
    img = np.ascontiguousarray(synth.c2_bc7(64, 21)[:61, :58])
    got, _ = _full_gpu(zt, img)
    assert np.array_equal(got, oracle.bc7_full_encode(img))


def test_full_search_never_worse_than_reference_search(zt, oracle):
    img = synth.c2_bc7(256, 2)
    full, _ = _full_gpu(zt, img)
    ref = zt.bc_encode(img, "bc7")
    src = img.astype(np.int64)
    df = (oracle.bc_decode(full, 256, 256, "bc7").astype(np.int64) - src) ** 2
    dr = (oracle.bc_decode(ref, 256, 256, "bc7").astype(np.int64) - src) ** 2
    bf = df.reshape(64, 4, 64, 4, 4).sum(axis=(1, 3, 4))
    br = dr.reshape(64, 4, 64, 4, 4).sum(axis=(1, 3, 4))
    assert (bf <= br).all() and bf.sum() < br.sum()


def test_full_search_candidate_deltas(zt, oracle):
    """Every used slot's delta is min(65535, its SSE - the block's best SSE) by the oracle's encoding of that
    candidate; unused slots hold 65535; the chosen candidate has delta 0."""
    import torch

    from oracle import mlp_ref
    blocks = synth.random_blocks(64, 29)
    img = synth.blocks_to_image(blocks)
    h, w = img.shape[:2]
    src = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    dst = torch.zeros(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda")
    cand = torch.full((64,), -1, dtype=torch.int32, device="cuda")
    deltas = torch.zeros((64, 512), dtype=torch.int16, device="cuda")
    zt.bc7_full_dev(zt.images([(src, dst, w, h, w * 4, "bc7")]), cand, deltas)
    torch.cuda.synchronize()
    d = deltas.cpu().numpy().view(np.uint16).astype(np.int64)
    c = cand.cpu().numpy()
    assert (d[:, ~mlp_ref.VALID] == 65535).all()
    assert (d[np.arange(64), c] == 0).all()
    valid = np.nonzero(mlp_ref.VALID)[0]
    for b in range(0, 64, 7):
        errs = np.array([oracle.bc7_candidate(blocks[b], int(s))[0] for s in valid], np.int64)
        best = errs[errs >= 0].min()
        exp = np.where(errs < 0, 65535, np.minimum(65535, errs - best))
        assert np.array_equal(d[b, valid], exp), b
