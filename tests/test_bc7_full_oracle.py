"""Pins the exhaustive all-mode BC7 search (project-defined, no reference counterpart) on the CPU:
every candidate's block decodes, through the reference-pinned decoder, to exactly the SSE it reports;
modes 6 / 1 / 5 are the reference's own encoders byte for byte; and the search never does worse than the
reference's mode-6/5/1 search on any block."""
import numpy as np

from oracle import loader
from zap_amd import synth

VALID = [c for c in range(512) if (c >> 6) not in (0, 4, 5, 6) or (c & 63) < {0: 16, 4: 8, 5: 4, 6: 1}[c >> 6]]


def test_candidate_numbering():
    assert len(VALID) == 285


def test_every_candidate_decodes_to_its_sse(oracle):
    blocks = synth.random_blocks(24, 31)
    for b, blk in enumerate(blocks):
        for c in VALID[b % 3::3]:
            e, out = oracle.bc7_candidate(blk, c)
            dec = oracle.bc_decode(out, 4, 4, "bc7").reshape(16, 4).astype(np.int64)
            assert e == int(((dec - blk.astype(np.int64)) ** 2).sum()), (b, c)
            mode = 0
            while not (out[0] >> mode) & 1:
                mode += 1
            assert mode == c >> 6


def test_reference_encoders_are_candidates(oracle):
    blocks = synth.random_blocks(64, 32)
    for blk in blocks:
        img = synth.blocks_to_image(blk[None])
        opaque = blk.copy(); opaque[:, 3] = 255
        ref_block = oracle.bc_encode(img, "bc7")
        b0 = int(ref_block[0])
        e, c, out = oracle.bc7_full_block(blk)
        assert e <= int(((oracle.bc_decode(ref_block, 4, 4, "bc7").reshape(16, 4).astype(np.int64) - blk) ** 2).sum())
        # mode 6 candidate == the reference's mode-6 block whenever the reference keeps mode 6
        if (b0 & 0x7F) == 0x40:
            assert np.array_equal(oracle.bc7_candidate(blk, 6 * 64)[1], ref_block)
        if (b0 & 3) == 2:  # reference chose mode 1: that partition's candidate is the same block
            part = (b0 >> 2) & 63
            assert np.array_equal(oracle.bc7_candidate(blk, 64 + part)[1], ref_block)
        if (b0 & 0x3F) == 0x20:  # mode 5 with its rotation
            rot = b0 >> 6
            assert np.array_equal(oracle.bc7_candidate(blk, 5 * 64 + rot)[1], ref_block)


def test_full_search_never_worse_than_reference_search(oracle):
    img = synth.c2_bc7(32, 77)
    full = oracle.bc7_full_encode(img)
    ref = oracle.bc_encode(img, "bc7")
    dfull = oracle.bc_decode(full, 32, 32, "bc7").astype(np.int64)
    dref = oracle.bc_decode(ref, 32, 32, "bc7").astype(np.int64)
    src = img.astype(np.int64)
    for by in range(8):
        for bx in range(8):
            sl = (slice(by * 4, by * 4 + 4), slice(bx * 4, bx * 4 + 4))
            assert ((dfull[sl] - src[sl]) ** 2).sum() <= ((dref[sl] - src[sl]) ** 2).sum()
    assert ((dfull - src) ** 2).sum() < ((dref - src) ** 2).sum()
