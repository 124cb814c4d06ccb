"""The oracle reproduces the committed golden hashes (made from the compiled reference) on the workloads."""
import hashlib

import numpy as np
import pytest

from zap_amd import synth


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_inputs_are_the_fixture_inputs(golden):
    assert sha(synth.c1_bc1(1024, 1)) == golden["inputs"]["c1_1024"]
    assert sha(synth.blocks_to_image(synth.random_blocks(8192, 7))) == golden["inputs"]["blocks_8192"]


def test_small_fixtures(oracle, golden):
    for key, ent in golden["small"].items():
        w, h = map(int, key.split("x"))
        img = np.frombuffer(bytes.fromhex(ent["input"]), np.uint8).reshape(h, w, 4)
        for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
            assert oracle.bc_encode(img, f).tobytes().hex() == ent[f], (key, f)
        assert oracle.astc_encode(img).tobytes().hex() == ent["astc"], key
        hf = np.frombuffer(bytes.fromhex(ent["hdr_input"]), np.float32).reshape(h, w, 4)
        assert oracle.bc6h_encode(hf).tobytes().hex() == ent["bc6h"], key
        assert oracle.mip_next(img, 0).tobytes().hex() == ent["mip"]
        assert oracle.mip_next(img, 1).tobytes().hex() == ent["mip_srgb"]


def test_blocks_and_c1(oracle, golden):
    blocks = synth.blocks_to_image(synth.random_blocks(8192, 7))
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        assert sha(oracle.bc_encode(blocks, f)) == golden["outputs"][f"blocks_8192_{f}"], f
    assert sha(oracle.astc_encode(blocks)) == golden["outputs"]["blocks_8192_astc"]
    c1 = synth.c1_bc1(1024, 1)
    for f in ["bc1", "bc4"]:
        assert sha(oracle.bc_encode(c1, f)) == golden["outputs"][f"c1_1024_{f}"], f


def test_c2_mips_and_bc7_lower_levels(oracle, golden):
    c2 = synth.c2_bc7(4096, 2)
    assert sha(c2) == golden["inputs"]["c2_4096"]
    chain = oracle.mip_chain(c2, 0)
    assert [sha(l) for l in chain] == golden["outputs"]["c2_mips_srgb0"]
    # BC7 of levels 2.. (1024^2 and below) keeps the CPU suite short; the GPU suite checks all 13 levels
    for lv in range(2, len(chain)):
        assert sha(oracle.bc_encode(chain[lv], "bc7")) == golden["outputs"]["c2_bc7_chain"][lv], lv


def test_near_singular_fixture(oracle, golden):
    ent = golden["near_singular"]
    assert ent["blocks"] >= 32
    img = np.frombuffer(bytes.fromhex(ent["input"]), np.uint8).reshape(4, 4 * ent["blocks"], 4)
    for f in ("bc1", "bc3"):
        assert oracle.bc_encode(img, f).tobytes().hex() == ent[f], f
    assert oracle.astc_encode(img).tobytes().hex() == ent["astc"]


def test_c4_batch_hashes_of_the_smallest_textures(oracle):
    """The batch workload's golden file agrees with the oracle (generator + mip chain + encoder) on its 12
    smallest textures; the GPU test checks all 512."""
    import hashlib
    import json
    import os

    from zap_amd import batch
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c4_batch.json")))["textures"]
    specs = batch.c4_specs(512, 4)
    assert len(gold) == 512
    for i in sorted(range(512), key=lambda i: specs[i][0] * specs[i][1])[:12]:
        w, h, fmt, srgb, kind = specs[i]
        base = batch.texture_cpu(w, h, kind, 400000 + i)
        chain = oracle.mip_chain(base, srgb)
        assert sha(base) == gold[i]["base"], i
        assert hashlib.sha256(b"".join(l.tobytes() for l in chain[1:])).hexdigest() == gold[i]["mips"], i
        assert hashlib.sha256(b"".join(oracle.bc_encode(l, fmt).tobytes() for l in chain)).hexdigest() == gold[i]["blocks"], i
