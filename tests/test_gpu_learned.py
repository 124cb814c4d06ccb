"""The committed proposal network (zap_amd/models/bc7_proposal.npy) on a texture it never saw: seed 77 (training
used 1000 and up, threshold calibration 500..502, the benchmark 2), at a side (1536) neither training nor
calibration used, so its noise scale (3-pixel finest cells) is new too.  At its calibrated threshold the learned encode
stays within 0.05 dB of the exhaustive all-mode search while most blocks skip that search (4 exact candidates
instead of 285); with no fallback at all, the trained network's top-4 lose far less PSNR than an untrained
network's top-4, which is what the training buys."""
import pytest

from zap_amd import learned, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(zt):
    return learned.load(zt)


@pytest.fixture(scope="module")
def chain(zt):
    return learned.Chain(zt, synth.c2_bc7(1536, 77, rects=96))


def test_learned_within_gap_at_calibrated_threshold(zt, model, chain):
    state, meta = model
    r = learned.evaluate(zt, chain, state, meta["topk"], [meta["threshold"]], steps=1, fallback_k=int(meta.get("fallback_k", 0)))
    lr = r["learned"][0]
    assert lr["gap_db"] <= 0.05, lr
    assert lr["fallback"] <= 0.35, lr                      # >= 65% of blocks: 4 candidates instead of 285
    assert lr["psnr"] > r["plain"]["psnr"] - 0.05           # close to or better than the reference search too


def test_trained_network_beats_untrained(zt, model, chain):
    state, meta = model
    trained = learned.evaluate(zt, chain, state, meta["topk"], [1e30], steps=1)["learned"][0]
    untrained_state = zt.mlp_state(seed=123, features=meta["features"])
    untrained = learned.evaluate(zt, chain, untrained_state, meta["topk"], [1e30], steps=1)["learned"][0]
    assert trained["fallback"] == 0.0 and untrained["fallback"] == 0.0
    assert trained["gap_db"] < 0.3, trained
    assert trained["gap_db"] < 0.25 * untrained["gap_db"], (trained, untrained)


@pytest.fixture(scope="module")
def chain4096(zt):
    return learned.Chain(zt, synth.c2_bc7(4096, 2))


@pytest.mark.parametrize("fallback_k", [64, 0])
def test_learned_gap_on_the_4096_benchmark_chain(zt, model, chain4096, fallback_k):
    """The learned-proposal workload itself: the 4096^2 seed-2 texture and its 13-level chain (never trained or
    calibrated on), at the model's calibrated threshold, with the shipped fallback (the network's top 64) and with
    the exhaustive search as the fallback; the chain's PSNR stays within 0.05 dB of the exhaustive search's."""
    state, meta = model
    r = learned.evaluate(zt, chain4096, state, meta["topk"], [meta["threshold"]], steps=1, fallback_k=fallback_k)
    lr = r["learned"][0]
    assert lr["gap_db"] <= 0.05, lr
    assert 0 < lr["fallback"] <= 0.05, lr                  # >= 95% of blocks: 4 exact candidates instead of 285
    assert lr["psnr"] > r["plain"]["psnr"] + 0.5, (lr, r["plain"])   # far above the reference's own search
