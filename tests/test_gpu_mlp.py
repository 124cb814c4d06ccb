"""The proposal network's own GPU kernels against the CPU oracle (oracle/mlp_ref.py): forward logits,
gradients of every parameter tensor, a loss that falls under training, and the learned BC7 encode's
contract (all-fallback = the exhaustive search; otherwise every block is the exact encoding of one of the
proposed candidates, never better than the exhaustive optimum)."""
import numpy as np
import pytest

from oracle import mlp_ref
from zap_amd import binding, synth

pytestmark = pytest.mark.gpu


def _setup(zt, B, seed, feat=0):
    import torch
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(mlp_ref.NPARAM) * 0.08).astype(np.float32)
    blocks = synth.random_blocks(B, seed).reshape(B, 64)
    labels = rng.choice(np.nonzero(mlp_ref.VALID)[0], B).astype(np.int32)
    st = zt.mlp_state(params=params, features=feat)
    tb = torch.from_numpy(np.ascontiguousarray(blocks)).cuda()
    tl = torch.from_numpy(labels).cuda()
    return params, blocks, labels, st, tb, tl


@pytest.mark.parametrize("feat", [0, 1, 2, 3, 4])
def test_forward_matches_oracle(zt, feat):
    import torch
    B = 300
    params, blocks, labels, st, tb, tl = _setup(zt, B, 1, feat)
    ws = zt.mlp_workspace(B)
    z = torch.empty((B, 512), dtype=torch.float32, device="cuda")
    zt.mlp_forward(st, tb, B, z, ws)
    torch.cuda.synchronize()
    got = z.cpu().numpy().astype(np.float64)
    exp = mlp_ref.forward(params, blocks, feat=feat)[3]
    err = np.abs(got - exp).max() / np.abs(exp).max()
    assert err < 2e-2, err


@pytest.mark.parametrize("feat,topk,risk_shift,B", [(0, 4, 0.0, 1000), (4, 1, 30.0, 37), (3, 16, -30.0, 517), (2, 7, 0.0, 5)])
def test_propose_matches_forward_topk(zt, feat, topk, risk_shift, B):
    """The fused proposal kernel equals the three-GEMM forward followed by a top-k over the used slots: same
    candidates (ranked by mlp_ref.propose_topk's keys: largest logit first, logits equal above their low 9 bits by
    ascending slot) and risk = expm1 of the risk logit, clamped."""
    import torch
    params, blocks, labels, st, tb, tl = _setup(zt, B, 11 + topk, feat)
    params[mlp_ref.OFF["b3"][0] + mlp_ref.RISK] += risk_shift   # push the risk logit into either clamp
    st = zt.mlp_state(params=params, features=feat)
    ws = zt.mlp_workspace(B)
    z = torch.empty((B, 512), dtype=torch.float32, device="cuda")
    zt.mlp_forward(st, tb, B, z, ws)
    cand = torch.full((B, topk), -1, dtype=torch.int32, device="cuda")
    risk = torch.full((B,), -1.0, dtype=torch.float32, device="cuda")
    zt.mlp_propose(st, tb, B, topk, cand, risk)
    torch.cuda.synchronize()
    lg = torch.full((B, 288), float("nan"), dtype=torch.float32, device="cuda")
    zt.mlp_propose_logits(st, tb, B, lg)
    torch.cuda.synchronize()
    zn = z.cpu().numpy()
    slots = np.nonzero(mlp_ref.VALID)[0]
    # the fused kernel's logits: bit-identical to the GEMM path on the kept rows (candidate slots, then the risk slot)
    kept = zn[:, np.concatenate([slots, [mlp_ref.RISK]])]
    lgn = lg.cpu().numpy()
    assert np.array_equal(lgn[:, :286].view(np.uint32), kept.view(np.uint32))
    assert (lgn[:, 286:] == 0).all()
    assert np.array_equal(cand.cpu().numpy(), mlp_ref.propose_topk(zn, topk))
    r = zn[:, mlp_ref.RISK].astype(np.float64)
    exp = np.where(r <= 0, 0.0, np.where(r >= 13.9, 1048576.0, np.expm1(r)))
    got = risk.cpu().numpy().astype(np.float64)
    assert np.allclose(got, exp, rtol=1e-5, atol=2e-6), np.abs(got - exp).max()
    if risk_shift > 0:
        assert (got == 1048576.0).all()
    if risk_shift < 0:
        assert (got == 0.0).all()


def test_gradients_match_oracle(zt):
    import torch
    B = 512
    params, blocks, labels, st, tb, tl = _setup(zt, B, 2)
    ws = zt.mlp_workspace(B)
    stats = torch.zeros(2, dtype=torch.float32, device="cuda")
    zt.mlp_gradients(st, tb, tl, B, ws, stats)
    torch.cuda.synchronize()
    g = zt.mlp_grads(st).astype(np.float64)
    loss, ge = mlp_ref.loss_and_grads(params, blocks, labels)
    assert abs(float(stats[0].item()) - loss) < 2e-2 * abs(loss)
    for k, (o, s) in mlp_ref.OFF.items():
        n = int(np.prod(s))
        a, b = g[o:o + n], ge[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 5e-2, (k, rel)


def test_soft_target_gradients_match_oracle(zt):
    import torch
    B = 384
    params, blocks, labels, st, tb, tl = _setup(zt, B, 4)
    rng = np.random.default_rng(4)
    deltas = rng.integers(0, 600, (B, 512)).astype(np.uint16)
    deltas[:, ~mlp_ref.VALID] = 65535
    deltas[np.arange(B), labels] = 0
    deltas[:7, 128:192] = 65535
    td = torch.from_numpy(deltas.view(np.int16)).cuda()
    ws = zt.mlp_workspace(B)
    stats = torch.zeros(2, dtype=torch.float32, device="cuda")
    zt.mlp_gradients(st, tb, None, B, ws, stats, deltas=td, temperature=50.0)
    torch.cuda.synchronize()
    g = zt.mlp_grads(st).astype(np.float64)
    loss, ge = mlp_ref.loss_and_grads(params, blocks, None, deltas=deltas, T=50.0)
    assert abs(float(stats[0].item()) - loss) < 2e-2 * abs(loss)
    for k, (o, s) in mlp_ref.OFF.items():
        n = int(np.prod(s))
        a, b = g[o:o + n], ge[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 5e-2, (k, rel)


def test_risk_head_gradients_match_oracle(zt):
    import torch
    B = 512
    params, blocks, labels, st, tb, tl = _setup(zt, B, 6, feat=4)
    rng = np.random.default_rng(6)
    deltas = rng.integers(0, 4000, (B, 512)).astype(np.uint16)
    deltas[:, ~mlp_ref.VALID] = 65535
    td = torch.from_numpy(deltas.view(np.int16)).cuda()
    ws = zt.mlp_workspace(B)
    stats = torch.zeros(2, dtype=torch.float32, device="cuda")
    zt.mlp_gradients(st, tb, tl, B, ws, stats, deltas=td, risk_k=4, risk_weight=0.3)
    torch.cuda.synchronize()
    g = zt.mlp_grads(st).astype(np.float64)
    loss, ge = mlp_ref.loss_and_grads(params, blocks, labels, deltas=deltas, risk_k=4, risk_w=0.3, feat=4)
    assert abs(float(stats[0].item()) - loss) < 2e-2 * abs(loss)
    for k, (o, s) in mlp_ref.OFF.items():
        n = int(np.prod(s))
        a, b = g[o:o + n], ge[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 5e-2, (k, rel)
    rb = mlp_ref.OFF["b3"][0] + mlp_ref.RISK
    assert abs(g[rb] - ge[rb]) < 2e-2 * abs(ge[rb])


def test_training_reduces_loss(zt):
    import torch
    B = 1024
    params, blocks, labels, st, tb, tl = _setup(zt, B, 3)
    labels = (np.arange(B) % 7 * 64).astype(np.int32)  # learnable: 7 classes
    labels[labels == 256] = 320
    tl = torch.from_numpy(labels).cuda()
    st = zt.mlp_state(seed=5)
    ws = zt.mlp_workspace(B)
    losses = []
    for step in range(1, 301):
        stats = torch.zeros(2, dtype=torch.float32, device="cuda")
        zt.mlp_train_step(st, tb, tl, B, 3e-3, step, ws, stats)
        if step in (1, 300):
            losses.append(float(stats[0].item()))
    assert losses[1] < 0.5 * losses[0], losses


def _decode_sse(oracle, bc, img):
    h, w = img.shape[:2]
    d = oracle.bc_decode(bc, w, h, "bc7").astype(np.int64) - img.astype(np.int64)
    return (d * d).reshape(h // 4, 4, w // 4, 4, 4).sum(axis=(1, 3, 4)).reshape(-1)


def _cand_of(block: np.ndarray) -> int:
    b0 = int(block[0])
    mode = (b0 & -b0).bit_length() - 1
    bits = int.from_bytes(block.tobytes(), "little")
    if mode in (0, 1, 2, 3, 7):
        return mode * 64 + ((bits >> (mode + 1)) & (15 if mode == 0 else 63))
    if mode == 4:
        return 256 + ((bits >> 5) & 3) * 2 + ((bits >> 7) & 1)
    if mode == 5:
        return 320 + ((bits >> 6) & 3)
    return 384


def test_learned_encode_contract(zt, oracle):
    import torch
    img = synth.c2_bc7(64, 33)
    src = torch.from_numpy(img).cuda()
    outs = []
    full = torch.zeros(binding.bc_size(64, 64, "bc7"), dtype=torch.uint8, device="cuda")
    imgs_full = zt.images([(src, full, 64, 64, 256, "bc7")])
    zt.bc7_full_dev(imgs_full)
    st = zt.mlp_state(seed=9)
    # (threshold, fallback_k): every block falls back to the exhaustive search; none falls back; some do; every
    # block falls back to the network's top 285 (all candidates: the exhaustive result again), top 16, top 16 for some
    runs = [(-1.0, 0), (1e30, 0), (1.0, 0), (-1.0, 285), (-1.0, 16), (1.0, 16)]
    for thr, fk in runs:
        dst = torch.zeros_like(full)
        imgs = zt.images([(src, dst, 64, 64, 256, "bc7")])
        ws = zt.bc7_learned_workspace(imgs, 4, fk)
        zt.bc7_learned_dev(imgs, st, 4, thr, ws, fallback_k=fk)
        torch.cuda.synchronize()
        outs.append(dst.cpu().numpy())
    f = full.cpu().numpy()
    assert np.array_equal(outs[0], f)                      # everything fell back to the exhaustive search
    assert np.array_equal(outs[3], f)                      # the wide fallback over all 285 candidates is that search
    sse_f = _decode_sse(oracle, f, img)
    sse = [_decode_sse(oracle, o, img) for o in outs]
    assert (sse[4] <= sse[1]).all()                        # top 16 contain the top 4
    assert (sse[5] <= sse[1]).all()
    for learned, sse_l in zip(outs[1:], sse[1:]):
        assert (sse_l >= sse_f).all()
        for b in range(0, 256, 5):
            by, bx = divmod(b, 16)
            blk = img[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4].reshape(16, 4)
            c = _cand_of(learned[16 * b:16 * b + 16])
            e, exp = oracle.bc7_candidate(blk, c)
            assert np.array_equal(exp, learned[16 * b:16 * b + 16]) and e == sse_l[b], (b, c)
