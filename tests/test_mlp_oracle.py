"""The network oracle's analytic gradients equal central finite differences of its own loss (exact arithmetic),
so a dropped term, wrong sign or transposed operand in the back-propagation fails here."""
import numpy as np

from oracle import mlp_ref


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = (rng.standard_normal(mlp_ref.NPARAM) * 0.05).astype(np.float64)
    blocks = rng.integers(0, 256, (6, 64), dtype=np.uint8)
    labels = rng.choice(np.nonzero(mlp_ref.VALID)[0], 6)
    _, g = mlp_ref.loss_and_grads(p, blocks, labels, bf16_points=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 12, replace=False) for o, s in mlp_ref.OFF.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (mlp_ref.loss_and_grads(pp, blocks, labels, False)[0] - mlp_ref.loss_and_grads(pm, blocks, labels, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e-3], np.float32)
    b = mlp_ref.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == np.float32(1.015625) and b[3] == -2.5
    assert abs(b[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_soft_target_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = (rng.standard_normal(mlp_ref.NPARAM) * 0.05).astype(np.float64)
    blocks = rng.integers(0, 256, (5, 64), dtype=np.uint8)
    deltas = rng.integers(0, 400, (5, 512)).astype(np.uint16)
    deltas[:, ~mlp_ref.VALID] = 65535
    deltas[np.arange(5), rng.choice(np.nonzero(mlp_ref.VALID)[0], 5)] = 0
    deltas[0, 64:128] = 65535                       # saturated candidates take no target mass
    _, g = mlp_ref.loss_and_grads(p, blocks, None, False, deltas, 40.0)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 10, replace=False) for o, s in mlp_ref.OFF.values()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (mlp_ref.loss_and_grads(pp, blocks, None, False, deltas, 40.0)[0]
              - mlp_ref.loss_and_grads(pm, blocks, None, False, deltas, 40.0)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_soft_targets_limit_to_one_hot():
    # a small temperature puts all mass on the delta-0 slot, a large one spreads it evenly over the used slots
    d = np.full((1, 512), 300, np.uint16)
    d[0, ~mlp_ref.VALID] = 65535
    d[0, 70] = 0
    d[0, 71] = 65535
    t = mlp_ref.targets(1, deltas=d, T=1.0)
    assert t[0, 70] == 1.0 and t.sum() == 1.0
    t = mlp_ref.targets(1, deltas=d, T=1e9)
    assert np.allclose(t[0, mlp_ref.VALID & (np.arange(512) != 71)], 1.0 / 284) and t[0, 71] == 0.0


def test_risk_head_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    p = (rng.standard_normal(mlp_ref.NPARAM) * 0.05).astype(np.float64)
    blocks = rng.integers(0, 256, (6, 64), dtype=np.uint8)
    labels = rng.choice(np.nonzero(mlp_ref.VALID)[0], 6)
    deltas = rng.integers(0, 3000, (6, 512)).astype(np.uint16)
    deltas[:, ~mlp_ref.VALID] = 65535
    _, g = mlp_ref.loss_and_grads(p, blocks, labels, False, deltas, None, 4, 0.3)
    o3 = mlp_ref.OFF["W3"][0]
    idx = list(o3 + np.arange(128) * 512 + mlp_ref.RISK)[:16] + [mlp_ref.OFF["b3"][0] + mlp_ref.RISK]
    idx += list(rng.choice(np.arange(0, mlp_ref.OFF["W2"][0]), 12, replace=False))
    h = 1e-6
    f = lambda q: mlp_ref.loss_and_grads(q, blocks, labels, False, deltas, None, 4, 0.3)[0]
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h; pm[i] -= h
        fd = (f(pp) - f(pm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    assert abs(g[mlp_ref.OFF["b3"][0] + mlp_ref.RISK]) > 1e-3   # the risk term is live


def test_topk_order_and_ties():
    z = np.zeros((1, 512))
    z[0, 70] = 3.0; z[0, 20] = 5.0   # slot 20 (mode 0, partition 20) is unused
    z[0, 100] = 2.0; z[0, 99] = 2.0; z[0, 385] = 9.0
    assert list(mlp_ref.topk(z, 4)[0]) == [70, 99, 100, 0]


def test_propose_topk_keys():
    """The fused kernel's ranking (mlp_ref.propose_topk) is the float order of the logits wherever they differ
    above their low 9 bits, and the lowest slot first inside such a tie, for either sign."""
    z = np.zeros((1, 512))
    z[0, 70] = 3.0; z[0, 20] = 5.0; z[0, 100] = 2.0; z[0, 99] = 2.0; z[0, 385] = 9.0
    assert list(mlp_ref.propose_topk(z, 4)[0]) == [70, 99, 100, 0]           # unused slots 20 and 385 never rank
    rng = np.random.default_rng(5)
    used = np.nonzero(mlp_ref.VALID)[0]
    zs = np.zeros((64, 512), np.float32)
    zs[:, used] = rng.permutation(np.linspace(-20, 20, 64 * used.size, dtype=np.float32)).reshape(64, used.size)
    assert np.array_equal(mlp_ref.propose_topk(zs, 16), mlp_ref.topk(zs, 16))  # spacing 1.4e-4: far above the 9 bits
    near = np.zeros((1, 512), np.float32)
    for lo, hi, bits in [(used[5], used[9], 0x3FC00100), (used[40], used[41], 0xBFC00100)]:   # +-1.50003
        near[:] = -50.0
        a = np.array([bits], np.uint32).view(np.float32)[0]
        b = np.nextafter(a, np.float32(np.inf), dtype=np.float32)             # larger by one ulp, same 9-bit bucket
        near[0, lo], near[0, hi] = a, b
        assert list(mlp_ref.propose_topk(near, 2)[0]) == [lo, hi]            # tie: lowest slot first
        near[0, hi] = a + np.float32(1e-3)                                    # above the 9 bits: the larger first
        assert list(mlp_ref.propose_topk(near, 2)[0]) == [hi, lo]


def test_feature_normalisations():
    rng = np.random.default_rng(3)
    blocks = rng.integers(0, 256, (40, 64), dtype=np.uint8)
    blocks[0] = 77                                   # flat block: R = 0 -> divided by 1, all features 0
    raw = mlp_ref.features(blocks, False, 0)
    assert np.allclose(raw, blocks / 255.0, atol=1e-7)
    for mode in (1, 2, 3):
        x = mlp_ref.features(blocks, False, mode).reshape(-1, 16, 4)
        assert np.allclose(x.mean(1), 0.0, atol=1e-6)   # centred per channel
        assert (x[0] == 0).all()
    x2 = mlp_ref.features(blocks, False, 2).reshape(-1, 16, 4)
    b = blocks.reshape(-1, 16, 4).astype(np.int64)
    r = (b.max(1) - b.min(1)).max(1)
    spread = (x2.max(1) - x2.min(1)).max(1)
    assert np.allclose(spread[1:], 1.0, atol=1e-6) and np.all(r[1:] > 0)   # the widest channel spans exactly 1
    x3 = mlp_ref.features(blocks, False, 3).reshape(-1, 16, 4)
    assert np.allclose((x3.max(1) - x3.min(1)).max(1)[1:], np.sqrt(r[1:] / 64.0), rtol=1e-6)
    # mode 4: colour as mode 3 over the colour range; alpha (a - 255) / 255, so opaque pixels give exactly 0
    x4 = mlp_ref.features(blocks, False, 4).reshape(-1, 16, 4)
    rc = np.maximum(1, (b.max(1) - b.min(1))[:, :3].max(1))
    assert np.allclose(x4[:, :, 3], (b[:, :, 3] - 255) / 255.0, atol=1e-7)
    assert np.allclose((x4[:, :, :3].max(1) - x4[:, :, :3].min(1)).max(1)[1:], np.sqrt(rc[1:] / 64.0), rtol=1e-6)
    opaque = blocks.copy().reshape(-1, 16, 4)
    opaque[:, :, 3] = 255
    assert (mlp_ref.features(opaque, False, 4).reshape(-1, 16, 4)[:, :, 3] == 0).all()
