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
