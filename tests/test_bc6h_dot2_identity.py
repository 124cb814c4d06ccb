"""The BC6H fit's squared errors (zap_amd/csrc/zt_bc6h.hip `fit`): for half values t, p <= 0x7BFF (what
`half_of` and `finish` produce), |t|^2 + |p|^2 - t . 2p formed with two v_dot2_u32_u16 (a.lo * b.lo + a.hi * b.hi + c,
modulo 2^32) equals the exact sum of squared differences that the reference compares in double
(zap_tex.h:854-871).  Checked at the extremes of the range and on random samples, with the channel packing the
kernel uses (t01 = t0 | t1 << 16, t2; palette doubled: 2 p0 | 2 p1 << 16, 2 p2)."""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)
HMAX = 0x7BFF


def dot2(a, b, c):
    a = a.astype(np.uint64)
    b = b.astype(np.uint64)
    return (((a & np.uint64(0xFFFF)) * (b & np.uint64(0xFFFF)) + (a >> np.uint64(16)) * (b >> np.uint64(16)) + c.astype(np.uint64)) & M32)


def kernel_sse(t, p):
    t01 = (t[:, 0] | (t[:, 1] << 16)).astype(np.uint64)
    t2 = t[:, 2].astype(np.uint64)
    p01 = ((2 * p[:, 0]) | ((2 * p[:, 1]) << 16)).astype(np.uint64)
    p2 = (2 * p[:, 2]).astype(np.uint64)
    pp = ((p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1] + p[:, 2] * p[:, 2]).astype(np.uint64)) & M32
    zero = np.zeros_like(t2)
    tt = dot2(t01, t01, dot2(t2, t2, zero))
    x = dot2(t01, p01, dot2(t2, p2, zero))
    return (tt + pp - x) & M32


def test_dot2_sse_is_exact():
    rng = np.random.default_rng(6)
    edge = np.array([0, 1, 2, 0x3FFF, 0x4000, HMAX - 1, HMAX], dtype=np.int64)
    grid = np.stack(np.meshgrid(*[edge] * 6, indexing="ij"), -1).reshape(-1, 6)   # every extreme combination
    rnd = rng.integers(0, HMAX + 1, size=(2_000_000, 6), dtype=np.int64)
    for v in (grid, rnd):
        t, p = v[:, :3], v[:, 3:]
        exact = ((t - p) ** 2).sum(axis=1)
        assert exact.max() < 2 ** 32
        assert np.array_equal(kernel_sse(t, p).astype(np.int64), exact)
    # the cross term t . 2p really wraps at the top of the range (so the test exercises the modular argument)
    t = np.full(1, HMAX, dtype=np.uint64)
    t01, t2 = t | (t << np.uint64(16)), t
    x = dot2(t01, np.uint64(2) * t01, dot2(t2, np.uint64(2) * t2, np.zeros(1, dtype=np.uint64)))
    assert int(x[0]) != 6 * HMAX * HMAX and 6 * HMAX * HMAX >= 2 ** 32
