"""Integer identities the kernels use in place of a division, checked over every input they can see."""


def test_bc4_division_by_seven():
    """zt_bc14.hip bc4_palette8: the BC4 interpolant numerators (8 - i) a0 + (i - 1) a1 + 3 with a0, a1 <= 255
    are at most 1788; (n * 9363) >> 16 equals n // 7 for every one of them (and up to 13109)."""
    assert max((8 - i) * 255 + (i - 1) * 255 + 3 for i in range(2, 8)) == 1788
    assert all((n * 9363) >> 16 == n // 7 for n in range(13110))
    assert (13110 * 9363) >> 16 != 13110 // 7


def test_srgb_bucket_encode_equals_threshold_search(oracle):
    """zt_mip.hip srgb_encode: the reference's sRGB encode of a linear value f (the first k with f < thr[k]) as
    bucket[floor(f * 4096)] + (thr[c] <= f), bucket[b] = the thresholds below b / 4096 and thr[255] = +inf.  Exact
    when no 1/4096-wide bucket holds two thresholds; checked on every threshold, its float neighbours, every bucket
    edge and 200,000 random values in [0, 1]."""
    import numpy as np
    _lin, thr = oracle.srgb_tables()
    thr = np.asarray(thr, dtype=np.float32)
    idx = np.floor(thr.astype(np.float32) * np.float32(4096)).astype(np.int64)
    per = np.bincount(idx, minlength=4097)
    assert per.max() == 1
    bucket = np.concatenate([[0], np.cumsum(per)])[:4097].astype(np.int64)
    thr_pad = np.concatenate([thr, np.array([np.inf], dtype=np.float32)])
    rng = np.random.default_rng(5)
    f = np.concatenate([thr, np.nextafter(thr, np.float32(0)), np.nextafter(thr, np.float32(2)),
                        np.arange(4097, dtype=np.float32) / np.float32(4096), rng.random(200000, dtype=np.float32),
                        np.array([0.0, 1.0], dtype=np.float32)]).astype(np.float32)
    f = f[(f >= 0) & (f <= 1)]
    b = np.floor(f * np.float32(4096)).astype(np.int64)
    c = bucket[b]
    got = c + (thr_pad[c] <= f)
    want = np.searchsorted(thr, f, side="right")   # the number of thresholds <= f = the first k with f < thr[k]
    assert np.array_equal(got, want)
