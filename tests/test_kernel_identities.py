"""Integer identities the kernels use in place of a division, checked over every input they can see."""


def test_bc4_division_by_seven():
    """zt_bc14.hip bc4_palette8: the BC4 interpolant numerators (8 - i) a0 + (i - 1) a1 + 3 with a0, a1 <= 255
    are at most 1788; (n * 9363) >> 16 equals n // 7 for every one of them (and up to 13109)."""
    assert max((8 - i) * 255 + (i - 1) * 255 + 3 for i in range(2, 8)) == 1788
    assert all((n * 9363) >> 16 == n // 7 for n in range(13110))
    assert (13110 * 9363) >> 16 != 13110 // 7
