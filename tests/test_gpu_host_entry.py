"""Host entry points as the reference's callers use them (zap_tex.h:1474-1506, :1442): images with a side beyond
65,536 pixels (the reference takes any size), and many threads calling at once, as the reference viewer's banded
encoder does (viewer.cpp:455-471).  Each host call leases its own stream and staging buffers, up to 8 per device;
the outputs must equal the oracle's and the calls made one at a time."""
import threading

import numpy as np
import pytest

from zap_amd import synth

pytestmark = pytest.mark.gpu


def strip(w, h, seed):
    """A gradient with noise and a few hard edges, any shape."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.empty((h, w, 4), np.uint8)
    img[..., 0] = (x * 7 + y * 3) & 255
    img[..., 1] = ((x // 37 + y // 5) * 40) & 255
    img[..., 2] = rng.integers(0, 256, (h, w))
    img[..., 3] = np.where((x // 101) % 3 == 0, 255, (x + y) & 255)
    return img


@pytest.mark.parametrize("w,h", [(70001, 5), (3, 66000)])
def test_sides_beyond_65536(zt, oracle, w, h):
    img = strip(w, h, w + h)
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        assert np.array_equal(zt.bc_encode(img, f), oracle.bc_encode(img, f)), f
    assert np.array_equal(zt.mip_next(img, 1), oracle.mip_next(img, 1))
    bc = zt.bc_encode(img, "bc7")
    assert np.array_equal(zt.bc_decode(bc, w, h, "bc7"), oracle.bc_decode(bc, w, h, "bc7"))


def test_concurrent_host_calls_match_serial(zt):
    fmts = ["bc1", "bc3", "bc4", "bc5", "bc7"]
    imgs = [np.ascontiguousarray(synth.c2_bc7(160, s)[: 90 + 3 * s, : 120 + 2 * s]) for s in range(15)]
    jobs = [(i, fmts[i % 5], 30.0 if i % 4 == 3 else 0.0) for i in range(len(imgs))]
    serial = {j: zt.bc_encode(imgs[j[0]], j[1], rdo=j[2]) for j in jobs}
    serial_mips = [zt.mip_next(im, 1) for im in imgs]
    got, mips, errors = {}, {}, []

    def work(k):
        try:
            for rep in range(3):
                for j in jobs[k % 5::5]:
                    got[(k, rep, j)] = zt.bc_encode(imgs[j[0]], j[1], rdo=j[2])
                i = (k + rep) % len(imgs)
                mips[(k, rep)] = (i, zt.mip_next(imgs[i], 1))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(12)]   # more threads than host slots
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not any(t.is_alive() for t in threads), "a host call did not return"
    assert not errors, errors
    assert len(got) == 12 * 3 * 3
    for (k, rep, j), out in got.items():
        assert np.array_equal(out, serial[j]), (k, rep, j)
    for (k, rep), (i, m) in mips.items():
        assert np.array_equal(m, serial_mips[i]), (k, rep, i)
