"""The reference casts unclamped floats to int in three places: BC1's zap__to565 (zap_tex.h:46, fed by the
least-squares endpoints of :142-143), BC6H's quantiser (:842) and ASTC's zap__astc_quant (:1187).  On x86 a value
whose truncation leaves int32, or a NaN, becomes INT_MIN (cvttss2si); the kernels reproduce that in cvt_i32.

These tests pin whether any input reaches that case.  The oracle's counting build records, per cast site, the
largest |argument| that fits int32 and the number that do not, and counts least-squares steps that pass the
reference's |det| >= 1e-6 guard with a nearly singular system.  The inputs are built to drive the solves towards
singularity (zap_amd.synth.degenerate_blocks: near-solid, two-colour, collinear, one-outlier blocks) plus HDR
extremes for BC6H.

What the numbers show, and why:
  * BC1: det = (sum of w_i^2)(sum of v_i^2) - (sum of w_i v_i)^2 is 1/2 sum_ij (w_i - w_j)^2 in exact arithmetic, so
    it is 0 (every index alike) or >= 15/9.  With every index 2 or 3 the float det is -3.8e-6 and passes the guard,
    which the degenerate blocks do reach; its numerators are then pure rounding residue, and the endpoints stay
    within a few hundred.
  * ASTC: weights are k/64, so every sum is exact and an all-alike system gives det = 0 exactly; otherwise
    det >= 2^-12 and |endpoint| <= 130560 / 2^-12 < 2^31.  No nearly singular step passes the guard.
  * BC6H: the quantiser sees PCA and least-squares endpoints of finished values (at most 65504 * 64 / 31 after
    zap__f2h's clamp, whatever the float input: NaN, infinities, negatives); the 4-bit weights are at least 4/64
    apart, so a least-squares endpoint extrapolates the data at most about 16 times and the argument stays below
    about 4e4.
So no input reaches INT_MIN; the x86 semantics in cvt_i32 are kept for fidelity, not exercised."""
import ctypes as C

import numpy as np
import pytest

from oracle import loader
from zap_amd import synth


@pytest.fixture(scope="module")
def counting():
    import os
    import subprocess
    if not os.path.exists(os.path.join(loader.HERE, "_lib", "libzto_count.so")):
        subprocess.run(["make", "-s", "-C", loader.HERE, "oracle"], check=True)
    o = loader.oracle(counting=True)
    o.lib.zto_cast_stats.argtypes = [C.c_void_p] * 3
    return o


def _stats(o):
    m, bad, near = np.zeros(3), np.zeros(3, np.uint64), np.zeros(2, np.uint64)
    o.lib.zto_cast_stats(m.ctypes.data, bad.ctypes.data, near.ctypes.data)
    return m, bad, near


def test_bc1_near_singular_least_squares_stays_far_from_int_min(counting):
    img = synth.blocks_to_image(synth.degenerate_blocks(60000, 21))
    counting.lib.zto_cast_reset()
    counting.bc_encode(img, "bc1")
    counting.bc_encode(img, "bc3")
    m, bad, near = _stats(counting)
    assert bad[0] == 0
    assert near[0] > 100, near          # the nearly singular branch (det = -3.8e-6) is reached ...
    assert m[0] < 1000, m               # ... and its casts stay within a few hundred


def test_astc_least_squares_is_never_nearly_singular(counting):
    img = synth.blocks_to_image(synth.degenerate_blocks(12000, 22))
    counting.lib.zto_cast_reset()
    counting.astc_encode(img)
    m, bad, near = _stats(counting)
    assert bad[2] == 0 and near[1] == 0, (bad, near)
    assert 255 < m[2] < 5000, m          # endpoints overshoot [0, 255] (the clamp matters) but stay small


def test_bc6h_quantiser_is_bounded_for_any_float(counting):
    h = np.float32
    # blocks of specials (NaN, infinities, negatives, denormals, beyond the half range) mixed with ordinary values
    specials = np.array([np.nan, np.inf, -np.inf, -1.0, 0.0, 1e-45, 6e-8, 65504, 65520, 1e30, -1e30, 3.4e38], np.float32)
    k = np.arange(4096 * 16)
    vals = specials[synth.hash32(31, k) % np.uint32(len(specials))].astype(h)
    noise = (synth.hash32(32, k) & np.uint32(0xFFFF)).astype(h) / h(8.0)
    mix = np.where((synth.hash32(33, k) & np.uint32(3)) == 0, vals, noise).reshape(4096, 16)
    blocks = np.repeat(mix[:, :, None], 4, axis=2).reshape(4096, 4, 4, 4).transpose(1, 0, 2, 3).reshape(4, 4096 * 4, 4)
    counting.lib.zto_cast_reset()
    counting.bc6h_encode(np.ascontiguousarray(blocks, dtype=np.float32))
    m, bad, _ = _stats(counting)
    assert bad[1] == 0 and m[1] < 2000, (m, bad)


def test_degenerate_blocks_match_the_reference(oracle, ref):
    img = synth.blocks_to_image(synth.degenerate_blocks(6000, 23))
    for f in ["bc1", "bc3", "bc4", "bc5", "bc7"]:
        assert np.array_equal(oracle.bc_encode(img, f), ref.bc_encode(img, f)), f
    assert np.array_equal(oracle.astc_encode(img[:, :4 * 2000]), ref.astc_encode(img[:, :4 * 2000]))
