"""The C ABI library loads, exports every function include/zaptex.h declares, and rejects bad calls."""
import ctypes as C

import numpy as np
import pytest

from zap_amd import binding


def test_exports_every_declared_symbol(zt):
    names = binding.declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(zt.lib, n)]
    assert not missing, missing


def test_dropin_header_wraps_every_reference_entry_point():
    text = open(binding.HEADER.replace("zaptex.h", "zap_tex.h")).read()
    for fn in ["zap_bc_encode", "zap_bc_decode", "zap_bc6h_encode", "zap_bc6h_decode", "zap_astc_encode", "zap_astc_decode",
               "zap_mip_levels", "zap_mip_next", "zap_mip_next_f", "zap_bc_split", "zap_bc_merge", "zap_bc_size", "zap_bc_block_bytes"]:
        assert f" {fn}(" in text, fn


def test_rdo_runs_on_the_device_only(zt):
    """rdo > 0 is a device path like the rest: without a device it fails, it never runs on the CPU; a non-BCn image is
    refused before any device work."""
    img = np.zeros((4, 4, 4), np.uint8)
    if zt.device_count() == 0:
        with pytest.raises(binding.ZapTexError, match="device"):
            zt.bc_encode(img, "bc1", rdo=10.0)
    bad = (binding.ZapImage * 1)(binding.ZapImage(img.ctypes.data, img.ctypes.data, 16, 4, 4, binding.ASTC))
    assert zt.lib.zaptex_encode_rdo_dev(bad, 1, 10.0, None, 0, None) == 1


def test_bad_arguments(zt):
    out = np.zeros(16, np.uint8)
    assert zt.lib.zaptex_bc_encode(None, 4, 4, 16, 4, 0.0, out.ctypes.data_as(C.c_void_p)) == 1
    assert zt.lib.zaptex_bc_encode(out.ctypes.data_as(C.c_void_p), 4, 4, 16, 9, 0.0, out.ctypes.data_as(C.c_void_p)) == 1
    assert zt.lib.zaptex_set_option(1, 0) == 1 and zt.lib.zaptex_set_option(1, 65) == 1
    assert zt.lib.zaptex_set_option(99, 1) == 1
    assert zt.get_option(1) == 4


def test_no_cpu_fallback_without_device(zt):
    if zt.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(binding.ZapTexError, match="device"):
        zt.bc_encode(np.zeros((4, 4, 4), np.uint8), "bc7")


def test_sizes_match_reference_formulas():
    assert binding.bc_size(5, 3, "bc1") == 2 * 1 * 8 and binding.bc_size(5, 3, "bc7") == 32
    assert binding.mip_levels(1, 1) == 1 and binding.mip_levels(256, 128) == 9 and binding.mip_levels(5, 3) == 3
    assert binding.mip_sizes(4096, 4096)[-1] == (1, 1) and len(binding.mip_sizes(4096, 4096)) == 13


def test_batch_arena_sizes_match_the_documented_layout(zt):
    from zap_amd import batch
    for (w, h, fmt, levels) in [(1, 1, "bc7", 1), (300, 200, "bc1", 9), (4096, 16, "bc5", 13), (5, 3, "bc4", 2), (1000, 750, "bc3", 10)]:
        sizes, mo, oo = batch.layout(w, h, fmt, levels)
        al = lambda v: (v + 255) // 256 * 256
        assert zt.batch_out_bytes(w, h, fmt, levels) == oo[-1] + al(binding.bc_size(*sizes[-1], fmt))
        exp_m = (mo[-1] + al(sizes[-1][0] * sizes[-1][1] * 4)) if levels > 1 else 0
        assert zt.batch_mip_bytes(w, h, levels) == exp_m
        assert sizes == binding.mip_sizes(w, h)[:levels]


def test_batch_rejects_bad_textures_before_touching_a_device(zt):
    buf = np.zeros(1 << 16, np.uint8)
    p = (buf.ctypes.data + 255) // 256 * 256
    good = binding.ZapTexture(p, 64 * 4, 64, 64, binding.fmt_id("bc7"), 0, 7, p, p)
    h = C.c_void_p()
    for field, bad in [("fmt", 6), ("levels", 8), ("levels", 0), ("w", 0), ("stride", 100), ("out", p + 16)]:
        t = binding.ZapTexture.from_buffer_copy(good)
        setattr(t, field, bad)
        arr = (binding.ZapTexture * 1)(t)
        assert zt.lib.zaptex_batch_create(arr, 1, p, 1 << 12, C.byref(h)) == 1, field
    arr = (binding.ZapTexture * 1)(good)
    assert zt.lib.zaptex_batch_create(arr, 1, p, 16, C.byref(h)) == 1        # workspace too small
    assert zt.lib.zaptex_batch_workspace_size(arr, 1) > 0


def test_launch_shapes_are_validated(zt):
    assert zt.lib.zaptex_set_launch(b"zt_bc7_stage6", 96, 10) == 1
    assert zt.lib.zaptex_set_launch(b"zt_bc7_stage6", 128, 0) == 1
    assert zt.lib.zaptex_set_launch(b"no_such_kernel", 128, 10) == 1
    wg, cap = zt.get_launch("zt_bc7_stage6")
    zt.set_launch("zt_bc7_stage6", 64, 77)
    assert zt.get_launch("zt_bc7_stage6") == (64, 77)
    zt.set_launch("zt_bc7_stage6", wg, cap)


def test_gpu_texture_generator_matches_host_generator():
    from zap_amd import batch
    for (w, h, kind, seed) in [(300, 200, "alpha", 7), (64, 256, "normal", 8), (129, 67, "colour", 9), (5, 3, "alpha", 10),
                               (1, 1, "normal", 11), (517, 33, "alpha", 12)]:
        assert np.array_equal(batch.texture_gpu(w, h, kind, seed, dev="cpu").numpy(), batch.texture_cpu(w, h, kind, seed)), (w, h, kind)


def test_c4_mix():
    from zap_amd import batch
    specs = batch.c4_specs(512, 4)
    fmts = [s[2] for s in specs]
    assert specs[0][:2] == (300, 200) and specs[1][:2] == (1000, 750)
    for fmt, pct in batch.FORMAT_MIX:
        assert abs(fmts.count(fmt) / 512 - pct / 100) < 0.07, fmt
    assert all(64 <= min(w, h) and max(w, h) <= 4096 and max(w, h) <= 8 * min(w, h) for w, h, *_ in specs)
    assert any((w & (w - 1)) or (h & (h - 1)) for w, h, *_ in specs[2:])   # non-powers of two beyond the two fixed ones
