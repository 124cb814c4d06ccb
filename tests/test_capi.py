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


def test_rdo_is_rejected_not_emulated(zt):
    img = np.zeros((4, 4, 4), np.uint8)
    with pytest.raises(binding.ZapTexError, match="unsupported"):
        zt.bc_encode(img, "bc1", rdo=10.0)


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
