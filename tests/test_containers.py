"""The `zap tex` file fixtures (tests/golden/golden.json "containers", written by the reference CLI zap.c built
from source) hold what the DDS and .astc layouts fix: magic, sizes, mip counts, FourCC / DXGI codes, caps, the
payload size summed over the chain.  Where oracle/_ref/zapcli is built (this container), the small cases are
re-run to show the fixture is current."""
import hashlib
import os
import struct
import subprocess

import numpy as np
import pytest

from zap_amd import binding, synth

DXGI = {"bc1": 71, "bc3": 77, "bc4": 80, "bc5": 83, "bc7": 98}
FOURCC = {"bc1": b"DXT1", "bc3": b"DXT5", "bc4": b"BC4U", "bc5": b"BC5U"}
SIZES = {"c1_1024": (1024, 1024), "c2_4096": (4096, 4096), "c3_astc_2048": (2048, 2048), "small_37x21": (37, 21),
         "small_5x3": (5, 3), "small_1x1": (1, 1)}


def test_headers_follow_the_dds_and_astc_layouts(golden):
    runs = golden["containers"]
    assert len(runs) >= 12
    for name, run in runs.items():
        a = run["args"]
        fmt, mips, srgb = a[a.index("-f") + 1], "-m" in a, "-S" in a
        w, h = SIZES[run["image"]]
        hd = bytes.fromhex(run["header"])
        if fmt == "astc":
            assert hd[:7] == bytes([0x13, 0xAB, 0xA1, 0x5C, 4, 4, 1]), name
            assert int.from_bytes(hd[7:10], "little") == w and int.from_bytes(hd[10:13], "little") == h and hd[13:16] == b"\1\0\0"
            assert run["bytes"] == 16 + binding.bc_size(w, h, "bc7"), name
            continue
        srgb = srgb and fmt not in ("bc4", "bc5")
        dx10 = fmt == "bc7" or srgb
        levels = binding.mip_levels(w, h) if mips else 1
        assert hd[:4] == b"DDS " and len(hd) == (148 if dx10 else 128), name
        size, flags, hh, ww, linear, _depth, nlev = struct.unpack_from("<7I", hd, 4)
        assert (size, hh, ww, linear, nlev) == (124, h, w, binding.bc_size(w, h, fmt), levels), name
        assert flags == 0x1 | 0x2 | 0x4 | 0x1000 | 0x80000 | (0x20000 if mips else 0), name
        assert struct.unpack_from("<2I", hd, 76) == (32, 4) and hd[84:88] == (b"DX10" if dx10 else FOURCC[fmt]), name
        assert struct.unpack_from("<I", hd, 108)[0] == 0x1000 | (0x400008 if mips else 0), name
        if dx10:
            assert struct.unpack_from("<5I", hd, 128) == (DXGI[fmt] + int(srgb), 3, 0, 1, 0), name
        payload = sum(binding.bc_size(lw, lh, fmt) for lw, lh in binding.mip_sizes(w, h)[:levels])
        assert run["bytes"] == len(hd) + payload, name


def test_small_cases_reproduce_with_the_reference_cli(golden, tmp_path):
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "zapcli")
    if not os.path.exists(cli):
        pytest.skip("oracle/_ref/zapcli (the reference CLI built from zap.c) is not built here")
    for name, run in golden["containers"].items():
        if not run["image"].startswith("small_"):
            continue
        w, h = SIZES[run["image"]]
        img = np.ascontiguousarray(synth.c2_bc7(max(w, h, 4), 40 + w * 7 + h)[:h, :w])
        img.tofile(tmp_path / "in.rgba")
        dst = tmp_path / "out.bin"
        subprocess.run([cli, "tex", *run["args"], str(w), str(h), str(tmp_path / "in.rgba"), str(dst)], check=True, capture_output=True)
        assert hashlib.sha256(dst.read_bytes()).hexdigest() == run["sha256"], name


def test_library_headers_and_no_overrun_without_a_device(zt, golden):
    """libzaptex writes each file's header before it touches the device, so on a host without one the call fails
    with ZAPTEX_E_DEVICE after writing exactly the header: it equals the reference CLI's, and nothing past the
    file's size changes (a 1x1 BC1 file is 136 bytes, shorter than a DX10 header)."""
    import ctypes as C
    if zt.device_count() > 0:
        pytest.skip("a device is present: tests/test_gpu_dropin.py checks the whole files")
    for name, run in golden["containers"].items():
        a = run["args"]
        fmt = a[a.index("-f") + 1]
        if fmt == "astc":
            continue
        w, h = SIZES[run["image"]]
        n = zt.lib.zaptex_tex_file_size(w, h, binding.fmt_id(fmt), int("-m" in a), int("-S" in a))
        assert n == run["bytes"], name
        buf = np.full(n + 64, 0xEE, np.uint8)
        img = np.zeros((h, w, 4), np.uint8)
        rc = zt.lib.zaptex_tex_file(img.ctypes.data_as(C.c_void_p), w, h, w * 4, binding.fmt_id(fmt), 0.0, int("-m" in a),
                                    int("-S" in a), buf.ctypes.data_as(C.c_void_p), n)
        assert rc == 2, (name, rc)
        hd = bytes.fromhex(run["header"])
        assert buf[:len(hd)].tobytes() == hd, name
        assert (buf[n:] == 0xEE).all(), name
