"""The drop-in boundary as a C program sees it, on the GPU.

* tests/dropin_caller.c, compiled against include/zap_tex.h (C11 and C++17) and linked with -lzaptex, runs the
  reference CLI's texture loop (zap.c:331-343: zap_bc_encode, zap_mip_next, repeat) and the other entry points;
  its bytes equal the compiled reference's (tests/golden/golden.json), the decoder and split equal the oracle's.
* zap_amd/_lib/zaptex-tex, the `zap tex` command on libzaptex, writes the DDS and .astc files the reference CLI
  (zap.c, built from source into oracle/_ref/zapcli when the golden file was made) writes, byte for byte.
"""
import ctypes as C
import hashlib
import os
import subprocess

import numpy as np
import pytest

from zap_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "zap_amd", "_lib", "zaptex-tex")
FMTS = {"bc1": 0, "bc3": 1, "bc4": 2, "bc5": 3, "bc7": 4}


def sha_file(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def _small(w, h):
    return np.ascontiguousarray(synth.c2_bc7(max(w, h, 4), 40 + w * 7 + h)[:h, :w])


@pytest.fixture(scope="module")
def images():
    imgs = {"c1_1024": synth.c1_bc1(1024, 1), "c2_4096": synth.c2_bc7(4096, 2), "c3_astc_2048": synth.c3_astc(2048, 3)}
    for w, h in [(37, 21), (5, 3), (1, 1)]:
        imgs[f"small_{w}x{h}"] = _small(w, h)
    return imgs


def test_tex_tool_writes_the_reference_cli_files(golden, images, tmp_path):
    for name, run in golden["containers"].items():
        img = images[run["image"]]
        h, w = img.shape[:2]
        src = tmp_path / (run["image"] + ".rgba")
        if not src.exists():
            img.tofile(src)
        dst = tmp_path / (name + (".astc" if "astc" in run["args"] else ".dds"))
        r = subprocess.run([TOOL, *run["args"], str(w), str(h), str(src), str(dst)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (name, r.stdout, r.stderr)
        data = dst.read_bytes()
        assert data[:len(bytes.fromhex(run["header"]))].hex() == run["header"], name
        assert len(data) == run["bytes"] and hashlib.sha256(data).hexdigest() == run["sha256"], name


@pytest.mark.parametrize("lang", [0, 1], ids=["c11", "cxx17"])
def test_dropin_caller_runs_the_cli_loop(dropin_caller, golden, images, tmp_path, lang):
    exe = dropin_caller[lang]
    cases = [("c1_1024", "bc1", 0, 0, golden["outputs"]["c1_1024_bc1"]),
             ("c1_1024", "bc3", 0, 1, golden["containers"]["c1_bc3_mips"]["payload_sha256"]),
             ("small_37x21", "bc3", 1, 1, golden["containers"]["odd_bc3_mips_srgb"]["payload_sha256"]),
             ("small_1x1", "bc1", 0, 1, golden["containers"]["odd_bc1_mips"]["payload_sha256"])]
    if lang == 0:
        cases.append(("c2_4096", "bc7", 0, 1, golden["containers"]["c2_bc7_mips"]["payload_sha256"]))
    for key, fmt, srgb, mips, want in cases:
        img = images[key]
        h, w = img.shape[:2]
        src, dst = tmp_path / f"{key}.rgba", tmp_path / f"{key}_{fmt}_{srgb}_{mips}.bin"
        img.tofile(src)
        r = subprocess.run([exe, "chain", str(FMTS[fmt]), str(srgb), str(mips), str(w), str(h), str(src), str(dst)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (key, fmt, r.stdout, r.stderr)
        assert sha_file(dst) == want, (key, fmt, srgb, mips)


def test_dropin_caller_other_entry_points(dropin_caller, golden, oracle, tmp_path):
    ent = golden["small"]["67x45"]
    w, h = 67, 45
    img = np.frombuffer(bytes.fromhex(ent["input"]), np.uint8).reshape(h, w, 4)
    hdr = np.frombuffer(bytes.fromhex(ent["hdr_input"]), np.float32).reshape(h, w, 4)
    img.tofile(tmp_path / "in.rgba")
    hdr.tofile(tmp_path / "in.f32")
    prefix = str(tmp_path / "out")
    r = subprocess.run([dropin_caller[0], "small", str(w), str(h), str(tmp_path / "in.rgba"), str(tmp_path / "in.f32"), prefix],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "merge restores the blocks" in r.stdout, (r.stdout, r.stderr)
    read = lambda ext: np.fromfile(prefix + "." + ext, np.uint8)
    assert read("astc").tobytes().hex() == ent["astc"]
    assert read("bc6h").tobytes().hex() == ent["bc6h"]
    bc7 = read("bc7")
    assert bc7.tobytes().hex() == ent["bc7"]
    assert np.array_equal(read("dec").reshape(h, w, 4), oracle.bc_decode(bc7, w, h, "bc7"))
    planes = np.zeros_like(bc7)
    oracle.lib.zto_bc_shuffle(bc7.ctypes.data_as(C.c_void_p), planes.ctypes.data_as(C.c_void_p), bc7.size, 4, 1)
    assert np.array_equal(read("split"), planes)
