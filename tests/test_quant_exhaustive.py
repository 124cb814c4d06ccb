"""The kernels' BC7 endpoint quantisers (an FMA-corrected v/255, candidates from a pre-clamped g, packed keys) give
the reference's (zap_tex.h:426-438, :613-624) result for every float.  tests/quant_exhaustive.c restates both;
here it runs over every 37th float bit pattern (about 116 M values); `./quant_exhaustive 1` covers all 2^32
(a few minutes on 8 cores) and found no mismatch for |v| <= 2048."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_quantisers_match_reference_arithmetic():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "quant_exhaustive")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", os.path.join(HERE, "quant_exhaustive.c"), "-o", exe, "-lm"],
                       check=True, capture_output=True)
        r = subprocess.run([exe, "37"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
