"""Shared fixtures.  Tests marked `gpu` need a HIP device (MI355X); everything else runs on the CPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import loader  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(os.path.join(loader.HERE, "_lib", "libzto.so")):
        subprocess.run(["make", "-s", "-C", loader.HERE, "oracle"], check=True)
    return loader.oracle()


@pytest.fixture(scope="session")
def ref():
    if not loader.reference_available():
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    return loader.reference()


@pytest.fixture(scope="session")
def ref64():
    if not loader.reference_available():
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    return loader.reference(m1k64=True)


@pytest.fixture(scope="session")
def zt():
    import zap_amd
    return zap_amd.load()


def _build_c(src: str, out: str, lang: str) -> str:
    """Compile a C test program against include/ and libzaptex when it is missing or older than its inputs."""
    lib = os.path.join(ROOT, "zap_amd", "_lib", "libzaptex.so")
    deps = [src, lib, os.path.join(ROOT, "include", "zap_tex.h"), os.path.join(ROOT, "include", "zaptex.h")]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(d) for d in deps if os.path.exists(d)):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        cc = ["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++17", "-x", "c++"]
        subprocess.run([*cc, "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-L", os.path.dirname(lib),
                        "-lzaptex", "-Wl,-rpath," + os.path.dirname(lib), "-Wl,-rpath-link,/opt/rocm/lib", "-o", out], check=True)
    return out


@pytest.fixture(scope="session")
def dropin_caller():
    """tests/dropin_caller.c built as C11 and as C++17 against the drop-in zap_tex.h: (c binary, c++ binary)."""
    src = os.path.join(ROOT, "tests", "dropin_caller.c")
    return (_build_c(src, os.path.join(ROOT, "tests", "_bin", "dropin_caller"), "c"),
            _build_c(src, os.path.join(ROOT, "tests", "_bin", "dropin_caller_cxx"), "c++"))
