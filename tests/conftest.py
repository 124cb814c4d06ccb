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
