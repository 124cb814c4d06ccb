"""The drop-in include/zap_tex.h compiles unchanged in a C11 and a C++17 caller that uses only the reference's
names, links against libzaptex, and routes a failed call to the caller's zap_tex_on_error handler instead of
aborting.  The failing call is an argument error, rejected before any device work, so this runs without a GPU."""
import subprocess


def test_caller_compiles_as_c_and_cxx_and_handles_errors(dropin_caller):
    for exe in dropin_caller:
        r = subprocess.run([exe, "onerror"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (exe, r.stdout, r.stderr)
        assert "handled zap_bc_encode status 1" in r.stdout, r.stdout
