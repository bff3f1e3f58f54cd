"""The host factorisation (pyjjasim_amd/csrc/jja_factor.cpp) built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on a driver that
factorises grid Laplacians, checks L L^T = K, and feeds it an indefinite and
malformed matrices (tests/native/factor_sanitize.cpp). CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_factor_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = tmp_path / "factor_san"
    build = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", os.path.join(ROOT, "tests", "native", "factor_sanitize.cpp"),
             os.path.join(ROOT, "pyjjasim_amd", "csrc", "jja_factor.cpp"), "-o", str(exe)]
    subprocess.run(build, check=True)
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([str(exe)], capture_output=True, text=True, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "sanitized factor checks passed" in out.stdout
