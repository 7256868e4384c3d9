"""ASan + UBSan builds of the host C and C++ (CPU).

tests/san/Makefile compiles the C oracle (FP64 and FP32) and the library's
ASCII geometry reader (taichi_lbm3d_amd/csrc/geo_ascii.h, the code behind
lbm3d_read_geo_ascii) with -fsanitize=address,undefined and no recovery, each
under a driver that walks the branches: dense and sparse storage, thin grids,
every face mode and force convention for the oracle; well-formed, malformed,
oversized-token, empty and non-regular inputs for the reader.  A sanitizer
report aborts the driver, so a zero exit status and its final "ok" are the
pass condition.
"""
import os
import shutil
import subprocess

import pytest

SAN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "san")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc and g++")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-j4"], cwd=SAN, check=True, capture_output=True, text=True)
    return os.path.join(SAN, "_build")


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, env=env)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("exe", ["oracle64_san", "oracle32_san"])
def test_oracle_under_sanitizers(built, exe):
    rc, out = run(os.path.join(built, exe))
    assert rc == 0 and out.rstrip().endswith("ok"), out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out


def test_geo_reader_under_sanitizers(built, tmp_path):
    rc, out = run(os.path.join(built, "geo_ascii_san"), str(tmp_path))
    assert rc == 0 and out.rstrip().endswith("ok"), out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out
