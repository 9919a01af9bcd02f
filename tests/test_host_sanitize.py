"""The host-only code under AddressSanitizer + UndefinedBehaviorSanitizer
(with LeakSanitizer): tests/cpp/sanitize_host.cpp drives the parameter sets
and file reader (csrc/sf_params.cpp), the neighbour list and padding producer
(csrc/neighbor_list.cpp) and the oracle (oracle/sf_oracle.c) through the cases
the CPU tests give them.  Built with g++/gcc only (no device code), so it runs
here; the reference forces a Release build (CMakeLists.txt:12) and has none.
CPU only."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-fopenmp"]


@pytest.fixture(scope="module")
def driver():
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("needs g++ and gcc")
    out = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out, exist_ok=True)
    obj = os.path.join(out, "sf_oracle_san.o")
    exe = os.path.join(out, "sanitize_host")
    csrc = os.path.join(ROOT, "libdescriptor_amd", "csrc")
    subprocess.check_call(["gcc", "-std=c99", "-ffp-contract=off", *FLAGS, "-c",
                           os.path.join(ROOT, "oracle", "sf_oracle.c"), "-o", obj])
    subprocess.check_call(["g++", "-std=c++17", *FLAGS, "-I", os.path.join(ROOT, "include"),
                           "-I", csrc, os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
                           os.path.join(csrc, "sf_params.cpp"),
                           os.path.join(csrc, "neighbor_list.cpp"), obj, "-o", exe, "-lm"])
    return exe


def test_host_code_is_sanitizer_clean(driver):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([driver, GOLDEN], capture_output=True, text=True, timeout=300, env=env)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "0 failed checks" in r.stdout
    for marker in ("AddressSanitizer", "LeakSanitizer", "runtime error"):
        assert marker not in report, report[-4000:]
