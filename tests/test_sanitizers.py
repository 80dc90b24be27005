"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
planner (pure C++, fyfft_amd/csrc/planner.cpp) and the CPU oracle
(oracle/oracle.c) are built with g++/gcc -fsanitize=address,undefined and
their check drivers (tests/cpp/) run; any sanitizer report fails the test.
Host compilers only: no device code is built here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _run(cmd, env=None):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (" ".join(cmd), r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    return env


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_planner_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "planner_check")
    _run(["g++", "-std=c++17", *SAN, "-I" + os.path.join(ROOT, "fyfft_amd", "csrc"),
          os.path.join(ROOT, "tests", "cpp", "planner_check.cpp"),
          os.path.join(ROOT, "fyfft_amd", "csrc", "planner.cpp"), "-o", exe])
    out = _run([exe], env=_env())
    print(out)
    assert "0 failures" in out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_check")
    _run(["gcc", "-std=gnu11", "-fcx-limited-range", *SAN, "-I" + os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "cpp", "oracle_check.c"), os.path.join(ROOT, "oracle", "oracle.c"),
          "-o", exe, "-lm", "-lpthread"])
    out = _run([exe], env=_env())
    print(out)
    assert "0 failures" in out
