"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code:

* the oracle (oracle/*.c, `make -C oracle san`), driven by
  tests/sanitize/oracle_exercise.py in a child process with libasan
  preloaded and SCFT_ORACLE_LIB pointing at the sanitized library;
* the host C++ of libscft_hip.so that needs no device (host_math.cpp and the
  Broyden iteration of broyden.cpp through its callback entry), compiled with
  g++ into tests/_build/san/host_check (tests/sanitize/host_check.cpp).

Any report aborts the child (-fno-sanitize-recover=undefined, ASan's default)
and fails the test.  The device code and the HIP host code in capi.hip are
outside this build (hipcc's sanitizer runtime is a different one)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "tests", "_build", "san")
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
            "-fno-omit-frame-pointer", "-g", "-O1"]


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


pytestmark = pytest.mark.skipif(shutil.which("g++") is None or _libasan() is None,
                                reason="gcc with libasan needed")


def test_host_code_under_asan_ubsan():
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "host_check")
    src = [os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
           os.path.join(ROOT, "scft_amd", "csrc", "host_math.cpp"),
           os.path.join(ROOT, "scft_amd", "csrc", "broyden.cpp")]
    subprocess.run(["g++", "-std=c++17"] + SANITIZE + src + ["-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    p = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all checks passed" in p.stdout


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle_san.so")
    env = dict(os.environ, SCFT_ORACLE_LIB=lib, LD_PRELOAD=_libasan(),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", OMP_NUM_THREADS="2")
    script = os.path.join(ROOT, "tests", "sanitize", "oracle_exercise.py")
    p = subprocess.run([sys.executable, script], capture_output=True, text=True, env=env,
                       timeout=600)
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-4000:]
    assert "oracle_exercise: done" in p.stdout
