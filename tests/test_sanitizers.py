"""CPU sanitizer builds: the oracle with GCC's AddressSanitizer and UndefinedBehaviorSanitizer
(tests/san/oracle_scenarios.py over every oracle entry point the tests use), and the host
build of the library's adjoint pieces (tests/adjoint_harness.hip) with clang's, running
tests/test_adjoint_pieces.py. Each runs in a child process with the sanitizer runtime
preloaded; a report aborts the child and fails the test. Host code only: no GPU code is built
with a sanitizer."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")

SAN_ENV = {
    "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1:detect_odr_violation=0",
    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
}


def _runtime(cmd):
    try:
        path = subprocess.run(cmd, capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return path if os.path.isfile(path) else None


def _run(argv, env_extra, timeout):
    env = dict(os.environ, **SAN_ENV, **env_extra)
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run(argv, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-6000:]
    return out


def test_oracle_under_asan_ubsan():
    asan = _runtime(["g++", "-print-file-name=libasan.so"])
    if asan is None or shutil.which("make") is None:
        pytest.skip("GCC's AddressSanitizer runtime is not available")
    lib = os.path.join(BUILD, "liboracle_san.so")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "san", f"SAN_OUT={lib}"], check=True,
                   capture_output=True)
    out = _run([sys.executable, os.path.join(HERE, "san", "oracle_scenarios.py")],
               {"LD_PRELOAD": asan, "MIST_ORACLE_LIB": lib}, timeout=900)
    assert "all scenarios done" in out and f"oracle library: {lib}" in out


def test_adjoint_harness_under_asan_ubsan():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    rt = _runtime([hipcc, "-print-file-name=libclang_rt.asan-x86_64.so"])
    if rt is None:
        clang = os.path.join(os.path.dirname(os.path.realpath(hipcc)), "..", "llvm", "bin", "clang")
        rt = _runtime([clang, "--print-file-name=libclang_rt.asan-x86_64.so"])
    if rt is None:
        pytest.skip("clang's AddressSanitizer runtime is not available")
    os.makedirs(BUILD, exist_ok=True)
    lib = os.path.join(BUILD, "libadjh_san.so")
    # sanitizers on the host pass only: each -fsanitize= follows -Xarch_host
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC",
                    "-shared", "-ffp-contract=off",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-Xarch_host", "-fno-sanitize-recover=undefined",
                    os.path.join(HERE, "adjoint_harness.hip"), "-o", lib + ".tmp"], check=True, capture_output=True)
    os.replace(lib + ".tmp", lib)
    out = _run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                os.path.join(HERE, "test_adjoint_pieces.py")],
               {"LD_PRELOAD": rt, "MIST_ADJH_LIB": lib}, timeout=900)
    assert " passed" in out
