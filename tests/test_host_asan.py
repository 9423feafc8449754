"""Host-side AddressSanitizer run (CPU only): builds the HIP library with its host code
instrumented (device code is not: GPU sanitizers are not used here) and the CPU oracle, both with
clang's AddressSanitizer, then runs the C-ABI tests (argument and workspace validation, sizing,
error paths) and the oracle tests against those builds in a child process with the sanitizer
runtime preloaded. Any invalid host memory access in either library fails the child."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "_build", "asan")
CLANG = "/opt/rocm/llvm/bin/clang"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


@pytest.mark.skipif(_runtime() is None or not os.path.exists(CLANG), reason="clang AddressSanitizer runtime not found")
def test_abi_and_oracle_under_host_asan():
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libdiso_hip.so")
    ora = os.path.join(OUT, "liboracle.so")
    san = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    host_san = [a for f in san for a in ("-Xarch_host", f)]  # host code only
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-fPIC", "-shared", "-std=c++17",
                    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *host_san,
                    "-shared-libasan", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "diso_amd", "csrc"), "-o", lib,
                    os.path.join(ROOT, "diso_amd", "csrc", "diso_hip.hip")], check=True, capture_output=True)
    subprocess.run([CLANG, "-O1", "-g", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", *san, "-shared-libasan",
                    "-shared", "-o", ora, os.path.join(ROOT, "oracle", "oracle.c")], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=_runtime(), DISO_AMD_LIB=lib, DISO_ORACLE_LIB=ora,
               ASAN_OPTIONS="detect_leaks=0:alloc_dealloc_mismatch=0:detect_odr_violation=0:abort_on_error=0")
    code = ("import ctypes, sys, pytest; from diso_amd import _lib; from oracle import oracle as O; "
            "_lib.lib(); O.lib(); "
            "assert ctypes.CDLL(_lib.LIB_PATH).__asan_report_load4 and ctypes.CDLL(O._LIB_PATH).__asan_report_load4; "
            "sys.exit(pytest.main(['-q', '-p', 'no:cacheprovider', 'tests/test_abi.py', 'tests/test_oracle.py']))")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "AddressSanitizer" not in r.stderr
