"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

Builds the oracle and the AVX2 CPU port (`make -C oracle sanitize`) and the library's host code --
C-ABI parameter checks, key generation, wire format, serving scheduler, drop-in classes -- with both
sanitizers (fhetoothpaste_amd.build.build_sanitized; device code is built as usual), then runs the
CPU tests that call into them in a child process with the sanitizer runtime preloaded.  Any report
(out-of-bounds access, use after free, signed overflow, misaligned access, ...) aborts the child and
fails this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"
SUBSET = ["tests/test_wire.py", "tests/test_keygen_streams.py", "tests/test_params.py", "tests/test_oracle.py",
          "tests/test_cpu_port.py", "tests/test_sap_host.py", "tests/test_golden.py", "tests/test_scheduler.py",
          "tests/test_capi_symbols.py"]


def _asan_runtime():
    if not os.path.exists(CLANG):
        return None
    path = subprocess.check_output([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], text=True).strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_host_code_under_asan_ubsan():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("no clang AddressSanitizer runtime on this machine")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    sys.path.insert(0, ROOT)
    from fhetoothpaste_amd import build as B
    lib = B.build_sanitized(verbose=False)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:verify_asan_link_order=0:detect_odr_violation=0",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "TOOTHPASTE_AMD_LIB": lib,
        "TOOTHPASTE_ORACLE_LIB_DIR": os.path.join(ROOT, "oracle", "_san"),
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + SUBSET,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
