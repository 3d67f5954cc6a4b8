"""CPU test: the tuning cache file of tp_autotune_cached (fhetoothpaste_amd/csrc/tuning_cache.cpp),
compiled with g++ beside tests/cpp/tuning_cache_check.cpp: write / read round trip of every field,
the lookup key, the default of a key older files lack, a missing file, malformed files."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "fhetoothpaste_amd", "csrc")


def test_tuning_cache_file(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = str(tmp_path / "check")
    # the parser under AddressSanitizer + UndefinedBehaviorSanitizer (malformed input must not read past
    # the text)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", CSRC, "-o", exe,
                           os.path.join(HERE, "cpp", "tuning_cache_check.cpp"), os.path.join(CSRC, "tuning_cache.cpp")])
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
