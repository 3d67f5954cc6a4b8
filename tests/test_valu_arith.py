"""CPU test: the instruction-level Goldilocks primitives of fhetoothpaste_amd/csrc/gl_valu.hpp.

The header's host build runs the same op sequences as its device (inline-asm) build, lane by lane;
tests/cpp/valu_arith_check.cpp compares every primitive (scalar and group forms) with 128-bit integer
arithmetic on edge values that reach each carry/borrow case and on random values.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "valu_arith_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_valu_primitives_match_int128(tmp_path):
    exe = str(tmp_path / "valu_arith_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", exe, SRC])
    out = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok ")
