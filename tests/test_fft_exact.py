"""CPU test: exactness of the floating-point external product (fhetoothpaste_amd/csrc/fft_key.cpp).

tests/cpp/fft_check.cpp builds the transform-domain key of random and of adversarial key rows,
multiplies it by the folded transforms of digit polynomials in double precision (the device's
factorisation, position order and key layout) and checks that every limb convolution rounds to
its exact integer value (schoolbook negacyclic convolution over int64), that the limbs recombine to
the exact product mod Q (Goldilocks NTT), that the observed rounding error stays below the proven
bound and that the bound itself is below 1/2 for the gadget 2^12 x 3.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "fft_check.cpp")
LIB_SRC = os.path.join(ROOT, "fhetoothpaste_amd", "csrc", "fft_key.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fft_product_is_exact(tmp_path):
    exe = str(tmp_path / "fft_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-o", exe, SRC, LIB_SRC, "-lpthread"])
    out = subprocess.run([exe, "3"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok ")
    worst, bound_any, bound_key = (float(x) for x in out.stdout.split()[1:4])
    assert worst < 1e-3 and bound_any < 0.5 and bound_key <= bound_any
