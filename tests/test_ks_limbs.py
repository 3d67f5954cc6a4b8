"""CPU test: the key-switching key as matrix-core operands (fhetoothpaste_amd/csrc/ks_limbs.hpp).

tests/cpp/ks_limbs_dump.cpp writes the int8 operand matrix the key-switch kernel reads
(kernels_ks_mfma.hip).  This file replays the kernel's arithmetic with numpy integers: per 64-lane
wave and k-step, the digit operand of lane l is row l % 32, input word 2 s + l // 32, its 16 slots
as the kernel packs them (b = 4: even nibbles of (x + C) >> (64 - 4 d) in slots 0..7, odd ones in
8..15; other bases: slot p = digit p); the key operand is the dumped 16 bytes of lane l; the 32x32
products are summed over k, the bias 128 sum(u) added back, the eight bytes recombined and reduced
mod Q, then the correction and negation of the vector kernel.  The result must equal the oracle's
key switch (lwe_conversion.cpp:205-239) on the same key and inputs, which is computed without any
of this layout; a planted error in the operand matrix must make it differ.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "ks_limbs_dump.cpp")
CSRC = os.path.join(HERE, "..", "fhetoothpaste_amd", "csrc")
Q = 2**64 - 2**32 + 1


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("ks_limbs")
    exe = str(d / "dump")
    # host code under AddressSanitizer + UndefinedBehaviorSanitizer (the operand builder walks a large
    # index space; any out-of-bounds access or overflow aborts the dump)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", CSRC, "-o", exe, SRC])

    def run(ksk, n_in, n_out, b, dg):
        kin, kout = str(d / "ksk.bin"), str(d / "kl.bin")
        np.ascontiguousarray(ksk, dtype=np.uint64).tofile(kin)
        subprocess.check_call([exe, kin, kout, str(n_in), str(n_out), str(b), str(dg)])
        raw = np.fromfile(kout, dtype=np.int8)
        slots = raw[-64:].view(np.int32)
        return raw[:-64], slots
    return run


def corrector(b):
    """Signed-to-unsigned corrector C (gadget_decomp.h:16-28), as the context computes it."""
    C, shift = 0, 64 - b
    while shift > 0:
        C += 1 << (shift + b - 1)
        shift -= b
    return (C + (1 << (shift + b - 1))) % Q


def correction(ksk, n_in, n_out, b, dg):
    """corr[col] = sum_ij delta_j K[i][j][col]: the signed digits are u_ij - delta_j."""
    D = 64 // b + (1 if 64 % b else 0)
    least = 1 << ((64 - b * D) + b - 1)
    W = n_out + 1
    K = ksk.reshape(n_in, dg, W).astype(object)
    corr = [0] * W
    for j in range(dg):
        delta = (1 << (b - 1)) if j != D - 1 else least
        for col in range(W):
            corr[col] = (corr[col] + delta * int(sum(K[:, j, col]))) % Q
    return corr


def operand_digits(x, b, dg):
    """The 16 slot bytes of a word x + C as the kernel packs them."""
    if b == 4:
        v = x >> (64 - 4 * dg)
        lo, hi = v & 0x0F0F0F0F0F0F0F0F, (v >> 4) & 0x0F0F0F0F0F0F0F0F
        return [(lo >> (8 * p)) & 255 for p in range(8)] + [(hi >> (8 * p)) & 255 for p in range(8)]
    return [((x >> (64 - (j + 1) * b)) & ((1 << b) - 1)) if j < dg else 0 for j in range(16)]


def model(kl, lwes, n_in, n_out, b, dg, corr):
    W, groups, steps = n_out + 1, (n_out + 1 + 31) // 32, n_in // 2
    C = corrector(b)
    count = lwes.shape[0]
    ops = kl.reshape(groups, steps, 8, 64, 16).astype(np.int64)
    # digit operands of every ciphertext: [count][steps][half][16]
    A = np.array([[[operand_digits((int(lwes[c][2 * s + h]) + C) % Q, b, dg) for h in range(2)]
                   for s in range(steps)] for c in range(count)], dtype=np.int64)
    U = A.sum(axis=(1, 2, 3))
    out = np.zeros((count, W), dtype=np.uint64)
    for g in range(groups):
        for c in range(count):
            T = [0] * 32
            for t in range(8):
                # lane l: column l % 32, slots of word 2 s + l // 32
                St = np.zeros(32, dtype=np.int64)
                for h in range(2):
                    St += np.einsum("sp,slp->l", A[c, :, h, :], ops[g, :, t, 32 * h:32 * h + 32, :])
                for l in range(32):
                    T[l] += (int(St[l]) + 128 * int(U[c])) << (8 * t)
            for l in range(32):
                col = 32 * g + l
                if col >= W:
                    continue
                v = (T[l] - corr[col]) % Q
                out[c, col] = (Q - v) % Q if col < n_out else (int(lwes[c][n_in]) - v) % Q
    return out


@pytest.mark.parametrize("n_in,n_out,b,dg", [(16, 40, 4, 14), (16, 33, 5, 12), (8, 31, 4, 16), (8, 20, 7, 9),
                                            (8, 20, 1, 16), (8, 63, 2, 16), (8, 32, 3, 16), (8, 64, 6, 10),
                                            (16, 7, 4, 5)])
def test_operands_reproduce_the_key_switch(dump, oracle, n_in, n_out, b, dg):
    rng = np.random.default_rng(n_in + 7 * b + dg)
    ksk = rng.integers(0, Q, size=n_in * dg * (n_out + 1), dtype=np.uint64)
    lwes = rng.integers(0, Q, size=(4, n_in + 1), dtype=np.uint64)
    lwes[1] = Q - 1
    lwes[2] = 0
    kl, slots = dump(ksk, n_in, n_out, b, dg)
    assert sorted(j for j in slots if j >= 0) == list(range(dg))
    corr = correction(ksk, n_in, n_out, b, dg)
    got = model(kl, lwes, n_in, n_out, b, dg, corr)
    for c in range(lwes.shape[0]):
        assert np.array_equal(got[c], oracle.lwe_keyswitch(lwes[c], ksk, n_in, n_out, b, dg)), c
    # a planted error (one key byte of one operand, in a slot that holds a digit) changes an output
    bad = kl.copy()
    used = [p for p in range(16) if slots[p] >= 0]
    bad[3 * 16 + used[len(used) // 2]] ^= 1
    assert not np.array_equal(model(bad, lwes, n_in, n_out, b, dg, corr), got)
