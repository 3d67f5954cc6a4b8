"""CPU test: the tables of the twisted and two-wave N = 1024 transforms (host_ntt.hpp), through a
model of the kernels' data flow, against the plain negacyclic transform.

tests/cpp/ntt_tables_dump.cpp writes the tables; this file replays, with Python integers, the
order of operations of ntt_forward_t64 / ntt_inverse_t64 (ntt_wave.hpp) and of the two-wave
transforms (ntt_half.hpp: level 0 from both halves, levels 1..3 in the H1 layout, the twist, the
cyclic 64-point levels t0..t2 in H2 with power-of-two twiddles, t3..t5 in H3 with the per-lane
table, and the mirror inverse with the halves' last level; N^-1 is in the inverse's untwist table,
so no level scales) on the dumped tables, and compares with
a Cooley-Tukey transform whose twiddles are psi^brv(i), psi = 7^((Q-1)/2N) computed here.  The
cyclic twiddle exponents of the kernels (w = 2^39) are checked against w^brv(m) as well.
"""
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "ntt_tables_dump.cpp")
CSRC = os.path.join(HERE, "..", "fhetoothpaste_amd", "csrc")

Q = 2**64 - 2**32 + 1
N = 1024
LOGN = 10
# kDft64Pow2 (ntt_wave.hpp): exponents of 2 of the cyclic transform's twiddles, levels t0..t3
K = [[0], [0, 48], [0, 48, 120, 168], [0, 48, 120, 168, 156, 12, 84, 132]]


def brv(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


PSI = pow(7, (Q - 1) >> (LOGN + 1), Q)
PSI_REV = [pow(PSI, brv(i, LOGN), Q) for i in range(N)]
PSI_INV_REV = [pow(pow(PSI, Q - 2, Q), brv(i, LOGN), Q) for i in range(N)]


def fwd(a):
    a = list(a)
    m, t = 1, N // 2
    while m < N:
        for i in range(m):
            w = PSI_REV[m + i]
            for j in range(2 * i * t, 2 * i * t + t):
                u, v = a[j], a[j + t] * w % Q
                a[j], a[j + t] = (u + v) % Q, (u - v) % Q
        m, t = 2 * m, t // 2
    return a


def fwd_levels(a, levels):
    a = list(a)
    m, t = 1, N // 2
    for _ in range(levels):
        for i in range(m):
            w = PSI_REV[m + i]
            for j in range(2 * i * t, 2 * i * t + t):
                u, v = a[j], a[j + t] * w % Q
                a[j], a[j + t] = (u + v) % Q, (u - v) % Q
        m, t = 2 * m, t // 2
    return a


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("ntt_tables")
    exe, out = str(d / "dump"), str(d / "tables.bin")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", CSRC, "-o", exe, SRC])
    subprocess.check_call([exe, out])
    w = [int(x) for x in np.fromfile(out, dtype=np.uint64)]
    assert len(w) == 2 * N + 4 * 2 * N
    return {"psi_rev": w[:N], "psi_inv_rev": w[N:2 * N], "t64": w[2 * N:4 * N], "it64": w[4 * N:6 * N],
            "h64": w[6 * N:8 * N], "ih64": w[8 * N:10 * N]}


def test_psi_tables(tables):
    assert tables["psi_rev"] == PSI_REV
    assert tables["psi_inv_rev"] == PSI_INV_REV
    # the cyclic transform's root and its twiddle exponents
    w = pow(PSI, 32, Q)
    assert w == pow(2, 39, Q) and pow(w, 32, Q) == Q - 1
    for t in range(4):
        for m in range(1 << t):
            assert pow(2, K[t][m], Q) == pow(w, brv(m, t) * (32 >> t), Q), (t, m)


def test_twisted_transform_tables(tables):
    t64, it64 = tables["t64"], tables["it64"]
    rng = random.Random(7)
    a = [rng.randrange(Q) for _ in range(N)]
    ref = fwd(a)
    x = fwd_levels(a, 4)
    for g in range(N):  # twist: pass-2 position of lane L, word k = 2c + q
        j, i = g >> 6, g & 63
        L, k = (j << 2) | (i & 3), i >> 2
        x[g] = x[g] * t64[2 * (64 * (k >> 1) + L) + (k & 1)] % Q
    for t in range(4):
        d = 32 >> t
        for g in range(N):
            i = g & 63
            if i & d:
                continue
            tw = pow(2, K[t][i >> (6 - t)], Q)
            u, v = x[g], x[g + d] * tw % Q
            x[g], x[g + d] = (u + v) % Q, (u - v) % Q
    for lv, base in ((8, 256), (9, 512)):
        d = N >> (lv + 1)
        for g in range(N):
            if g & d:
                continue
            u, v = x[g], x[g + d] * t64[N + base + (g >> (10 - lv))] % Q
            x[g], x[g + d] = (u + v) % Q, (u - v) % Q
    assert x == ref
    # inverse
    y = list(ref)
    for lv, base in ((9, 512), (8, 256)):
        d = N >> (lv + 1)
        for g in range(N):
            if g & d:
                continue
            u, v = y[g], y[g + d]
            y[g], y[g + d] = (u + v) % Q, (u - v) * it64[N + base + (g >> (10 - lv))] % Q
    for t in (3, 2, 1, 0):
        d = 32 >> t
        for g in range(N):
            i = g & 63
            if i & d:
                continue
            tw = pow(2, (192 - K[t][i >> (6 - t)]) % 192, Q)
            u, v = y[g], y[g + d]
            y[g], y[g + d] = (u + v) % Q, (u - v) * tw % Q
    for g in range(N):
        j, i = g >> 6, g & 63
        L, k = (j << 2) | (i & 3), i >> 2
        y[g] = y[g] * it64[2 * (64 * (k >> 1) + L) + (k & 1)] % Q
    for lv in (3, 2, 1, 0):
        m, t = 1 << lv, N >> (lv + 1)
        for blk in range(m):
            w = PSI_INV_REV[m + blk]
            for jj in range(2 * blk * t, 2 * blk * t + t):
                u, v = y[jj], y[jj + t]
                y[jj], y[jj + t] = (u + v) % Q, (u - v) * w % Q
    # N^-1 rides on the untwist table: no scaling after the last level
    assert y == a


def bu(H, b):
    return (b // H) * 2 * H + b % H


P1 = lambda l, k: l | (k << 6)
P2 = lambda l, k: (l & 7) | (k << 3) | ((l >> 3) << 6)
P3 = lambda l, k: k | (l << 3)


def relayout(x, A, B):
    d = {A(l, k): x[l][k] for l in range(64) for k in range(8)}
    return [[d[B(l, k)] for k in range(8)] for l in range(64)]


def ct4(x, H, tw):
    for l in range(64):
        for b in range(4):
            lo = bu(H, b)
            u, v = x[l][lo], x[l][lo + H] * tw(l, lo) % Q
            x[l][lo], x[l][lo + H] = (u + v) % Q, (u - v) % Q


def gs4(x, H, tw):
    for l in range(64):
        for b in range(4):
            lo = bu(H, b)
            u, v = x[l][lo], x[l][lo + H]
            x[l][lo], x[l][lo + H] = (u + v) % Q, (u - v) * tw(l, lo) % Q


def test_two_wave_transform_tables(tables):
    h, ih = tables["h64"], tables["ih64"]
    rng = random.Random(11)
    digits = [rng.randrange(1 << 12) for _ in range(N)]
    ref = fwd(digits)
    out = [0] * N
    for q in range(2):
        base = lambda l: (64 * q + l) * 8
        # level 0 from both halves (the wave keeps its half's output), levels 1..3 in H1
        x = [[(digits[P1(l, k)] + (1 - 2 * q) * PSI_REV[1] * digits[P1(l, k) + 512]) % Q for k in range(8)]
             for l in range(64)]
        for T in (1, 2, 3):
            ct4(x, 8 >> T, lambda l, lo, T=T: PSI_REV[(1 << T) + ((q << (T - 1)) | (lo >> (4 - T)))])
        x = relayout(x, P1, P2)
        x = [[x[l][k] * h[base(l) + k] % Q for k in range(8)] for l in range(64)]
        ct4(x, 4, lambda l, lo: 1)
        ct4(x, 2, lambda l, lo: pow(2, K[1][lo >> 2], Q))
        ct4(x, 1, lambda l, lo: pow(2, K[2][lo >> 1], Q))
        x = relayout(x, P2, P3)
        ct4(x, 4, lambda l, lo: h[N + base(l)])
        ct4(x, 2, lambda l, lo: h[N + base(l) + 1 + (lo >> 2)])
        ct4(x, 1, lambda l, lo: h[N + base(l) + 3 + (lo >> 1)])
        for l in range(64):
            for k in range(8):
                out[512 * q + P3(l, k)] = x[l][k]
    assert out == ref
    # inverse: both halves up to the last level, then the halves' exchange
    r = [rng.randrange(Q) for _ in range(N)]
    halves = []
    for q in range(2):
        base = lambda l: (64 * q + l) * 8
        x = [[r[512 * q + P3(l, k)] for k in range(8)] for l in range(64)]
        gs4(x, 1, lambda l, lo: ih[N + base(l) + 3 + (lo >> 1)])
        gs4(x, 2, lambda l, lo: ih[N + base(l) + 1 + (lo >> 2)])
        gs4(x, 4, lambda l, lo: ih[N + base(l)])
        x = relayout(x, P3, P2)
        gs4(x, 1, lambda l, lo: pow(2, (192 - K[2][lo >> 1]) % 192, Q))
        gs4(x, 2, lambda l, lo: pow(2, (192 - K[1][lo >> 2]) % 192, Q))
        gs4(x, 4, lambda l, lo: 1)
        x = [[x[l][k] * ih[base(l) + k] % Q for k in range(8)] for l in range(64)]
        x = relayout(x, P2, P1)
        for T in (3, 2, 1):
            gs4(x, 8 >> T, lambda l, lo, T=T: PSI_INV_REV[(1 << T) + ((q << (T - 1)) | (lo >> (4 - T)))])
        halves.append(x)
    coef = [0] * N
    for l in range(64):
        for k in range(8):
            a, b = halves[0][l][k], halves[1][l][k]   # N^-1 came with the untwist
            coef[P1(l, k)] = (a + b) % Q
            coef[P1(l, k) + 512] = (a - b) * PSI_INV_REV[1] % Q
    assert fwd(coef) == r
