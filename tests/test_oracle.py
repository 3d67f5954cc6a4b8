"""Pins the CPU oracle against facts that do not come from the oracle itself.

The reference ships no golden vectors (its tests are randomized and check
invariants), so the oracle is pinned by closed forms, textbook arithmetic,
brute force and the reference tests' own noise-free invariants
(tests/cggi_blind_rotation_test.cpp:118-123, tests/backend/solinas_test.cpp:35-172,
tests/glwe_conversion_test.cpp:82-107).
"""
import random

import numpy as np
import pytest

Q = 0xFFFFFFFF00000001


def brv(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def poly_phase_rotate(m, e, N):
    """Coefficients of X^e * m(X) mod X^N + 1 (pure Python)."""
    e %= 2 * N
    out = [0] * N
    for i, c in enumerate(m):
        k = i + e
        sign = 1
        while k >= N:
            k -= N
            sign = -sign
        out[k] = (sign * int(c)) % Q
    return out


# --------------------------------------------------------------------------
# Field arithmetic (solinas_internal.h:39-111; solinas_test.cpp:35-138)
# --------------------------------------------------------------------------

EDGE = [0, 1, 2, Q - 1, Q - 2, 2**32 - 1, 2**32, 2**32 + 1, 2**63, 2**63 - 1, 2**64 - 2**32, Q // 2, 7]


def test_field_ops_vs_bigint(oracle):
    L = oracle.lib()
    rng = random.Random(1)
    vals = EDGE + [rng.randrange(Q) for _ in range(300)]
    for a in vals:
        for b in vals[:40]:
            assert L.or_add(a, b) == (a + b) % Q
            assert L.or_sub(a, b) == (a - b) % Q
            assert L.or_mul(a, b) == (a * b) % Q
            assert L.or_mul_textbook(a, b) == (a * b) % Q


def test_reduce128_any_input(oracle):
    L = oracle.lib()
    rng = random.Random(2)
    cases = [(2**64 - 1, 2**64 - 1), (0, 2**64 - 1), (2**64 - 1, 0), (2**32, 0), (Q, Q)]
    cases += [(rng.getrandbits(64), rng.getrandbits(64)) for _ in range(2000)]
    for hi, lo in cases:
        assert L.or_reduce128(hi, lo) == ((hi << 64) | lo) % Q


def test_inverse_and_pow(oracle):
    L = oracle.lib()
    assert L.or_inv(0) == 0
    rng = random.Random(3)
    for x in EDGE[1:] + [rng.randrange(1, Q) for _ in range(50)]:
        if x % Q:
            assert (L.or_inv(x) * x) % Q == 1
    assert pow(7, (Q - 1) // 2, Q) == Q - 1  # 7 is a non-residue: generator check of the reference root choice


# --------------------------------------------------------------------------
# NTT (solinas.cpp:168-313): closed form, convolution theorem, round trip
# --------------------------------------------------------------------------

@pytest.mark.parametrize("N", [128, 256])
def test_ntt_closed_form_bruteforce(oracle, N):
    R = oracle.Ring(N)
    rng = np.random.default_rng(N)
    a = rng.integers(0, Q, size=N, dtype=np.uint64)
    out = R.ntt(a)
    logN = N.bit_length() - 1
    psi = pow(7, (Q - 1) // (2 * N), Q)
    assert R.psi == psi
    av = [int(x) for x in a]
    for k in range(N):
        x = pow(psi, 2 * brv(k, logN) + 1, Q)
        ref = 0
        for j in reversed(range(N)):
            ref = (ref * x + av[j]) % Q
        assert int(out[k]) == ref
    assert np.array_equal(R.ntt_closed_form(a), out)


@pytest.mark.parametrize("N,psi", [(1024, 455906449640507599), (2048, 17492915097719143606)])
def test_ntt_monomials_large(oracle, N, psi):
    R = oracle.Ring(N)
    assert R.psi == psi
    logN = N.bit_length() - 1
    for i in [0, 1, 5, N // 2 + 3, N - 1]:
        a = np.zeros(N, dtype=np.uint64)
        a[i] = 1
        out = R.ntt(a)
        for k in [0, 1, 2, 17, N // 2, N - 1]:
            assert int(out[k]) == pow(psi, (2 * brv(k, logN) + 1) * i, Q)


@pytest.mark.parametrize("N", [128, 1024, 2048])
def test_ntt_roundtrip_and_zero(oracle, N):
    R = oracle.Ring(N)
    rng = np.random.default_rng(7)
    a = rng.integers(0, Q, size=3 * N, dtype=np.uint64)
    assert np.array_equal(R.intt(R.ntt(a)), a)
    z = np.zeros(N, dtype=np.uint64)
    assert not R.ntt(z).any()


def test_convolution_theorem(oracle):
    N = 128
    R = oracle.Ring(N)
    rng = np.random.default_rng(11)
    a = rng.integers(0, Q, size=N, dtype=np.uint64)
    b = rng.integers(0, Q, size=N, dtype=np.uint64)
    # schoolbook in pure Python
    c = [0] * N
    for i in range(N):
        for j in range(N):
            p = int(a[i]) * int(b[j])
            if i + j < N:
                c[i + j] += p
            else:
                c[i + j - N] -= p
    c = np.array([x % Q for x in c], dtype=np.uint64)
    A, B = R.ntt(a), R.ntt(b)
    prod = np.array([(int(x) * int(y)) % Q for x, y in zip(A, B)], dtype=np.uint64)
    assert np.array_equal(R.intt(prod), c)
    assert np.array_equal(R.negacyclic_mul(a, b), c)


# --------------------------------------------------------------------------
# Gadget decomposition (mux_operator.cpp:23-84, gadget_decomp.h:16-205)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("b,C", [(16, 0x8000800080008000), (14, 0x8002000800200080), (12, 0x8008008008008008),
                                 (5, 0x8421084210842108), (4, 0x8888888888888888)])
def test_signed_corrector_constants(oracle, b, C):
    assert oracle.lib().or_signed_corrector(b, None) == C


@pytest.mark.parametrize("b,l,shifts", [(16, 4, [48, 32, 16, 0]), (12, 3, [60, 48, 36]), (14, 3, [56, 42, 28])])
def test_unsigned_decomposition(oracle, b, l, shifts):
    rng = np.random.default_rng(b)
    x = np.concatenate([rng.integers(0, Q, size=200, dtype=np.uint64),
                        np.array([0, 1, Q - 1, 2**63, 2**64 - 2**32], dtype=np.uint64)])
    out = np.zeros(l * x.size, dtype=np.uint64)
    oracle.lib().or_decompose_unsigned(oracle.P(out), oracle.P(x), x.size, b, l)
    out = out.reshape(l, x.size)
    for i, xv in enumerate(x):
        xv = int(xv)
        rec = sum(int(out[j, i]) << shifts[j] for j in range(l))
        assert rec == (xv >> shifts[-1]) << shifts[-1]  # top bits, low shifts[-1] bits dropped
        assert all(int(out[j, i]) < (1 << b) for j in range(l))


@pytest.mark.parametrize("b,l", [(16, 4), (16, 3), (12, 3), (14, 3), (8, 8), (12, 6)])
def test_signed_decomposition(oracle, b, l):
    rng = np.random.default_rng(100 + b)
    x = np.concatenate([rng.integers(0, Q, size=300, dtype=np.uint64),
                        np.array([0, 1, Q - 1, Q // 2, 2**63, 2**32 - 1], dtype=np.uint64)])
    out = np.zeros(l * x.size, dtype=np.uint64)
    oracle.lib().or_decompose_signed(oracle.P(out), oracle.P(x), x.size, b, l)
    out = out.reshape(l, x.size)
    D = -(-64 // b)
    w = [1 << max(64 - b * (j + 1), 0) for j in range(l)]
    for i, xv in enumerate(x):
        digs = [int(out[j, i]) if int(out[j, i]) < Q // 2 else int(out[j, i]) - Q for j in range(l)]
        for j, d in enumerate(digs):
            half = 1 << (b - 1) if j != D - 1 else 1 << ((64 - b * D) + b - 1)
            assert -half <= d < half
        rec = sum(wj * d for wj, d in zip(w, digs)) % Q
        err = (int(xv) - rec) % Q
        err = err if err < Q // 2 else err - Q
        if l == D:
            assert err == 0  # full signed decomposition recomposes exactly
        else:
            assert abs(err) <= (1 << (64 - b * l))


# --------------------------------------------------------------------------
# Blind rotation: the reference tests' noise-free invariant
# (cggi_blind_rotation_test.cpp:118-123, 169-174; bmmp_blind_rotation_tests.cpp:103-108)
# --------------------------------------------------------------------------

def _noise_free_br(oracle, N, n, b, l, method, dist, seed):
    lwe_key = oracle.secret(n, dist, seed, 1)
    rlwe_key = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lwe_key, rlwe_key, n, b, l, method, 0 if method == oracle.BINARY else 1, 0.0, seed)
    rng = np.random.default_rng(seed)
    lwe = rng.integers(0, 2 * N, size=n + 1, dtype=np.uint64)
    msg = [i for i in range(N)]
    acc_ntt = oracle.encrypt_rlwe(N, np.array(msg, dtype=np.uint64), rlwe_key, 0.0, seed, 9)
    R = oracle.Ring(N)
    acc_coef = R.intt(acc_ntt)
    out = oracle.blind_rotate(N, lwe, acc_coef, brk, n, b, l, method, False)
    phase = oracle.phase_rlwe(N, out, rlwe_key)
    ks = [int(k) if int(k) < 2 else -1 for k in lwe_key]
    e = (int(lwe[n]) - sum(int(a) * s for a, s in zip(lwe[:n], ks))) % (2 * N)
    expect = poly_phase_rotate(msg, e, N)
    return phase, expect, out, R


@pytest.mark.parametrize("method,dist,n", [(0, 0, 10), (1, 1, 10), (2, 0, 10)])
def test_blind_rotate_noise_free(oracle, method, dist, n):
    phase, expect, out, R = _noise_free_br(oracle, 128, n, 16, 4, method, dist, seed=5 + method)
    assert [int(x) for x in phase] == expect


def test_blind_rotate_coef_output_matches_ntt_output(oracle):
    N, n = 128, 6
    _, _, out_ntt, R = _noise_free_br(oracle, N, n, 16, 4, 0, 0, seed=21)
    lwe_key = oracle.secret(n, 0, 21, 1)
    rlwe_key = oracle.secret(N, 0, 21, 2)
    brk = oracle.brk(N, lwe_key, rlwe_key, n, 16, 4, 0, 0, 0.0, 21)
    rng = np.random.default_rng(21)
    lwe = rng.integers(0, 2 * N, size=n + 1, dtype=np.uint64)
    acc_coef = R.intt(oracle.encrypt_rlwe(N, np.arange(N, dtype=np.uint64), rlwe_key, 0.0, 21, 9))
    out_coef = oracle.blind_rotate(N, lwe, acc_coef, brk, n, 16, 4, 0, True)
    assert np.array_equal(R.intt(out_ntt), out_coef)


def test_sample_extract_phase(oracle):
    N = 128
    R = oracle.Ring(N)
    s = oracle.secret(N, 0, 3, 2)
    rng = np.random.default_rng(3)
    m = rng.integers(0, Q, size=N, dtype=np.uint64)
    ct = oracle.encrypt_rlwe(N, m, s, 0.0, 3, 4)
    coef = R.intt(ct)
    lwe = oracle.sample_extract(coef, N)
    assert oracle.phase_lwe(lwe, s, N) == int(m[0])


@pytest.mark.parametrize("b,d,exact", [(4, 16, True), (4, 14, False), (5, 12, False), (8, 8, True)])
def test_lwe_keyswitch_noise_free(oracle, b, d, exact):
    n_in, n_out = 64, 20
    src = oracle.secret(n_in, 0, 8, 1)
    dst = oracle.secret(n_out, 1, 8, 2)
    ksk = oracle.lwe_ksk(src, dst, n_in, n_out, b, d, 0.0, 8)
    rng = np.random.default_rng(8)
    a = rng.integers(0, Q, size=n_in, dtype=np.uint64)
    m = (Q // 8) * 3
    bb = (m + sum(int(x) * int(s) for x, s in zip(a, src))) % Q
    lwe_in = np.concatenate([a, np.array([bb], dtype=np.uint64)])
    out = oracle.lwe_keyswitch(lwe_in, ksk, n_in, n_out, b, d)
    ph = oracle.phase_lwe(out, dst, n_out)
    err = (ph - m) % Q
    err = err if err < Q // 2 else err - Q
    if exact:
        assert err == 0
    else:
        assert abs(err) <= n_in * (1 << (64 - b * d))


def test_gaussian_sampler_moments(oracle):
    out = np.zeros(200000, dtype=np.uint64)
    oracle.lib().or_gauss_samples(oracle.P(out), out.size, 3.2, 1, 77)
    v = np.where(out > Q // 2, out.astype(np.float64) - float(Q), out.astype(np.float64))
    v = np.array([int(x) - Q if int(x) > Q // 2 else int(x) for x in out[:200000]], dtype=np.float64)
    assert abs(v.mean()) < 0.05
    assert abs(v.std() - 3.2) < 0.05
    assert np.abs(v).max() <= 39


def test_gadget_kinds(oracle):
    assert list(oracle.gadget(12, 3, 0)) == [2**60, 2**48, 2**36]
    assert list(oracle.gadget(12, 3, 1)) == [2**52, 2**40, 2**28]
    assert list(oracle.gadget(16, 4, 0)) == list(oracle.gadget(16, 4, 1)) == [2**48, 2**32, 2**16, 1]
    assert list(oracle.gadget(14, 3, 0)) == [2**56, 2**42, 2**28]


# --------------------------------------------------------------------------
# RLWE key switching, automorphism, trace, packing: the reference tests' noise-free invariants
# (tests/automorphism_test.cpp:80-82, trace_tests.cpp:71-74, packing_tests.cpp:65-67,99-104)
# --------------------------------------------------------------------------

def _phase_coef(oracle, N, ct_ntt, s):
    return [int(x) for x in oracle.phase_rlwe(N, ct_ntt, s)]


def test_automorphism_poly_closed_form(oracle):
    N = 64
    rng = np.random.default_rng(1)
    a = rng.integers(0, Q, size=N, dtype=np.uint64)
    for t in (3, 5, 17, 2 * N - 1, 257 % (2 * N)):
        out = oracle.automorphism_poly(a, t, N)
        # tau_t(a)(X) = a(X^t) mod X^N + 1, checked monomial by monomial
        ref = [0] * N
        for i in range(N):
            e = (i * t) % (2 * N)
            if e < N:
                ref[e] = (ref[e] + int(a[i])) % Q
            else:
                ref[e - N] = (ref[e - N] - int(a[i])) % Q
        assert [int(x) for x in out] == ref


@pytest.mark.parametrize("t", [3, 5, 129, 257])
def test_automorphism_noise_free(oracle, t):
    N, b, d = 128, 16, 4
    s = oracle.secret(N, 0, 4, 2)
    R = oracle.Ring(N)
    s_t = oracle.automorphism_poly(s, t, N)
    ksk = oracle.rlwe_ksk(N, s_t, s, b, d, 0.0, 4)
    rng = np.random.default_rng(t)
    m = rng.integers(0, Q, size=N, dtype=np.uint64)
    ct = R.intt(oracle.encrypt_rlwe(N, m, s, 0.0, 4, 8))  # coefficients
    out = oracle.automorphism(N, ct, t, ksk, b, d)
    assert _phase_coef(oracle, N, out, s) == [int(x) for x in oracle.automorphism_poly(m, t, N)]


def test_trace_noise_free(oracle):
    N, b, d = 128, 16, 4
    s = oracle.secret(N, 0, 9, 2)
    keys = oracle.auto_keys(N, s, b, d, 0.0, 9)
    rng = np.random.default_rng(9)
    m = rng.integers(0, Q, size=N, dtype=np.uint64)
    ct = oracle.Ring(N).intt(oracle.encrypt_rlwe(N, m, s, 0.0, 9, 8))
    out = oracle.trace(N, ct, keys, b, d)
    ph = _phase_coef(oracle, N, out, s)
    assert ph[0] == int(m[0]) and not any(ph[1:])


@pytest.mark.parametrize("length", [128, 16, 2])
def test_pack_noise_free(oracle, length):
    N, b, d = 128, 16, 4
    s = oracle.secret(N, 0, 12, 2)
    keys = oracle.auto_keys(N, s, b, d, 0.0, 12)
    rng = np.random.default_rng(length)
    msgs = rng.integers(0, Q, size=length, dtype=np.uint64)
    lwes = np.zeros((length, N + 1), dtype=np.uint64)
    for i in range(length):
        a = rng.integers(0, Q, size=N, dtype=np.uint64)
        bb = (int(msgs[i]) + sum(int(x) * int(y) for x, y in zip(a, s))) % Q
        lwes[i, :N] = a
        lwes[i, N] = bb
    out = oracle.pack(N, lwes, length, keys, b, d)
    R = oracle.Ring(N)
    ph = _phase_coef(oracle, N, np.concatenate([R.ntt(out[:N]), R.ntt(out[N:])]), s)
    stride = N // length
    for j in range(N):
        want = int(msgs[j // stride]) if j % stride == 0 else 0
        assert ph[j] == want, j


def _noise_free_lwes(oracle, N, s, msgs, seed):
    rng = np.random.default_rng(seed)
    lwes = np.zeros((len(msgs), N + 1), dtype=np.uint64)
    for i, m in enumerate(msgs):
        a = rng.integers(0, Q, size=N, dtype=np.uint64)
        lwes[i, :N] = a
        lwes[i, N] = (int(m) + sum(int(x) * int(y) for x, y in zip(a, s))) % Q
    return lwes


@pytest.mark.parametrize("length", [1, 3, 16, 77, 128])
def test_pack_consecutive_noise_free(oracle, length):
    """PackConsecutively: coefficient i holds message i for i < len and 0 beyond (packing_tests.cpp:99-104)."""
    N, b, d = 128, 16, 4
    s = oracle.secret(N, 0, 13, 2)
    keys = oracle.auto_keys(N, s, b, d, 0.0, 13)
    msgs = np.random.default_rng(length).integers(0, Q, size=length, dtype=np.uint64)
    lwes = _noise_free_lwes(oracle, N, s, msgs, length + 100)
    out = oracle.pack_consecutive(N, lwes, length, keys, b, d)
    R = oracle.Ring(N)
    ph = _phase_coef(oracle, N, np.concatenate([R.ntt(out[:N]), R.ntt(out[N:])]), s)
    for j in range(N):
        assert ph[j] == (int(msgs[j]) if j < length else 0), j


def test_pack_consecutive_full_equals_pack(oracle):
    """With len = N nothing is pruned: the stack walk equals Pack's level order, bit for bit (with noise)."""
    N, b, d = 128, 16, 4
    s = oracle.secret(N, 0, 14, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 14)
    lwes = np.random.default_rng(9).integers(0, Q, size=(N, N + 1), dtype=np.uint64)
    assert np.array_equal(oracle.pack_consecutive(N, lwes, N, keys, b, d), oracle.pack(N, lwes, N, keys, b, d))


@pytest.mark.parametrize("basebits,digits", [(16, 4), (10, 4)])
def test_lwe_to_rgsw_noise_free(oracle, basebits, digits):
    """LWE(m q/2) -> RGSW(m): with noise-free keys, lev_right row i has phase m B^(i + D - digits) and
    lev_left row i phase -s m B^(i + D - digits), D = ceil(64 / log2 B) (scheme_switch_tests.cpp:100-119,
    170-188)."""
    N, n, b, l = 1024, 16, 16, 4
    lk = oracle.secret(n, 0, 21, 1)
    s = oracle.secret(N, 0, 21, 2)
    brk = oracle.brk(N, lk, s, n, b, l, oracle.BINARY, 0, 0.0, 21)
    tkeys = oracle.auto_keys(N, s, 16, 4, 0.0, 21)
    R = oracle.Ring(N)
    s_ntt = R.ntt(s)
    s2 = R.intt(np.array([int(x) * int(x) % Q for x in s_ntt], dtype=np.uint64))
    sq = oracle.rlwe_ksk(N, s2, s, 16, 4, 0.0, 21)
    D = -(-64 // basebits)
    rng = np.random.default_rng(basebits)
    for m in (0, 1):
        a = rng.integers(0, 2 * N, size=n)
        bb = (int(sum(int(x) * int(y) for x, y in zip(a, lk))) + m * N) % (2 * N)
        lwe = np.array(list(a) + [bb], dtype=np.uint64)
        out = oracle.lwe_to_rgsw(N, lwe, brk, n, b, l, oracle.BINARY, tkeys, 16, 4, sq, 16, 4, digits, basebits)
        for i in range(2 * digits):
            row = out[i * 2 * N:(i + 1) * 2 * N]
            ph = oracle.phase_rlwe(N, row, s)
            scal = (m * pow(2, basebits * (i % digits + D - digits), Q)) % Q
            if i >= digits:
                want = [scal] + [0] * (N - 1)
            else:
                want = [(Q - scal * int(x)) % Q for x in s]
            for j in range(N):
                err = (int(ph[j]) - want[j]) % Q
                err = err if err < Q // 2 else err - Q
                assert abs(err) <= 1, (m, i, j, err)


def _sap_keys(oracle, N, n, seed, sigma=0.0):
    lk = oracle.secret(n, 0, seed, 1)
    s = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lk, s, n, 16, 4, oracle.BINARY, 0, sigma, seed)
    keys = oracle.auto_keys(N, s, 16, 4, sigma, seed)
    return lk, s, brk, keys


def test_sap_homtrunc_noise_free(oracle):
    """HomTrunc of a trivial RLWE with phase X^e keeps X^(e / radix) (sap_decomposer_tests.cpp:62-126)."""
    N, radix = 1024, 32
    _, s, _, keys = _sap_keys(oracle, N, 4, 41)
    R = oracle.Ring(N)
    for e in (0, 5, 31, 32, 100, 517, 1023):
        x = np.zeros(2 * N, dtype=np.uint64)
        x[N + e] = 1
        x[N:] = R.ntt(x[N:])
        out = oracle.sap_homtrunc(N, x, radix, keys, 16, 4)
        ph = oracle.phase_rlwe(N, out, s)
        want = np.zeros(N, dtype=np.uint64)
        want[e // radix] = 1
        assert np.array_equal(ph, want), e


def test_sap_reset_noise_free(oracle):
    """ResetAccumulator (sap_decomposer.cpp:314-373) of the trivial RLWE Q/2N X^e with noise-free keys.

    The extraction poly leaves coefficient 0 = e' Q/2N with e' = e - e mod radix; key switching a
    trivial sample keeps it trivial; FLOOR to Z_2N gives floor(2N e' floor(Q/2N) / Q), which is
    e' - 1 for e' > 0 (exact rationals and the x87 quotient agree here); the noise-free blind
    rotation of Q/2N X^0 then returns exactly Q/2N X^(that).  (The reference's TestResetAccumulator,
    sap_decomposer_tests.cpp:128-165, expects the block of e itself; on this noise-free trivial
    input its own code lands one coefficient lower, in the block below, as here.)"""
    N, n, radix = 1024, 10, 32
    lk, s, brk, _ = _sap_keys(oracle, N, n, 42)
    ksk = oracle.lwe_ksk(s, lk, N, n, 16, 4, 0.0, 42)
    R = oracle.Ring(N)
    step = Q // (2 * N)
    for e in (0, 5, 37, 300, 777, 1023):
        x = np.zeros(2 * N, dtype=np.uint64)
        x[N + e] = step
        x[N:] = R.ntt(x[N:])
        acc = oracle.sap_reset(N, x, radix, ksk, n, 16, 4, Q, brk, n, 16, 4, oracle.BINARY)
        ep = e - e % radix
        idx = (2 * N * ep * step) // Q
        assert idx == (ep - 1 if ep else 0)
        want = np.zeros(N, dtype=np.uint64)
        want[idx] = step
        assert np.array_equal(oracle.phase_rlwe(N, acc, s), want), e


def test_sap_decompose_noise_free(oracle):
    """Digit i of the output decrypts to digit i (radix 32) of the message of an LWE modulo 32^8
    (sap_decomposer_tests.cpp:167-226)."""
    N, n, radix = 1024, 10, 32
    lk, s, brk, keys = _sap_keys(oracle, N, n, 43)
    beta = int(np.count_nonzero(lk))
    qin = radix ** 8
    rng = np.random.default_rng(43)
    for _ in range(2):
        msg = int(rng.integers(0, qin))
        a = [int(v) for v in rng.integers(0, qin, size=n)]
        b = (sum(x * int(y) for x, y in zip(a, lk)) + msg) % qin
        lwe = np.array(a + [b], dtype=np.uint64)
        out = oracle.sap_decompose(N, lwe, radix, beta, brk, n, 16, 4, oracle.BINARY, keys, 16, 4)
        assert out.shape[0] >= 8
        for i in range(8):
            ph = oracle.phase_lwe(out[i], s, N)
            got = int(oracle.modswitch(np.array([ph], dtype=np.uint64), Q, 2 * N, 0)[0])
            assert got == (msg >> (5 * i)) % radix, (i, got)
