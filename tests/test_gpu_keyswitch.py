"""GPU parity of the RLWE key-switch family, consecutive packing and elementwise arithmetic.

RLWEtoRLWEConverter::Convert, AutomorphismEvaluator::Eval / EvalSwitchOnly (any odd t, overwrite
and the reference's subtract-into-output semantics), LWEtoRLWEPacker::PackConsecutively and the
MathWorker elementwise ops, each bit-exact against the oracle (or exact integer arithmetic).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 0xFFFFFFFF00000001


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tp():
    import fhetoothpaste_amd as tp
    assert tp.device_count() >= 1, "no gfx950 device"
    return tp


def _canon(rng, shape):
    return rng.integers(0, Q, size=shape, dtype=np.uint64)


def _ctx(tp, N, b, d):
    return tp.Context(N, 2, 12, 3, 0, 0, 0, 0, 0, b, d)


@pytest.mark.parametrize("N,b,d", [(1024, 16, 4), (2048, 12, 3)])
def test_rlwe_ksk_keygen_matches_oracle(tp, oracle, N, b, d):
    say("rlwe ksk keygen", N, b, d)
    src = oracle.secret(N, 0, 31, 2)
    dst = oracle.secret(N, 0, 32, 2)
    assert np.array_equal(tp.keygen_rlwe_ksk(src, dst, N, b, d, 3.2, 33), oracle.rlwe_ksk(N, src, dst, b, d, 3.2, 33))


@pytest.mark.parametrize("N,b,d", [(1024, 16, 4), (2048, 16, 4), (2048, 12, 3)])
def test_rlwe_convert_vs_oracle(tp, oracle, N, b, d):
    say("rlwe convert", N, b, d)
    rng = np.random.default_rng(N + b)
    key = oracle.rlwe_ksk(N, oracle.secret(N, 0, 1, 2), oracle.secret(N, 0, 2, 2), b, d, 3.2, 3)
    ctx = _ctx(tp, N, b, d)
    ctx.import_rlwe_ksk(key)
    x = _canon(rng, (5, 2 * N))  # [a coef | b NTT]
    x[1, :N] = 0
    x[2, :N] = Q - 1
    prev = _canon(rng, (5, 2 * N))
    got_over = ctx.rlwe_keyswitch(x, 1, 0, 0)
    got_acc = ctx.rlwe_keyswitch(x, 1, 0, 0, out=prev)
    for c in range(5):
        assert np.array_equal(got_over[c], oracle.rlwe_keyswitch(N, np.zeros(2 * N, np.uint64), x[c], key, b, d)), c
        assert np.array_equal(got_acc[c], oracle.rlwe_keyswitch(N, prev[c], x[c], key, b, d)), c


@pytest.mark.parametrize("N,t", [(1024, 3), (1024, 2047), (2048, 5), (2048, 1025), (2048, 77)])
def test_automorphism_modes_vs_oracle(tp, oracle, N, t):
    b, d = 16, 4
    say("automorphism modes", N, t)
    rng = np.random.default_rng(t)
    s = oracle.secret(N, 0, 5, 2)
    key = oracle.rlwe_ksk(N, oracle.automorphism_poly(s, t, N), s, b, d, 3.2, 6)
    ctx = _ctx(tp, N, b, d)
    ctx.import_rlwe_ksk(key)
    R = oracle.Ring(N)
    x = _canon(rng, (3, 2 * N))
    prev = _canon(rng, (3, 2 * N))
    ev = ctx.rlwe_keyswitch(x, t, 1, 0)
    ev_acc = ctx.rlwe_keyswitch(x, t, 1, 0, out=prev)
    sw = ctx.rlwe_keyswitch(x, t, 2, 0)
    sw_acc = ctx.rlwe_keyswitch(x, t, 2, 0, out=prev)
    zero = np.zeros(2 * N, np.uint64)
    for c in range(3):
        ta = oracle.automorphism_poly(x[c, :N], t, N)
        tb = R.ntt(oracle.automorphism_poly(x[c, N:], t, N))
        buf = np.concatenate([ta, tb])
        buf0 = np.concatenate([ta, zero[:N]])
        assert np.array_equal(ev[c], oracle.automorphism(N, x[c], t, key, b, d)), c
        assert np.array_equal(ev_acc[c], oracle.rlwe_keyswitch(N, prev[c], buf, key, b, d)), c
        assert np.array_equal(sw[c], oracle.rlwe_keyswitch(N, zero, buf0, key, b, d)), c
        assert np.array_equal(sw_acc[c], oracle.rlwe_keyswitch(N, prev[c], buf0, key, b, d)), c


def test_automorphism_key_index_matches_automorphism_entry(tp, oracle):
    N, b, d = 1024, 16, 4
    say("automorphism key index")
    s = oracle.secret(N, 0, 8, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 8)
    ctx = _ctx(tp, N, b, d)
    ctx.import_auto_keys(keys)
    x = _canon(np.random.default_rng(8), (2, 2 * N))
    for j in (1, 4, 10):
        t = (1 << j) + 1
        got = ctx.rlwe_keyswitch(x, t, 1, j)
        assert np.array_equal(got, ctx.automorphism(x, t)), j
        kj = keys[(j - 1) * d * 2 * N: j * d * 2 * N]
        assert np.array_equal(got[0], oracle.automorphism(N, x[0], t, kj, b, d)), j


def test_rlwe_keyswitch_rejects(tp):
    from fhetoothpaste_amd import ToothpasteError
    say("rlwe keyswitch rejects")
    ctx = _ctx(tp, 1024, 16, 4)
    x = np.zeros((1, 2048), np.uint64)
    with pytest.raises(ToothpasteError):
        ctx.rlwe_keyswitch(x, 1, 0, 0)  # no key imported
    ctx.import_rlwe_ksk(np.zeros(4 * 2048, np.uint64))
    for t, mode, key in [(2, 1, 0), (3, 0, 0), (2049, 1, 0), (3, 1, 11), (3, 3, 0)]:
        with pytest.raises(ToothpasteError):
            ctx.rlwe_keyswitch(x, t, mode, key)


@pytest.mark.parametrize("N,length", [(1024, 1), (1024, 5), (1024, 64), (1024, 1000), (2048, 2048), (2048, 300)])
def test_pack_consecutive_vs_oracle(tp, oracle, N, length):
    b, d = 16, 4
    say("pack consecutive", N, length)
    s = oracle.secret(N, 0, 40, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 40)
    ctx = _ctx(tp, N, b, d)
    ctx.import_auto_keys(keys)
    rng = np.random.default_rng(length)
    msgs = rng.integers(0, 16, size=length).astype(np.uint64) * np.uint64(Q // 16)
    a = _canon(rng, (length, N))
    e = rng.integers(-8, 9, size=length)
    lwes = np.zeros((length, N + 1), np.uint64)
    lwes[:, :N] = a
    for i in range(length):
        lwes[i, N] = (int(msgs[i]) + int(e[i]) + sum(int(x) * int(y) for x, y in zip(a[i], s))) % Q
    got = ctx.pack_consecutive(lwes)
    ref = oracle.pack_consecutive(N, lwes, length, keys, b, d)
    assert np.array_equal(got, ref)
    # decryption: coefficient i = message i (i < len), 0 beyond
    R = oracle.Ring(N)
    ph = oracle.phase_rlwe(N, np.concatenate([R.ntt(got[:N]), R.ntt(got[N:])]), s)
    for i in range(N):
        want = int(msgs[i]) if i < length else 0
        err = (int(ph[i]) - want) % Q
        err = err if err < Q // 2 else err - Q
        assert abs(err) < Q // 64, i


def test_eltwise_ops_exact(tp):
    say("eltwise")
    ctx = _ctx(tp, 1024, 16, 4)
    rng = np.random.default_rng(1)
    n = 100003  # not a multiple of the work-group size
    a, b, c = (_canon(rng, n) for _ in range(3))
    a[:4] = [0, 1, Q - 1, Q - 2]
    b[:4] = [Q - 1, Q - 1, Q - 1, 3]
    s = 0xFEDCBA9876543210 % Q
    A, B, C = (x.astype(object) for x in (a, b, c))
    want = {
        0: (A + B) % Q, 1: (A - B) % Q, 2: (A * B) % Q, 3: (A * B + C) % Q, 4: (A * s) % Q,
    }
    for op, w in want.items():
        got = ctx.eltwise(op, a, b if op in (0, 1, 2, 3) else None, c if op == 3 else None, s if op == 4 else 0)
        assert np.array_equal(got.astype(object), w.astype(object)), op
    raw = rng.integers(0, 2 ** 64, size=n, dtype=np.uint64)
    raw[:3] = [Q, Q + 5, 2 ** 64 - 1]
    got = ctx.eltwise(5, raw)
    assert np.array_equal(got.astype(object), raw.astype(object) % Q)


@pytest.mark.parametrize("tM", [2048, 4096, 1 << 20])
def test_modswitch_vs_long_double(tp, oracle, tM):
    """ModulusSwitch on the GPU == the reference's x87 long double arithmetic (oracle), near ties included."""
    say("modswitch", tM)
    ctx = _ctx(tp, 1024, 16, 4)
    rng = np.random.default_rng(tM)
    v = list(rng.integers(0, Q, size=20000, dtype=np.uint64))
    for k in rng.integers(1, tM, size=3000):
        k = int(k)
        for d in (-1, 0, 1):
            v.append((((2 * k + 1) * Q) // (2 * tM) + d) % Q)  # near half-integers
            v.append(((k * Q + tM - 1) // tM + d) % Q)          # near integers
    v += [0, 1, Q - 1, Q - 2, (1 << 63), (1 << 63) - 1]
    v = np.array(v, dtype=np.uint64)
    for kind in range(3):
        assert np.array_equal(ctx.modswitch(v, Q, tM, kind), oracle.modswitch(v, Q, tM, kind)), kind
    r = ctx.modswitch(v, Q, tM, 3, seed=5)
    assert np.all((r == oracle.modswitch(v, Q, tM, 1)) | (r == oracle.modswitch(v, Q, tM, 2)))
    assert np.array_equal(r, ctx.modswitch(v, Q, tM, 3, seed=5))  # the seeded coin is reproducible


@pytest.mark.parametrize("digits,basebits,b,d", [(4, 10, 16, 4), (2, 20, 16, 4), (4, 14, 14, 4)])
def test_lwe_to_rgsw_batch_vs_oracle(tp, oracle, digits, basebits, b, d):
    """Batched LWEtoRGSWConverter::Convert (lwe_to_rgsw_conversion.cpp:163-231): every output word
    equals the oracle's, for each input of the batch."""
    N, n = 1024, 16
    say("lwe to rgsw batch", digits, basebits, b, d)
    lk = oracle.secret(n, 0, 120 + digits, 1)
    s = oracle.secret(N, 0, 120 + digits, 2)
    brk = oracle.brk(N, lk, s, n, b, d, oracle.BINARY, 0, 3.2, 121)
    tkeys = oracle.auto_keys(N, s, b, d, 3.2, 122)
    R = oracle.Ring(N)
    s2 = R.intt(np.array([int(x) * int(x) % Q for x in R.ntt(s)], dtype=np.uint64))
    sq = oracle.rlwe_ksk(N, s2, s, b, d, 3.2, 123)
    ctx = tp.Context(N, n, b, d, 0, 0, 0, 0, 0, b, d)
    ctx.import_brk(brk)
    ctx.import_auto_keys(tkeys)
    ctx.import_rlwe_ksk(sq)
    rng = np.random.default_rng(digits + basebits)
    lwes = rng.integers(0, 2 * N, size=(3, n + 1), dtype=np.uint64)
    got = ctx.lwe_to_rgsw(lwes, digits, basebits)
    for c in range(3):
        ref = oracle.lwe_to_rgsw(N, lwes[c], brk, n, b, d, oracle.BINARY, tkeys, b, d, sq, b, d, digits, basebits)
        assert np.array_equal(got[c], ref), c
