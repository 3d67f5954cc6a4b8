"""GPU parity of the RLWE key switching path: automorphism, trace and n-LWE -> RLWE packing,
bit for bit against the oracle, plus decryption of the packed slots."""
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


def _ctx(tp, N, b, d):
    return tp.Context(N, 2, 16, 4, 0, auto_base_log2=b, auto_digits=d)


def test_auto_keygen_matches_oracle(tp, oracle):
    say("auto keygen")
    N = 1024
    s = oracle.secret(N, 0, 3, 2)
    assert np.array_equal(tp.keygen_auto_keys(s, N, 16, 4, 3.2, 5), oracle.auto_keys(N, s, 16, 4, 3.2, 5))


@pytest.mark.parametrize("N,b,d", [(1024, 16, 4), (2048, 16, 4), (2048, 12, 3)])
def test_automorphism_vs_oracle(tp, oracle, N, b, d):
    say("automorphism", N, b, d)
    s = oracle.secret(N, 0, 7, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 7)
    ctx = _ctx(tp, N, b, d)
    ctx.import_auto_keys(keys)
    rng = np.random.default_rng(N + b)
    x = rng.integers(0, Q, size=(5, 2 * N), dtype=np.uint64)
    logN = N.bit_length() - 1
    kw = d * 2 * N
    for j in (1, 2, logN):
        t = (1 << j) + 1
        got = ctx.automorphism(x, t)
        for c in range(x.shape[0]):
            ref = oracle.automorphism(N, x[c], t, keys[(j - 1) * kw: j * kw], b, d)
            assert np.array_equal(got[c], ref), (t, c)


@pytest.mark.parametrize("N", [1024, 2048])
def test_trace_vs_oracle(tp, oracle, N):
    say("trace", N)
    b, d = 16, 4
    s = oracle.secret(N, 0, 8, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 8)
    ctx = _ctx(tp, N, b, d)
    ctx.import_auto_keys(keys)
    rng = np.random.default_rng(N)
    x = rng.integers(0, Q, size=(3, 2 * N), dtype=np.uint64)
    got = ctx.trace(x)
    for c in range(x.shape[0]):
        assert np.array_equal(got[c], oracle.trace(N, x[c], keys, b, d)), c


@pytest.mark.parametrize("N,length", [(1024, 1024), (1024, 64), (2048, 2048), (1024, 2)])
def test_pack_vs_oracle_and_decrypts(tp, oracle, N, length):
    say("pack", N, length)
    b, d = 16, 4
    s = oracle.secret(N, 0, 10, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 10)
    ctx = _ctx(tp, N, b, d)
    ctx.import_auto_keys(keys)
    rng = np.random.default_rng(length)
    delta = Q // 16
    msgs = rng.integers(0, 16, size=length, dtype=np.uint64)
    a = rng.integers(0, Q, size=(length, N), dtype=np.uint64)
    s_obj = [int(v) for v in s]
    lwes = np.zeros((length, N + 1), dtype=np.uint64)
    lwes[:, :N] = a
    e = rng.integers(-4, 5, size=length)
    for i in range(length):
        inner = sum(int(x) * y for x, y in zip(a[i], s_obj))
        lwes[i, N] = (inner + int(msgs[i]) * delta + int(e[i])) % Q
    perm = rng.permutation(length).astype(np.uint32)
    got = ctx.pack(lwes, perm)
    ref = oracle.pack(N, np.ascontiguousarray(lwes[perm]), length, keys, b, d)
    assert np.array_equal(got, ref)
    R = oracle.Ring(N)
    ph = oracle.phase_rlwe(N, np.concatenate([R.ntt(got[:N]), R.ntt(got[N:])]), s)
    stride = N // length
    for j in range(length):
        v = int(ph[j * stride])
        dec = ((v + delta // 2) // delta) % 16
        assert dec == int(msgs[perm[j]]), j
