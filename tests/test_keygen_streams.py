"""CPU tests of the key-generation word streams and noise sampler (fhetoothpaste_amd/csrc/keygen.cpp).

* The entropy stream (seed TP_SEED_OS_ENTROPY) is ChaCha20: its block function is checked against
  the test vector of RFC 8439 section 2.3.2 and against an independent Python restatement of the
  RFC's algorithm on random keys, nonces and counters.
* Keys drawn from the entropy stream differ from call to call and keep their distribution.
* The discrete Gaussian keeps its variance for a small sigma (inversion table) and for a large one
  (Canonne-Kamath-Steinke rejection sampler), and its seeded samples equal the oracle's, draw for
  draw, on both paths.
* The Python key-generation wrappers reject arrays of the wrong length before calling the library.
"""
import numpy as np
import pytest

import fhetoothpaste_amd as tp
from fhetoothpaste_amd import operators as ops
from oracle import sample_gaussian as oracle_gaussian

MASK = 0xFFFFFFFF


def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & MASK


def _chacha_block_py(key, counter, nonce):
    """RFC 8439 section 2.3, restated from the RFC text."""
    state = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574] + list(key) + [counter] + list(nonce)
    x = list(state)

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]) & MASK; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & MASK; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & MASK; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & MASK; x[b] = _rotl(x[b] ^ x[c], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return [(x[i] + state[i]) & MASK for i in range(16)]


def test_chacha20_rfc8439_vector():
    key = np.frombuffer(bytes(range(32)), dtype="<u4")
    nonce = np.frombuffer(bytes.fromhex("000000090000004a00000000"), dtype="<u4")
    out = ops.chacha20_block(key, 1, nonce)
    serialized = out.astype("<u4").tobytes()
    assert serialized.hex() == (
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")


def test_chacha20_matches_python_restatement():
    rng = np.random.default_rng(5)
    for _ in range(32):
        key = rng.integers(0, 2**32, 8, dtype=np.uint64).astype(np.uint32)
        nonce = rng.integers(0, 2**32, 3, dtype=np.uint64).astype(np.uint32)
        ctr = int(rng.integers(0, 2**32))
        got = ops.chacha20_block(key, ctr, nonce)
        want = _chacha_block_py([int(k) for k in key], ctr, [int(n) for n in nonce])
        assert [int(v) for v in got] == want


def test_entropy_stream_differs_per_call_and_keeps_distribution():
    a = tp.keygen_secret(4096, 0, ops.SEED_OS_ENTROPY, 1)
    b = tp.keygen_secret(4096, 0, ops.SEED_OS_ENTROPY, 1)
    assert not np.array_equal(a, b)
    for s in (a, b):
        assert set(np.unique(s)) <= {0, 1}
        assert abs(float(s.mean()) - 0.5) < 0.05
    t = tp.keygen_secret(8192, 1, ops.SEED_OS_ENTROPY, 1)
    frac = {v: float((t == v).mean()) for v in (0, 1, tp.Q - 1)}
    assert abs(frac[0] - 0.5) < 0.04 and abs(frac[1] - 0.25) < 0.03 and abs(frac[tp.Q - 1] - 0.25) < 0.03
    # keys under the entropy stream: uniform words of two calls differ, and stay below Q
    N, n = 1024, 4
    lk = tp.keygen_secret(n, 0, 7, 1)
    rk = tp.keygen_secret(N, 0, 7, 2)
    k1 = tp.keygen_brk(lk, rk, N, n, 16, 4, 0, 0, 3.2, ops.SEED_OS_ENTROPY)
    k2 = tp.keygen_brk(lk, rk, N, n, 16, 4, 0, 0, 3.2, ops.SEED_OS_ENTROPY)
    assert not np.array_equal(k1, k2)
    assert int(k1.max()) < tp.Q and int(k2.max()) < tp.Q


@pytest.mark.parametrize("sigma,count", [(3.19, 400000), (30.5, 200000), (2.0 ** 20, 100000)])
def test_gaussian_variance(sigma, count):
    x = ops.sample_gaussian(count, sigma, 99, 3).astype(np.float64)
    se_mean = sigma / np.sqrt(count)
    assert abs(x.mean()) < 6 * se_mean
    # variance of a sample variance ~ 2 sigma^4 / count
    assert abs(x.var() / sigma ** 2 - 1.0) < 6 * np.sqrt(2.0 / count)
    if sigma < 21:
        assert np.abs(x).max() <= np.ceil(12 * sigma)


@pytest.mark.parametrize("sigma", [0.0, 3.2, 21.0, 30.5, 4096.0, 2.0 ** 20])
def test_gaussian_equals_oracle(sigma):
    got = ops.sample_gaussian(5000, sigma, 1234, 7)
    want = oracle_gaussian(5000, sigma, 1234, 7)
    assert np.array_equal(got, want)


def test_entropy_gaussian_differs_per_call():
    a = ops.sample_gaussian(1000, 3.2, ops.SEED_OS_ENTROPY, 0)
    b = ops.sample_gaussian(1000, 3.2, ops.SEED_OS_ENTROPY, 0)
    assert not np.array_equal(a, b)


def test_keygen_rejects_short_arrays():
    N, n = 1024, 8
    lk = tp.keygen_secret(n, 0, 1, 1)
    rk = tp.keygen_secret(N, 0, 1, 2)
    with pytest.raises(ValueError):
        tp.keygen_brk(lk[:-1], rk, N, n, 16, 4, 0, 0, 3.2, 1)
    with pytest.raises(ValueError):
        tp.keygen_brk(lk, rk[: N // 2], N, n, 16, 4, 0, 0, 3.2, 1)
    with pytest.raises(ValueError):
        tp.keygen_auto_keys(rk[:-3], N, 16, 4, 3.2, 1)
    with pytest.raises(ValueError):
        tp.keygen_rlwe_ksk(rk, rk[:100], N, 16, 4, 3.2, 1)
    with pytest.raises(ValueError):
        tp.keygen_lwe_ksk(rk, lk[:4], N, n, 4, 14, 3.2, 1)
