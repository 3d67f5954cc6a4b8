"""GPU parity at the full sizes the throughput figures are quoted on, one launch each.

* Binary bootstrap n=630 N=1024 (b=2^12 l=3, key switch 1024->630 b=2^4 d=14), 4096 ciphertexts in
  one call: rows either side of 2048 (where a launch of this size no longer fits the resident
  waves at once and the second half of the grid starts), and the first and last rows, are bit-exact
  against the oracle; every row decrypts to its sign-LUT value.
* Ternary bootstrap n=800 N=2048 (b=2^16 l=4, key switch 2048->800 b=2^5 d=12), one call over 16
  request batches with lognormal sizes (median 256, at most 4096), each batch with its own random
  3-bit LUT (tv_stride = 2N): the first and last row of several batches are bit-exact against the
  oracle; every row decrypts to its batch's LUT value.
* BMMP n=1024 N=2048 (b=2^14 l=3) over 2048 ciphertexts, then packing the 2048 extracted LWEs
  under a seeded permutation (automorphism keys b=2^16 d=4): sampled extracted rows are bit-exact
  against the oracle's rotation + extraction, the oracle's packing of the GPU's own extracted LWEs
  equals the GPU's packing word for word, and all 2048 slots decrypt.

The noiseless output of a bootstrap is the constant coefficient of X^p * TV (p the input's phase
in Z_2N), computed here from the test vector itself; decryption errors must stay below q/64.
Every case prints what it runs before touching the GPU.
"""
import sys

import numpy as np
import pytest

from fhetoothpaste_amd.workload import lognormal_counts, random_3bit_lut, sign_test_vector

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


def phases(lwes, key, N):
    """b - <a, s> mod 2N for every row of lwes [count, n + 1] (key entries 0, 1, Q-1)."""
    n = lwes.shape[1] - 1
    s = np.where(np.asarray(key) == np.uint64(Q - 1), -1, np.asarray(key).astype(np.int64))
    inner = lwes[:, :n].astype(np.int64) @ s
    return (lwes[:, n].astype(np.int64) - inner) % (2 * N)


def noiseless(tv, N, p):
    """Constant coefficient of X^p * tv_b in Z_Q[X]/(X^N + 1)."""
    b = tv[N:]
    if p == 0:
        return int(b[0])
    if p <= N:
        return (Q - int(b[N - p])) % Q
    return int(b[2 * N - p])


def assert_close(ph, want, what):
    err = (int(ph) - int(want)) % Q
    err = err if err < Q // 2 else err - Q
    assert abs(err) < Q // 64, (what, err)


def test_binary_bootstrap_4096_one_launch(tp, oracle):
    N, n, b, l, n_out, kb, kd, count = 1024, 630, 12, 3, 630, 4, 14, 4096
    seed = 201
    say("full size binary bootstrap", N, n, count)
    lk = tp.keygen_secret(n, 0, seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 0, 0, 3.2, seed)
    ksk = tp.keygen_lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, seed)
    ctx = tp.Context(N, n, b, l, 0, n_out, kb, kd)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    lwes = np.random.default_rng(seed).integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    lwes[1, :n] = 0  # phase = b alone
    tv = sign_test_vector(N)
    got = ctx.bootstrap_batch(lwes, tv)
    assert got.shape == (count, n_out + 1)
    for r in (0, 1, 2046, 2047, 2048, 2049, count - 2, count - 1):
        ref, _ = oracle.bootstrap_batch(N, np.ascontiguousarray(lwes[r]), tv, brk, n, b, l, oracle.BINARY, ksk,
                                        n_out, kb, kd, 1)
        assert np.array_equal(got[r], ref[0]), r
    p = phases(lwes, lk, N)
    for r in range(count):
        assert_close(oracle.phase_lwe(got[r], lk, n_out), noiseless(tv, N, int(p[r])), r)
    ctx.close()


def test_ternary_bootstrap_lognormal_batches_one_launch(tp, oracle):
    N, n, b, l, n_out, kb, kd = 2048, 800, 16, 4, 800, 5, 12
    seed = 202
    rng = np.random.default_rng(seed)
    sizes = lognormal_counts(rng, 16, median=256, sigma_ln=1.0, lo=1, hi=4096)
    total = int(sizes.sum())
    say("full size ternary bootstrap, 16 batches", N, n, total, list(sizes))
    lk = tp.keygen_secret(n, 1, seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 1, 1, 3.2, seed)
    ksk = tp.keygen_lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, seed)
    ctx = tp.Context(N, n, b, l, 1, n_out, kb, kd)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    lwes = rng.integers(0, 2 * N, size=(total, n + 1), dtype=np.uint64)
    luts = [random_3bit_lut(N, rng)[0] for _ in range(16)]
    batch_of = np.repeat(np.arange(16), sizes)
    tvs = np.stack([luts[k] for k in batch_of])  # [total, 2N]: one LUT per batch, repeated per row
    got = ctx.bootstrap_batch(lwes, tvs)
    assert got.shape == (total, n_out + 1)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    big = int(np.argmax(sizes))
    rows = sorted({0, total - 1, int(starts[big]), int(starts[big] + sizes[big] - 1),
                   int(starts[7]), int(starts[7] + sizes[7] - 1), int(starts[15]), int(starts[big] + sizes[big] // 2)})
    for r in rows:
        k = int(batch_of[r])
        ref, _ = oracle.bootstrap_batch(N, np.ascontiguousarray(lwes[r]), luts[k], brk, n, b, l, oracle.TERNARY, ksk,
                                        n_out, kb, kd, 1)
        assert np.array_equal(got[r], ref[0]), (r, k)
    p = phases(lwes, lk, N)
    for r in range(total):
        assert_close(oracle.phase_lwe(got[r], lk, n_out), noiseless(luts[int(batch_of[r])], N, int(p[r])), r)
    ctx.close()


def test_bmmp_2048_then_pack(tp, oracle):
    N, n, b, l, count = 2048, 1024, 14, 3, 2048
    ab, ad = 16, 4
    seed = 203
    rng = np.random.default_rng(seed)
    say("full size BMMP then pack", N, n, count)
    lk = tp.keygen_secret(n, 0, seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 2, 1, 3.2, seed)
    keys = tp.keygen_auto_keys(rk, N, ab, ad, 3.2, seed)
    ctx = tp.Context(N, n, b, l, 2, 0, 0, 0, auto_base_log2=ab, auto_digits=ad)
    ctx.import_brk(brk)
    ctx.import_auto_keys(keys)
    lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    tv = random_3bit_lut(N, rng)[0]
    perm = rng.permutation(count).astype(np.uint32)
    ext = ctx.eval_func_batch(lwes, tv)
    assert ext.shape == (count, N + 1)
    for r in (0, 1, 1023, 1024, 1500, count - 2, count - 1):
        acc = oracle.blind_rotate(N, lwes[r], tv, brk, n, b, l, oracle.BMMP, True)
        assert np.array_equal(ext[r], oracle.sample_extract(acc, N)), r
    say("pack 2048 extracted LWEs", N)
    packed = ctx.pack(ext, perm)
    ref = oracle.pack(N, np.ascontiguousarray(ext[perm]), count, keys, ab, ad)
    assert np.array_equal(packed, ref)
    R = oracle.Ring(N)
    ph = oracle.phase_rlwe(N, np.concatenate([R.ntt(packed[:N]), R.ntt(packed[N:])]), rk)
    p = phases(lwes, lk, N)
    for j in range(count):
        assert_close(ph[j], noiseless(tv, N, int(p[perm[j]])), j)
    ctx.close()
