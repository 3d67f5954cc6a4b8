"""GPU parity of the exact floating-point external product (kernels_br_fft.hip) for binary CGGI at
N = 1024: the floating-point engine's rotated accumulators (coefficients and NTT domain) and
extracted LWEs equal the oracle's and the Goldilocks-NTT engine's, bit for bit, at one and two
waves per SIMD, for gadgets whose rounding bound certifies exact results; gadgets it does not
certify stay on the NTT engine.

Every case prints what it runs before touching the GPU.
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


def _ctx(tp, n, b, l, seed):
    N = 1024
    lk = tp.keygen_secret(n, 0, seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 0, 0, 3.2, seed)
    ctx = tp.Context(N, n, b, l, 0)
    ctx.import_brk(brk)
    ctx.set_br_split_max(0)  # keep every launch on the one-work-group-per-ciphertext kernels
    return ctx, brk, lk, rk


@pytest.mark.parametrize("n,b,l", [(630, 12, 3), (40, 12, 2), (24, 8, 4), (30, 11, 4), (36, 10, 3)])
def test_fft_engine_matches_oracle_and_ntt(tp, oracle, n, b, l):
    N = 1024
    say("fft engine", n, b, l)
    ctx, brk, lk, rk = _ctx(tp, n, b, l, 300 + n + b)
    assert ctx.br_engine() == 0   # the Goldilocks NTT unless the caller opts in
    ctx.set_br_engine(1)
    assert ctx.br_engine() == 1
    rng = np.random.default_rng(n * b)
    count = 7
    lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    lwes[1, ::2] = 0          # skipped steps
    lwes[2, :n] = N           # X^N = -1 every step
    lwes[3, :n] = 1
    lwes[4, :n] = 2 * N - 1
    tv = rng.integers(0, Q, size=(count, 2 * N), dtype=np.uint64)
    tv[5] = Q - 1
    results = {}
    for engine, waves in [(1, 1), (1, 2), (0, 2)]:
        ctx.set_br_engine(engine)
        ctx.set_fft_waves(waves)
        assert ctx.br_engine() == engine
        ext, rl = ctx.eval_func_batch(lwes, tv, want_rlwe=True, rlwe_ntt=False)
        _, rln = ctx.eval_func_batch(lwes, tv, want_rlwe=True, rlwe_ntt=True)
        results[(engine, waves)] = (ext, rl, rln)
    base = results[(0, 2)]
    for key, val in results.items():
        for a, b_ in zip(val, base):
            assert np.array_equal(a, b_), key
    ext, rl, rln = results[(1, 1)]
    for c in (range(count) if n <= 40 else (0, 2, count - 1)):
        ref = oracle.blind_rotate(N, lwes[c], tv[c], brk, n, b, l, oracle.BINARY, True)
        assert np.array_equal(rl[c], ref), c
        assert np.array_equal(ext[c], oracle.sample_extract(ref, N)), c
        R = oracle.Ring(N)
        assert np.array_equal(rln[c], np.concatenate([R.ntt(ref[:N]), R.ntt(ref[N:])])), c
    ctx.close()


def test_fft_engine_bootstrap_with_keyswitch(tp, oracle):
    """The full bootstrap (key switch 1024 -> 630) on the floating-point engine equals the oracle."""
    N, n, b, l, n_out, kb, kd = 1024, 630, 12, 3, 630, 4, 14
    say("fft engine bootstrap + key switch")
    lk = oracle.secret(n, 0, 17, 1)
    rk = oracle.secret(N, 0, 17, 2)
    brk = oracle.brk(N, lk, rk, n, b, l, oracle.BINARY, 0, 3.2, 17)
    ksk = oracle.lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 17)
    ctx = tp.Context(N, n, b, l, 0, n_out, kb, kd)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    ctx.set_br_split_max(0)
    ctx.set_br_engine(1)
    assert ctx.br_engine() == 1
    lwes = oracle.random_mod_pow2(5 * (n + 1), 2 * N, 17, 18).reshape(5, n + 1)
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    got = ctx.bootstrap_batch(lwes, tv)
    ref, _ = oracle.bootstrap_batch(N, lwes.reshape(-1), tv, brk, n, b, l, oracle.BINARY, ksk, n_out, kb, kd, 4)
    assert np.array_equal(got, ref)
    ctx.close()


def test_uncertified_gadget_stays_on_ntt(tp):
    """2^16 x 4 at N = 1024: the rounding bound exceeds 1/2, so the NTT engine runs; other
    methods and ring dimensions have no floating-point kernel."""
    from fhetoothpaste_amd import ToothpasteError
    say("fft engine eligibility")
    ctx, _, _, _ = _ctx(tp, 8, 16, 4, 5)
    assert ctx.br_engine() == 0
    ctx.set_br_engine(1)      # accepted (built) but not certified for this key
    assert ctx.br_engine() == 0
    ctx.close()
    ter = tp.Context(1024, 8, 16, 4, 1)
    with pytest.raises(ToothpasteError):
        ter.set_br_engine(1)
    ter.close()
    big = tp.Context(2048, 8, 12, 3, 0)
    with pytest.raises(ToothpasteError):
        big.set_br_engine(1)
    big.close()
