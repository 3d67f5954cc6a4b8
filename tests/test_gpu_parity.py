"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle, bit for bit.

Every case prints what it runs before touching the GPU, so a failing run names its case.
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


def _keys(oracle, N, n, b, l, method, dist, sigma, seed):
    lwe_key = oracle.secret(n, dist, seed, 1)
    rlwe_key = oracle.secret(N, 0, seed, 2)
    kind = 0 if method == oracle.BINARY else 1
    brk = oracle.brk(N, lwe_key, rlwe_key, n, b, l, method, kind, sigma, seed)
    return lwe_key, rlwe_key, brk


@pytest.mark.parametrize("N", [1024, 2048, 4096])
def test_ntt_forward_inverse(tp, oracle, N):
    say("ntt", N)
    ctx = tp.Context(N, 2, 16, 4)
    rng = np.random.default_rng(N)
    a = rng.integers(0, Q, size=(37, N), dtype=np.uint64)
    a[0] = 0
    a[1] = Q - 1
    R = oracle.Ring(N)
    f = ctx.ntt_forward(a)
    assert np.array_equal(f, R.ntt(a.reshape(-1)).reshape(a.shape))
    inv = ctx.ntt_inverse(f)
    assert np.array_equal(inv, a)
    # inverse of arbitrary values against the oracle
    assert np.array_equal(ctx.ntt_inverse(a), R.intt(a.reshape(-1)).reshape(a.shape))


def test_keygen_matches_oracle(tp, oracle):
    say("keygen bytes")
    N, n = 1024, 8
    for dist in (0, 1):
        assert np.array_equal(tp.keygen_secret(n, dist, 77, 1), oracle.secret(n, dist, 77, 1))
    lk = oracle.secret(n, 0, 5, 1)
    rk = oracle.secret(N, 0, 5, 2)
    for method, kind in [(0, 0), (1, 1), (2, 1)]:
        a = tp.keygen_brk(lk, rk, N, n, 12, 3, method, kind, 3.2, 9)
        b = oracle.brk(N, lk, rk, n, 12, 3, method, kind, 3.2, 9)
        assert np.array_equal(a, b), method
    src = oracle.secret(64, 0, 4, 2)
    dst = oracle.secret(20, 1, 4, 1)
    assert np.array_equal(tp.keygen_lwe_ksk(src, dst, 64, 20, 4, 14, 3.2, 6),
                          oracle.lwe_ksk(src, dst, 64, 20, 4, 14, 3.2, 6))


@pytest.mark.parametrize("N,n,b,l", [(1024, 12, 16, 4), (1024, 20, 12, 3), (2048, 10, 16, 4)])
@pytest.mark.parametrize("as_coef", [False, True])
def test_blind_rotate_single(tp, oracle, N, n, b, l, as_coef):
    say("blind_rotate binary", N, n, b, l, as_coef)
    lk, rk, brk = _keys(oracle, N, n, b, l, oracle.BINARY, 0, 3.2, 31 + n)
    ctx = tp.Context(N, n, b, l, 0)
    ctx.import_brk(brk)
    rng = np.random.default_rng(n)
    for trial in range(3):
        lwe = rng.integers(0, 2 * N, size=n + 1, dtype=np.uint64)
        if trial == 1:
            lwe[::3] = 0  # zero coefficients are skipped by the reference loop
        acc = rng.integers(0, Q, size=2 * N, dtype=np.uint64)
        got = ctx.blind_rotate(lwe, acc, as_coef)
        ref = oracle.blind_rotate(N, lwe, acc, brk, n, b, l, oracle.BINARY, as_coef)
        assert np.array_equal(got, ref), trial


def _sign_tv(N):
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    return tv


@pytest.mark.parametrize("N,n,b,l,count", [(1024, 500, 16, 4, 3), (1024, 630, 12, 3, 5)])
@pytest.mark.parametrize("split", [False, True], ids=["one-wave-per-ct", "digit-parallel"])
def test_eval_func_batch_vs_oracle(tp, oracle, N, n, b, l, count, split):
    say("eval_func batch", N, n, b, l, count, split)
    lk, rk, brk = _keys(oracle, N, n, b, l, oracle.BINARY, 0, 3.2, 7)
    ctx = tp.Context(N, n, b, l, 0)
    ctx.import_brk(brk)
    ctx.set_br_split_max(count if split else 0)
    lwes = oracle.random_mod_pow2(count * (n + 1), 2 * N, 7, 50).reshape(count, n + 1)
    tv = _sign_tv(N)
    got, rl = ctx.eval_func_batch(lwes, tv, want_rlwe=True, rlwe_ntt=False)
    for c in range(count):
        ref_acc = oracle.blind_rotate(N, lwes[c], tv, brk, n, b, l, oracle.BINARY, True)
        assert np.array_equal(rl[c], ref_acc), c
        assert np.array_equal(got[c], oracle.sample_extract(ref_acc, N)), c
    # decryption of the evaluated function: sign of phase
    for c in range(count):
        ph = oracle.phase_lwe(got[c], rk, N)
        ks = [int(x) for x in lk]
        p = (int(lwes[c][n]) - sum(int(a) * s for a, s in zip(lwes[c][:n], ks))) % (2 * N)
        want = Q // 8 if (p == 0 or p > N) else Q - Q // 8
        err = (ph - want) % Q
        err = err if err < Q // 2 else err - Q
        assert abs(err) < Q // 16, (c, p)


@pytest.mark.parametrize("count", [37, 70])
def test_keyswitch_batch_vs_oracle(tp, oracle, count):
    """37 and 70 ciphertexts run the matrix-core kernel (the row-split kernel takes at most 16)."""
    N, n_out, b, d = 1024, 630, 4, 14
    say("lwe keyswitch", N, n_out, b, d, count)
    src = oracle.secret(N, 0, 3, 2)
    dst = oracle.secret(n_out, 0, 3, 1)
    ksk = oracle.lwe_ksk(src, dst, N, n_out, b, d, 3.2, 3)
    ctx = tp.Context(N, 4, 12, 3, 0, n_out, b, d)
    ctx.import_lwe_ksk(ksk)
    rng = np.random.default_rng(5)
    lwes = rng.integers(0, Q, size=(count, N + 1), dtype=np.uint64)
    lwes[0] = Q - 1
    lwes[1] = 0
    got = ctx.lwe_keyswitch_batch(lwes)
    for c in range(count):
        assert np.array_equal(got[c], oracle.lwe_keyswitch(lwes[c], ksk, N, n_out, b, d)), c


@pytest.mark.parametrize("N,n_out,b,d", [(1024, 40, 3, 20), (1024, 33, 2, 32), (1024, 17, 16, 4)],
                         ids=["partial-last-chunk", "64-row-chunks", "partial-last-tile"])
@pytest.mark.parametrize("count", [100, 300])
def test_keyswitch_chunk_shapes(tp, oracle, N, n_out, b, d, count):
    """The batched kernel stages min(4, 64 // d) input words (their d key rows each) per chunk: d = 20
    gives 3 words and 1024 = 341 * 3 + 1 leaves a last chunk of one word; d = 32 fills the 64 staged
    rows exactly; n_out + 1 = 18 ends in a partial 4-column tile.  Every output equals the oracle."""
    say("lwe keyswitch chunk shapes", N, n_out, b, d, count)
    src = oracle.secret(N, 0, 21, 2)
    dst = oracle.secret(n_out, 0, 21, 1)
    ksk = oracle.lwe_ksk(src, dst, N, n_out, b, d, 3.2, 21)
    ctx = tp.Context(N, 4, 12, 3, 0, n_out, b, d)
    ctx.import_lwe_ksk(ksk)
    lwes = np.random.default_rng(count + d).integers(0, Q, size=(count, N + 1), dtype=np.uint64)
    lwes[3] = Q - 1
    lwes[4] = 0
    got = ctx.lwe_keyswitch_batch(lwes)
    for c in range(count):
        assert np.array_equal(got[c], oracle.lwe_keyswitch(lwes[c], ksk, N, n_out, b, d)), c


def test_keyswitch_tile_widths_agree(tp, oracle):
    """Vector-unit kernel (matrix cores off): a launch large enough for the 8-column tiles (4000
    ciphertexts, 1024 -> 630, 16 blocks of 256 lanes) equals two launches of 2000 on the 4-column
    tiles (8 blocks, the last one partial), word for word, and sampled rows equal the oracle."""
    N, n_out, b, d = 1024, 630, 4, 14
    say("lwe keyswitch tile widths", N, n_out, b, d)
    src = tp.keygen_secret(N, 0, 12, 2)
    dst = tp.keygen_secret(n_out, 0, 12, 1)
    ksk = tp.keygen_lwe_ksk(src, dst, N, n_out, b, d, 3.2, 12)
    ctx = tp.Context(N, 4, 12, 3, 0, n_out, b, d)
    ctx.import_lwe_ksk(ksk)
    ctx.set_ks_mfma(0)
    assert ctx.ks_mfma() == 0
    count = 4000
    lwes = np.random.default_rng(13).integers(0, Q, size=(count, N + 1), dtype=np.uint64)
    lwes[7] = Q - 1
    whole = ctx.lwe_keyswitch_batch(lwes)
    halves = np.concatenate([ctx.lwe_keyswitch_batch(lwes[:count // 2]), ctx.lwe_keyswitch_batch(lwes[count // 2:])])
    assert np.array_equal(whole, halves)
    for c in (0, 7, 1999, 2000, count - 1):
        assert np.array_equal(whole[c], oracle.lwe_keyswitch(lwes[c], ksk, N, n_out, b, d)), c


@pytest.mark.parametrize("N,n_out,b,d", [(1024, 630, 4, 14), (2048, 800, 5, 12), (1024, 40, 6, 10), (1024, 33, 7, 9),
                                         (1024, 70, 3, 16), (1024, 31, 4, 16), (1024, 64, 2, 16), (1024, 20, 1, 16),
                                         (1024, 95, 4, 5)])
@pytest.mark.parametrize("count", [17, 300])
def test_keyswitch_matrix_cores_match(tp, oracle, N, n_out, b, d, count):
    """More than 16 ciphertexts run the matrix-core key switch (int8 digit and key-byte operands,
    32-bit sums recombined mod Q): word for word equal to the vector-unit kernel, sampled rows and
    the edge inputs (all Q - 1, all 0) equal to the oracle.  Counts 17 and 300 end in partial
    128-row tiles; n_out + 1 = 41, 34, 71, 32, 65, 21, 96 give partial and whole 32-column groups;
    b = 4 takes the nibble path, the other bases the bit-field one."""
    say("lwe keyswitch matrix cores", N, n_out, b, d, count)
    src = tp.keygen_secret(N, 0, 14 + b, 2)
    dst = tp.keygen_secret(n_out, 0, 14 + b, 1)
    ksk = tp.keygen_lwe_ksk(src, dst, N, n_out, b, d, 3.2, 14 + d)
    ctx = tp.Context(N, 4, 12, 3, 0, n_out, b, d)
    ctx.import_lwe_ksk(ksk)
    assert ctx.ks_mfma() == 1
    lwes = np.random.default_rng(count + 7 * b + d).integers(0, Q, size=(count, N + 1), dtype=np.uint64)
    lwes[1] = Q - 1
    lwes[2] = 0
    got = ctx.lwe_keyswitch_batch(lwes)
    ctx.set_ks_mfma(0)
    assert ctx.ks_mfma() == 0
    vec = ctx.lwe_keyswitch_batch(lwes)
    assert np.array_equal(got, vec)
    for c in (0, 1, 2, count // 2, count - 1):
        assert np.array_equal(got[c], oracle.lwe_keyswitch(lwes[c], ksk, N, n_out, b, d)), c


@pytest.mark.parametrize("N,n_out,b,d", [(1024, 630, 4, 14), (2048, 800, 5, 12), (1024, 17, 16, 4)])
def test_keyswitch_single_ciphertext(tp, oracle, N, n_out, b, d):
    """count == 1 runs the row-split kernel (partial sums over key rows meet in LDS)."""
    say("lwe keyswitch single", N, n_out, b, d)
    src = oracle.secret(N, 0, 8, 2)
    dst = oracle.secret(n_out, 0, 8, 1)
    ksk = oracle.lwe_ksk(src, dst, N, n_out, b, d, 3.2, 8)
    ctx = tp.Context(N, 4, 12, 3, 0, n_out, b, d)
    ctx.import_lwe_ksk(ksk)
    rng = np.random.default_rng(N + b)
    for trial in range(3):
        lwe = rng.integers(0, Q, size=(1, N + 1), dtype=np.uint64)
        if trial == 1:
            lwe[0] = Q - 1
        got = ctx.lwe_keyswitch_batch(lwe)
        assert np.array_equal(got[0], oracle.lwe_keyswitch(lwe[0], ksk, N, n_out, b, d)), trial


def test_bootstrap_batch_n630_N1024_shape(tp, oracle):
    """n=630, N=1024, b=12 l=3, KS 1024->630 b=4 d=14: bit-exact and decrypts."""
    N, n, b, l, n_out, kb, kd = 1024, 630, 12, 3, 630, 4, 14
    say("bootstrap batch n=630 N=1024 shape")
    lk, rk, brk = _keys(oracle, N, n, b, l, oracle.BINARY, 0, 3.2, 11)
    ksk = oracle.lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 11)
    ctx = tp.Context(N, n, b, l, 0, n_out, kb, kd)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    count = 6
    lwes = oracle.random_mod_pow2(count * (n + 1), 2 * N, 11, 60).reshape(count, n + 1)
    tv = _sign_tv(N)
    got = ctx.bootstrap_batch(lwes, tv)
    ref, _ = oracle.bootstrap_batch(N, lwes.reshape(-1), tv, brk, n, b, l, oracle.BINARY, ksk, n_out, kb, kd, 1)
    assert np.array_equal(got, ref)
    ks = [int(x) for x in lk]
    for c in range(count):
        ph = oracle.phase_lwe(got[c], lk, n_out)
        p = (int(lwes[c][n]) - sum(int(a) * s for a, s in zip(lwes[c][:n], ks))) % (2 * N)
        want = Q // 8 if (p == 0 or p > N) else Q - Q // 8
        err = (ph - want) % Q
        err = err if err < Q // 2 else err - Q
        assert abs(err) < Q // 16, (c, p)


@pytest.mark.parametrize("method,N,n,b,l", [(1, 1024, 14, 16, 4), (1, 2048, 10, 16, 4), (2, 1024, 12, 14, 3),
                                            (2, 2048, 10, 14, 3), (2, 1024, 8, 16, 4)])
def test_blind_rotate_ternary_bmmp(tp, oracle, method, N, n, b, l):
    say("blind_rotate method", method, N, n, b, l)
    dist = 1 if method == oracle.TERNARY else 0
    lk, rk, brk = _keys(oracle, N, n, b, l, method, dist, 3.2, 40 + n)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(brk)
    rng = np.random.default_rng(n + method)
    for trial in range(3):
        lwe = rng.integers(0, 2 * N, size=n + 1, dtype=np.uint64)
        if trial == 1:
            lwe[::2] = 0
            lwe[1] = N
        acc = rng.integers(0, Q, size=2 * N, dtype=np.uint64)
        for as_coef in (False, True):
            got = ctx.blind_rotate(lwe, acc, as_coef)
            ref = oracle.blind_rotate(N, lwe, acc, brk, n, b, l, method, as_coef)
            assert np.array_equal(got, ref), (trial, as_coef)


@pytest.mark.parametrize("method,N,n,b,l,count", [(1, 2048, 800, 16, 4, 2), (2, 2048, 1024, 14, 3, 2)])
def test_eval_func_ternary_bmmp_full_size(tp, oracle, method, N, n, b, l, count):
    """Configs 3 and 4 parameter sets (ternary n=800 N=2048; BMMP n=1024 N=2048 b=14 l=3)."""
    say("eval_func full size method", method, N, n)
    dist = 1 if method == oracle.TERNARY else 0
    lk, rk, brk = _keys(oracle, N, n, b, l, method, dist, 3.2, 90 + method)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(brk)
    lwes = oracle.random_mod_pow2(count * (n + 1), 2 * N, 13, 70 + method).reshape(count, n + 1)
    tv = _sign_tv(N)
    got, rl = ctx.eval_func_batch(lwes, tv, want_rlwe=True, rlwe_ntt=True)
    for c in range(count):
        ref = oracle.blind_rotate(N, lwes[c], tv, brk, n, b, l, method, False)
        assert np.array_equal(rl[c], ref), c
    ks = [int(x) if int(x) < 2 else -1 for x in lk]
    for c in range(count):
        ph = oracle.phase_lwe(got[c], rk, N)
        p = (int(lwes[c][n]) - sum(int(a) * s for a, s in zip(lwes[c][:n], ks))) % (2 * N)
        want = Q // 8 if (p == 0 or p > N) else Q - Q // 8
        err = (ph - want) % Q
        err = err if err < Q // 2 else err - Q
        assert abs(err) < Q // 16, (c, p)


@pytest.mark.parametrize("method,N,n,b,l", [(0, 1024, 24, 12, 3), (1, 2048, 12, 16, 4), (2, 1024, 12, 14, 3),
                                            (2, 2048, 8, 14, 3), (0, 4096, 6, 16, 4)])
def test_one_wave_variant_matches_oracle(tp, oracle, method, N, n, b, l):
    """The 512-register (one wave per SIMD) build of the blind rotation is bit-identical."""
    say("one-wave variant", method, N, n, b, l)
    dist = 1 if method == oracle.TERNARY else 0
    lk, rk, brk = _keys(oracle, N, n, b, l, method, dist, 3.2, 60 + n)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(brk)
    if N == 1024:
        ctx.set_br_split_max(0)  # keep these launches on br_kernel
    lwes = oracle.random_mod_pow2(3 * (n + 1), 2 * N, 61, 62 + method).reshape(3, n + 1)
    tv = np.random.default_rng(n).integers(0, Q, size=2 * N, dtype=np.uint64)
    ctx.set_br_waves(1)
    got1, rl1 = ctx.eval_func_batch(lwes, tv, want_rlwe=True)
    ctx.set_br_waves(2)
    got2, rl2 = ctx.eval_func_batch(lwes, tv, want_rlwe=True)
    assert np.array_equal(rl1, rl2) and np.array_equal(got1, got2)
    for c in range(3):
        assert np.array_equal(rl1[c], oracle.blind_rotate(N, lwes[c], tv, brk, n, b, l, method, True)), c


def test_autotune_picks_a_variant(tp, oracle):
    N, n, b, l = 1024, 16, 16, 4
    say("autotune")
    lk, rk, brk = _keys(oracle, N, n, b, l, oracle.TERNARY, 1, 3.2, 70)
    ctx = tp.Context(N, n, b, l, oracle.TERNARY)
    ctx.import_brk(brk)
    default_split = ctx.br_split_max()
    assert default_split >= 2
    r = ctx.autotune(512)
    assert r["waves"] in (1, 2) and r["ms_two_waves"] > 0 and r["ms_one_wave"] > 0 and r["ms_digit_parallel"] > 0
    chosen = r["ms_one_wave"] if r["waves"] == 1 else r["ms_two_waves"]
    assert chosen <= min(r["ms_one_wave"], r["ms_two_waves"])
    if r["ms_digit_parallel"] < chosen:
        assert r["split_max"] >= 512
    else:
        assert r["split_max"] < 512
    # the chosen variant still computes the same rotation
    lwes = oracle.random_mod_pow2(2 * (n + 1), 2 * N, 71, 72).reshape(2, n + 1)
    tv = np.random.default_rng(5).integers(0, Q, size=2 * N, dtype=np.uint64)
    _, rl = ctx.eval_func_batch(lwes, tv, want_rlwe=True)
    for c in range(2):
        assert np.array_equal(rl[c], oracle.blind_rotate(N, lwes[c], tv, brk, n, b, l, oracle.TERNARY, True)), c
    from fhetoothpaste_amd import ToothpasteError
    ctx2 = tp.Context(N, n, 12, 2, 0)
    with pytest.raises(ToothpasteError):
        ctx2.set_br_waves(1)  # not built for 2 digits
    ctx3 = tp.Context(2048, n, b, l, 0)
    assert ctx3.br_split_max() == 0
    with pytest.raises(ToothpasteError):
        ctx3.set_br_split_max(4)  # digit-parallel kernel is built for N = 1024 only
    ctx3.import_brk(tp.keygen_brk(tp.keygen_secret(n, 0, 71, 1), tp.keygen_secret(2048, 0, 71, 2), 2048, n, b, l,
                                  0, 0, 3.2, 71))
    r3 = ctx3.autotune(4)
    assert r3["ms_digit_parallel"] == -1 and r3["split_max"] == 0


@pytest.mark.parametrize("method,n,b,l", [(0, 20, 16, 2), (0, 24, 12, 3), (0, 16, 16, 4), (1, 14, 16, 4),
                                          (1, 12, 14, 3), (2, 12, 14, 3), (2, 8, 16, 2)])
@pytest.mark.parametrize("rlwe_ntt", [False, True])
def test_single_ciphertext_latency_kernel(tp, oracle, method, n, b, l, rlwe_ntt):
    """One ciphertext at N=1024 runs on the digit-parallel kernel (2 l waves); extracted LWE and
    rotated accumulator (both domains) bit-exact against the oracle."""
    N = 1024
    say("latency kernel", method, n, b, l, rlwe_ntt)
    dist = 1 if method == oracle.TERNARY else 0
    lk, rk, brk = _keys(oracle, N, n, b, l, method, dist, 3.2, 60 + n + l)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(brk)
    rng = np.random.default_rng(n * 7 + l)
    for trial in range(2):
        lwe = rng.integers(0, 2 * N, size=(1, n + 1), dtype=np.uint64)
        if trial == 1:
            lwe[0, ::3] = 0
        tv = rng.integers(0, Q, size=2 * N, dtype=np.uint64)
        got, rl = ctx.eval_func_batch(lwe, tv, want_rlwe=True, rlwe_ntt=rlwe_ntt)
        ref_rl = oracle.blind_rotate(N, lwe[0], tv, brk, n, b, l, method, not rlwe_ntt)
        assert np.array_equal(rl[0], ref_rl), trial
        ref_coef = ref_rl if not rlwe_ntt else oracle.blind_rotate(N, lwe[0], tv, brk, n, b, l, method, True)
        assert np.array_equal(got[0], oracle.sample_extract(ref_coef, N)), trial


@pytest.mark.parametrize("method,n,b,l", [(0, 24, 12, 3), (1, 14, 16, 4), (2, 12, 14, 3), (0, 18, 16, 2)])
def test_digit_parallel_batch_matches_oracle(tp, oracle, method, n, b, l):
    """A batch on the digit-parallel kernel (2 l waves per ciphertext, one test vector per
    ciphertext, both output domains) equals br_kernel's launch and the oracle, bit for bit."""
    N, count = 1024, 6
    say("digit-parallel batch", method, n, b, l)
    dist = 1 if method == oracle.TERNARY else 0
    lk, rk, brk = _keys(oracle, N, n, b, l, method, dist, 3.2, 90 + n + l)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(brk)
    rng = np.random.default_rng(91 + n)
    lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    lwes[2, ::2] = 0
    tvs = rng.integers(0, Q, size=(count, 2 * N), dtype=np.uint64)
    outs = {}
    for split in (0, count):
        ctx.set_br_split_max(split)
        for rlwe_ntt in (False, True):
            outs[split, rlwe_ntt] = ctx.eval_func_batch(lwes, tvs, want_rlwe=True, rlwe_ntt=rlwe_ntt)
    for rlwe_ntt in (False, True):
        assert np.array_equal(outs[0, rlwe_ntt][0], outs[count, rlwe_ntt][0])
        assert np.array_equal(outs[0, rlwe_ntt][1], outs[count, rlwe_ntt][1])
    for c in range(count):
        ref = oracle.blind_rotate(N, lwes[c], tvs[c], brk, n, b, l, method, True)
        assert np.array_equal(outs[count, False][1][c], ref), c
        assert np.array_equal(outs[count, False][0][c], oracle.sample_extract(ref, N)), c
        ref_ntt = oracle.blind_rotate(N, lwes[c], tvs[c], brk, n, b, l, method, False)
        assert np.array_equal(outs[count, True][1][c], ref_ntt), c


@pytest.mark.parametrize("n,b,l", [(24, 12, 3), (18, 16, 2), (20, 20, 3), (630, 12, 3)])
def test_split_halves_matches_one_wave_per_digit(tp, oracle, n, b, l):
    """Binary CGGI at N=1024: the digit-parallel launch with every transform split over two waves
    (4 l waves per ciphertext) and the one with one wave per digit transform (2 l waves) give the
    same extracted LWEs and accumulators in both domains, equal to the oracle (first and last
    ciphertext at n=630, every ciphertext otherwise).  b = 20 takes digits above 16 bits (the
    generic first level) and a digit straddling the two 32-bit words."""
    N, count = 1024, 5
    say("split halves", n, b, l)
    lk, rk, brk = _keys(oracle, N, n, b, l, oracle.BINARY, 0, 3.2, 120 + n + l)
    ctx = tp.Context(N, n, b, l, 0)
    ctx.import_brk(brk)
    ctx.set_br_split_max(count)
    assert ctx.br_split_halves() == 1
    rng = np.random.default_rng(121 + n)
    lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    lwes[1, ::2] = 0
    lwes[3, :n] = N  # every rotation by X^N = -1
    tvs = rng.integers(0, Q, size=(count, 2 * N), dtype=np.uint64)
    outs = {}
    for halves in (1, 0):
        ctx.set_br_split_halves(halves)
        for rlwe_ntt in (False, True):
            outs[halves, rlwe_ntt] = ctx.eval_func_batch(lwes, tvs, want_rlwe=True, rlwe_ntt=rlwe_ntt)
    assert ctx.br_split_halves() == 0
    for rlwe_ntt in (False, True):
        assert np.array_equal(outs[0, rlwe_ntt][0], outs[1, rlwe_ntt][0])
        assert np.array_equal(outs[0, rlwe_ntt][1], outs[1, rlwe_ntt][1])
    for c in ([0, count - 1] if n > 100 else range(count)):
        ref = oracle.blind_rotate(N, lwes[c], tvs[c], brk, n, b, l, oracle.BINARY, True)
        assert np.array_equal(outs[1, False][1][c], ref), c
        assert np.array_equal(outs[1, False][0][c], oracle.sample_extract(ref, N)), c
