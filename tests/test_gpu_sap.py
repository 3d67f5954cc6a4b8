"""GPU parity of the SAP digit decomposition (sap_decomposer.cpp:190-494): HomTrunc,
ResetAccumulator and Decompose over batches, bit-exact against the oracle, plus decryption of
every digit (the reference's TestDecomp invariant, sap_decomposer_tests.cpp:167-237)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 0xFFFFFFFF00000001
N = 1024
BB, BD = 16, 4      # blind-rotation gadget
AB, AD = 16, 4      # automorphism keys
KB, KD = 4, 14      # LWE key switch N -> n (reset)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tp():
    import fhetoothpaste_amd as tp
    assert tp.device_count() >= 1, "no gfx950 device"
    return tp


def _setup(tp, oracle, n, method, dist, seed):
    lk = oracle.secret(n, dist, seed, 1)
    s = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lk, s, n, BB, BD, method, 0 if method == oracle.BINARY else 1, 3.2, seed)
    keys = oracle.auto_keys(N, s, AB, AD, 3.2, seed + 1)
    ksk = oracle.lwe_ksk(s, lk, N, n, KB, KD, 3.2, seed + 2)
    ctx = tp.Context(N, n, BB, BD, method, n, KB, KD, 0, AB, AD)
    ctx.import_brk(brk)
    ctx.import_auto_keys(keys)
    ctx.import_lwe_ksk(ksk)
    return dict(ctx=ctx, lk=lk, s=s, brk=brk, keys=keys, ksk=ksk, n=n, method=method)


@pytest.fixture(scope="module")
def binary(tp, oracle):
    return _setup(tp, oracle, 16, oracle.BINARY, 0, 201)


@pytest.mark.parametrize("radix,count", [(32, 3), (4, 2), (2, 1), (1024, 2)])
def test_sap_homtrunc_vs_oracle(binary, oracle, radix, count):
    say("sap homtrunc", radix, count)
    rng = np.random.default_rng(radix)
    x = rng.integers(0, Q, size=(count, 2 * N), dtype=np.uint64)
    got = binary["ctx"].sap_homtrunc(x, radix)
    for c in range(count):
        assert np.array_equal(got[c], oracle.sap_homtrunc(N, x[c], radix, binary["keys"], AB, AD)), c


@pytest.mark.parametrize("radix", [32, 0, 8])
def test_sap_reset_vs_oracle(binary, oracle, radix):
    say("sap reset", radix)
    rng = np.random.default_rng(7 + radix)
    x = rng.integers(0, Q, size=(2, 2 * N), dtype=np.uint64)
    got = binary["ctx"].sap_reset(x, radix)
    n = binary["n"]
    for c in range(2):
        want = oracle.sap_reset(N, x[c], radix, binary["ksk"], n, KB, KD, Q, binary["brk"], n, BB, BD,
                                binary["method"])
        assert np.array_equal(got[c], want), c


def _lwe(lk, msg, qin, rng):
    n = lk.size
    a = [int(v) for v in rng.integers(0, qin, size=n, dtype=np.uint64)]
    sk = [int(y) if int(y) < Q // 2 else int(y) - Q for y in lk]  # ternary -1 is stored as Q - 1
    b = (sum(x * y for x, y in zip(a, sk)) + msg) % qin
    return np.array(a + [b], dtype=np.uint64)


def _decrypt_digit(oracle, ct, s):
    ph = oracle.phase_lwe(ct, s, N)
    return int(oracle.modswitch(np.array([ph], dtype=np.uint64), Q, 2 * N, 0)[0])


def _check_decompose(setup, oracle, lwes, radix, restart=None):
    ctx, n = setup["ctx"], setup["n"]
    beta = int(np.count_nonzero(setup["lk"]))
    got, counts = ctx.sap_decompose(lwes, radix, beta, restart=restart)
    for c in range(lwes.shape[0]):
        want = oracle.sap_decompose(N, lwes[c], radix, beta, setup["brk"], n, BB, BD, setup["method"], setup["keys"],
                                    AB, AD, restart=restart, ksk=setup["ksk"], kb=KB, kd=KD)
        assert counts[c] == want.shape[0], c
        assert np.array_equal(got[c, :counts[c]], want), c
        assert not got[c, counts[c]:].any(), c  # digits beyond an input's count are not written
    return got, counts


def test_sap_decompose_vs_oracle_and_decrypts(binary, oracle):
    """Inputs mod 32^8 (8 digits), mod 2^20 (4) and one with a_0 > 2^63 (the x86 shift: modulus 1
    then doubled by the values) in one batch; digit i decrypts to digit i of the message."""
    say("sap decompose binary")
    rng = np.random.default_rng(11)
    radix, qin = 32, 32 ** 8
    msgs = [int(rng.integers(0, qin)) for _ in range(3)]
    rows = [_lwe(binary["lk"], m, qin, rng) for m in msgs]
    rows.append(_lwe(binary["lk"], 12345, 1 << 20, rng))
    odd = rng.integers(0, 1 << 63, size=binary["n"] + 1, dtype=np.uint64)
    odd[0] = (1 << 63) + 5
    rows.append(odd)
    lwes = np.stack(rows)
    got, counts = _check_decompose(binary, oracle, lwes, radix)
    assert list(counts[:3]) == [8, 8, 8]
    for c, m in enumerate(msgs):
        for i in range(8):
            assert _decrypt_digit(oracle, got[c, i], binary["s"]) == (m >> (5 * i)) % radix, (c, i)


def test_sap_decompose_radix_4(binary, oracle):
    say("sap decompose radix 4")
    rng = np.random.default_rng(12)
    qin = 1 << 12
    msgs = [int(rng.integers(0, qin)) for _ in range(2)]
    lwes = np.stack([_lwe(binary["lk"], m, qin, rng) for m in msgs])
    got, counts = _check_decompose(binary, oracle, lwes, 4)
    for c, m in enumerate(msgs):
        for i in range(counts[c]):
            assert _decrypt_digit(oracle, got[c, i], binary["s"]) == (m >> (2 * i)) % 4, (c, i)


def test_sap_decompose_with_restart(binary, oracle):
    """The periodic accumulator reset (never triggered by the reference's own restart estimate)."""
    say("sap decompose restart")
    rng = np.random.default_rng(13)
    lwes = np.stack([_lwe(binary["lk"], int(rng.integers(0, 1 << 25)), 1 << 25, rng) for _ in range(2)])
    _check_decompose(binary, oracle, lwes, 32, restart=2)


def test_sap_decompose_ternary(tp, oracle):
    say("sap decompose ternary")
    setup = _setup(tp, oracle, 16, oracle.TERNARY, 1, 301)
    rng = np.random.default_rng(14)
    lwes = np.stack([_lwe(setup["lk"], int(rng.integers(0, 1 << 30)), 1 << 30, rng) for _ in range(3)])
    _check_decompose(setup, oracle, lwes, 32)
    setup["ctx"].close()


def test_sap_rejects(tp, binary):
    from fhetoothpaste_amd import ToothpasteError
    say("sap rejects")
    ctx = binary["ctx"]
    x = np.zeros((1, binary["n"] + 1), np.uint64)
    x[0, 0] = 1 << 40
    for radix in (3, 0, 2048):
        with pytest.raises(ToothpasteError):
            ctx.sap_decompose(x, radix, 4)
    with pytest.raises(ToothpasteError):
        ctx.sap_decompose(x, 32, 4, max_digits=4)  # needs 8 digits
    with pytest.raises(ToothpasteError):
        ctx.sap_decompose(x, 32, 4, restart=0)
    with pytest.raises(ToothpasteError):
        ctx.sap_homtrunc(np.zeros((1, 2 * N), np.uint64), 48)
    bare = tp.Context(N, binary["n"], BB, BD)
    with pytest.raises(ToothpasteError):
        bare.sap_homtrunc(np.zeros((1, 2 * N), np.uint64), 32)  # no automorphism keys


def test_sap_decompose_reference_parameters(tp, oracle):
    """The reference test's own setting (sap_decomposer_tests.cpp:25-42, 167-237): N=4096, n=10,
    binary keys, gadget 2^16 x 4, radix 32, inputs mod 32^8; bit-exact against the oracle and every
    digit decrypts."""
    say("sap decompose N=4096")
    n, NN = 10, 4096
    lk = oracle.secret(n, 0, 401, 1)
    s = oracle.secret(NN, 0, 401, 2)
    brk = oracle.brk(NN, lk, s, n, BB, BD, oracle.BINARY, 0, 3.2, 401)
    keys = oracle.auto_keys(NN, s, AB, AD, 3.2, 402)
    ctx = tp.Context(NN, n, BB, BD, oracle.BINARY, 0, 0, 0, 0, AB, AD)
    ctx.import_brk(brk)
    ctx.import_auto_keys(keys)
    rng = np.random.default_rng(15)
    qin = 32 ** 8
    msgs = [int(rng.integers(0, qin)) for _ in range(2)]
    lwes = np.stack([_lwe(lk, m, qin, rng) for m in msgs])
    beta = int(np.count_nonzero(lk))
    got, counts = ctx.sap_decompose(lwes, 32, beta)
    for c, m in enumerate(msgs):
        want = oracle.sap_decompose(NN, lwes[c], 32, beta, brk, n, BB, BD, oracle.BINARY, keys, AB, AD)
        assert counts[c] == want.shape[0] == 8
        assert np.array_equal(got[c], want), c
        for i in range(8):
            ph = oracle.phase_lwe(got[c, i], s, NN)
            dig = int(oracle.modswitch(np.array([ph], dtype=np.uint64), Q, 2 * NN, 0)[0])
            assert dig == (m >> (5 * i)) % 32, (c, i)
    ctx.close()
