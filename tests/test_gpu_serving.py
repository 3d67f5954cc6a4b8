"""Serving loop on the GPU: every request's output is the same alone and in the mix, bit for bit.

Two parameter sets stay resident: binary n=630 N=1024 (b=12 l=3, key switch 1024->630 b=4 d=14)
and ternary n=800 N=2048 (b=16 l=4, key switch 2048->800 b=5 d=12).  Requests carry random
3-bit LUTs (values uniform mod 8 scaled by q/16) and uneven ciphertext counts; the server runs
them with a small slot pool and a chunk length that divides neither n, so ciphertexts are
admitted mid-flight and park their accumulators between launches.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 0xFFFFFFFF00000001

SETS = {
    "binary630": dict(N=1024, n=630, b=12, l=3, method=0, dist=0, n_out=630, kb=4, kd=14, seed=21),
    "ternary800": dict(N=2048, n=800, b=16, l=4, method=1, dist=1, n_out=800, kb=5, kd=12, seed=22),
}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tp():
    import fhetoothpaste_amd as tp
    assert tp.device_count() >= 1, "no gfx950 device"
    return tp


@pytest.fixture(scope="module")
def sets(tp, oracle):
    out = {}
    for name, P in SETS.items():
        N, n, seed = P["N"], P["n"], P["seed"]
        lk = oracle.secret(n, P["dist"], seed, 1)
        rk = oracle.secret(N, 0, seed, 2)
        kind = 0 if P["method"] == 0 else 1
        brk = oracle.brk(N, lk, rk, n, P["b"], P["l"], P["method"], kind, 3.2, seed)
        ksk = oracle.lwe_ksk(rk, lk, N, P["n_out"], P["kb"], P["kd"], 3.2, seed)
        ctx = tp.Context(N, n, P["b"], P["l"], P["method"], P["n_out"], P["kb"], P["kd"])
        ctx.import_brk(brk)
        ctx.import_lwe_ksk(ksk)
        out[name] = dict(P, ctx=ctx, lk=lk, rk=rk, brk=brk, ksk=ksk)
    return out


def lut_tv(N, rng):
    """b-half of the test vector: 8 blocks of N/8 coefficients, values v * q/16 with v uniform mod 8."""
    v = rng.integers(0, 8, size=8)
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = np.repeat(v.astype(np.uint64) * np.uint64(Q // 16), N // 8)
    return tv


def expected_value(tv, N, p):
    """Constant coefficient of X^p * tv_b in Z_Q[X]/(X^N + 1): the noiseless bootstrap output."""
    b = tv[N:]
    if p == 0:
        return int(b[0])
    if p <= N:
        return (Q - int(b[N - p])) % Q
    return int(b[2 * N - p])


def check_decrypts(S, lwes, tv, out):
    N, n = S["N"], S["n"]
    key = [int(x) if int(x) < 2 else -1 for x in S["lk"]]
    import oracle as O
    for c in range(lwes.shape[0]):
        p = (int(lwes[c][n]) - sum(int(a) * s for a, s in zip(lwes[c][:n], key))) % (2 * N)
        ph = O.phase_lwe(out[c], S["lk"], S["n_out"])
        err = (ph - expected_value(tv, N, p)) % Q
        err = err if err < Q // 2 else err - Q
        assert abs(err) < Q // 64, (c, p)


def make_requests(sets, oracle, counts, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    names = list(SETS)
    for i, cnt in enumerate(counts):
        name = names[i % 2]
        S = sets[name]
        lwes = oracle.random_mod_pow2(cnt * (S["n"] + 1), 2 * S["N"], seed, 100 + i).reshape(cnt, S["n"] + 1)
        reqs.append((name, lwes, lut_tv(S["N"], rng)))
    return reqs


def test_bootstrap_ternary_set_vs_oracle(sets, oracle):
    """Anchor of the second set: bootstrap (blind rotation + key switch 2048->800) bit-exact to the oracle."""
    S = sets["ternary800"]
    say("ternary800 bootstrap vs oracle")
    lwes = oracle.random_mod_pow2(3 * (S["n"] + 1), 2 * S["N"], 5, 7).reshape(3, S["n"] + 1)
    tv = lut_tv(S["N"], np.random.default_rng(3))
    got = S["ctx"].bootstrap_batch(lwes, tv)
    ref, _ = oracle.bootstrap_batch(S["N"], lwes.reshape(-1), tv, S["brk"], S["n"], S["b"], S["l"], S["method"],
                                    S["ksk"], S["n_out"], S["kb"], S["kd"], 3)
    assert np.array_equal(got, ref)
    check_decrypts(S, lwes, tv, got)


@pytest.mark.parametrize("chunk,slots", [(37, 48), (1000, 4096)])
@pytest.mark.parametrize("split", [True, False], ids=["digit-parallel", "one-wave-per-ct"])
def test_mix_equals_alone(tp, sets, oracle, chunk, slots, split):
    """Each request's result in the serving mix equals its result alone.  The references run alone
    on br_kernel in one launch; the server's chunks of the N = 1024 set run on the digit-parallel
    kernel (at most twice the CU count in flight) or, with its threshold at 0, on br_kernel."""
    from fhetoothpaste_amd.serving import Server
    counts = [1, 37, 5, 64, 2, 19, 90, 3, 11, 40]
    say("serving mix chunk", chunk, "slots", slots, "split", split, "counts", counts)
    reqs = make_requests(sets, oracle, counts, 31 + chunk)
    names = list(SETS)
    defaults = {name: sets[name]["ctx"].br_split_max() for name in names}
    for name in names:
        sets[name]["ctx"].set_br_split_max(0)
    alone = [sets[name]["ctx"].bootstrap_batch(lwes, tv) for name, lwes, tv in reqs]
    for name in names:
        sets[name]["ctx"].set_br_split_max(defaults[name] if split else 0)
    ctxs = [sets[name]["ctx"] for name in names]
    try:
        _serve_and_compare(Server, ctxs, reqs, names, alone, counts, chunk, slots)
    finally:
        for name in names:
            sets[name]["ctx"].set_br_split_max(defaults[name])
    for i in (1, 2):
        name, lwes, tv = reqs[i]
        check_decrypts(sets[name], lwes, tv, alone[i])


def _serve_and_compare(Server, ctxs, reqs, names, alone, counts, chunk, slots):
    with Server(ctxs, chunk_steps=chunk, max_slots_per_set=slots) as srv:
        handles = []
        for i, (name, lwes, tv) in enumerate(reqs):
            handles.append(srv.submit(names.index(name), lwes, tv))
            if i == 3:
                handles[0].wait(120)  # later requests arrive while earlier ones are on the device
        for h in handles:
            h.wait(120)
        for i, h in enumerate(handles):
            assert np.array_equal(h.out, alone[i]), (i, reqs[i][0])
            t_sub, t_done = h.times()
            assert t_done >= t_sub
            h.release()
        st = srv.stats()
        assert st["ciphertexts"] == sum(counts)
        if chunk < 630:
            assert st["launches"] > 2 * (630 // chunk)


def test_cancel_and_reuse(tp, sets, oracle):
    from fhetoothpaste_amd.serving import Server
    say("serving cancel")
    S = sets["binary630"]
    reqs = make_requests(sets, oracle, [300, 4], 77)
    big_lwes, big_tv = reqs[0][1], reqs[0][2]
    lwes = oracle.random_mod_pow2(4 * (S["n"] + 1), 2 * S["N"], 9, 9).reshape(4, S["n"] + 1)
    want = S["ctx"].bootstrap_batch(lwes, big_tv)
    with Server([S["ctx"]], chunk_steps=100, max_slots_per_set=32) as srv:
        big = srv.submit(0, big_lwes, big_tv)
        big.cancel()
        big.wait(120)
        assert big.done()
        small = srv.submit(0, lwes, big_tv)
        small.wait(120)
        assert np.array_equal(small.out, want)
        big.release()
        small.release()


def test_server_rejects_bad_input(tp, sets):
    from fhetoothpaste_amd import ToothpasteError
    from fhetoothpaste_amd.serving import Server
    say("serving rejects")
    S = sets["binary630"]
    with Server([S["ctx"]], chunk_steps=64, max_slots_per_set=8) as srv:
        bad = np.zeros((1, S["n"] + 1), dtype=np.uint64)
        bad[0, 3] = 2 * S["N"]
        with pytest.raises(ToothpasteError):
            srv.submit(0, bad, np.zeros(2 * S["N"], dtype=np.uint64))
        with pytest.raises(ValueError):
            srv.submit(0, bad, np.zeros(S["N"], dtype=np.uint64))
    bare = tp.Context(1024, 16, 12, 3, 0)
    with pytest.raises(ToothpasteError):
        Server([bare])
