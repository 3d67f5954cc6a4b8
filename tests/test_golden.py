"""The oracle reproduces the frozen vectors of tests/golden/ (written by tests/golden/make_golden.py)."""
import numpy as np
import pytest

from golden_vectors import Q, load, manifest


def test_manifest_lists_every_vector():
    assert set(manifest()) == {"ntt_N1024", "ntt_N2048", "br_binary", "br_ternary", "br_bmmp", "bootstrap_n630",
                               "pack_N1024", "modswitch_2048"}


@pytest.mark.parametrize("name", ["ntt_N1024", "ntt_N2048"])
def test_ntt(oracle, name):
    m, p = load(name)
    assert np.array_equal(oracle.Ring(m["N"]).ntt(p["input"]), p["output"])


@pytest.mark.parametrize("name", ["br_binary", "br_ternary", "br_bmmp"])
def test_blind_rotation(oracle, name):
    m, p = load(name)
    N, n, b, l, method, seed = m["N"], m["n"], m["base_log2"], m["digits"], m["method"], m["seed"]
    lk = oracle.secret(n, m["key_dist"], seed, 1)
    rk = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lk, rk, n, b, l, method, m["gadget_kind"], m["sigma"], seed)
    assert np.array_equal(oracle.random_mod_pow2(n + 1, 2 * N, *m["lwe_stream"]), p["lwe"])
    acc = oracle.blind_rotate(N, p["lwe"], p["tv"], brk, n, b, l, method, True)
    assert np.array_equal(acc, p["acc"])
    assert np.array_equal(oracle.sample_extract(acc, N), p["extracted"])


def test_bootstrap(oracle):
    m, p = load("bootstrap_n630")
    N, n, seed = m["N"], m["n"], m["seed"]
    lk = oracle.secret(n, 0, seed, 1)
    rk = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lk, rk, n, m["base_log2"], m["digits"], 0, 0, m["sigma"], seed)
    ksk = oracle.lwe_ksk(rk, lk, N, m["ks_out"], m["ks_base_log2"], m["ks_digits"], m["sigma"], seed)
    out, _ = oracle.bootstrap_batch(N, p["lwes"], p["tv"], brk, n, m["base_log2"], m["digits"], 0, ksk, m["ks_out"],
                                    m["ks_base_log2"], m["ks_digits"], 3)
    assert np.array_equal(out.reshape(-1), p["out"])


def test_packing(oracle):
    m, p = load("pack_N1024")
    N, b, d, seed = m["N"], m["base_log2"], m["digits"], m["seed"]
    keys = oracle.auto_keys(N, oracle.secret(N, 0, seed, 2), b, d, m["sigma"], seed)
    lwes = p["lwes"].reshape(4, N + 1)
    assert np.array_equal(oracle.pack(N, lwes, 4, keys, b, d), p["pack"])
    assert np.array_equal(oracle.pack_consecutive(N, lwes[:3], 3, keys, b, d), p["pack_consecutive"])


def test_modswitch(oracle):
    _, p = load("modswitch_2048")
    for kind, field in enumerate(("round", "floor", "ceil")):
        assert np.array_equal(oracle.modswitch(p["in"], Q, 2048, kind), p[field]), field
