"""The GPU path reproduces the frozen vectors of tests/golden/ bit for bit."""
import sys

import numpy as np
import pytest

from golden_vectors import Q, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tp():
    import fhetoothpaste_amd as tp
    assert tp.device_count() >= 1, "no gfx950 device"
    return tp


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["ntt_N1024", "ntt_N2048"])
def test_ntt(tp, name):
    say("golden", name)
    m, p = load(name)
    ctx = tp.Context(m["N"], 2, 16, 4)
    assert np.array_equal(ctx.ntt_forward(p["input"][None, :])[0], p["output"])


@pytest.mark.parametrize("name", ["br_binary", "br_ternary", "br_bmmp"])
def test_blind_rotation(tp, name):
    say("golden", name)
    m, p = load(name)
    N, n, b, l, method, seed = m["N"], m["n"], m["base_log2"], m["digits"], m["method"], m["seed"]
    lk = tp.keygen_secret(n, m["key_dist"], seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    ctx = tp.Context(N, n, b, l, method)
    ctx.import_brk(tp.keygen_brk(lk, rk, N, n, b, l, method, m["gadget_kind"], m["sigma"], seed))
    ext, acc = ctx.eval_func_batch(p["lwe"][None, :], p["tv"], want_rlwe=True)
    assert np.array_equal(acc[0], p["acc"])
    assert np.array_equal(ext[0], p["extracted"])


def test_bootstrap(tp):
    say("golden bootstrap_n630")
    m, p = load("bootstrap_n630")
    N, n, seed = m["N"], m["n"], m["seed"]
    lk = tp.keygen_secret(n, 0, seed, 1)
    rk = tp.keygen_secret(N, 0, seed, 2)
    ctx = tp.Context(N, n, m["base_log2"], m["digits"], 0, m["ks_out"], m["ks_base_log2"], m["ks_digits"])
    ctx.import_brk(tp.keygen_brk(lk, rk, N, n, m["base_log2"], m["digits"], 0, 0, m["sigma"], seed))
    ctx.import_lwe_ksk(tp.keygen_lwe_ksk(rk, lk, N, m["ks_out"], m["ks_base_log2"], m["ks_digits"], m["sigma"], seed))
    out = ctx.bootstrap_batch(p["lwes"].reshape(3, n + 1), p["tv"])
    assert np.array_equal(out.reshape(-1), p["out"])


def test_packing(tp):
    say("golden pack_N1024")
    m, p = load("pack_N1024")
    N, b, d, seed = m["N"], m["base_log2"], m["digits"], m["seed"]
    ctx = tp.Context(N, 2, 12, 3, 0, 0, 0, 0, 0, b, d)
    ctx.import_auto_keys(tp.keygen_auto_keys(tp.keygen_secret(N, 0, seed, 2), N, b, d, m["sigma"], seed))
    lwes = p["lwes"].reshape(4, N + 1)
    assert np.array_equal(ctx.pack(lwes), p["pack"])
    assert np.array_equal(ctx.pack_consecutive(lwes[:3]), p["pack_consecutive"])


def test_modswitch(tp):
    say("golden modswitch_2048")
    _, p = load("modswitch_2048")
    ctx = tp.Context(1024, 2, 16, 4)
    for kind, field in enumerate(("round", "floor", "ceil")):
        assert np.array_equal(ctx.modswitch(p["in"], Q, 2048, kind), p[field]), field
