"""GPU: the tuning cache (tp_autotune_cached).  A miss times the variants and writes the JSON file;
a second context of the same parameter set and count hits it (no timing) and takes the same
settings, the digit-parallel form (one or two waves per transform) included; another count misses; a malformed file is rejected; launches on the tuned settings still
equal the oracle."""
import json
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
Q = 0xFFFFFFFF00000001


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_tuning_cache_roundtrip(oracle, tmp_path):
    import fhetoothpaste_amd as tp
    from fhetoothpaste_amd import ToothpasteError
    assert tp.device_count() >= 1, "no gfx950 device"
    N, n, b, l = 1024, 24, 12, 3
    say("tuning cache", tmp_path)
    lk = tp.keygen_secret(n, 0, 41, 1)
    rk = tp.keygen_secret(N, 0, 41, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 0, 0, 3.2, 41)
    path = str(tmp_path / "tuning.json")
    settings = []
    for attempt in range(2):
        ctx = tp.Context(N, n, b, l, 0)
        ctx.import_brk(brk)
        if attempt == 1:
            ctx.set_br_split_halves(1 - settings[0][4])   # a hit restores the stored form
        r = ctx.autotune(1500, cache=path)
        assert r["cache_hit"] == (attempt == 1)
        settings.append((r["waves"], r["split_max"], r["engine"], r["grid"], r["split_halves"]))
        assert r["ms_two_waves"] > 0 and r["ms_digit_parallel"] > 0 and r["split_halves"] in (0, 1)
        if attempt == 0:
            doc = json.load(open(path))
            (e,) = doc["entries"]
            assert (e["N"], e["n"], e["b"], e["l"], e["method"], e["count"]) == (N, n, b, l, 0, 1500)
            assert e["arch"].startswith("gfx950") and len(e["ms"]) == 6
            assert e["split_halves"] == r["split_halves"]
        lwes = oracle.random_mod_pow2(3 * (n + 1), 2 * N, 42, attempt).reshape(3, n + 1)
        tv = np.random.default_rng(attempt).integers(0, Q, size=2 * N, dtype=np.uint64)
        _, rl = ctx.eval_func_batch(lwes, tv, want_rlwe=True)
        for c in range(3):
            assert np.array_equal(rl[c], oracle.blind_rotate(N, lwes[c], tv, brk, n, b, l, oracle.BINARY, True)), c
        if attempt == 1:
            r2 = ctx.autotune(64, cache=path)      # another count: a miss, appended
            assert not r2["cache_hit"]
            assert len(json.load(open(path))["entries"]) == 2
        ctx.close()
    assert settings[0] == settings[1]
    bad = tmp_path / "bad.json"
    bad.write_text('{"entries": [{"arch": "gfx950", "N": }]}')
    ctx = tp.Context(N, n, b, l, 0)
    ctx.import_brk(brk)
    with pytest.raises(ToothpasteError):
        ctx.autotune(1500, cache=str(bad))
    ctx.close()
