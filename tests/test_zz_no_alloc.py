"""GPU: tp_bootstrap_batch_device_noalloc (scratch in a static device array, no allocation) equals
tp_bootstrap_batch_device word for word, on two calls in a row, the second with a smaller input
(so the shared array holds the first call's rows when the second runs), for a shared and a
per-ciphertext test vector; counts beyond the array's limit are rejected before any launch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    import fhetoothpaste_amd as tp
    assert tp.device_count() >= 1, "no gfx950 device"
    N, n, b, l, n_out, kb, kd = 1024, 630, 12, 3, 630, 4, 14
    lk = tp.keygen_secret(n, 0, 21, 1)
    rk = tp.keygen_secret(N, 0, 21, 2)
    ctx = tp.Context(N, n, b, l, 0, n_out, kb, kd)
    ctx.import_brk(tp.keygen_brk(lk, rk, N, n, b, l, 0, 0, 3.2, 21))
    ctx.import_lwe_ksk(tp.keygen_lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 21))
    yield tp, ctx, (N, n, n_out)
    ctx.close()


def _run_both(tp, ctx, dims, count, per_ct_tv, seed):
    N, n, n_out = dims
    rng = np.random.default_rng(seed)
    lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
    tv = rng.integers(0, tp.Q, size=(count if per_ct_tv else 1) * 2 * N, dtype=np.uint64)
    want = ctx.bootstrap_batch(lwes, tv)
    d_in = tp.DeviceBuffer.from_host(ctx, lwes)
    d_tv = tp.DeviceBuffer.from_host(ctx, tv)
    d_out = tp.DeviceBuffer(ctx, count * (n_out + 1))
    ctx.bootstrap_batch_noalloc_device(count, d_in, d_tv, 2 * N if per_ct_tv else 0, d_out)
    got = d_out.to_host().reshape(count, n_out + 1)
    return got, want


@pytest.mark.parametrize("per_ct_tv", [False, True])
def test_noalloc_equals_allocating_entry_two_calls(setup, per_ct_tv):
    tp, ctx, dims = setup
    for count, seed in [(700, 1 + per_ct_tv), (130, 3 + per_ct_tv)]:
        say("noalloc bootstrap", count, "per-ct tv" if per_ct_tv else "shared tv")
        got, want = _run_both(tp, ctx, dims, count, per_ct_tv, seed)
        assert np.array_equal(got, want), count


def test_noalloc_rejects_counts_beyond_scratch(setup):
    tp, ctx, dims = setup
    N, n, n_out = dims
    say("noalloc limit")
    limit = (36 << 17) // (N + 1)
    d_in = tp.DeviceBuffer(ctx, 8)
    d_out = tp.DeviceBuffer(ctx, 8)
    with pytest.raises(tp.ToothpasteError) as e:
        ctx.bootstrap_batch_noalloc_device(limit + 1, d_in, d_in, 0, d_out)
    assert e.value.code == -1  # TP_E_INVALID, raised on the host before any launch
