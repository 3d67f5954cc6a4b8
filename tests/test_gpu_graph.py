"""GPU: the graph-captured bootstrap (tp_graph_bootstrap_create / tp_graph_launch) equals the eager
tp_bootstrap_batch_device word for word on every replay: several replays of one graph, with eager
bootstraps on other buffers and host synchronisations between them, new inputs written into the
captured buffers between replays, and both kernel routes (digit-parallel for a small count, the
one-wave-per-ciphertext kernel above the threshold).  Sampled rows also equal the oracle."""
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
    N, n, b, l, n_out, kb, kd = 1024, 100, 12, 3, 100, 4, 14
    lk = tp.keygen_secret(n, 0, 31, 1)
    rk = tp.keygen_secret(N, 0, 31, 2)
    brk = tp.keygen_brk(lk, rk, N, n, b, l, 0, 0, 3.2, 31)
    ksk = tp.keygen_lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 31)
    ctx = tp.Context(N, n, b, l, 0, n_out, kb, kd)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    yield tp, ctx, brk, ksk, (N, n, b, l, n_out, kb, kd)
    ctx.close()


@pytest.mark.parametrize("count", [40, 700])
def test_graph_replays_equal_eager(setup, oracle, count):
    tp, ctx, brk, ksk, (N, n, b, l, n_out, kb, kd) = setup
    say("graph bootstrap", count)
    rng = np.random.default_rng(count)
    tv = rng.integers(0, tp.Q, size=2 * N, dtype=np.uint64)
    d_in = tp.DeviceBuffer(ctx, count * (n + 1))
    d_tv = tp.DeviceBuffer.from_host(ctx, tv)
    d_out = tp.DeviceBuffer(ctx, count * (n_out + 1))
    g = ctx.bootstrap_graph(count, d_in, d_tv, 0, d_out)
    try:
        for rep in range(4):
            lwes = rng.integers(0, 2 * N, size=(count, n + 1), dtype=np.uint64)
            d_in.copy_from_host(lwes)
            g.launch()
            ctx.sync()
            got = d_out.to_host().reshape(count, n_out + 1)
            # eager work between replays: the same inputs through the allocating entry point
            want = ctx.bootstrap_batch(lwes, tv)
            assert np.array_equal(got, want), rep
            if rep == 0:
                rows = [0, count // 2, count - 1]
                ref, _ = oracle.bootstrap_batch(N, np.ascontiguousarray(lwes[rows]).reshape(-1), tv, brk, n, b, l,
                                                oracle.BINARY, ksk, n_out, kb, kd, 1)
                assert np.array_equal(got[rows], ref)
    finally:
        g.close()
        for buf in (d_in, d_tv, d_out):
            buf.free()


def test_graph_rejects(setup):
    tp, ctx, *_ = setup
    say("graph rejects")
    from fhetoothpaste_amd import ToothpasteError
    d = tp.DeviceBuffer(ctx, 16)
    with pytest.raises(ToothpasteError):
        ctx.bootstrap_graph(0, d, d, 0, d)
    d.free()
