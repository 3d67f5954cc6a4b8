"""The vectorised CPU port (oracle/cpu_port.cpp, the CPU baseline bench.py times) equals the scalar
oracle word for word: transforms at N = 1024/2048, and full bootstraps (binary key, blind rotation
+ sample extraction + key switch) at the benchmark's gadgets."""
import numpy as np
import pytest

Q = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def port(oracle):
    if oracle.port() is None:
        pytest.skip("no AVX2 on this CPU")
    return oracle


@pytest.mark.parametrize("N", [16, 1024, 2048])
def test_port_ntt_matches_oracle(port, N):
    rng = np.random.default_rng(N)
    a = rng.integers(0, Q, size=(3, N), dtype=np.uint64)
    a[0, :] = Q - 1
    R = port.Ring(N)
    fwd = port.port_ntt(N, a)
    for c in range(3):
        assert np.array_equal(fwd[c], R.ntt(a[c])), c
    back = port.port_ntt(N, fwd, inverse=True)
    assert np.array_equal(back, a)


@pytest.mark.parametrize("n,b,l,n_out,kb,kd", [(24, 12, 3, 16, 4, 14), (20, 16, 4, 12, 4, 16), (8, 10, 2, 8, 5, 12)])
def test_port_bootstrap_matches_oracle(port, n, b, l, n_out, kb, kd):
    N = 1024
    lk = port.secret(n, 0, 3, 1)
    rk = port.secret(N, 0, 3, 2)
    brk = port.brk(N, lk, rk, n, b, l, port.BINARY, 0, 3.2, 3)
    ksk = port.lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 3)
    count = 5
    lwes = port.random_mod_pow2(count * (n + 1), 2 * N, 3, 4).reshape(count, n + 1)
    lwes[1, ::2] = 0      # skipped steps
    lwes[2, n] = N + 5    # X^b with b >= N
    tv = np.random.default_rng(n).integers(0, Q, size=2 * N, dtype=np.uint64)
    ref, _ = port.bootstrap_batch(N, lwes.reshape(-1), tv, brk, n, b, l, port.BINARY, ksk, n_out, kb, kd, 1)
    got, _, per = port.port_bootstrap_batch(N, lwes, tv, brk, n, b, l, ksk, n_out, kb, kd, threads=2, per_ct=True)
    assert np.array_equal(got, ref)
    assert (per > 0).all()
    ext_ref, _ = port.bootstrap_batch(N, lwes.reshape(-1), tv, brk, n, b, l, port.BINARY)
    ext, _ = port.port_bootstrap_batch(N, lwes, tv, brk, n, b, l)
    assert np.array_equal(ext, ext_ref)


def test_port_bootstrap_headline_shape(port):
    """One ciphertext at the benchmark's shape (n = 630, N = 1024, 2^12 x 3, key switch 1024 -> 630
    with 2^4 x 14)."""
    N, n, b, l, n_out, kb, kd = 1024, 630, 12, 3, 630, 4, 14
    lk = port.secret(n, 0, 11, 1)
    rk = port.secret(N, 0, 11, 2)
    brk = port.brk(N, lk, rk, n, b, l, port.BINARY, 0, 3.2, 11)
    ksk = port.lwe_ksk(rk, lk, N, n_out, kb, kd, 3.2, 11)
    lwes = port.random_mod_pow2(n + 1, 2 * N, 11, 12).reshape(1, n + 1)
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    ref, _ = port.bootstrap_batch(N, lwes.reshape(-1), tv, brk, n, b, l, port.BINARY, ksk, n_out, kb, kd, 1)
    got, _ = port.port_bootstrap_batch(N, lwes, tv, brk, n, b, l, ksk, n_out, kb, kd, threads=1)
    assert np.array_equal(got, ref)
