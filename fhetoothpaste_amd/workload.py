"""Synthetic workloads of the bootstrapping path: LUT test vectors, request streams.

Test vectors follow NegacyclicFunctionEvaluator::EvalFunc's builder
(negacyclic_function_evaluator.cpp:70-89) as intended: b = [f(0), -f(N-1), ..., -f(1)], a = 0,
so the constant coefficient of X^p * TV is f(p) for p in [0, N) and -f(p - N) beyond.  (The
reference writes f(0) into a[0] instead of b[0]; with a = 0 kept, f(0) lands in b[0] here.)
"""
import numpy as np

Q = 0xFFFFFFFF00000001


def lut_test_vector(f_vals):
    """f_vals: f(0..N-1) as uint64 mod Q -> test vector [a = 0 | b] of 2N coefficients."""
    f = np.ascontiguousarray(f_vals, dtype=np.uint64)
    N = f.size
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N] = f[0]
    rev = f[1:][::-1]  # f(N-1), ..., f(1)
    tv[N + 1:] = np.where(rev == 0, np.uint64(0), np.uint64(Q) - rev)
    return tv


def random_3bit_lut(N, rng):
    """f(p) = v_{floor(8p/N)} * floor(Q/16), v_k uniform mod 8; returns (test vector, v)."""
    v = rng.integers(0, 8, size=8).astype(np.uint64)
    f = np.repeat(v * np.uint64(Q // 16), N // 8)
    return lut_test_vector(f), v


def sign_test_vector(N):
    """Sign LUT of the plain and batched workloads: b = floor(Q/8) everywhere, a = 0."""
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    return tv


def lognormal_counts(rng, size, median=32, sigma_ln=1.0, lo=1, hi=512):
    """Ciphertext counts per request: round(lognormal(ln median, sigma_ln)) clipped to [lo, hi]."""
    c = np.rint(rng.lognormal(np.log(median), sigma_ln, size=size))
    return np.clip(c, lo, hi).astype(np.int64)


def poisson_arrivals(rng, rate, duration):
    """Arrival instants (seconds from 0) of a Poisson stream of `rate` per second over `duration`."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.array(out)
        out.append(t)


def eval_phase(lwe, key, N):
    """b - <a, s> mod 2N of an input LWE in Z_2N (key entries 0, 1 or Q-1 for -1)."""
    n = lwe.size - 1
    s = np.where(np.asarray(key) == np.uint64(Q - 1), -1, np.asarray(key).astype(np.int64))
    return int((int(lwe[n]) - int(np.dot(lwe[:n].astype(np.int64), s))) % (2 * N))
