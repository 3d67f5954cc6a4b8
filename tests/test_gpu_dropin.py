"""The C++ drop-in classes (include/toothpaste_amd.hpp), driven by tests/cpp/dropin_parity.cpp the
way the reference's own tests drive the reference classes, checked bit for bit against the oracle:
CGGI (binary, ternary) and BMMP rotators incl. batch and result-aliases-accumulator, the function
evaluator (LUT, std::function with the reference's test-vector builder, batch), LWE and RLWE
converters with the reference's subtract-into-output semantics, automorphism Eval/EvalSwitchOnly,
trace Eval/EvalAuto (both overloads' evaluator indexing), Pack/PackConsecutively, the MathWorker
and SampleExtract.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 0xFFFFFFFF00000001
N = 1024
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_parity")


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin")
    print("[case] dropin_parity", out, flush=True)
    sys.stdout.flush()
    assert os.path.exists(BIN), "tests/cpp/dropin_parity is not built (python -m fhetoothpaste_amd.build)"
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr

    def load(name):
        return np.fromfile(os.path.join(out, name + ".bin"), dtype=np.uint64)
    return load


def _rotator(oracle, load, tag, n, dist, b, l, method, seed):
    lk = oracle.secret(n, dist, seed, 1)
    rk = oracle.secret(N, 0, seed, 2)
    brk = oracle.brk(N, lk, rk, n, b, l, method, 0 if method == oracle.BINARY else 1, 3.2, seed)
    assert np.array_equal(load(tag + "_brk"), brk)
    lwe, acc = load(tag + "_lwe"), load(tag + "_acc")
    ref_ntt = oracle.blind_rotate(N, lwe, acc, brk, n, b, l, method, False)
    ref_coef = oracle.blind_rotate(N, lwe, acc, brk, n, b, l, method, True)
    assert np.array_equal(load(tag + "_out_ntt"), ref_ntt)
    assert np.array_equal(load(tag + "_out_coef"), ref_coef)
    assert np.array_equal(load(tag + "_alias"), ref_coef)
    lwes = load(tag + "_batch_lwes").reshape(3, n + 1)
    batch = load(tag + "_batch").reshape(3, 2 * N)
    for c in range(3):
        assert np.array_equal(batch[c], oracle.blind_rotate(N, lwes[c], acc, brk, n, b, l, method, True)), c
    return lk, rk, brk


def test_cggi_binary_and_function_evaluator(run, oracle):
    n, b, l = 32, 12, 3
    lk, rk, brk = _rotator(oracle, run, "cggi_bin", n, 0, b, l, oracle.BINARY, 101)
    lwe, lut = run("fe_lwe"), run("fe_lut")
    ext = oracle.sample_extract(oracle.blind_rotate(N, lwe, lut, brk, n, b, l, oracle.BINARY, True), N)
    assert np.array_equal(run("fe_out_lut"), ext)
    # std::function overload: word 0 = f(0), word N + i = -f(N - i) (negacyclic_function_evaluator.cpp:70-89)
    f = [((x * 7 + 3) % 16) * (Q // 16) for x in range(N)]
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[0] = f[0]
    for i in range(1, N):
        tv[N + i] = (Q - f[N - i]) % Q
    ext_f = oracle.sample_extract(oracle.blind_rotate(N, lwe, tv, brk, n, b, l, oracle.BINARY, True), N)
    assert np.array_equal(run("fe_out_fn"), ext_f)
    lwes = run("fe_batch_lwes").reshape(4, n + 1)
    got = run("fe_batch").reshape(4, N + 1)
    for c in range(4):
        ref = oracle.sample_extract(oracle.blind_rotate(N, lwes[c], lut, brk, n, b, l, oracle.BINARY, True), N)
        assert np.array_equal(got[c], ref), c


def test_cggi_ternary(run, oracle):
    _rotator(oracle, run, "cggi_ter", 16, 1, 16, 4, oracle.TERNARY, 102)


def test_bmmp(run, oracle):
    _rotator(oracle, run, "bmmp", 16, 0, 14, 3, oracle.BMMP, 103)


def test_lwe_converter(run, oracle):
    n_out, b, d = 64, 4, 14
    src = oracle.secret(N, 0, 104, 2)
    dst = oracle.secret(n_out, 0, 104, 1)
    ksk = oracle.lwe_ksk(src, dst, N, n_out, b, d, 3.2, 104)
    assert np.array_equal(run("lks_key"), ksk)
    res = oracle.lwe_keyswitch(run("lks_in"), ksk, N, n_out, b, d)  # (0.., b) - acc
    prev = run("lks_prev")
    want = np.array([(int(prev[k]) + int(res[k])) % Q for k in range(n_out)] + [int(res[n_out])], dtype=np.uint64)
    assert np.array_equal(run("lks_out"), want)
    ins = run("lks_batch_in").reshape(3, N + 1)
    outs = run("lks_batch").reshape(3, n_out + 1)
    for c in range(3):
        assert np.array_equal(outs[c], oracle.lwe_keyswitch(ins[c], ksk, N, n_out, b, d)), c


def test_rlwe_converter(run, oracle):
    src = oracle.secret(N, 0, 105, 3)
    dst = oracle.secret(N, 0, 105, 2)
    key = oracle.rlwe_ksk(N, src, dst, 16, 4, 3.2, 105)
    assert np.array_equal(run("rks_key"), key)
    assert np.array_equal(run("rks_out"), oracle.rlwe_keyswitch(N, run("rks_prev"), run("rks_in"), key, 16, 4))


def test_automorphism_evaluator(run, oracle):
    t = 5
    s = oracle.secret(N, 0, 106, 2)
    key = oracle.rlwe_ksk(N, oracle.automorphism_poly(s, t, N), s, 16, 4, 3.2, 106)
    x, prev = run("auto_in"), run("auto_prev")
    R = oracle.Ring(N)
    ta = oracle.automorphism_poly(x[:N], t, N)
    buf = np.concatenate([ta, R.ntt(oracle.automorphism_poly(x[N:], t, N))])
    assert np.array_equal(run("auto_out"), oracle.rlwe_keyswitch(N, prev, buf, key, 16, 4))
    buf0 = np.concatenate([ta, np.zeros(N, np.uint64)])
    assert np.array_equal(run("auto_switch_only"), oracle.rlwe_keyswitch(N, prev, buf0, key, 16, 4))


def test_trace_evaluator(run, oracle):
    b, d = 16, 4
    s = oracle.secret(N, 0, 107, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 107)
    x = run("trace_in")
    assert np.array_equal(run("trace_out"), oracle.trace(N, x, keys, b, d))
    kw = d * 2 * N
    # EvalAuto(ptr, 9): evaluator index 2 -> t = 9, key 3; EvalAuto(vector, 9): evaluator 3 -> t = 17, key 4
    assert np.array_equal(run("trace_auto_ptr"), oracle.automorphism(N, x, 9, keys[2 * kw:3 * kw], b, d))
    assert np.array_equal(run("trace_auto_vec"), oracle.automorphism(N, x, 17, keys[3 * kw:4 * kw], b, d))


def test_packer(run, oracle):
    b, d = 16, 4
    s = oracle.secret(N, 0, 108, 2)
    keys = oracle.auto_keys(N, s, b, d, 3.2, 108)
    lwes = run("pack_lwes").reshape(16, N + 1)
    assert np.array_equal(run("pack_out"), oracle.pack(N, lwes, 16, keys, b, d))
    assert np.array_equal(run("pack_consecutive"), oracle.pack_consecutive(N, lwes[:5], 5, keys, b, d))


def test_math_worker(run, oracle):
    R = oracle.Ring(N)
    x = run("mw_x")
    assert np.array_equal(run("mw_ntt"), R.ntt(x))
    assert np.array_equal(run("mw_back"), x)
    a, b, c = (run(k).astype(object) for k in ("mw_a", "mw_b", "mw_c"))
    L = a.size
    s = 0x123456789ABCDEF
    elw = run("mw_elw").astype(object).reshape(6, L)
    want = [(a + b) % Q, (a - b) % Q, (a * b) % Q, (a * b + c) % Q, (a * s) % Q, run("mw_raw").astype(object) % Q]
    for i, w in enumerate(want):
        assert np.array_equal(elw[i], w), i
    assert np.array_equal(run("mw_addeq").astype(object), (a + b) % Q)
    sc = [int(v) for v in run("mw_scalar")]
    a0, b0, c0 = int(a[0]), int(b[0]), int(c[0])
    assert sc == [(a0 + b0) % Q, (a0 - b0) % Q, a0 * b0 % Q, (a0 * b0 + c0) % Q, pow(12345, Q - 2, Q),
                  pow(7, 1000, Q), 0, Q, N]


def test_sample_extract_template(run):
    x = run("se_in")
    for idx, name in ((0, "se_0"), (5, "se_5")):
        want = np.zeros(N + 1, dtype=np.uint64)
        for i in range(idx + 1):
            want[i] = x[idx - i]
        for i in range(idx + 1, N):
            if int(x[N - i]) != 0:
                want[i] = Q - int(x[N - i])
        want[N] = x[N + idx]
        assert np.array_equal(run(name), want), idx


def test_modulus_switch(run, oracle):
    x = run("ms_in")
    for kind in range(3):
        assert np.array_equal(run("ms_out%d" % kind), oracle.modswitch(x, Q, 2048, kind)), kind
    fl, ce, rnd = oracle.modswitch(x, Q, 2048, 1), oracle.modswitch(x, Q, 2048, 2), run("ms_random")
    assert np.all((rnd == fl) | (rnd == ce))


def test_lwe_to_rgsw(run, oracle):
    n = 16
    lk = oracle.secret(n, 0, 109, 1)
    s = oracle.secret(N, 0, 109, 2)
    brk = oracle.brk(N, lk, s, n, 16, 4, oracle.BINARY, 0, 3.2, 109)
    tkeys = oracle.auto_keys(N, s, 16, 4, 3.2, 110)
    R = oracle.Ring(N)
    s2 = R.intt(np.array([int(x) * int(x) % Q for x in R.ntt(s)], dtype=np.uint64))
    sq = oracle.rlwe_ksk(N, s2, s, 16, 4, 3.2, 111)
    ref = oracle.lwe_to_rgsw(N, run("rgsw_lwe"), brk, n, 16, 4, oracle.BINARY, tkeys, 16, 4, sq, 16, 4, 4, 10)
    assert np.array_equal(run("rgsw_out"), ref)


def test_sap_decomposer(run, oracle):
    n = 16
    lk = oracle.secret(n, 0, 112, 1)
    s = oracle.secret(N, 0, 112, 2)
    brk = oracle.brk(N, lk, s, n, 16, 4, oracle.BINARY, 0, 3.2, 112)
    keys = oracle.auto_keys(N, s, 16, 4, 3.2, 113)
    ksk = oracle.lwe_ksk(s, lk, N, n, 4, 14, 3.2, 114)
    beta = int(np.count_nonzero(lk))
    lwe, msg = run("sap_lwe"), int(run("sap_msg")[0])
    want = oracle.sap_decompose(N, lwe, 32, beta, brk, n, 16, 4, oracle.BINARY, keys, 16, 4)
    out = run("sap_out").reshape(8, N + 1)
    assert want.shape[0] == 8
    assert np.array_equal(out, want)
    for i in range(8):
        ph = oracle.phase_lwe(out[i], s, N)
        assert int(oracle.modswitch(np.array([ph], dtype=np.uint64), Q, 2 * N, 0)[0]) == (msg >> (5 * i)) % 32, i
    assert np.array_equal(run("sap_ht"), oracle.sap_homtrunc(N, run("sap_ht_in"), 32, keys, 16, 4))
    assert np.array_equal(run("sap_reset"),
                          oracle.sap_reset(N, run("sap_reset_in"), 32, ksk, n, 4, 14, Q, brk, n, 16, 4, oracle.BINARY))
