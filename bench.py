"""Benchmark: CGGI bootstraps/s on one MI355X, batched persistent path vs plain path.

Workload (one "step" = one batch): 4096 independent LWE inputs (n=630, binary key, entries
uniform in [0, 2048)), blind rotation over N=1024 with a 61.9 MB NTT-domain bootstrapping
key (gadget base 2^12, 3 digits, unsigned digits as in the reference binary path), sample
extraction, then LWE->LWE key switching 1024 -> 630 (base 2^4, 14 digits).  Keys are
seeded and synthetic; inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this path runs on one GPU.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

Q = 0xFFFFFFFF00000001
N, n, BB, L = 1024, 630, 12, 3
KS_OUT, KS_B, KS_D = 630, 4, 14
BATCH = 4096
SEED = 2026


def algorithmic_mulmods():
    """Mulmods per bootstrap (every NTT butterfly counts as one), binary CGGI + key switch."""
    T = (N // 2) * (N.bit_length() - 1)
    w_br = n * ((2 + 2 * L) * T + 2 * N + 4 * L * N) + 4 * T + 2 * N
    w_ks = N * KS_D * (KS_OUT + 1)
    return w_br, w_ks


def read_traffic(batch):
    """HBM bytes per blind-rotation launch from the committed profiler summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "br_kernel_n630_N1024_batch4096_traffic.txt")
    if batch != BATCH or not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return int(float(vals["hbm_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# second resident parameter set of the serving leg: ternary n=800, N=2048, exact gadget 2^16 x 4,
# key switch 2048 -> 800 with base 2^5 and 12 digits
TER = dict(N=2048, n=800, b=16, l=4, method=1, n_out=800, kb=5, kd=12)


def serving_leg(args, tp, ctx_bin):
    """Poisson stream over {binary n=630 N=1024, ternary n=800 N=2048}, random 3-bit LUTs,
    lognormal counts (median 32, sigma_ln 1.0, clipped to [1, 512]); lambda = 0.8 x peak."""
    from fhetoothpaste_amd.serving import Server, replay
    from fhetoothpaste_amd.workload import lognormal_counts, poisson_arrivals, random_3bit_lut

    t0 = time.time()
    seed = SEED + 1
    lk3 = tp.keygen_secret(TER["n"], 1, seed, 1)
    rk3 = tp.keygen_secret(TER["N"], 0, seed, 2)
    ctx_ter = tp.Context(TER["N"], TER["n"], TER["b"], TER["l"], TER["method"], TER["n_out"], TER["kb"], TER["kd"])
    ctx_ter.import_brk(tp.keygen_brk(lk3, rk3, TER["N"], TER["n"], TER["b"], TER["l"], TER["method"], 1, 3.2, seed))
    ctx_ter.import_lwe_ksk(tp.keygen_lwe_ksk(rk3, lk3, TER["N"], TER["n_out"], TER["kb"], TER["kd"], 3.2, seed))
    tune = ctx_ter.autotune(4096)
    log("serving keys ready (%.1f s); ternary set variant %d wave(s)/SIMD (two: %.1f ms, one: %.1f ms)" % (
        time.time() - t0, tune["waves"], tune["ms_two_waves"], tune["ms_one_wave"]))

    rng = np.random.default_rng(SEED + 5)
    dims = [(N, n), (TER["N"], TER["n"])]
    pools = [rng.integers(0, 2 * NN, size=(4096, nn + 1), dtype=np.uint64) for NN, nn in dims]

    def draw(k):
        sets = rng.integers(0, 2, size=k)
        counts = lognormal_counts(rng, k)
        reqs = []
        for si, cnt in zip(sets, counts):
            off = int(rng.integers(0, 4096 - cnt + 1))
            tv, _ = random_3bit_lut(dims[si][0], rng)
            reqs.append([0.0, int(si), pools[si][off:off + cnt], tv])
        return reqs

    def summarize(records):
        lat = np.array([r["done"] - r["arrival"] for r in records])
        cts = sum(r["count"] for r in records)
        span = max(r["done"] for r in records) - min(r["arrival"] for r in records)
        return lat, cts, span

    mean_count = float(lognormal_counts(np.random.default_rng(7), 200000).mean())
    with Server([ctx_bin, ctx_ter], chunk_steps=args.serve_chunk) as srv:
        # peak: a burst of the same request mix, all arriving at once
        burst = draw(400)
        rec = replay(srv, burst)
        _, cts, span = summarize(rec)
        peak = cts / span
        log("serving peak %.0f bootstraps/s (%d ciphertexts in %.2f s)" % (peak, cts, span))
        lam = 0.8 * peak / mean_count
        arrivals = poisson_arrivals(rng, lam, args.serve_seconds)
        stream = draw(len(arrivals))
        for r, t in zip(stream, arrivals):
            r[0] = float(t)
        log("serving stream: %d requests over %.0f s (lambda %.2f/s)" % (len(stream), args.serve_seconds, lam))
        rec = replay(srv, stream, progress=log)
        st = srv.stats()
    ctx_ter.close()
    lat, cts, span = summarize(rec)
    out = {
        "sets": ["binary n=630 N=1024 b=2^12 l=3, KS 1024->630 b=2^4 d=14",
                 "ternary n=800 N=2048 b=2^16 l=4, KS 2048->800 b=2^5 d=12"],
        "requests": len(rec), "ciphertexts": int(cts), "duration_s": round(span, 3),
        "peak_bootstraps_per_s": round(peak, 1), "lambda_requests_per_s": round(lam, 3),
        "mean_count": round(mean_count, 2),
        "offered_bootstraps_per_s": round(lam * mean_count, 1),
        "throughput_bootstraps_per_s": round(cts / span, 1),
        "p50_ms": round(1000 * float(np.percentile(lat, 50)), 2),
        "p99_ms": round(1000 * float(np.percentile(lat, 99)), 2),
        "ternary_br_waves_per_simd": tune["waves"],
        "chunk_steps": args.serve_chunk, "launches": st["launches"], "iterations": st["iterations"],
        "counts": "lognormal median 32 sigma_ln 1.0 clipped to [1, 512]",
        "lut": "random 3-bit, f(p) = v_floor(8p/N) * floor(q/16)",
    }
    for si in (0, 1):
        l_s = np.array([r["done"] - r["arrival"] for r in rec if r["set"] == si])
        if l_s.size:
            out["set%d_p50_ms" % si] = round(1000 * float(np.percentile(l_s, 50)), 2)
            out["set%d_p99_ms" % si] = round(1000 * float(np.percentile(l_s, 99)), 2)
    return out


def timed(ctx, lib, chk, fn, reps):
    """Average ms of fn() on the context stream (HIP events), after one warm-up call."""
    import ctypes
    fn()
    ctx.sync()
    chk(lib.tp_timer_mark(ctx.h, 3))
    for _ in range(reps):
        fn()
    chk(lib.tp_timer_mark(ctx.h, 4))
    ms = ctypes.c_float()
    chk(lib.tp_timer_elapsed(ctx.h, 3, 4, ctypes.byref(ms)))
    return ms.value / reps


def extra_workloads(tp, reps=2):
    """The other workloads: uneven ternary batches (n=800, N=2048) in one launch with a random 3-bit
    LUT per batch, BMMP (n=1024, N=2048) followed by packing all outputs, and SAP digit
    decomposition (N=4096, radix 32) over a batch."""
    from fhetoothpaste_amd import _lib
    from fhetoothpaste_amd.operators import DeviceBuffer
    from fhetoothpaste_amd.workload import lognormal_counts, random_3bit_lut
    lib, chk = _lib.lib(), _lib.check
    out = {}
    rng = np.random.default_rng(SEED + 9)

    # ternary, 16 lognormal batches (median 256, sigma_ln 1.0, clipped to [1, 4096]) concatenated
    P = TER
    seed = SEED + 1
    lk = tp.keygen_secret(P["n"], 1, seed, 1)
    rk = tp.keygen_secret(P["N"], 0, seed, 2)
    ctx = tp.Context(P["N"], P["n"], P["b"], P["l"], P["method"], P["n_out"], P["kb"], P["kd"])
    ctx.import_brk(tp.keygen_brk(lk, rk, P["N"], P["n"], P["b"], P["l"], P["method"], 1, 3.2, seed))
    ctx.import_lwe_ksk(tp.keygen_lwe_ksk(rk, lk, P["N"], P["n_out"], P["kb"], P["kd"], 3.2, seed))
    sizes = lognormal_counts(rng, 16, median=256, sigma_ln=1.0, lo=1, hi=4096)
    total = int(sizes.sum())
    tvs = np.concatenate([np.tile(random_3bit_lut(P["N"], rng)[0], (int(c), 1)) for c in sizes])
    lwes = rng.integers(0, 2 * P["N"], size=(total, P["n"] + 1), dtype=np.uint64)
    tune = ctx.autotune(total)
    d_lwe, d_tv = DeviceBuffer.from_host(ctx, lwes), DeviceBuffer.from_host(ctx, tvs)
    d_out = DeviceBuffer(ctx, total * (P["n_out"] + 1))
    ms = timed(ctx, lib, chk, lambda: chk(lib.tp_bootstrap_batch_device(ctx.h, total, d_lwe.p, d_tv.p, 2 * P["N"],
                                                                         d_out.p)), reps)
    out["ternary_uneven_batches"] = {
        "workload": "16 lognormal batches (median 256, max 4096) in one launch, ternary n=800 N=2048 b=2^16 l=4, "
                    "random 3-bit LUT per batch, KS 2048->800 b=2^5 d=12",
        "ciphertexts": total, "batch_sizes": [int(c) for c in sizes], "ms": round(ms, 2),
        "bootstraps_per_s": round(total / ms * 1000, 1), "br_waves_per_simd": tune["waves"]}
    log("ternary uneven batches: %d ciphertexts, %.1f ms" % (total, ms))
    for b in (d_lwe, d_tv, d_out):
        b.free()
    ctx.close()

    # BMMP n=1024 N=2048 b=2^14 l=3, batch 2048, then packing the 2048 extracted LWEs (permuted)
    N4, n4, B4 = 2048, 1024, 2048
    seed = SEED + 2
    lk = tp.keygen_secret(n4, 0, seed, 1)
    rk = tp.keygen_secret(N4, 0, seed, 2)
    ctx = tp.Context(N4, n4, 14, 3, 2, 0, 0, 0, 0, 16, 4)
    ctx.import_brk(tp.keygen_brk(lk, rk, N4, n4, 14, 3, 2, 1, 3.2, seed))
    ctx.import_auto_keys(tp.keygen_auto_keys(rk, N4, 16, 4, 3.2, seed))
    tune = ctx.autotune(B4)
    lwes = rng.integers(0, 2 * N4, size=(B4, n4 + 1), dtype=np.uint64)
    tv = random_3bit_lut(N4, rng)[0]
    perm = rng.permutation(B4).astype(np.uint32)
    d_lwe, d_tv = DeviceBuffer.from_host(ctx, lwes), DeviceBuffer.from_host(ctx, tv)
    d_ext = DeviceBuffer(ctx, B4 * (N4 + 1))
    d_perm = DeviceBuffer.from_host(ctx, np.frombuffer(perm.tobytes(), dtype=np.uint64))
    d_pack = DeviceBuffer(ctx, 2 * N4)
    ms_br = timed(ctx, lib, chk, lambda: chk(lib.tp_eval_func_batch_device(ctx.h, B4, d_lwe.p, d_tv.p, 0, d_ext.p,
                                                                            None, 0)), reps)
    ms_pack = timed(ctx, lib, chk, lambda: chk(lib.tp_pack_device(ctx.h, B4, d_ext.p, d_perm.p, d_pack.p)), reps)
    out["bmmp_and_packing"] = {
        "workload": "BMMP n=1024 N=2048 b=2^14 l=3 batch 2048 (blind rotation + sample extract), then packing the "
                    "2048 outputs under a random permutation (automorphism keys b=2^16 d=4)",
        "br_ms": round(ms_br, 2), "pack_ms": round(ms_pack, 3),
        "bootstraps_per_s": round(B4 / (ms_br + ms_pack) * 1000, 1), "br_waves_per_simd": tune["waves"]}
    log("BMMP batch %.1f ms, packing %.2f ms" % (ms_br, ms_pack))
    for b in (d_lwe, d_tv, d_ext, d_perm, d_pack):
        b.free()
    ctx.close()

    # SAP digit decomposition at the reference's test parameters (N=4096, n=10, radix 32, inputs
    # mod 32^8 -> 8 digits), a batch of inputs per call
    N5, n5, B5, radix = 4096, 10, 1024, 32
    seed = SEED + 3
    lk = tp.keygen_secret(n5, 0, seed, 1)
    rk = tp.keygen_secret(N5, 0, seed, 2)
    ctx = tp.Context(N5, n5, 16, 4, 0, n5, 16, 4, 0, 16, 4)
    ctx.import_brk(tp.keygen_brk(lk, rk, N5, n5, 16, 4, 0, 0, 3.2, seed))
    ctx.import_auto_keys(tp.keygen_auto_keys(rk, N5, 16, 4, 3.2, seed))
    lwes = rng.integers(0, radix ** 8, size=(B5, n5 + 1), dtype=np.uint64)
    lwes[:, 0] |= np.uint64(1 << 39)  # modulus 2^40 read off the values: 8 digits each
    beta = int(np.count_nonzero(lk))
    d_lwe = DeviceBuffer.from_host(ctx, lwes)
    d_out = DeviceBuffer(ctx, B5 * 8 * (N5 + 1))
    d_cnt = DeviceBuffer(ctx, B5 // 2)
    ms_sap = timed(ctx, lib, chk, lambda: chk(lib.tp_sap_decompose_device(ctx.h, B5, radix, beta, (1 << 64) - 1,
                                                                         d_lwe.p, d_out.p, 8, d_cnt.p)), reps)
    out["sap_decomposition"] = {
        "workload": "SAP digit decomposition, binary n=10 N=4096 b=2^16 l=4, automorphism keys b=2^16 d=4, radix 32, "
                    "%d inputs mod 2^40 (8 digits each) per call" % B5,
        "inputs": B5, "ms": round(ms_sap, 2), "decompositions_per_s": round(B5 / ms_sap * 1000, 1),
        "digits_per_s": round(8 * B5 / ms_sap * 1000, 1)}
    log("SAP decomposition: %d inputs, %.1f ms" % (B5, ms_sap))
    for b in (d_lwe, d_out, d_cnt):
        b.free()
    ctx.close()

    # LWE -> RGSW scheme switching (the reference test's shape: n=100, N=1024, gadget 2^14 x 4 for the
    # rotation, the trace and the s^2 -> s switch, 4 output digits of base 2^14), 1024 inputs per call
    N6, n6, B6 = 1024, 100, 1024
    seed = SEED + 4
    lk = tp.keygen_secret(n6, 0, seed, 1)
    rk = tp.keygen_secret(N6, 0, seed, 2)
    ctx = tp.Context(N6, n6, 14, 4, 0, 0, 0, 0, 0, 14, 4)
    ctx.import_brk(tp.keygen_brk(lk, rk, N6, n6, 14, 4, 0, 0, 3.2, seed))
    ctx.import_auto_keys(tp.keygen_auto_keys(rk, N6, 14, 4, 3.2, seed))
    # s^2 mod X^N + 1 (binary s, small integer coefficients) for the squaring key
    full = np.convolve(rk.astype(np.int64), rk.astype(np.int64))
    sq_int = full[:N6].copy()
    sq_int[:N6 - 1] -= full[N6:]
    sq = np.array([v % Q for v in sq_int.tolist()], dtype=np.uint64)
    ctx.import_rlwe_ksk(tp.keygen_rlwe_ksk(sq, rk, N6, 14, 4, 3.2, seed))
    lwes = rng.integers(0, 2 * N6, size=(B6, n6 + 1), dtype=np.uint64)
    d_lwe = DeviceBuffer.from_host(ctx, lwes)
    d_out = DeviceBuffer(ctx, B6 * 4 * N6 * 4)
    ms_rg = timed(ctx, lib, chk, lambda: chk(lib.tp_lwe_to_rgsw_device(ctx.h, B6, 4, 14, d_lwe.p, d_out.p)), reps)
    out["lwe_to_rgsw"] = {
        "workload": "LWE -> RGSW scheme switching, binary n=100 N=1024, gadget 2^14 x 4 (rotation, trace, "
                    "s^2 -> s switch), 4 output digits of base 2^14, %d inputs per call" % B6,
        "inputs": B6, "ms": round(ms_rg, 2), "conversions_per_s": round(B6 / ms_rg * 1000, 1)}
    log("LWE->RGSW: %d inputs, %.1f ms" % (B6, ms_rg))
    for b in (d_lwe, d_out):
        b.free()
    ctx.close()

    # modulus switch Q -> 2N (ROUND, the reference's long double rounding) of a key-switched batch:
    # 4096 LWEs of dimension 630, the link from a bootstrap's output back to a blind-rotation input
    ctx = tp.Context(1024, 2, 16, 4)
    words = 4096 * 631
    d_src = DeviceBuffer.from_host(ctx, rng.integers(0, Q, size=words, dtype=np.uint64))
    d_dst = DeviceBuffer(ctx, words)
    ms_ms = timed(ctx, lib, chk, lambda: chk(lib.tp_modswitch_device(ctx.h, d_dst.p, d_src.p, words, Q, 2048, 0, 0)),
                  max(reps, 5))
    out["modulus_switch"] = {
        "workload": "ModulusSwitch Q -> 2048 (ROUND) of 4096 LWEs of dimension 630 (2,584,576 words)",
        "ms": round(ms_ms, 4), "words_per_s": round(words / ms_ms * 1000, 1),
        "GBps": round(16 * words / ms_ms / 1e6, 1), "hbm_peak_GBps": 8000}
    log("modulus switch: %d words, %.3f ms" % (words, ms_ms))
    for b in (d_src, d_dst):
        b.free()
    ctx.close()

    # batched negacyclic NTT (forward and inverse, in place) of 16384 polynomials at N = 1024 and 2048
    out["ntt_batch"] = {}
    for NN in (1024, 2048):
        P = 16384
        ctx = tp.Context(NN, 4, 12, 3, 0)
        d = DeviceBuffer.from_host(ctx, rng.integers(0, Q, size=P * NN, dtype=np.uint64))
        ms_f = timed(ctx, lib, chk, lambda: chk(lib.tp_ntt_forward_device(ctx.h, d.p, d.p, P)), reps)
        ms_i = timed(ctx, lib, chk, lambda: chk(lib.tp_ntt_inverse_device(ctx.h, d.p, d.p, P)), reps)
        butterflies = P * (NN // 2) * (NN.bit_length() - 1)
        out["ntt_batch"][str(NN)] = {
            "polynomials": P, "forward_ms": round(ms_f, 3), "inverse_ms": round(ms_i, 3),
            "forward_per_s": round(P / ms_f * 1000, 1), "inverse_per_s": round(P / ms_i * 1000, 1),
            "forward_GBps": round(16 * NN * P / ms_f / 1e6, 1),
            "forward_butterflies_per_s": round(butterflies / ms_f * 1000, 1)}
        log("NTT N=%d: %d polynomials, forward %.3f ms, inverse %.3f ms" % (NN, P, ms_f, ms_i))
        d.free()
        ctx.close()

    # wire format (host code): encode and decode of a blind-rotation key of the headline shape
    # (binary n=630, N=1024, l=3: 7,741,440 words), header checks and FNV-1a checksum included
    from fhetoothpaste_amd import wire
    params = _lib.Params(N, n, BB, L, 0, KS_OUT, KS_B, KS_D, 0, 0)
    key = rng.integers(0, Q, size=n * 4 * N * L, dtype=np.uint64)
    t0 = time.perf_counter()
    blob = wire.encode(wire.BRK, params, key, wire.NTT_DOMAIN)
    t1 = time.perf_counter()
    kind, _, _, back = wire.decode(blob)
    t2 = time.perf_counter()
    if kind != wire.BRK or not np.array_equal(back, key):
        raise RuntimeError("wire round trip changed the key")
    out["wire_format"] = {
        "workload": "TPW1 encode + decode (host, one thread) of a binary n=630 N=1024 l=3 blind-rotation key "
                    "(%d bytes)" % len(blob),
        "encode_ms": round((t1 - t0) * 1e3, 2), "decode_ms": round((t2 - t1) * 1e3, 2),
        "encode_GBps": round(len(blob) / (t1 - t0) / 1e9, 2), "decode_GBps": round(len(blob) / (t2 - t1) / 1e9, 2)}
    log("wire format: encode %.1f ms, decode %.1f ms" % ((t1 - t0) * 1e3, (t2 - t1) * 1e3))
    return out


def cpu_baseline_leg(args, lwes, tv, brk, ksk):
    """The headline workload on the host cores: the AVX2 CPU port (oracle/cpu_port.cpp, equal word for
    word to the scalar oracle) on one thread and on every thread this process may use, per-bootstrap
    latency percentiles from the one-thread run, and NTT calibration; the scalar oracle's one-thread
    rate beside it.  Bounded samples (about args.cpu_seconds in all)."""
    import oracle as O
    nproc = os.cpu_count()
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = affinity
    out = {"unit": "bootstraps/s", "kind": "port", "nproc": nproc, "affinity": affinity, "cpu_model": cpu_model()}
    port = O.port()
    if port is None:
        # no AVX2 here: the scalar oracle on every thread
        cnt = threads
        _, secs = O.bootstrap_batch(N, np.ascontiguousarray(lwes[:cnt]).reshape(-1), tv, brk, n, BB, L, O.BINARY,
                                    ksk, KS_OUT, KS_B, KS_D, threads)
        out.update({"value": round(cnt / secs, 3), "cores": threads, "isa": "scalar",
                    "sample": "%d bootstraps on %d threads (scalar oracle), %.1f s" % (cnt, threads, secs)})
        return out
    budget = max(args.cpu_seconds, 3.0)
    # one thread: per-bootstrap latencies (p50/p99) and the single-core rate
    probe, secs1, _ = O.port_bootstrap_batch(N, lwes[:1], tv, brk, n, BB, L, ksk, KS_OUT, KS_B, KS_D, 1, per_ct=True)
    cnt1 = int(min(len(lwes), max(4, 0.3 * budget / max(secs1, 1e-3))))
    _, secs1, per = O.port_bootstrap_batch(N, lwes[:cnt1], tv, brk, n, BB, L, ksk, KS_OUT, KS_B, KS_D, 1, per_ct=True)
    # every thread: the aggregate rate
    cntn = int(min(len(lwes), max(threads, threads * 0.5 * budget / max(secs1 / cnt1, 1e-3))))
    cntn -= cntn % threads
    _, secsn = O.port_bootstrap_batch(N, lwes[:cntn], tv, brk, n, BB, L, ksk, KS_OUT, KS_B, KS_D, threads)
    f1, i1 = O.port_ntt_time(1024, 2000)
    f2, i2 = O.port_ntt_time(2048, 1000)
    # the scalar oracle (the plain C restatement), one bootstrap on one thread
    _, secs_or = O.bootstrap_batch(N, np.ascontiguousarray(lwes[:1]).reshape(-1), tv, brk, n, BB, L, O.BINARY, ksk,
                                   KS_OUT, KS_B, KS_D, 1)
    out.update({
        "value": round(cntn / secsn, 3), "cores": threads, "isa": "avx2",
        "one_thread_bootstraps_per_s": round(cnt1 / secs1, 3),
        "p50_ms": round(1e3 * float(np.percentile(per, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(per, 99)), 3),
        "ntt_us": {"fwd_1024": round(f1 * 1e6, 2), "inv_1024": round(i1 * 1e6, 2),
                   "fwd_2048": round(f2 * 1e6, 2), "inv_2048": round(i2 * 1e6, 2)},
        "scalar_oracle_one_thread_bootstraps_per_s": round(1.0 / secs_or, 3),
        "sample": "AVX2 port (oracle/cpu_port.cpp): %d bootstraps on 1 thread in %.1f s, %d on %d threads in %.1f s "
                  "(%s; nproc %d, affinity %d)" % (cnt1, secs1, cntn, threads, secsn, cpu_model(), nproc, affinity)})
    log("CPU baseline: %.1f bootstraps/s on %d threads, %.2f on one (p50 %.1f ms, p99 %.1f ms)"
        % (out["value"], threads, out["one_thread_bootstraps_per_s"], out["p50_ms"], out["p99_ms"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson request stream of the serving leg (0 skips it)")
    ap.add_argument("--no-extra", action="store_true", help="skip the ternary and BMMP+packing workloads")
    ap.add_argument("--serve-chunk", type=int, default=64, help="blind-rotation steps per serving launch")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="issue one device-to-device copy of this size first (profiler byte-counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported (ciphertexts shard across "
              "GPUs with replicated keys, which this build does not run)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import fhetoothpaste_amd as tp
    from fhetoothpaste_amd import _lib
    from fhetoothpaste_amd.operators import DeviceBuffer

    if tp.device_count() < 1:
        print("bench.py: no gfx950 device", file=sys.stderr)
        sys.exit(1)

    t0 = time.time()
    lk = tp.keygen_secret(n, 0, SEED, 1)
    rk = tp.keygen_secret(N, 0, SEED, 2)
    brk = tp.keygen_brk(lk, rk, N, n, BB, L, 0, 0, 3.2, SEED)
    ksk = tp.keygen_lwe_ksk(rk, lk, N, KS_OUT, KS_B, KS_D, 3.2, SEED)
    keygen_s = time.time() - t0
    ctx = tp.Context(N, n, BB, L, 0, KS_OUT, KS_B, KS_D)
    ctx.import_brk(brk)
    ctx.import_lwe_ksk(ksk)
    tune = ctx.autotune(args.batch)
    log("blind-rotation variant: engine %s, %d wave(s)/SIMD, grid %s (NTT two: %.1f ms, one: %.1f ms, digit-parallel: "
        "%.1f ms, persistent grid: %.1f ms); digit-parallel up to %d ciphertexts"
        % ("floating point" if tune["engine"] else "NTT", tune["fft_waves"] if tune["engine"] else tune["waves"],
           tune["grid"] or "one work-group per ciphertext", tune["ms_two_waves"], tune["ms_one_wave"],
           tune["ms_digit_parallel"], tune["ms_persistent_grid"], tune["split_max"]))

    B = args.batch
    rng = np.random.default_rng(SEED)
    lwes = rng.integers(0, 2 * N, size=(B, n + 1), dtype=np.uint64)
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    d_lwe = DeviceBuffer.from_host(ctx, lwes)
    d_tv = DeviceBuffer.from_host(ctx, tv)
    d_ext = DeviceBuffer(ctx, B * (N + 1))
    d_out = DeviceBuffer(ctx, B * (KS_OUT + 1))
    lib = _lib.lib()
    chk = _lib.check

    def step(count=B, lwe_p=None, out_p=None, ext_p=None, mark=False):
        lwe_p = lwe_p or d_lwe.p
        if mark:
            chk(lib.tp_timer_mark(ctx.h, 0))
        chk(lib.tp_eval_func_batch_device(ctx.h, count, lwe_p, d_tv.p, 0, ext_p or d_ext.p, None, 0))
        if mark:
            chk(lib.tp_timer_mark(ctx.h, 1))
        chk(lib.tp_lwe_keyswitch_batch_device(ctx.h, count, ext_p or d_ext.p, N + 1, out_p or d_out.p))
        if mark:
            chk(lib.tp_timer_mark(ctx.h, 2))

    if args.copy_calibration_mb:
        words = args.copy_calibration_mb * (1 << 20) // 8
        c_src, c_dst = DeviceBuffer(ctx, words), DeviceBuffer(ctx, words)
        chk(lib.tp_copy_on_device(ctx.h, c_dst.p, c_src.p, words))
        ctx.sync()
        c_src.free()
        c_dst.free()

    for _ in range(args.warmup):
        step()
    ctx.sync()

    # timed region: K batches, inputs already in HBM
    import ctypes
    br_ms, ks_ms = [], []
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step(mark=True)
        ms_a, ms_b = ctypes.c_float(), ctypes.c_float()
        chk(lib.tp_timer_elapsed(ctx.h, 0, 1, ctypes.byref(ms_a)))
        chk(lib.tp_timer_elapsed(ctx.h, 1, 2, ctypes.byref(ms_b)))
        br_ms.append(ms_a.value)
        ks_ms.append(ms_b.value)
    ctx.sync()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    value = B * args.steps / elapsed

    # plain path: one bootstrap per launch (blind rotation grid of 1 + key switch of 1)
    plain_lat = []
    for i in range(args.plain_steps + 1):
        ctx.sync()
        t = time.perf_counter()
        step(count=1)
        ctx.sync()
        if i:
            plain_lat.append(time.perf_counter() - t)
    plain_bs = 1.0 / (sum(plain_lat) / len(plain_lat)) if plain_lat else float("nan")

    # batch-size sweep: one full bootstrap launch pair (blind rotation + key switch) per batch size,
    # with the blind rotation on br_kernel (one wave per ciphertext) and on the digit-parallel
    # kernel (2l waves per ciphertext); HIP-event times, best of two after a warm-up
    sweep = []
    split_tuned = ctx.br_split_max()
    for cnt in (1, 16, 64, 256, 512, 1024):
        if cnt > B:
            break
        row = {"batch": cnt}
        for label, split in (("one_wave_per_ct_ms", 0), ("digit_parallel_ms", cnt)):
            ctx.set_br_split_max(split)
            step(count=cnt)
            best = None
            for _ in range(2):
                step(count=cnt, mark=True)
                t_ms = ctypes.c_float()
                chk(lib.tp_timer_elapsed(ctx.h, 0, 2, ctypes.byref(t_ms)))
                best = t_ms.value if best is None else min(best, t_ms.value)
            row[label] = round(best, 3)
        row["bootstraps_per_s"] = round(cnt / min(row["one_wave_per_ct_ms"], row["digit_parallel_ms"]) * 1000, 1)
        row["launched_on"] = "digit-parallel" if cnt <= split_tuned else "br_kernel"
        sweep.append(row)
    ctx.set_br_split_max(split_tuned)
    log("batch sweep: " + ", ".join("%d: %.1f/%.1f ms" % (r["batch"], r["one_wave_per_ct_ms"], r["digit_parallel_ms"])
                                    for r in sweep))

    # the batch's key switch on the vector units (matrix cores off; the same words), for comparison
    ks_vec_ms = None
    if ctx.ks_mfma():
        ctx.set_ks_mfma(0)
        chk(lib.tp_lwe_keyswitch_batch_device(ctx.h, B, d_ext.p, N + 1, d_out.p))
        chk(lib.tp_timer_mark(ctx.h, 1))
        for _ in range(3):
            chk(lib.tp_lwe_keyswitch_batch_device(ctx.h, B, d_ext.p, N + 1, d_out.p))
        chk(lib.tp_timer_mark(ctx.h, 2))
        t_ms = ctypes.c_float()
        chk(lib.tp_timer_elapsed(ctx.h, 1, 2, ctypes.byref(t_ms)))
        ks_vec_ms = t_ms.value / 3
        ctx.set_ks_mfma(1)
        log("key switch of the batch: %.3f ms on the matrix cores, %.3f ms on the vector units"
            % (sum(ks_ms) / len(ks_ms), ks_vec_ms))

    # batch latency percentiles: every bootstrap of a batch completes with its batch
    lat_ms = sorted(br + ks for br, ks in zip(br_ms, ks_ms))
    p50 = float(np.percentile(lat_ms, 50))
    p99 = float(np.percentile(lat_ms, 99))

    w_br, w_ks = algorithmic_mulmods()
    br_avg_s = 1e-3 * sum(br_ms) / len(br_ms)
    achieved = (w_br * B) / br_avg_s / 1e12  # Tmulmod/s in the dominant kernel
    # INT64 multiply bound of the chip: 256 CUs x 64 lanes x 2.4 GHz of full-rate v_mad_u64_u32,
    # four per 64x64->128 product
    peak = 256 * 64 * 2.4e9 / 4 / 1e12
    traffic = read_traffic(B)
    key_bytes = (n * 4 * N * L + N * KS_D * (KS_OUT + 1)) * 8
    result = {
        "metric": "CGGI bootstraps/sec at n=630,N=1024 (batched vs plain); p50/p99 latency",
        "value": round(value, 2),
        "unit": "bootstraps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (mod 2^64-2^32+1)",
        "data": "synthetic (seeded keys and inputs)",
        "config": {"workload": "batched CGGI bootstrap: blind rotation n=630 N=1024 b=2^12 l=3 + sample extract + "
                               "LWE key switch 1024->630 b=2^4 d=14", "batch": B, "n": n, "N": N,
                   "parallelism": "one GPU, one work-group per ciphertext"},
        "plain_bootstraps_per_s": round(plain_bs, 2),
        "batched_vs_plain": round(value / plain_bs, 2),
        "latency_ms": {"batch_p50": round(p50, 3), "batch_p99": round(p99, 3),
                       "plain_mean": round(1000 * sum(plain_lat) / max(len(plain_lat), 1), 3),
                       "plain_p50": round(1000 * float(np.percentile(plain_lat, 50)), 3) if plain_lat else None,
                       "plain_p99": round(1000 * float(np.percentile(plain_lat, 99)), 3) if plain_lat else None},
        "kernel_ms": {"blind_rotate_avg": round(1000 * br_avg_s, 3),
                      "keyswitch_avg": round(sum(ks_ms) / len(ks_ms), 3),
                      "keyswitch_engine": "matrix cores (int8, exact)" if ks_vec_ms is not None else "vector units",
                      "keyswitch_vector_units_avg": None if ks_vec_ms is None else round(ks_vec_ms, 3)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": round(peak, 4), "unit": "Tmulmod/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "kernel": "br_kernel<10,3,binary>", "work_per_launch_mulmod": w_br * B,
                     "hbm_key_stream_GBps": round(key_bytes / (br_avg_s + 1e-3 * sum(ks_ms) / len(ks_ms)) / 1e9, 2)},
        "keygen_s": round(keygen_s, 2),
        "tuning": {"br_waves_per_simd": tune["waves"], "br_ms_two_waves": round(tune["ms_two_waves"], 2),
                   "br_ms_one_wave": round(tune["ms_one_wave"], 2),
                   "br_ms_digit_parallel": round(tune["ms_digit_parallel"], 2),
                   "br_ms_persistent_grid": round(tune["ms_persistent_grid"], 2), "br_grid": tune["grid"],
                   "digit_parallel_up_to": split_tuned},
        "batch_sweep": sweep,
    }

    if not args.no_extra:
        result["other_workloads"] = extra_workloads(tp)

    if args.serve_seconds > 0:
        result["serving"] = serving_leg(args, tp, ctx)
        result["latency_ms"]["serve_p50"] = result["serving"]["p50_ms"]
        result["latency_ms"]["serve_p99"] = result["serving"]["p99_ms"]

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_leg(args, lwes, tv, brk, ksk)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
