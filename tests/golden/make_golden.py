"""Writes the frozen vectors in tests/golden/ (run from the repository root:
`python tests/golden/make_golden.py`).

The reference ships no known-answer vectors for this path (its tests are randomized) and cannot be
built here, so these are outputs of the oracle (oracle/), frozen once: they pin the oracle and the
GPU path against later drift, next to the checks that pin the oracle itself (tests/test_oracle.py).
Every key is regenerated from the seed recorded in manifest.json; each file is a TPW1 record
(fhetoothpaste_amd.wire) whose payload layout manifest.json gives.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from fhetoothpaste_amd import wire  # noqa: E402

Q = O.Q


def params(N=0, n=0, b=0, l=0, method=0, ks_out=0, ks_b=0, ks_d=0, rb=0, rd=0):
    return (N, n, b, l, method, ks_out, ks_b, ks_d, rb, rd)


def main():
    man = {}
    rng = np.random.default_rng(20261203)

    # forward NTT at N = 1024 and 2048: [input | output]
    for N in (1024, 2048):
        x = rng.integers(0, Q, size=N, dtype=np.uint64)
        y = O.Ring(N).ntt(x)
        name = "ntt_N%d" % N
        wire.save(os.path.join(HERE, name + ".tpw"), wire.RAW, params(N), np.concatenate([x, y]))
        man[name] = {"what": "forward negacyclic NTT", "layout": ["input[N]", "output[N]"], "N": N}

    # blind rotation + sample extraction, three methods: [lwe | tv | rotated acc (coef) | extracted]
    for tag, N, n, b, l, method, dist, seed in [("br_binary", 1024, 24, 12, 3, 0, 0, 11),
                                                ("br_ternary", 1024, 12, 16, 4, 1, 1, 12),
                                                ("br_bmmp", 2048, 8, 14, 3, 2, 0, 13)]:
        lk = O.secret(n, dist, seed, 1)
        rk = O.secret(N, 0, seed, 2)
        brk = O.brk(N, lk, rk, n, b, l, method, 0 if method == 0 else 1, 3.2, seed)
        lwe = O.random_mod_pow2(n + 1, 2 * N, seed, 30)
        tv = rng.integers(0, Q, size=2 * N, dtype=np.uint64)
        acc = O.blind_rotate(N, lwe, tv, brk, n, b, l, method, True)
        ext = O.sample_extract(acc, N)
        wire.save(os.path.join(HERE, tag + ".tpw"), wire.RAW, params(N, n, b, l, method),
                  np.concatenate([lwe, tv, acc, ext]))
        man[tag] = {"what": "blind rotation (coefficient output) + SampleExtract(0)",
                    "layout": ["lwe[n+1]", "tv[2N]", "acc[2N]", "extracted[N+1]"],
                    "N": N, "n": n, "base_log2": b, "digits": l, "method": method, "key_dist": dist,
                    "gadget_kind": 0 if method == 0 else 1, "sigma": 3.2, "seed": seed,
                    "lwe_stream": [seed, 30]}

    # full bootstrap at the batched workload's shape, 3 ciphertexts: [lwes | tv | outputs]
    N, n, b, l, ko, kb, kd, seed = 1024, 630, 12, 3, 630, 4, 14, 14
    lk = O.secret(n, 0, seed, 1)
    rk = O.secret(N, 0, seed, 2)
    brk = O.brk(N, lk, rk, n, b, l, 0, 0, 3.2, seed)
    ksk = O.lwe_ksk(rk, lk, N, ko, kb, kd, 3.2, seed)
    lwes = O.random_mod_pow2(3 * (n + 1), 2 * N, seed, 31)
    tv = np.zeros(2 * N, dtype=np.uint64)
    tv[N:] = Q // 8
    out, _ = O.bootstrap_batch(N, lwes, tv, brk, n, b, l, 0, ksk, ko, kb, kd, 1)
    wire.save(os.path.join(HERE, "bootstrap_n630.tpw"), wire.RAW, params(N, n, b, l, 0, ko, kb, kd),
              np.concatenate([lwes, tv, out.reshape(-1)]))
    man["bootstrap_n630"] = {"what": "blind rotation + SampleExtract + LWE key switch 1024->630, 3 ciphertexts",
                             "layout": ["lwes[3][n+1]", "tv[2N]", "out[3][n_out+1]"],
                             "N": N, "n": n, "base_log2": b, "digits": l, "ks_out": ko, "ks_base_log2": kb,
                             "ks_digits": kd, "sigma": 3.2, "seed": seed, "lwe_stream": [seed, 31]}

    # packing 4 LWEs (N = 1024) and PackConsecutively of 3: [lwes | pack | pack_consecutive]
    N, b, d, seed = 1024, 16, 4, 15
    s = O.secret(N, 0, seed, 2)
    keys = O.auto_keys(N, s, b, d, 3.2, seed)
    lwes = rng.integers(0, Q, size=(4, N + 1), dtype=np.uint64)
    pk = O.pack(N, lwes, 4, keys, b, d)
    pc = O.pack_consecutive(N, lwes[:3], 3, keys, b, d)
    wire.save(os.path.join(HERE, "pack_N1024.tpw"), wire.RAW, params(N, 0, 0, 0, 0, 0, 0, 0, b, d),
              np.concatenate([lwes.reshape(-1), pk, pc]))
    man["pack_N1024"] = {"what": "Pack(4) and PackConsecutively(3)", "layout": ["lwes[4][N+1]", "pack[2N]",
                                                                               "pack_consecutive[2N]"],
                         "N": N, "base_log2": b, "digits": d, "sigma": 3.2, "seed": seed}

    # modulus switch Q -> 2048, near half-integers and near integers: [in | round | floor | ceil]
    v = [(((2 * k + 1) * Q) // 4096 + dd) % Q for k in range(0, 2048, 16) for dd in (-1, 0, 1)]
    v += [(-(-(k * Q) // 2048) + dd) % Q for k in range(1, 2048, 16) for dd in (-1, 0, 1)]
    v = np.array(v, dtype=np.uint64)
    outs = [O.modswitch(v, Q, 2048, kind) for kind in range(3)]
    wire.save(os.path.join(HERE, "modswitch_2048.tpw"), wire.RAW, params(), np.concatenate([v] + outs))
    man["modswitch_2048"] = {"what": "ModulusSwitch Q -> 2048 in x87 long double", "count": int(v.size),
                             "layout": ["in[count]", "round[count]", "floor[count]", "ceil[count]"]}

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1, sort_keys=True)
    print("wrote", len(man), "vectors")


if __name__ == "__main__":
    main()
