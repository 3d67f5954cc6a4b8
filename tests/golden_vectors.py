"""Reader for the frozen vectors of tests/golden/ (TPW1 records described by manifest.json)."""
import json
import os

import numpy as np

from fhetoothpaste_amd import wire

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
Q = 0xFFFFFFFF00000001


def manifest():
    with open(os.path.join(HERE, "manifest.json")) as f:
        return json.load(f)


def load(name):
    """Payload of a vector, split along its manifest layout (sizes from the manifest's dimensions)."""
    m = manifest()[name]
    _, _, _, data = wire.load(os.path.join(HERE, name + ".tpw"))
    N, n = m.get("N", 0), m.get("n", 0)
    sizes = {
        "input[N]": N, "output[N]": N, "lwe[n+1]": n + 1, "tv[2N]": 2 * N, "acc[2N]": 2 * N,
        "extracted[N+1]": N + 1, "lwes[3][n+1]": 3 * (n + 1), "out[3][n_out+1]": 3 * (m.get("ks_out", 0) + 1),
        "lwes[4][N+1]": 4 * (N + 1), "pack[2N]": 2 * N, "pack_consecutive[2N]": 2 * N,
    }
    parts, pos = {}, 0
    for field in m["layout"]:
        k = sizes.get(field, m.get("count", 0))
        parts[field.split("[")[0]] = data[pos:pos + k]
        pos += k
    assert pos == data.size, name
    return m, parts
