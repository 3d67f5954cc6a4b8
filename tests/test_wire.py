"""Wire format (host code): round trips, parameter checks and corruption detection."""
import numpy as np
import pytest

from fhetoothpaste_amd import wire
from fhetoothpaste_amd._lib import ToothpasteError

Q = 0xFFFFFFFF00000001


def test_round_trip_file_and_bytes(tmp_path, oracle):
    N, n, b, l = 1024, 4, 12, 3
    lk = oracle.secret(n, 0, 1, 1)
    rk = oracle.secret(N, 0, 1, 2)
    brk = oracle.brk(N, lk, rk, n, b, l, 0, 0, 3.2, 1)
    params = (N, n, b, l, 0, 0, 0, 0, 0, 0)
    path = tmp_path / "brk.tpw"
    wire.save(path, wire.BRK, params, brk, wire.NTT_DOMAIN)
    kind, flags, p, data = wire.load(path)
    assert kind == wire.BRK and flags == wire.NTT_DOMAIN and p.ring_dim == N and p.lwe_dim == n
    assert np.array_equal(data, brk)
    blob = wire.encode(wire.LWE_BATCH, params, np.arange(3 * (n + 1), dtype=np.uint64))
    assert len(blob) == 72 + 8 * 3 * (n + 1) and blob[:4] == b"TPW1"
    k2, _, _, d2 = wire.decode(blob)
    assert k2 == wire.LWE_BATCH and np.array_equal(d2, np.arange(3 * (n + 1), dtype=np.uint64))


def test_rejects_bad_records():
    params = (1024, 4, 12, 3, 0, 0, 0, 0, 0, 0)
    with pytest.raises(ToothpasteError):  # BRK of the wrong size for its parameters
        wire.encode(wire.BRK, params, np.zeros(10, np.uint64))
    blob = bytearray(wire.encode(wire.RAW, params, np.arange(100, dtype=np.uint64)))
    wire.decode(bytes(blob))
    for pos in (0, 4, 8, 56, 72 + 13):  # magic, version, kind, word count, one payload byte
        bad = bytearray(blob)
        bad[pos] ^= 0x5A
        with pytest.raises(ToothpasteError):
            wire.decode(bytes(bad))
    with pytest.raises(ToothpasteError):  # truncated
        wire.decode(bytes(blob[:-8]))
