"""Versioned wire format for keys and ciphertext batches (tp_wire_* of libtoothpaste_amd.so).

A record is the reference-layout uint64 array plus a header naming its kind and parameters and
an FNV-1a checksum; see include/toothpaste_amd.h for the layout.  Host code only.
"""
import ctypes

import numpy as np

from . import _lib as T
from ._lib import check, lib, ptr

RAW, BRK, LWE_KSK, AUTO_KEYS, RLWE_KSK, LWE_BATCH, RLWE_BATCH = range(7)
NTT_DOMAIN = 1


def _params(p):
    if isinstance(p, T.Params):
        return p
    return T.Params(*p)


def encode(kind, params, data, flags=0):
    data = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
    n = T.SZ()
    check(lib().tp_wire_bytes(data.size, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    p = _params(params)
    check(lib().tp_wire_encode(buf, n.value, kind, flags, ctypes.byref(p), ptr(data), data.size))
    return buf.raw


def decode(blob):
    words, kind, flags, p = T.SZ(), T.INT(), T.U32(), T.Params()
    check(lib().tp_wire_decode(blob, len(blob), ctypes.byref(kind), ctypes.byref(flags), ctypes.byref(p), None, 0,
                               ctypes.byref(words)))
    out = np.empty(words.value, dtype=np.uint64)
    check(lib().tp_wire_decode(blob, len(blob), None, None, None, ptr(out), out.size, ctypes.byref(words)))
    return kind.value, flags.value, p, out


def save(path, kind, params, data, flags=0):
    data = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
    p = _params(params)
    check(lib().tp_wire_save(str(path).encode(), kind, flags, ctypes.byref(p), ptr(data), data.size))


def load(path):
    words, kind, flags, p = T.SZ(), T.INT(), T.U32(), T.Params()
    check(lib().tp_wire_load(str(path).encode(), ctypes.byref(kind), ctypes.byref(flags), ctypes.byref(p), None, 0,
                             ctypes.byref(words)))
    out = np.empty(words.value, dtype=np.uint64)
    check(lib().tp_wire_load(str(path).encode(), None, None, None, ptr(out), out.size, ctypes.byref(words)))
    return kind.value, flags.value, p, out
