"""ctypes view of the CPU oracle (oracle/toothpaste_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

Q = 0xFFFFFFFF00000001
BINARY, TERNARY, BMMP = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
# TOOTHPASTE_ORACLE_LIB_DIR selects another build of the oracle libraries (oracle/_san: the
# ASan/UBSan build of `make -C oracle sanitize`, which the CPU sanitizer test loads).
_LIB_DIR = os.environ.get("TOOTHPASTE_ORACLE_LIB_DIR") or _HERE
_LIB_PATH = os.path.join(_LIB_DIR, "liboracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
U64 = ctypes.c_uint64
U32 = ctypes.c_uint32


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_add": (U64, [U64, U64]),
            "or_sub": (U64, [U64, U64]),
            "or_mul": (U64, [U64, U64]),
            "or_mul_textbook": (U64, [U64, U64]),
            "or_reduce128": (U64, [U64, U64]),
            "or_pow": (U64, [U64, U64]),
            "or_inv": (U64, [U64]),
            "or_ring_create": (ctypes.c_void_p, [U32]),
            "or_ring_destroy": (None, [ctypes.c_void_p]),
            "or_ring_psi": (U64, [ctypes.c_void_p]),
            "or_ntt_forward": (None, [ctypes.c_void_p, u64p, u64p, U32]),
            "or_ntt_inverse": (None, [ctypes.c_void_p, u64p, u64p, U32]),
            "or_ntt_closed_form": (None, [ctypes.c_void_p, u64p, u64p]),
            "or_negacyclic_mul": (None, [U32, u64p, u64p, u64p]),
            "or_signed_corrector": (U64, [U32, u64p]),
            "or_decompose_unsigned": (None, [u64p, u64p, U32, U32, U32]),
            "or_decompose_signed": (None, [u64p, u64p, U32, U32, U32]),
            "or_external_product": (None, [ctypes.c_void_p, u64p, u64p, u64p, ctypes.c_int, U32, U32]),
            "or_binary_cmux": (None, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32]),
            "or_multi_mux": (None, [ctypes.c_void_p, u64p, U32, u64p, u64p, U32, U32]),
            "or_blind_rotate": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, u64p, u64p, U32, U32, U32,
                                               ctypes.c_int, ctypes.c_int]),
            "or_sample_extract": (None, [u64p, u64p, U32]),
            "or_lwe_keyswitch": (ctypes.c_int, [u64p, u64p, u64p, U32, U32, U32, U32]),
            "or_keygen_secret": (None, [u64p, U32, ctypes.c_int, U64, U64]),
            "or_gadget": (None, [u64p, U32, U32, ctypes.c_int]),
            "or_keygen_brk": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32, U32, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, U64]),
            "or_keygen_lwe_ksk": (None, [u64p, u64p, u64p, U32, U32, U32, U32, ctypes.c_double, U64]),
            "or_encrypt_rlwe": (None, [ctypes.c_void_p, u64p, u64p, u64p, ctypes.c_double, U64, U64]),
            "or_random_mod_pow2": (None, [u64p, ctypes.c_size_t, U64, U64, U64]),
            "or_gauss_samples": (None, [u64p, ctypes.c_size_t, ctypes.c_double, U64, U64]),
            "or_phase_rlwe": (None, [ctypes.c_void_p, u64p, u64p, u64p]),
            "or_phase_lwe": (U64, [u64p, u64p, U32]),
            "or_keygen_rlwe_ksk": (None, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32, ctypes.c_double, U64, U64]),
            "or_automorphism_poly": (None, [u64p, u64p, U64, U32]),
            "or_keygen_auto_keys": (None, [ctypes.c_void_p, u64p, u64p, U32, U32, ctypes.c_double, U64]),
            "or_rlwe_keyswitch": (None, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32]),
            "or_automorphism": (None, [ctypes.c_void_p, u64p, u64p, U64, u64p, U32, U32]),
            "or_trace": (None, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32]),
            "or_pack": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, U32, u64p, U32, U32]),
            "or_lwe_to_rgsw": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, u64p, U32, U32, U32, ctypes.c_int, u64p,
                                               U32, U32, u64p, U32, U32, U32, U32]),
            "or_sap_homtrunc": (None, [ctypes.c_void_p, u64p, u64p, U32, u64p, U32, U32]),
            "or_sap_reset": (ctypes.c_int, [ctypes.c_void_p, u64p, U32, u64p, U32, U32, U32, U64, u64p, U32, U32, U32,
                                            ctypes.c_int]),
            "or_sap_decompose": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, U32, U32, u64p, U32, U32, U32,
                                                ctypes.c_int, u64p, U32, U32, U64, u64p, U32, U32]),
            "or_modswitch": (None, [u64p, ctypes.c_uint64, U64, U64, ctypes.c_int]),
            "or_sample_gaussian": (None, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, U64, U64]),
            "or_pack_consecutive": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, U32, u64p, U32, U32]),
            "or_bootstrap_batch": (ctypes.c_double, [ctypes.c_void_p, u64p, u64p, ctypes.c_size_t, u64p, u64p,
                                                     U32, U32, U32, ctypes.c_int, u64p, U32, U32, U32,
                                                     ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a):
    """Pointer to a contiguous uint64 numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def arr(n):
    return np.zeros(n, dtype=np.uint64)


class Ring:
    """Z_Q[X]/(X^N+1) with the reference NTT convention (psi = 7^((Q-1)/2N), bit-reversed output)."""

    _cache = {}

    def __new__(cls, N):
        if N in cls._cache:
            return cls._cache[N]
        self = super().__new__(cls)
        self.N = N
        self.h = lib().or_ring_create(N)
        cls._cache[N] = self
        return self

    @property
    def psi(self):
        return lib().or_ring_psi(self.h)

    def ntt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = np.empty_like(a)
        lib().or_ntt_forward(self.h, P(out), P(a), a.size // self.N)
        return out

    def intt(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = np.empty_like(a)
        lib().or_ntt_inverse(self.h, P(out), P(a), a.size // self.N)
        return out

    def ntt_closed_form(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        out = arr(self.N)
        lib().or_ntt_closed_form(self.h, P(out), P(a))
        return out

    def negacyclic_mul(self, a, b):
        out = arr(self.N)
        lib().or_negacyclic_mul(self.N, P(out), P(np.ascontiguousarray(a, np.uint64)),
                                P(np.ascontiguousarray(b, np.uint64)))
        return out


def secret(length, dist, seed, tag):
    out = arr(length)
    lib().or_keygen_secret(P(out), length, dist, seed, tag)
    return out


def sample_gaussian(count, sigma, seed, tag):
    out = np.zeros(count, dtype=np.int64)
    lib().or_sample_gaussian(out.ctypes.data, count, sigma, seed, tag)
    return out


def gadget(b, l, kind):
    out = arr(l)
    lib().or_gadget(P(out), b, l, kind)
    return out


def brk(N, lwe_key, rlwe_key, n, b, l, method, gadget_kind, sigma, seed):
    R = Ring(N)
    per = {BINARY: 4 * N * l * n, TERNARY: 8 * N * l * n, BMMP: 12 * N * l * (n // 2)}[method]
    out = arr(per)
    rc = lib().or_keygen_brk(R.h, P(out), P(lwe_key), P(rlwe_key), n, b, l, method, gadget_kind, sigma, seed)
    assert rc == 0
    return out


def lwe_ksk(src, dst, n_in, n_out, b, d, sigma, seed):
    out = arr(n_in * d * (n_out + 1))
    lib().or_keygen_lwe_ksk(P(out), P(src), P(dst), n_in, n_out, b, d, sigma, seed)
    return out


def random_mod_pow2(length, bound, seed, tag):
    out = arr(length)
    lib().or_random_mod_pow2(P(out), length, bound, seed, tag)
    return out


def blind_rotate(N, lwe, acc_coef, brk_bytes, n, b, l, method, as_coef):
    R = Ring(N)
    out = arr(2 * N)
    rc = lib().or_blind_rotate(R.h, P(out), P(np.ascontiguousarray(lwe, np.uint64)),
                               P(np.ascontiguousarray(acc_coef, np.uint64)), P(brk_bytes), n, b, l, method,
                               int(as_coef))
    assert rc == 0, rc
    return out


def sample_extract(rlwe_coef, N):
    out = arr(N + 1)
    lib().or_sample_extract(P(out), P(np.ascontiguousarray(rlwe_coef, np.uint64)), N)
    return out


def lwe_keyswitch(lwe_in, ksk, n_in, n_out, b, d):
    out = arr(n_out + 1)
    rc = lib().or_lwe_keyswitch(P(out), P(np.ascontiguousarray(lwe_in, np.uint64)), P(ksk), n_in, n_out, b, d)
    assert rc == 0
    return out


def phase_rlwe(N, rlwe_ntt, s_coef):
    out = arr(N)
    lib().or_phase_rlwe(Ring(N).h, P(out), P(np.ascontiguousarray(rlwe_ntt, np.uint64)), P(s_coef))
    return out


def phase_lwe(lwe, s, n):
    return lib().or_phase_lwe(P(np.ascontiguousarray(lwe, np.uint64)), P(s), n)


def encrypt_rlwe(N, msg_coef, s_coef, sigma, seed, tag):
    out = arr(2 * N)
    lib().or_encrypt_rlwe(Ring(N).h, P(out), P(np.ascontiguousarray(msg_coef, np.uint64)), P(s_coef), sigma,
                          seed, tag)
    return out


def bootstrap_batch(N, lwes, tv, brk_bytes, n, b, l, method, ksk=None, n_out=0, ks_b=0, ks_d=0, threads=1):
    count = lwes.size // (n + 1)
    width = (n_out + 1) if ksk is not None else (N + 1)
    out = arr(count * width)
    secs = lib().or_bootstrap_batch(Ring(N).h, P(out), P(lwes), count, P(tv), P(brk_bytes), n, b, l, method,
                                    P(ksk), n_out, ks_b, ks_d, threads)
    return out.reshape(count, width), secs


def auto_keys(N, s, b, d, sigma, seed):
    """Automorphism keys for t = 2^j + 1, j = 1..log2 N: [log2 N][d][2N], NTT domain."""
    logN = N.bit_length() - 1
    out = arr(logN * d * 2 * N)
    lib().or_keygen_auto_keys(Ring(N).h, P(out), P(s), b, d, sigma, seed)
    return out


def rlwe_ksk(N, src, dst, b, d, sigma, seed, tag=6):
    out = arr(d * 2 * N)
    lib().or_keygen_rlwe_ksk(Ring(N).h, P(out), P(src), P(dst), b, d, sigma, seed, tag)
    return out


def automorphism_poly(a, t, N):
    out = arr(N)
    lib().or_automorphism_poly(P(out), P(np.ascontiguousarray(a, np.uint64)), t, N)
    return out


def automorphism(N, rlwe_coef, t, ksk, b, d):
    out = arr(2 * N)
    lib().or_automorphism(Ring(N).h, P(out), P(np.ascontiguousarray(rlwe_coef, np.uint64)), t, P(ksk), b, d)
    return out


def rlwe_keyswitch(N, out, inp, ksk, b, d):
    out = np.ascontiguousarray(out, np.uint64).copy()
    lib().or_rlwe_keyswitch(Ring(N).h, P(out), P(np.ascontiguousarray(inp, np.uint64)), P(ksk), b, d)
    return out


def trace(N, rlwe_coef, keys, b, d):
    out = arr(2 * N)
    lib().or_trace(Ring(N).h, P(out), P(np.ascontiguousarray(rlwe_coef, np.uint64)), P(keys), b, d)
    return out


def pack(N, lwes, length, keys, b, d):
    out = arr(2 * N)
    rc = lib().or_pack(Ring(N).h, P(out), P(np.ascontiguousarray(lwes, np.uint64)), length, P(keys), b, d)
    assert rc == 0
    return out


def pack_consecutive(N, lwes, length, keys, b, d):
    out = arr(2 * N)
    rc = lib().or_pack_consecutive(Ring(N).h, P(out), P(np.ascontiguousarray(lwes, np.uint64)), length, P(keys), b, d)
    assert rc == 0
    return out


def modswitch(vec, source_modulus, target_modulus, kind):
    """ModulusSwitch in x87 long double (kind 0 ROUND, 1 FLOOR, 2 CEIL)."""
    out = np.ascontiguousarray(vec, dtype=np.uint64).copy()
    lib().or_modswitch(P(out), out.size, source_modulus, target_modulus, kind)
    return out


def lwe_to_rgsw(N, lwe, brk, n, bl, ll, method, trace_keys, bt, dt, sq_key, bs, ds, digits, basebits):
    """LWEtoRGSWConverter::Convert: [lev_left | lev_right], digits x 2N each, NTT domain."""
    out = arr(4 * N * digits)
    rc = lib().or_lwe_to_rgsw(Ring(N).h, P(out), P(np.ascontiguousarray(lwe, np.uint64)), P(brk), n, bl, ll, method,
                              P(trace_keys), bt, dt, P(sq_key), bs, ds, digits, basebits)
    assert rc == 0, rc
    return out


def sap_homtrunc(N, rlwe_ntt, radix, keys, b, d):
    out = arr(2 * N)
    lib().or_sap_homtrunc(Ring(N).h, P(out), P(np.ascontiguousarray(rlwe_ntt, np.uint64)), radix, P(keys), b, d)
    return out


def sap_reset(N, acc_ntt, radix, ksk, n_out, kb, kd, qks, brk, n, bl, ll, method):
    acc = np.ascontiguousarray(acc_ntt, np.uint64).copy()
    rc = lib().or_sap_reset(Ring(N).h, P(acc), radix, P(ksk), n_out, kb, kd, qks, P(brk), n, bl, ll, method)
    assert rc == 0, rc
    return acc


def sap_decompose(N, lwe, radix, beta, brk, n, bl, ll, method, keys, b, d, max_digits=64, restart=None, ksk=None,
                  kb=0, kd=0):
    """Digits [D][N+1]; restart/ksk (N -> n, base 2^kb, kd digits) for the periodic accumulator reset."""
    out = arr(max_digits * (N + 1))
    r = (1 << 64) - 1 if restart is None else restart
    k = ksk if ksk is not None else arr(1)
    cnt = lib().or_sap_decompose(Ring(N).h, P(out), P(np.ascontiguousarray(lwe, np.uint64)), radix, beta, P(brk), n,
                                 bl, ll, method, P(keys), b, d, r, P(k), kb, kd)
    assert cnt >= 0
    return out[:cnt * (N + 1)].reshape(cnt, N + 1)


# ---------------------------------------------------------------- vectorised CPU port (cpu_port.cpp)
_port = None
_PORT_PATH = os.path.join(_LIB_DIR, "libcpuport.so")


def port():
    """ctypes view of libcpuport.so, the AVX2 CPU port timed as the CPU baseline (None if the CPU
    lacks AVX2)."""
    global _port
    if _port is None:
        if not os.path.exists(_PORT_PATH):
            build()
        L = ctypes.CDLL(_PORT_PATH)
        L.cp_available.restype = ctypes.c_int
        L.cp_available.argtypes = []
        L.cp_ntt.restype = None
        L.cp_ntt.argtypes = [U32, u64p, ctypes.c_size_t, ctypes.c_int]
        L.cp_ntt_time.restype = None
        L.cp_ntt_time.argtypes = [U32, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.cp_bootstrap_batch.restype = ctypes.c_double
        L.cp_bootstrap_batch.argtypes = [U32, u64p, u64p, ctypes.c_size_t, u64p, u64p, U32, U32, U32, u64p, U32, U32,
                                         U32, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        _port = L
    return _port if _port.cp_available() else None


def port_ntt(N, polys, inverse=False):
    """Forward (or inverse) transforms of the rows of polys (count x N) with the CPU port."""
    a = np.ascontiguousarray(polys, dtype=np.uint64).reshape(-1, N).copy()
    port().cp_ntt(N, P(a), a.shape[0], 1 if inverse else 0)
    return a


def port_ntt_time(N, reps):
    """Seconds per forward and per inverse transform of N words (one thread)."""
    f, i = ctypes.c_double(), ctypes.c_double()
    port().cp_ntt_time(N, reps, ctypes.byref(f), ctypes.byref(i))
    return f.value, i.value


def port_bootstrap_batch(N, lwes, tv, brk_bytes, n, b, l, ksk=None, n_out=0, ks_b=0, ks_d=0, threads=1,
                         per_ct=False):
    """Binary-key bootstrap batch on the CPU port: (outputs, seconds[, per-ciphertext seconds])."""
    lwes = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1)
    count = lwes.size // (n + 1)
    width = (n_out + 1) if ksk is not None else (N + 1)
    out = arr(count * width)
    times = np.zeros(count, dtype=np.float64) if per_ct else None
    tp = times.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if per_ct else None
    secs = port().cp_bootstrap_batch(N, P(out), P(lwes), count, P(np.ascontiguousarray(tv, dtype=np.uint64)),
                                     P(brk_bytes), n, b, l, P(ksk), n_out, ks_b, ks_d, threads, tp)
    if secs < 0:
        raise ValueError("cp_bootstrap_batch rejected its arguments")
    if per_ct:
        return out.reshape(count, width), secs, times
    return out.reshape(count, width), secs
