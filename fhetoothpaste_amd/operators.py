"""Python mirror of the reference operator API for the bootstrapping path.

Names, argument meanings and error behaviour follow LSchild/FHEToothpaste:
  CGGIBlindRotationContext / CGGIBlindRotator   lib/include/operators/cggi_blind_rotation.h:19-161
  BMMPBlindRotationContext / BMMPBlindRotator   lib/include/operators/bmmp_blind_rotation.h:18-176
  LWEConversionContext / LWEtoLWEConverter      lib/include/operators/endo_glwe_conversion.h:15-239
  NegacyclicFunctionEvaluator                   lib/include/operators/negacyclic_function_evaluator.h:17-114
  SolinasWorker (ForwardNTT / BackwardNTT)      lib/include/backend/solinas_worker.h:56-126
Every call runs on the GPU through libtoothpaste_amd.so; there is no CPU path.
Reference asserts and std::invalid_argument become ValueError / ToothpasteError.
"""
import ctypes
import os

import numpy as np

from . import _lib as T
from ._lib import Q, check, lib, ptr

BINARY, TERNARY = 0, 1


def _max_digits(b):
    return 64 // b + (1 if 64 % b else 0)


class DeviceBuffer:
    """Device array of uint64 owned by a context."""

    def __init__(self, ctx, words):
        self.ctx = ctx
        self.words = int(words)
        p = ctypes.c_void_p()
        check(lib().tp_malloc(ctx.h, ctypes.byref(p), self.words))
        self.p = p.value

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        b = cls(ctx, a.size)
        check(lib().tp_copy_to_device(ctx.h, b.p, ptr(a), a.size))
        return b

    def copy_from_host(self, a):
        """Overwrites the buffer's leading words with the host array a."""
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        if a.size > self.words:
            raise ValueError("host array larger than the device buffer")
        check(lib().tp_copy_to_device(self.ctx.h, self.p, ptr(a), a.size))

    def to_host(self, words=None):
        words = self.words if words is None else words
        out = np.empty(words, dtype=np.uint64)
        check(lib().tp_copy_to_host(self.ctx.h, ptr(out), self.p, words))
        return out

    def free(self):
        if self.p is not None and self.ctx.h:
            check(lib().tp_free(self.ctx.h, self.p))
        self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device context (stream, tables, resident keys) for one parameter set."""

    def __init__(self, ring_dim, lwe_dim, base_log2, digits, method=T.METHOD_CGGI_BINARY, ks_out_dim=0,
                 ks_base_log2=0, ks_digits=0, device=0, auto_base_log2=0, auto_digits=0):
        self.params = T.Params(ring_dim, lwe_dim, base_log2, digits, method, ks_out_dim, ks_base_log2, ks_digits,
                               auto_base_log2, auto_digits)
        h = ctypes.c_void_p()
        check(lib().tp_context_create(ctypes.byref(h), ctypes.byref(self.params), device))
        self.h = h.value
        self.N, self.n = ring_dim, lwe_dim

    def close(self):
        if self.h:
            lib().tp_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().tp_context_sync(self.h))

    def import_brk(self, brk):
        brk = np.ascontiguousarray(brk, dtype=np.uint64)
        check(lib().tp_import_brk(self.h, ptr(brk), brk.size))

    def import_lwe_ksk(self, ksk):
        ksk = np.ascontiguousarray(ksk, dtype=np.uint64)
        check(lib().tp_import_lwe_ksk(self.h, ptr(ksk), ksk.size))

    def ntt_forward(self, polys):
        polys = np.ascontiguousarray(polys, dtype=np.uint64)
        count = polys.size // self.N
        d = DeviceBuffer.from_host(self, polys)
        check(lib().tp_ntt_forward_device(self.h, d.p, d.p, count))
        return d.to_host().reshape(polys.shape)

    def ntt_inverse(self, polys):
        polys = np.ascontiguousarray(polys, dtype=np.uint64)
        count = polys.size // self.N
        d = DeviceBuffer.from_host(self, polys)
        check(lib().tp_ntt_inverse_device(self.h, d.p, d.p, count))
        return d.to_host().reshape(polys.shape)

    def blind_rotate(self, lwe, acc, as_coef=False):
        lwe = np.ascontiguousarray(lwe, dtype=np.uint64)
        acc = np.ascontiguousarray(acc, dtype=np.uint64)
        out = np.empty(2 * self.N, dtype=np.uint64)
        check(lib().tp_blind_rotate(self.h, ptr(out), ptr(lwe), ptr(acc), int(as_coef)))
        return out

    def eval_func_batch(self, lwes, tv, want_rlwe=False, rlwe_ntt=False):
        """lwes [count, n+1] in Z_2N; tv [2N] shared or [count, 2N]. Returns extracted LWEs (and RLWEs)."""
        lwes = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.n + 1)
        tv = np.ascontiguousarray(tv, dtype=np.uint64)
        count = lwes.shape[0]
        d_in = DeviceBuffer.from_host(self, lwes)
        d_tv = DeviceBuffer.from_host(self, tv)
        d_out = DeviceBuffer(self, count * (self.N + 1))
        d_rl = DeviceBuffer(self, count * 2 * self.N) if want_rlwe else None
        stride = 0 if tv.size == 2 * self.N else 2 * self.N
        check(lib().tp_eval_func_batch_device(self.h, count, d_in.p, d_tv.p, stride, d_out.p,
                                              d_rl.p if d_rl else None, int(rlwe_ntt)))
        out = d_out.to_host().reshape(count, self.N + 1)
        if want_rlwe:
            return out, d_rl.to_host().reshape(count, 2 * self.N)
        return out

    def lwe_keyswitch_batch(self, lwes):
        lwes = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.N + 1)
        count = lwes.shape[0]
        W = self.params.ks_out_dim + 1
        d_in = DeviceBuffer.from_host(self, lwes)
        d_out = DeviceBuffer(self, count * W)
        check(lib().tp_lwe_keyswitch_batch_device(self.h, count, d_in.p, self.N + 1, d_out.p))
        return d_out.to_host().reshape(count, W)

    def import_auto_keys(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        check(lib().tp_import_auto_keys(self.h, ptr(keys), keys.size))

    def automorphism(self, rlwes_coef, t):
        """Batched AutomorphismEvaluator::Eval: [count, 2N] coefficients -> [count, 2N] NTT domain."""
        x = np.ascontiguousarray(rlwes_coef, dtype=np.uint64).reshape(-1, 2 * self.N)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, x.size)
        check(lib().tp_automorphism_device(self.h, x.shape[0], t, d_in.p, d_out.p))
        return d_out.to_host().reshape(x.shape)

    def trace(self, rlwes_coef):
        x = np.ascontiguousarray(rlwes_coef, dtype=np.uint64).reshape(-1, 2 * self.N)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, x.size)
        check(lib().tp_trace_device(self.h, x.shape[0], d_in.p, d_out.p))
        return d_out.to_host().reshape(x.shape)

    def pack(self, lwes, perm=None):
        """LWEtoRLWEPacker::Pack of [len, N+1] LWEs (optionally reordered by perm) -> [2N] coefficients."""
        x = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.N + 1)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, 2 * self.N)
        d_perm = None
        if perm is not None:
            p32 = np.ascontiguousarray(perm, dtype=np.uint32)
            d_perm = DeviceBuffer.from_host(self, np.frombuffer(
                np.concatenate([p32, np.zeros(p32.size % 2, np.uint32)]).tobytes(), dtype=np.uint64))
        check(lib().tp_pack_device(self.h, x.shape[0], d_in.p, d_perm.p if d_perm else None, d_out.p))
        return d_out.to_host()

    def autotune(self, count, cache=None):
        """Time the blind-rotation variants on `count` synthetic ciphertexts (tp_autotune): keeps the
        fastest one-wave-per-ciphertext variant (engine and waves) and grid, the faster form of the
        digit-parallel kernel (one or two waves per transform, where both are built), and moves the
        digit-parallel threshold to the side of `count` that won; with `cache` (a JSON file path,
        tp_autotune_cached) a stored result for this device, parameter set and count is applied
        instead of timing.  Returns {"waves", "split_max", "engine", "fft_waves", "grid",
        "split_halves", "cache_hit", "ms_two_waves", "ms_one_wave", "ms_digit_parallel", "ms_fft_one_wave",
        "ms_fft_two_waves", "ms_persistent_grid"} (-1 for a variant that is not built, not usable
        or not timed)."""
        ms = (ctypes.c_float * 6)()
        hit = ctypes.c_int(0)
        if cache is None:
            check(lib().tp_autotune(self.h, count, ms))
        else:
            check(lib().tp_autotune_cached(self.h, count, os.fsencode(cache), ctypes.byref(hit), ms))
        return {"waves": self.br_waves(), "split_max": self.br_split_max(), "engine": self.br_engine(),
                "fft_waves": self._get(T.TP_OPT_BR_FFT_WAVES), "grid": self._get(T.TP_OPT_BR_GRID),
                "split_halves": self.br_split_halves(), "cache_hit": bool(hit.value), "ms_two_waves": ms[0], "ms_one_wave": ms[1],
                "ms_digit_parallel": ms[2], "ms_fft_one_wave": ms[3], "ms_fft_two_waves": ms[4],
                "ms_persistent_grid": ms[5]}

    def _get(self, option):
        v = ctypes.c_int()
        check(lib().tp_get_option(self.h, option, ctypes.byref(v)))
        return v.value

    def set_br_engine(self, engine):
        """1: exact floating-point external product (binary CGGI, N = 1024), 0: Goldilocks NTT."""
        check(lib().tp_set_option(self.h, T.TP_OPT_BR_ENGINE, engine))

    def br_engine(self):
        """Engine the next launches use (1 only when the key's rounding bound certifies it)."""
        return self._get(T.TP_OPT_BR_ENGINE)

    def set_fft_waves(self, waves):
        check(lib().tp_set_option(self.h, T.TP_OPT_BR_FFT_WAVES, waves))

    def set_br_waves(self, waves):
        check(lib().tp_set_option(self.h, T.TP_OPT_BR_WAVES, waves))

    def br_waves(self):
        w = ctypes.c_int()
        check(lib().tp_get_option(self.h, T.TP_OPT_BR_WAVES, ctypes.byref(w)))
        return w.value

    def set_br_split_max(self, count):
        """Launches of at most `count` ciphertexts use the digit-parallel kernel (0: never)."""
        check(lib().tp_set_option(self.h, T.TP_OPT_BR_SPLIT_MAX, count))

    def br_split_max(self):
        v = ctypes.c_int()
        check(lib().tp_get_option(self.h, T.TP_OPT_BR_SPLIT_MAX, ctypes.byref(v)))
        return v.value

    def set_br_split_halves(self, on):
        """1: digit-parallel launches split every transform over two waves (binary CGGI, N = 1024,
        2 or 3 digits), 0: one wave per digit transform."""
        check(lib().tp_set_option(self.h, T.TP_OPT_BR_SPLIT_HALVES, int(on)))

    def br_split_halves(self):
        return self._get(T.TP_OPT_BR_SPLIT_HALVES)

    def set_ks_mfma(self, on):
        """1: LWE key switches of more than 16 ciphertexts run on the matrix cores where the key fits
        that form, 0: on the vector units (same results)."""
        check(lib().tp_set_option(self.h, T.TP_OPT_KS_MFMA, int(on)))

    def ks_mfma(self):
        return self._get(T.TP_OPT_KS_MFMA)

    def import_rlwe_ksk(self, key):
        key = np.ascontiguousarray(key, dtype=np.uint64)
        check(lib().tp_import_rlwe_ksk(self.h, ptr(key), key.size))

    def rlwe_keyswitch(self, rlwes, t=1, mode=0, key=0, out=None):
        """[count][2N] -> [count][2N] NTT domain (mode and key as tp_rlwe_keyswitch_device); with
        out given, the result is accumulated into a copy of it (the reference's semantics)."""
        x = np.ascontiguousarray(rlwes, dtype=np.uint64).reshape(-1, 2 * self.N)
        d_in = DeviceBuffer.from_host(self, x)
        if out is not None:
            d_out = DeviceBuffer.from_host(self, np.ascontiguousarray(out, dtype=np.uint64).reshape(x.shape))
            mode |= T.TP_RKS_ACCUMULATE
        else:
            d_out = DeviceBuffer(self, x.size)
        check(lib().tp_rlwe_keyswitch_device(self.h, x.shape[0], t, mode, key, d_in.p, d_out.p))
        return d_out.to_host().reshape(x.shape)

    def pack_consecutive(self, lwes):
        x = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.N + 1)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, 2 * self.N)
        check(lib().tp_pack_consecutive_device(self.h, x.shape[0], d_in.p, d_out.p))
        return d_out.to_host()

    def modswitch(self, vec, source_modulus, target_modulus, kind=0, seed=0):
        """ModulusSwitch on the GPU (kind 0 ROUND, 1 FLOOR, 2 CEIL, 3 RANDOM with a seeded coin)."""
        v = np.ascontiguousarray(vec, dtype=np.uint64)
        d = DeviceBuffer.from_host(self, v)
        check(lib().tp_modswitch_device(self.h, d.p, d.p, v.size, source_modulus, target_modulus, kind, seed))
        return d.to_host().reshape(v.shape)

    @staticmethod
    def sap_digit_counts(lwes, radix):
        """Digits per input as SAPDecomposer::Decompose reads the modulus off the values
        (sap_decomposer.cpp:427-433, with the x86 shift: 1 << 64 is 1); sizes the output."""
        x = np.ascontiguousarray(lwes, dtype=np.uint64)
        x = x.reshape(-1, x.shape[-1])
        a0 = [int(v) for v in x[:, 0]]
        m = np.array([1 << (((v - 1).bit_length() if v > 1 else 0) & 63) for v in a0], dtype=object)
        for j in range(1, x.shape[1]):
            col = x[:, j].astype(object)
            m = np.where(m <= col, m * 2, m)
        rl = (radix - 1).bit_length()
        counts = []
        for v in m:
            v = int(v) % (1 << 64)
            d = 0
            while v > 1:
                v >>= rl
                d += 1
            counts.append(d)
        return np.array(counts, dtype=np.int64)

    def sap_decompose(self, lwes, radix, beta, restart=None, max_digits=None):
        """SAPDecomposer::Decompose of [count][n+1] LWEs: returns (digits [count][D][N+1], counts)
        with D the largest digit count; rows beyond an input's own count are zero."""
        x = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.n + 1)
        count = x.shape[0]
        if max_digits is None:
            max_digits = max(1, int(self.sap_digit_counts(x, radix).max()) if count else 1)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer.from_host(self, np.zeros(count * max_digits * (self.N + 1), np.uint64))
        d_cnt = DeviceBuffer(self, (count + 1) // 2)
        r = (1 << 64) - 1 if restart is None else restart
        check(lib().tp_sap_decompose_device(self.h, count, radix, beta, r, d_in.p, d_out.p, max_digits, d_cnt.p))
        counts = np.frombuffer(d_cnt.to_host().tobytes(), dtype=np.uint32)[:count].copy()
        D = int(counts.max()) if count else 0
        out = d_out.to_host().reshape(count, max_digits, self.N + 1)[:, :D]
        return out, counts

    def sap_homtrunc(self, rlwes_ntt, radix):
        """SAPDecomposer::HomTrunc of [count][2N] NTT-domain RLWEs."""
        x = np.ascontiguousarray(rlwes_ntt, dtype=np.uint64).reshape(-1, 2 * self.N)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, x.size)
        check(lib().tp_sap_homtrunc_device(self.h, x.shape[0], radix, d_in.p, d_out.p))
        return d_out.to_host().reshape(x.shape)

    def sap_reset(self, acc_ntt, radix):
        """SAPDecomposer::ResetAccumulator of [count][2N] NTT-domain accumulators."""
        x = np.ascontiguousarray(acc_ntt, dtype=np.uint64).reshape(-1, 2 * self.N)
        d = DeviceBuffer.from_host(self, x)
        check(lib().tp_sap_reset_device(self.h, x.shape[0], radix, d.p))
        return d.to_host().reshape(x.shape)

    def lwe_to_rgsw(self, lwes, digits, basis_log2):
        """LWEtoRGSWConverter::Convert of [count][n+1] LWEs -> [count][4 N digits] (left | right Lev)."""
        x = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.n + 1)
        d_in = DeviceBuffer.from_host(self, x)
        d_out = DeviceBuffer(self, x.shape[0] * 4 * self.N * digits)
        check(lib().tp_lwe_to_rgsw_device(self.h, x.shape[0], digits, basis_log2, d_in.p, d_out.p))
        return d_out.to_host().reshape(x.shape[0], 4 * self.N * digits)

    def eltwise(self, op, a, b=None, c=None, scalar=0):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        bufs = [DeviceBuffer.from_host(self, np.ascontiguousarray(v, dtype=np.uint64)) if v is not None else None
                for v in (a, b, c)]
        d_out = DeviceBuffer(self, a.size)
        check(lib().tp_eltwise_device(self.h, op, d_out.p, bufs[0].p, bufs[1].p if bufs[1] else None,
                                      bufs[2].p if bufs[2] else None, scalar, a.size))
        return d_out.to_host().reshape(a.shape)

    def bootstrap_batch(self, lwes, tv):
        lwes = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, self.n + 1)
        tv = np.ascontiguousarray(tv, dtype=np.uint64)
        count = lwes.shape[0]
        W = self.params.ks_out_dim + 1
        d_in = DeviceBuffer.from_host(self, lwes)
        d_tv = DeviceBuffer.from_host(self, tv)
        d_out = DeviceBuffer(self, count * W)
        stride = 0 if tv.size == 2 * self.N else 2 * self.N
        check(lib().tp_bootstrap_batch_device(self.h, count, d_in.p, d_tv.p, stride, d_out.p))
        return d_out.to_host().reshape(count, W)

    def bootstrap_batch_noalloc_device(self, count, d_lwes, d_tv, tv_stride, d_out):
        """tp_bootstrap_batch_device_noalloc on device buffers (allocates nothing itself)."""
        check(lib().tp_bootstrap_batch_device_noalloc(self.h, count, d_lwes.p, d_tv.p, tv_stride, d_out.p))

    def bootstrap_graph(self, count, d_lwes, d_tv, tv_stride, d_out):
        """A BootstrapGraph replaying tp_bootstrap_batch_device's launches on these device buffers."""
        return BootstrapGraph(self, count, d_lwes, d_tv, tv_stride, d_out)


class BootstrapGraph:
    """tp_graph_bootstrap_create / tp_graph_launch: one captured full bootstrap per launch() (on the
    context's stream, asynchronous), reading the buffers' current contents."""

    def __init__(self, ctx, count, d_lwes, d_tv, tv_stride, d_out):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().tp_graph_bootstrap_create(ctx.h, count, d_lwes.p, d_tv.p, tv_stride, d_out.p,
                                              ctypes.byref(self.h)))

    def launch(self):
        check(lib().tp_graph_launch(self.h))

    def close(self):
        if self.h:
            lib().tp_graph_destroy(self.h)
            self.h = ctypes.c_void_p()


# ---------------------------------------------------------------- key generation (host)

# seed value that selects ChaCha20 keyed from the operating system (real keys); any other seed is
# a reproducible test stream whose outputs reveal its state
SEED_OS_ENTROPY = 0xFFFFFFFFFFFFFFFF


def _words(a, n, name):
    """a as a contiguous uint64 array of exactly n words (the C keygen reads n words)."""
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
    if a.size != n:
        raise ValueError(f"{name} holds {a.size} words, expected {n}")
    return a


def keygen_secret(length, dist, seed, tag):
    out = np.empty(length, dtype=np.uint64)
    check(lib().tp_keygen_secret(ptr(out), length, dist, seed, tag))
    return out


def keygen_brk(lwe_key, rlwe_key, N, n, base_log2, digits, method, gadget_kind, sigma, seed):
    if method not in (0, 1, 2):
        raise ValueError(f"unknown blind-rotation method {method}")
    per = {0: 4 * N * digits * n, 1: 8 * N * digits * n, 2: 12 * N * digits * (n // 2)}[method]
    lk, rk = _words(lwe_key, n, "lwe_key"), _words(rlwe_key, N, "rlwe_key")
    out = np.empty(per, dtype=np.uint64)
    check(lib().tp_keygen_brk(ptr(out), ptr(lk), ptr(rk), N, n, base_log2, digits, method, gadget_kind, sigma, seed))
    return out


def keygen_auto_keys(s, N, base_log2, digits, sigma, seed):
    logN = N.bit_length() - 1
    sk = _words(s, N, "s")
    out = np.empty(logN * digits * 2 * N, dtype=np.uint64)
    check(lib().tp_keygen_auto_keys(ptr(out), ptr(sk), N, base_log2, digits, sigma, seed))
    return out


def keygen_rlwe_ksk(src, dst, N, base_log2, digits, sigma, seed, tag=6):
    sk, dk = _words(src, N, "src"), _words(dst, N, "dst")
    out = np.empty(digits * 2 * N, dtype=np.uint64)
    check(lib().tp_keygen_rlwe_ksk(ptr(out), ptr(sk), ptr(dk), N, base_log2, digits, sigma, seed, tag))
    return out


def keygen_lwe_ksk(src, dst, n_in, n_out, base_log2, digits, sigma, seed):
    sk, dk = _words(src, n_in, "src"), _words(dst, n_out, "dst")
    out = np.empty(n_in * digits * (n_out + 1), dtype=np.uint64)
    check(lib().tp_keygen_lwe_ksk(ptr(out), ptr(sk), ptr(dk), n_in, n_out, base_log2, digits, sigma, seed))
    return out


def sample_gaussian(count, sigma, seed, tag=0):
    """count discrete Gaussian samples (the keygen noise sampler) as int64."""
    out = np.empty(count, dtype=np.int64)
    check(lib().tp_sample_gaussian(out.ctypes.data, count, sigma, seed, tag))
    return out


def chacha20_block(key_words, counter, nonce_words):
    """The ChaCha20 block function of the entropy stream (RFC 8439, section 2.3): 16 uint32 words."""
    key = np.ascontiguousarray(key_words, dtype=np.uint32)
    nonce = np.ascontiguousarray(nonce_words, dtype=np.uint32)
    if key.size != 8 or nonce.size != 3:
        raise ValueError("ChaCha20 takes an 8-word key and a 3-word nonce")
    out = np.empty(16, dtype=np.uint32)
    check(lib().tp_chacha20_block(key.ctypes.data, counter, nonce.ctypes.data, out.ctypes.data))
    return out


# ---------------------------------------------------------------- reference-named operators

class GenericKey:
    """GenericKey{label, data} (lib/include/static/keys.h:14-39)."""

    def __init__(self, label, data):
        self.label = label
        self.data = np.ascontiguousarray(data, dtype=np.uint64)


class CGGIBlindRotationContext:
    """CGGIBlindRotationContext(distribution, Q, N, n, basis, digits, std)
    (cggi_blind_rotation.cpp:14-25).  basis is L = 2^b as in the reference."""

    def __init__(self, distribution, modulus, ring_dim, lwe_dim, basis, digits, std, seed=0):
        if modulus != Q:
            raise ValueError("only the Goldilocks modulus runs on the GPU path")
        b = int(basis).bit_length() - 1
        if basis != 1 << b:
            raise ValueError("basis must be a power of two")
        self.distribution, self.N, self.n, self.b, self.l, self.std = distribution, ring_dim, lwe_dim, b, digits, std
        self.seed = seed

    def GetRingDimension(self):
        return self.N

    def GetLWEDimension(self):
        return self.n

    def GetModulus(self):
        return Q

    def ConstructOperator(self, bundle):
        """bundle = [lwe secret (n), RLWE secret (N, coefficients)] (cggi_blind_rotation.cpp:184-190)."""
        return CGGIBlindRotator(self, bundle[0].data, bundle[1].data)


class CGGIBlindRotator:
    def __init__(self, ctx, lwe_key, rlwe_key):
        self.ctx = ctx
        method = T.METHOD_CGGI_BINARY if ctx.distribution == BINARY else T.METHOD_CGGI_TERNARY
        self.dev = Context(ctx.N, ctx.n, ctx.b, ctx.l, method)
        # binary keys follow the unsigned decomposition (MakeRGSW weights); ternary the signed one
        kind = 0 if method == T.METHOD_CGGI_BINARY else 1
        self.brk = keygen_brk(lwe_key, rlwe_key, ctx.N, ctx.n, ctx.b, ctx.l, method, kind, ctx.std, ctx.seed)
        self.dev.import_brk(self.brk)

    def BlindRotate(self, lwe_vec, accumulator, output_as_coefficients=False):
        """BlindRotator::BlindRotate (blindrotation_operator.h:29-40); returns the result."""
        return self.dev.blind_rotate(lwe_vec, accumulator, output_as_coefficients)


class LWEtoLWEConverter:
    """LWEtoLWEConverter (lwe_conversion.cpp:148-239).  Convert follows the reference's
    accumulate-into-output semantics: output[:n_out] -= acc, output[n_out] = input[N] - acc."""

    def __init__(self, dev_ctx, ksk):
        self.dev = dev_ctx
        self.dev.import_lwe_ksk(ksk)

    def Convert(self, output, lwe_in):
        res = self.dev.lwe_keyswitch_batch(np.asarray(lwe_in, dtype=np.uint64)[None, :])[0]
        n_out = self.dev.params.ks_out_dim
        out = np.asarray(output, dtype=np.uint64)
        acc_neg = res[:n_out]
        for k in range(n_out):
            out[k] = (int(out[k]) + int(acc_neg[k])) % Q
        out[n_out] = res[n_out]
        return out


def SampleExtract(rlwe_coef, N):
    """SampleExtract(out, in, 0, N, Q) (static/sample_extraction.h:20-36), host helper."""
    rlwe_coef = np.asarray(rlwe_coef, dtype=np.uint64)
    out = np.zeros(N + 1, dtype=np.uint64)
    out[0] = rlwe_coef[0]
    a = rlwe_coef[1:N][::-1].astype(object)
    out[1:N] = np.array([(Q - int(v)) % Q for v in a], dtype=np.uint64)
    out[N] = rlwe_coef[N]
    return out
