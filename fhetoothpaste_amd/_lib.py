"""ctypes binding of libtoothpaste_amd.so (the C-ABI declared in include/toothpaste_amd.h).

The library is built in-tree by fhetoothpaste_amd.build.  There is no CPU fallback: if the
library or a gfx950 device is missing, calls raise ToothpasteError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TOOTHPASTE_AMD_LIB selects another build of the same library (the host-sanitised build of
# fhetoothpaste_amd.build.build_sanitized, which the CPU sanitizer test loads).
LIB_PATH = os.environ.get("TOOTHPASTE_AMD_LIB") or os.path.join(HERE, "libtoothpaste_amd.so")

Q = 0xFFFFFFFF00000001
TP_OK, TP_E_INVALID, TP_E_DEVICE, TP_E_STATE, TP_E_NOMEM, TP_E_TIMEOUT = 0, -1, -2, -3, -4, -5
METHOD_CGGI_BINARY, METHOD_CGGI_TERNARY, METHOD_BMMP = 0, 1, 2
TP_RKS_CONVERT, TP_RKS_AUTOMORPHISM, TP_RKS_SWITCH_ONLY, TP_RKS_ACCUMULATE = 0, 1, 2, 4
TP_OPT_BR_WAVES = 1
TP_OPT_BR_SPLIT_MAX = 2
TP_OPT_BR_ENGINE = 3
TP_OPT_BR_FFT_WAVES = 4
TP_OPT_BR_GRID = 5
TP_OPT_BR_SPLIT_HALVES = 6
TP_OPT_KS_MFMA = 7
TP_ELT_ADD, TP_ELT_SUB, TP_ELT_MUL, TP_ELT_FMA, TP_ELT_MULS, TP_ELT_REDUCE = 0, 1, 2, 3, 4, 5

u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p
U32, U64, SZ, INT = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int


class ToothpasteError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("toothpaste_amd error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("ring_dim", U32), ("lwe_dim", U32), ("br_base_log2", U32), ("br_digits", U32),
                ("method", U32), ("ks_out_dim", U32), ("ks_base_log2", U32), ("ks_digits", U32),
                ("rlwe_ks_base_log2", U32), ("rlwe_ks_digits", U32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("chunk_steps", U32), ("max_slots_per_set", U32), ("hbm_budget_bytes", SZ)]


# name -> argtypes (all return int status)
SIGNATURES = {
    "tp_device_count": [ctypes.POINTER(INT)],
    "tp_context_create": [ctypes.POINTER(vp), ctypes.POINTER(Params), INT],
    "tp_context_destroy": [vp],
    "tp_context_sync": [vp],
    "tp_context_params": [vp, ctypes.POINTER(Params)],
    "tp_import_brk": [vp, u64p, SZ],
    "tp_import_lwe_ksk": [vp, u64p, SZ],
    "tp_ntt_forward_device": [vp, vp, vp, SZ],
    "tp_ntt_inverse_device": [vp, vp, vp, SZ],
    "tp_blind_rotate": [vp, u64p, u64p, u64p, INT],
    "tp_eval_func_batch_device": [vp, SZ, vp, vp, SZ, vp, vp, INT],
    "tp_lwe_keyswitch_batch_device": [vp, SZ, vp, SZ, vp],
    "tp_bootstrap_batch_device": [vp, SZ, vp, vp, SZ, vp],
    "tp_bootstrap_batch_device_noalloc": [vp, SZ, vp, vp, SZ, vp],
    "tp_graph_bootstrap_create": [vp, SZ, vp, vp, SZ, vp, ctypes.POINTER(vp)],
    "tp_graph_launch": [vp],
    "tp_graph_destroy": [vp],
    "tp_malloc": [vp, ctypes.POINTER(vp), SZ],
    "tp_free": [vp, vp],
    "tp_copy_to_device": [vp, vp, u64p, SZ],
    "tp_copy_to_host": [vp, u64p, vp, SZ],
    "tp_copy_on_device": [vp, vp, vp, SZ],
    "tp_timer_mark": [vp, INT],
    "tp_timer_elapsed": [vp, INT, INT, ctypes.POINTER(ctypes.c_float)],
    "tp_keygen_secret": [u64p, U32, INT, U64, U64],
    "tp_keygen_brk": [u64p, u64p, u64p, U32, U32, U32, U32, INT, INT, ctypes.c_double, U64],
    "tp_keygen_lwe_ksk": [u64p, u64p, u64p, U32, U32, U32, U32, ctypes.c_double, U64],
    "tp_keygen_auto_keys": [u64p, u64p, U32, U32, U32, ctypes.c_double, U64],
    "tp_import_auto_keys": [vp, u64p, SZ],
    "tp_automorphism_device": [vp, SZ, U32, vp, vp],
    "tp_trace_device": [vp, SZ, vp, vp],
    "tp_pack_device": [vp, SZ, vp, vp, vp],
    "tp_import_rlwe_ksk": [vp, u64p, SZ],
    "tp_set_option": [vp, INT, INT],
    "tp_get_option": [vp, INT, ctypes.POINTER(INT)],
    "tp_autotune": [vp, SZ, ctypes.POINTER(ctypes.c_float)],
    "tp_autotune_cached": [vp, SZ, ctypes.c_char_p, ctypes.POINTER(INT), ctypes.POINTER(ctypes.c_float)],
    "tp_rlwe_keyswitch_device": [vp, SZ, U32, INT, INT, vp, vp],
    "tp_pack_consecutive_device": [vp, SZ, vp, vp],
    "tp_eltwise_device": [vp, INT, vp, vp, vp, vp, U64, SZ],
    "tp_modswitch_device": [vp, vp, vp, SZ, U64, U64, INT, U64],
    "tp_sap_decompose_device": [vp, SZ, U32, U32, U64, vp, vp, U32, vp],
    "tp_sap_homtrunc_device": [vp, SZ, U32, vp, vp],
    "tp_sap_reset_device": [vp, SZ, U64, vp],
    "tp_lwe_to_rgsw_device": [vp, SZ, U32, U32, vp, vp],
    "tp_keygen_rlwe_ksk": [u64p, u64p, u64p, U32, U32, U32, ctypes.c_double, U64, U64],
    "tp_sample_gaussian": [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, U64, U64],
    "tp_chacha20_block": [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p],
    "tp_wire_bytes": [SZ, ctypes.POINTER(SZ)],
    "tp_wire_encode": [ctypes.c_char_p, SZ, INT, U32, ctypes.POINTER(Params), u64p, SZ],
    "tp_wire_decode": [ctypes.c_char_p, SZ, ctypes.POINTER(INT), ctypes.POINTER(U32), ctypes.POINTER(Params), u64p,
                       SZ, ctypes.POINTER(SZ)],
    "tp_wire_save": [ctypes.c_char_p, INT, U32, ctypes.POINTER(Params), u64p, SZ],
    "tp_wire_load": [ctypes.c_char_p, ctypes.POINTER(INT), ctypes.POINTER(U32), ctypes.POINTER(Params), u64p, SZ,
                     ctypes.POINTER(SZ)],
    "tp_sched_create": [ctypes.POINTER(vp), U32, ctypes.POINTER(SZ), SZ, ctypes.POINTER(U32)],
    "tp_sched_destroy": [vp],
    "tp_sched_submit": [vp, U32, SZ, ctypes.POINTER(U64)],
    "tp_sched_admit": [vp, U32, ctypes.POINTER(U32), ctypes.POINTER(U64), ctypes.POINTER(U64), SZ,
                       ctypes.POINTER(SZ)],
    "tp_sched_release": [vp, U32, ctypes.POINTER(U32), ctypes.POINTER(U64), SZ, ctypes.POINTER(INT)],
    "tp_sched_cancel": [vp, U64],
    "tp_sched_forget": [vp, U64],
    "tp_sched_state": [vp, U32, ctypes.POINTER(SZ), ctypes.POINTER(SZ), ctypes.POINTER(SZ)],
    "tp_server_create": [ctypes.POINTER(vp), ctypes.POINTER(vp), U32, ctypes.POINTER(ServerConfig)],
    "tp_server_destroy": [vp],
    "tp_server_submit": [vp, U32, SZ, u64p, u64p, u64p, ctypes.POINTER(U64)],
    "tp_server_poll": [vp, U64, ctypes.POINTER(INT)],
    "tp_server_wait": [vp, U64, ctypes.c_double],
    "tp_server_cancel": [vp, U64],
    "tp_server_release": [vp, U64],
    "tp_server_times": [vp, U64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "tp_server_stats": [vp, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U64)],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ToothpasteError(TP_E_DEVICE, "libtoothpaste_amd.so is not built (python -m fhetoothpaste_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = INT
        L.tp_last_error.argtypes = []
        L.tp_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != TP_OK:
        raise ToothpasteError(rc, lib().tp_last_error().decode())
    return rc


def ptr(a):
    """uint64* of a C-contiguous numpy uint64 array."""
    if a is None:
        return None
    if a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous uint64 array")
    return a.ctypes.data_as(u64p)


def device_count():
    n = INT(0)
    check(lib().tp_device_count(ctypes.byref(n)))
    return n.value
