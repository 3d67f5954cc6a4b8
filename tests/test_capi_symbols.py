"""The C-ABI library loads and exports every symbol include/toothpaste_amd.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "toothpaste_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tp_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["tp_context_create", "tp_import_brk", "tp_blind_rotate", "tp_bootstrap_batch_device",
              "tp_lwe_keyswitch_batch_device", "tp_ntt_forward_device", "tp_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    from fhetoothpaste_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from fhetoothpaste_amd import _lib
    bound = set(_lib.SIGNATURES) | {"tp_last_error"}
    assert set(declared_symbols()) <= bound, set(declared_symbols()) - bound
