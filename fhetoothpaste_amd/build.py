"""Builds libtoothpaste_amd.so in-tree for gfx950 with hipcc (no CMake, no JIT cache).

    python -m fhetoothpaste_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libtoothpaste_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are switched off for all device code (none is wanted on this path).
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall"]

SOURCES = [
    ("kernels.hip", DEVICE_FLAGS),
    ("kernels_br1.hip", DEVICE_FLAGS),
    ("kernels_br_fft.hip", DEVICE_FLAGS),
    ("kernels_br_lat.hip", DEVICE_FLAGS),
    ("kernels_br_half.hip", DEVICE_FLAGS),
    ("kernels_ks_mfma.hip", DEVICE_FLAGS),
    ("kernels_pack.hip", DEVICE_FLAGS),
    ("kernels_eltwise.hip", DEVICE_FLAGS),
    ("kernels_sap.hip", DEVICE_FLAGS),
    ("kernels_rgsw.hip", DEVICE_FLAGS),
    ("noalloc.hip", DEVICE_FLAGS),
    ("capi.cpp", HOST_FLAGS),
    ("tuning_cache.cpp", HOST_FLAGS),
    ("fft_key.cpp", HOST_FLAGS),
    ("keygen.cpp", HOST_FLAGS),
    ("serve.cpp", HOST_FLAGS),
    ("wire.cpp", HOST_FLAGS),
    ("dropin.cpp", HOST_FLAGS),
]


def _compile(src, flags):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    subprocess.check_call(cmd)
    return obj


def build(verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [(s, f) for s, f in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(lambda sf: _compile(*sf), srcs))
    tmp = OUT + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    # C++ program that drives the drop-in classes for the parity tests (tests/test_gpu_dropin.py)
    parity = os.path.join(ROOT, "tests", "cpp", "dropin_parity")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-o", parity, os.path.join(ROOT, "tests", "cpp", "dropin_parity.cpp"),
                           "-L" + HERE, "-ltoothpaste_amd", "-Wl,-rpath,$ORIGIN/../../fhetoothpaste_amd"])
    if verbose:
        print("built", OUT)
    return OUT


# Host-sanitised variant for the CPU tests: the host sources (C-ABI, key generation, wire format,
# serving scheduler, drop-in classes) with AddressSanitizer + UndefinedBehaviorSanitizer, the device
# objects of the regular build unchanged.  -Xarch_host keeps every sanitizer off the GPU code.
SAN_DIR = os.path.join(HERE, "_san")
SAN_OUT = os.path.join(SAN_DIR, "libtoothpaste_amd.so")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-shared-libsan",
             "-fno-omit-frame-pointer", "-g", "-O1", "-std=c++17", "-fPIC"]


def build_sanitized(verbose=True):
    os.makedirs(SAN_DIR, exist_ok=True)
    objs, todo = [], []
    for src, flags in SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        if src.endswith(".hip"):
            obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
            if not os.path.exists(obj):
                todo.append((src, flags, obj))
        else:
            obj = os.path.join(SAN_DIR, os.path.splitext(src)[0] + ".o")
            todo.append((src, SAN_FLAGS, obj))
        objs.append(obj)
    os.makedirs(BUILD_DIR, exist_ok=True)

    def comp(t):
        src, flags, obj = t
        subprocess.check_call([HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=6) as ex:
        list(ex.map(comp, todo))
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address",
                           "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-shared-libsan", "-o", SAN_OUT] + objs)
    if verbose:
        print("built", SAN_OUT)
    return SAN_OUT


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build()
    sys.exit(0)
