"""MI355X (gfx950) implementation of the CGGI/BMMP bootstrapping path of LSchild/FHEToothpaste.

The compute runs in hand-written HIP kernels in libtoothpaste_amd.so (C-ABI in
include/toothpaste_amd.h); this package is a thin ctypes mirror of the reference
operator API.  Build with `python -m fhetoothpaste_amd.build`.
"""
from ._lib import Q, ToothpasteError, device_count, lib  # noqa: F401
from .operators import (  # noqa: F401
    BINARY, TERNARY, CGGIBlindRotationContext, CGGIBlindRotator, Context, DeviceBuffer, GenericKey,
    LWEtoLWEConverter, SampleExtract, keygen_auto_keys, keygen_brk, keygen_lwe_ksk, keygen_rlwe_ksk, keygen_secret)
