"""Parameter checks of tp_context_create, which run before any device call.

The checks guard what the kernels assume: ring dimensions with a built transform, digit counts with a
built blind-rotation kernel (2..4 at N = 1024 / 2048, 3..4 at N = 4096), gadgets whose
digits cover at most 64 bits (so a key-switch chunk of at most 64 key rows holds at least one input
word), and an even LWE dimension for BMMP's pairs.  Rejected parameters return TP_E_INVALID; accepted
ones go on to the device query (TP_E_DEVICE without a GPU, a context with one).
"""
import pytest

import fhetoothpaste_amd as tp
from fhetoothpaste_amd._lib import TP_E_INVALID, ToothpasteError

BASE = dict(ring_dim=1024, lwe_dim=16, base_log2=8, digits=2, method=0, ks_out_dim=0, ks_base_log2=0,
            ks_digits=0)


def create(**kw):
    p = dict(BASE, **kw)
    return tp.Context(p["ring_dim"], p["lwe_dim"], p["base_log2"], p["digits"], p["method"], p["ks_out_dim"],
                      p["ks_base_log2"], p["ks_digits"])


@pytest.mark.parametrize("kw", [
    dict(ring_dim=512), dict(ring_dim=8192), dict(ring_dim=1000),
    dict(method=3),
    dict(base_log2=0), dict(base_log2=33), dict(digits=0),
    dict(base_log2=16, digits=5),                       # 5 digits of 16 bits exceed 64
    dict(method=2, lwe_dim=15),                         # BMMP pairs the LWE coefficients
    dict(base_log2=16, digits=1),                       # no blind-rotation kernel for 1 digit
    dict(base_log2=12, digits=5),                       # ... nor for 5
    dict(ring_dim=4096, base_log2=16, digits=2),        # ... nor for 2 digits at N = 4096
    dict(ks_out_dim=40, ks_base_log2=0, ks_digits=4),
    dict(ks_out_dim=40, ks_base_log2=17, ks_digits=2),
    dict(ks_out_dim=40, ks_base_log2=4, ks_digits=0),
    dict(ks_out_dim=40, ks_base_log2=3, ks_digits=23),  # more than ceil(64 / 3) = 22 digits
    dict(ks_out_dim=40, ks_base_log2=3, ks_digits=22),  # 22 x 3 = 66 bits: the top digit would overflow
    dict(ks_out_dim=40, ks_base_log2=2, ks_digits=33),
], ids=lambda kw: ",".join("%s=%s" % x for x in kw.items()))
def test_rejects(kw):
    with pytest.raises(ToothpasteError) as e:
        create(**kw)
    assert e.value.code == TP_E_INVALID


@pytest.mark.parametrize("kw", [
    dict(),
    dict(ring_dim=2048, base_log2=16, digits=4),
    dict(ring_dim=4096, method=2, lwe_dim=10, base_log2=14, digits=3),
    dict(ks_out_dim=630, ks_base_log2=4, ks_digits=14),
    dict(ks_out_dim=33, ks_base_log2=2, ks_digits=32),  # 32 digits x 2 bits = 64: the most digits accepted
    dict(ks_out_dim=40, ks_base_log2=3, ks_digits=21),
    dict(ks_out_dim=17, ks_base_log2=16, ks_digits=4),
], ids=lambda kw: ",".join("%s=%s" % x for x in kw.items()) or "defaults")
def test_accepts(kw):
    try:
        create(**kw).close()
    except ToothpasteError as e:
        assert e.code != TP_E_INVALID, str(e)
