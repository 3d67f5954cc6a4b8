"""Serving scheduler core (tp_sched_*): FIFO admission per ciphertext under slot and HBM limits.

Host-only code of libtoothpaste_amd.so; runs without a GPU.
"""
import pytest

from fhetoothpaste_amd._lib import TP_E_INVALID, TP_E_NOMEM, ToothpasteError
from fhetoothpaste_amd.serving import Scheduler


def test_fifo_admission_and_release():
    s = Scheduler([100], 10_000, [4])
    r1 = s.submit(0, 3)
    r2 = s.submit(0, 2)
    got = s.admit(0, 16)
    # slots 0.. handed out in order; request 1 fully before request 2
    assert [(g[1], g[2]) for g in got] == [(r1, 0), (r1, 1), (r1, 2), (r2, 0)]
    assert sorted(g[0] for g in got) == [0, 1, 2, 3]
    assert s.state(0) == (0, 1, 400)
    assert s.admit(0, 16) == []  # no free slot
    # request 1 completes when its last ciphertext is released
    done = s.release(0, [got[0][0], got[1][0]], [r1, r1])
    assert done == [False, False]
    more = s.admit(0, 16)
    assert [(g[1], g[2]) for g in more] == [(r2, 1)]
    assert s.release(0, [got[2][0]], [r1]) == [True]
    assert s.release(0, [got[3][0], more[0][0]], [r2, r2]) == [False, True]
    assert s.state(0) == (4, 0, 0)


def test_budget_limits_admission_across_sets():
    # set 0 slots cost 300 B, set 1 slots 500 B; budget 1000 B shared
    s = Scheduler([300, 500], 1000, [8, 8])
    a = s.submit(0, 5)
    b = s.submit(1, 5)
    g0 = s.admit(0, 8)
    assert len(g0) == 3  # 900 B
    assert s.admit(1, 8) == []  # 500 more would exceed
    s.release(0, [g0[0][0], g0[1][0]], [a, a])  # 300 B in use
    g1 = s.admit(1, 8)
    assert len(g1) == 1 and g1[0][1] == b  # 800 B
    assert s.state(1)[2] == 800


def test_cap_limits_one_call():
    s = Scheduler([8], 1 << 20, [64])
    s.submit(0, 10)
    assert [g[2] for g in s.admit(0, 4)] == [0, 1, 2, 3]
    assert [g[2] for g in s.admit(0, 100)] == [4, 5, 6, 7, 8, 9]
    assert s.admit(0, 1) == []  # nothing left of request r
    assert s.state(0) == (54, 0, 80)


def test_cancel_drops_queued_ciphertexts():
    s = Scheduler([1], 100, [2])
    r1 = s.submit(0, 5)
    r2 = s.submit(0, 1)
    g = s.admit(0, 8)
    assert [(x[1], x[2]) for x in g] == [(r1, 0), (r1, 1)]
    s.cancel(r1)
    # the two admitted ciphertexts finish; the request completes after the second
    assert s.release(0, [g[0][0]], [r1]) == [False]
    assert s.release(0, [g[1][0]], [r1]) == [True]
    # request 2 is next in line
    assert [(x[1], x[2]) for x in s.admit(0, 8)] == [(r2, 0)]


def test_slot_reuse_keeps_pool_size():
    s = Scheduler([16], 1 << 30, [3])
    r = s.submit(0, 100)
    seen = set()
    left = 100
    while left:
        g = s.admit(0, 8)
        assert 0 < len(g) <= 3
        seen |= {x[0] for x in g}
        done = s.release(0, [x[0] for x in g], [x[1] for x in g])
        left -= len(g)
        assert done[-1] == (left == 0)
    assert seen == {0, 1, 2}
    assert s.state(0) == (3, 0, 0)
    s.forget(r)


def test_errors():
    with pytest.raises(ToothpasteError) as e:
        Scheduler([0], 10, [1])
    assert e.value.code == TP_E_INVALID
    s = Scheduler([64], 32, [4])
    with pytest.raises(ToothpasteError) as e:
        s.submit(0, 1)
    assert e.value.code == TP_E_NOMEM
    s2 = Scheduler([8], 64, [4])
    with pytest.raises(ToothpasteError):
        s2.submit(1, 1)
    with pytest.raises(ToothpasteError):
        s2.submit(0, 0)
    with pytest.raises(ToothpasteError):
        s2.cancel(12345)
