"""Continuous-batching bootstrap server (tp_server_* / tp_sched_* of libtoothpaste_amd.so).

The reference has no serving layer: it bootstraps one ciphertext per EvalFunc call
(negacyclic_function_evaluator.cpp:92-101) followed by LWEtoLWEConverter::Convert
(lwe_conversion.cpp:205-239).  Server keeps several parameter sets resident on one GPU and runs
the ciphertexts of concurrent requests in shared launches; each ciphertext's output equals the
one Context.bootstrap_batch gives for it alone.
"""
import ctypes

import numpy as np

from . import _lib as T
from ._lib import check, lib, ptr


class Scheduler:
    """Host-only FIFO slot scheduler (no GPU needed): admission per ciphertext under an HBM budget."""

    def __init__(self, slot_bytes, budget_bytes, max_slots):
        n = len(slot_bytes)
        self.h = ctypes.c_void_p()
        sb = (T.SZ * n)(*slot_bytes)
        ms = (T.U32 * n)(*max_slots)
        check(lib().tp_sched_create(ctypes.byref(self.h), n, sb, budget_bytes, ms))

    def close(self):
        if self.h:
            lib().tp_sched_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, set_index, count):
        rid = T.U64()
        check(lib().tp_sched_submit(self.h, set_index, count, ctypes.byref(rid)))
        return rid.value

    def admit(self, set_index, cap):
        slots = (T.U32 * max(cap, 1))()
        rids = (T.U64 * max(cap, 1))()
        idx = (T.U64 * max(cap, 1))()
        n = T.SZ()
        check(lib().tp_sched_admit(self.h, set_index, slots, rids, idx, cap, ctypes.byref(n)))
        return [(slots[i], rids[i], idx[i]) for i in range(n.value)]

    def release(self, set_index, slots, rids):
        n = len(slots)
        s = (T.U32 * max(n, 1))(*slots)
        r = (T.U64 * max(n, 1))(*rids)
        done = (T.INT * max(n, 1))()
        check(lib().tp_sched_release(self.h, set_index, s, r, n, done))
        return [bool(done[i]) for i in range(n)]

    def cancel(self, rid):
        check(lib().tp_sched_cancel(self.h, rid))

    def forget(self, rid):
        check(lib().tp_sched_forget(self.h, rid))

    def state(self, set_index):
        f, q, u = T.SZ(), T.SZ(), T.SZ()
        check(lib().tp_sched_state(self.h, set_index, ctypes.byref(f), ctypes.byref(q), ctypes.byref(u)))
        return f.value, q.value, u.value


class Request:
    """Handle of one submitted request; owns its input and output host buffers."""

    def __init__(self, server, rid, lwes, out):
        self.server, self.id, self.lwes, self.out = server, rid, lwes, out

    def done(self):
        d = T.INT()
        check(lib().tp_server_poll(self.server.h, self.id, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, timeout=60.0):
        check(lib().tp_server_wait(self.server.h, self.id, timeout))
        return self.out

    def cancel(self):
        check(lib().tp_server_cancel(self.server.h, self.id))

    def times(self):
        s, c = ctypes.c_double(), ctypes.c_double()
        check(lib().tp_server_times(self.server.h, self.id, ctypes.byref(s), ctypes.byref(c)))
        return s.value, c.value

    def release(self):
        check(lib().tp_server_release(self.server.h, self.id))


class Server:
    """Serving loop over contexts (one per parameter set, brk and LWE ksk imported)."""

    def __init__(self, contexts, chunk_steps=64, max_slots_per_set=4096, hbm_budget_bytes=0):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        cfg = T.ServerConfig(chunk_steps, max_slots_per_set, hbm_budget_bytes)
        self.h = ctypes.c_void_p()
        check(lib().tp_server_create(ctypes.byref(self.h), arr, len(self.contexts), ctypes.byref(cfg)))

    def close(self):
        if self.h:
            check(lib().tp_server_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, set_index, lwes, tv):
        ctx = self.contexts[set_index]
        lwes = np.ascontiguousarray(lwes, dtype=np.uint64).reshape(-1, ctx.n + 1)
        tv = np.ascontiguousarray(tv, dtype=np.uint64)
        if tv.size != 2 * ctx.N:
            raise ValueError("test vector must hold 2N coefficients")
        out = np.zeros((lwes.shape[0], ctx.params.ks_out_dim + 1), dtype=np.uint64)
        rid = T.U64()
        check(lib().tp_server_submit(self.h, set_index, lwes.shape[0], ptr(lwes), ptr(tv), ptr(out),
                                     ctypes.byref(rid)))
        return Request(self, rid.value, lwes, out)

    def stats(self):
        it, la, ct = T.U64(), T.U64(), T.U64()
        check(lib().tp_server_stats(self.h, ctypes.byref(it), ctypes.byref(la), ctypes.byref(ct)))
        return {"iterations": it.value, "launches": la.value, "ciphertexts": ct.value}


def replay(server, requests, progress=None):
    """Submit requests at their arrival instants and collect completion times.

    requests: list of (arrival_s, set_index, lwes, tv) sorted by arrival_s (seconds from the
    start of the replay).  Returns a list of dicts {set, count, arrival, submitted, done}, all in
    seconds of the steady clock the server stamps completions with (time.monotonic on Linux).
    Latency of a request = done - arrival (arrival is the scheduled instant, so a late submission
    counts against the server).  Completed requests are released as they are collected.
    """
    import queue
    import threading
    import time

    pending = queue.Queue()
    records = []
    errors = []

    def collect():
        while True:
            item = pending.get()
            if item is None:
                return
            h, rec = item
            try:
                h.wait(600.0)
                rec["submitted"], rec["done"] = h.times()
                h.release()
            except Exception as e:  # noqa: BLE001 - reported to the caller below
                errors.append(e)
                return
            records.append(rec)

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = time.monotonic()
    next_report = 10.0
    for arrival, set_index, lwes, tv in requests:
        now = time.monotonic() - t0
        if arrival > now:
            time.sleep(arrival - now)
        h = server.submit(set_index, lwes, tv)
        pending.put((h, {"set": set_index, "count": int(lwes.shape[0]), "arrival": t0 + arrival}))
        if progress and arrival >= next_report:
            progress("replay %.0f s: %d requests submitted, %d collected" % (arrival, len(records) + pending.qsize(),
                                                                            len(records)))
            next_report += 10.0
    pending.put(None)
    th.join()
    if errors:
        raise errors[0]
    return records
