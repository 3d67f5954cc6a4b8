// Continuous-batching bootstrap server.
//
// The reference evaluates one ciphertext per call (NegacyclicFunctionEvaluator::EvalFunc,
// negacyclic_function_evaluator.cpp:92-101, followed by LWEtoLWEConverter::Convert,
// lwe_conversion.cpp:205-239) and has no serving layer.  This one keeps the keys of every
// parameter set resident in HBM and schedules ciphertexts at blind-rotation-step granularity:
//
//   * Scheduler (host only): per parameter set a pool of slots; a slot holds one ciphertext's
//     input, accumulator, extracted LWE, output and test-vector row in HBM.  Ciphertexts of
//     queued requests are admitted FIFO whenever a slot is free and the HBM footprint of all
//     admitted slots stays within the budget.
//   * Server loop (one thread per parameter set, each on its set's stream so the sets overlap on
//     the device): every iteration admits new ciphertexts, advances every active slot of the set
//     by `chunk_steps` blind-rotation steps in one launch, key-switches the slots that finished,
//     copies their results out and frees the slots.  Requests that arrive while others are on
//     the device join at the next iteration.
//
// A ciphertext's result does not depend on its slot or on the iteration it was admitted in:
// every slot runs the same steps on its own data and the accumulator is parked in HBM between
// launches as the exact 64-bit words the kernel holds in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/toothpaste_amd.h"
#include "context_internal.hpp"

// ---------------------------------------------------------------- scheduler core (host only)

struct tp_sched {
    struct Request {
        uint64_t id = 0;
        uint32_t set = 0;
        size_t count = 0;
        size_t next = 0;  // next ciphertext index to admit
        size_t done = 0;  // ciphertexts released
        bool cancelled = false;
    };
    struct SetState {
        size_t slot_bytes = 0;
        std::vector<uint32_t> free_slots;
        std::deque<uint64_t> queue;  // requests with ciphertexts still to admit, FIFO
    };
    std::vector<SetState> sets;
    std::map<uint64_t, Request> reqs;
    size_t budget = 0, used = 0;
    uint64_t next_id = 1;
};

namespace {

bool request_complete(const tp_sched::Request& r) {
    return r.done == r.count || (r.cancelled && r.done == r.next);
}

}  // namespace

extern "C" {

int tp_sched_create(tp_sched** out, uint32_t nsets, const size_t* slot_bytes, size_t budget_bytes,
                    const uint32_t* max_slots) {
    if (!out || !slot_bytes || !max_slots || nsets == 0) return tp::set_error(TP_E_INVALID, "bad scheduler arguments");
    *out = nullptr;
    for (uint32_t s = 0; s < nsets; ++s)
        if (slot_bytes[s] == 0 || max_slots[s] == 0) return tp::set_error(TP_E_INVALID, "empty slot pool");
    tp_sched* S = new tp_sched;
    S->sets.resize(nsets);
    S->budget = budget_bytes;
    for (uint32_t s = 0; s < nsets; ++s) {
        S->sets[s].slot_bytes = slot_bytes[s];
        for (uint32_t i = max_slots[s]; i-- > 0;) S->sets[s].free_slots.push_back(i);  // slot 0 handed out first
    }
    *out = S;
    return TP_OK;
}

int tp_sched_destroy(tp_sched* s) {
    delete s;
    return TP_OK;
}

int tp_sched_submit(tp_sched* s, uint32_t set, size_t count, uint64_t* id) {
    if (!s || !id) return tp::set_error(TP_E_INVALID, "null argument");
    if (set >= s->sets.size()) return tp::set_error(TP_E_INVALID, "unknown parameter set");
    if (count == 0) return tp::set_error(TP_E_INVALID, "empty request");
    if (s->sets[set].slot_bytes > s->budget) return tp::set_error(TP_E_NOMEM, "one ciphertext exceeds the HBM budget");
    tp_sched::Request r;
    r.id = s->next_id++;
    r.set = set;
    r.count = count;
    s->reqs[r.id] = r;
    s->sets[set].queue.push_back(r.id);
    *id = r.id;
    return TP_OK;
}

int tp_sched_admit(tp_sched* s, uint32_t set, uint32_t* slots, uint64_t* req_ids, uint64_t* indices, size_t cap,
                   size_t* n_admitted) {
    if (!s || !n_admitted || (cap && (!slots || !req_ids || !indices)))
        return tp::set_error(TP_E_INVALID, "null argument");
    if (set >= s->sets.size()) return tp::set_error(TP_E_INVALID, "unknown parameter set");
    tp_sched::SetState& st = s->sets[set];
    size_t n = 0;
    while (n < cap && !st.queue.empty() && !st.free_slots.empty() && s->used + st.slot_bytes <= s->budget) {
        auto it = s->reqs.find(st.queue.front());
        if (it == s->reqs.end() || it->second.cancelled || it->second.next >= it->second.count) {
            st.queue.pop_front();
            continue;
        }
        tp_sched::Request& r = it->second;
        slots[n] = st.free_slots.back();
        st.free_slots.pop_back();
        req_ids[n] = r.id;
        indices[n] = r.next++;
        s->used += st.slot_bytes;
        ++n;
        if (r.next >= r.count) st.queue.pop_front();
    }
    *n_admitted = n;
    return TP_OK;
}

int tp_sched_release(tp_sched* s, uint32_t set, const uint32_t* slots, const uint64_t* req_ids, size_t n,
                     int* completed) {
    if (!s || (n && (!slots || !req_ids))) return tp::set_error(TP_E_INVALID, "null argument");
    if (set >= s->sets.size()) return tp::set_error(TP_E_INVALID, "unknown parameter set");
    tp_sched::SetState& st = s->sets[set];
    for (size_t i = 0; i < n; ++i) {
        st.free_slots.push_back(slots[i]);
        s->used -= st.slot_bytes;
        auto it = s->reqs.find(req_ids[i]);
        int done = 0;
        if (it != s->reqs.end()) {
            it->second.done++;
            done = request_complete(it->second);
        }
        if (completed) completed[i] = done;
    }
    return TP_OK;
}

int tp_sched_cancel(tp_sched* s, uint64_t id) {
    if (!s) return tp::set_error(TP_E_INVALID, "null argument");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    it->second.cancelled = true;
    return TP_OK;
}

int tp_sched_forget(tp_sched* s, uint64_t id) {
    if (!s) return tp::set_error(TP_E_INVALID, "null argument");
    s->reqs.erase(id);
    return TP_OK;
}

int tp_sched_state(tp_sched* s, uint32_t set, size_t* free_slots, size_t* queued_requests, size_t* used_bytes) {
    if (!s || set >= s->sets.size()) return tp::set_error(TP_E_INVALID, "bad arguments");
    if (free_slots) *free_slots = s->sets[set].free_slots.size();
    if (queued_requests) *queued_requests = s->sets[set].queue.size();
    if (used_bytes) *used_bytes = s->used;
    return TP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- server

namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point t) { return std::chrono::duration<double>(t.time_since_epoch()).count(); }

}  // namespace

struct tp_server {
    struct Req {
        uint64_t id = 0;
        uint32_t set = 0;
        size_t count = 0;
        const uint64_t* lwes = nullptr;  // caller-owned until completion
        uint64_t* out = nullptr;         // caller-owned
        std::vector<uint64_t> tv;
        int tv_row = -1;  // row of the device test-vector arena while the request has live slots
        size_t live_slots = 0;
        bool done = false, cancelled = false;
        Clock::time_point t_submit, t_done;
    };
    struct Set {
        tp_context* ctx = nullptr;
        uint32_t N = 0, n = 0, W = 0, total_steps = 0, capacity = 0;
        // device slot pool
        uint64_t *d_lwe = nullptr, *d_state = nullptr, *d_ext = nullptr, *d_out = nullptr, *d_tv = nullptr;
        uint32_t* d_ctrl = nullptr;  // [step | tv_index | active | rows], capacity words each
        // pinned host mirrors, rewritten only while the set's stream is idle
        uint32_t* h_ctrl = nullptr;
        uint64_t* h_in = nullptr;   // capacity x (n + 1)
        uint64_t* h_out = nullptr;  // capacity x W
        // host bookkeeping
        std::vector<uint32_t> step;
        std::vector<uint64_t> slot_req, slot_idx;  // request / ciphertext index per slot (0 = empty)
        std::vector<uint32_t> fresh, active, finished;
        std::vector<int> tv_free;
    };
    std::vector<Set> sets;
    tp_sched* sched = nullptr;
    uint32_t chunk_steps = 64;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<uint64_t, std::unique_ptr<Req>> reqs;
    std::vector<std::thread> workers;
    std::atomic<bool> stop{false};
    int failed = 0;
    std::string fail_msg;
    std::atomic<uint64_t> launches{0};  // advanced outside the lock by the set workers
    uint64_t iterations = 0, ciphertexts = 0;
};

namespace {

#define SRV_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return tp::set_error(TP_E_DEVICE, (std::string(#call ": ") + hipGetErrorString(e_)).c_str()); \
    } while (0)

void free_tv_row(tp_server::Set& st, tp_server::Req& r) {
    if (r.tv_row >= 0) {
        st.tv_free.push_back(r.tv_row);
        r.tv_row = -1;
    }
}

// Admission (lock held): assign slots, stage inputs, upload test vectors of new requests.
int admit(tp_server* S, uint32_t si) {
    tp_server::Set& st = S->sets[si];
    st.fresh.clear();
    std::vector<uint32_t> slots(st.capacity);
    std::vector<uint64_t> rid(st.capacity), idx(st.capacity);
    size_t n = 0;
    int rc = tp_sched_admit(S->sched, si, slots.data(), rid.data(), idx.data(), st.capacity, &n);
    if (rc) return rc;
    hipStream_t stream = tp::context_stream(st.ctx);
    for (size_t i = 0; i < n; ++i) {
        tp_server::Req& r = *S->reqs.at(rid[i]);
        if (r.tv_row < 0) {
            r.tv_row = st.tv_free.back();
            st.tv_free.pop_back();
            SRV_HIP(hipMemcpyAsync(st.d_tv + (size_t)r.tv_row * 2 * st.N, r.tv.data(), (size_t)2 * st.N * 8,
                                   hipMemcpyHostToDevice, stream));
        }
        r.live_slots++;
        const uint32_t sl = slots[i];
        st.slot_req[sl] = rid[i];
        st.slot_idx[sl] = idx[i];
        st.step[sl] = 0;
        st.h_ctrl[st.capacity + sl] = (uint32_t)r.tv_row;
        std::memcpy(st.h_in + i * (st.n + 1), r.lwes + idx[i] * (st.n + 1), (size_t)(st.n + 1) * 8);
        st.fresh.push_back(sl);
    }
    return TP_OK;
}

// Device work of one iteration for one set (no lock: touches only this set's buffers).
int advance(tp_server* S, tp_server::Set& st) {
    hipStream_t stream = tp::context_stream(st.ctx);
    for (size_t i = 0; i < st.fresh.size(); ++i)
        SRV_HIP(hipMemcpyAsync(st.d_lwe + (size_t)st.fresh[i] * (st.n + 1), st.h_in + i * (st.n + 1),
                               (size_t)(st.n + 1) * 8, hipMemcpyHostToDevice, stream));
    st.active.clear();
    st.finished.clear();
    for (uint32_t sl = 0; sl < st.capacity; ++sl)
        if (st.slot_req[sl]) {
            st.active.push_back(sl);
            if (st.step[sl] + S->chunk_steps >= st.total_steps) st.finished.push_back(sl);
        }
    if (st.active.empty()) return TP_OK;
    const uint32_t C = st.capacity;
    std::copy(st.step.begin(), st.step.end(), st.h_ctrl);
    std::copy(st.active.begin(), st.active.end(), st.h_ctrl + 2 * C);
    std::copy(st.finished.begin(), st.finished.end(), st.h_ctrl + 3 * C);
    SRV_HIP(hipMemcpyAsync(st.d_ctrl, st.h_ctrl, (size_t)4 * C * 4, hipMemcpyHostToDevice, stream));
    int rc = tp::launch_chunk(st.ctx, st.active.size(), st.d_ctrl + 2 * C, st.d_ctrl, st.d_state, st.d_lwe, st.d_tv,
                              st.d_ctrl + C, st.d_ext, S->chunk_steps);
    if (rc) return rc;
    S->launches++;
    if (!st.finished.empty()) {
        rc = tp::launch_ks_rows(st.ctx, st.finished.size(), st.d_ext, st.N + 1, st.d_ctrl + 3 * C, st.d_out);
        if (rc) return rc;
        SRV_HIP(hipMemcpyAsync(st.h_out, st.d_out, st.finished.size() * st.W * 8, hipMemcpyDeviceToHost, stream));
    }
    for (uint32_t sl : st.active) st.step[sl] = std::min(st.step[sl] + S->chunk_steps, st.total_steps);
    return TP_OK;
}

// Delivery (lock held, set's stream idle): copy results out, free slots, complete requests.
bool deliver(tp_server* S, uint32_t si) {
    tp_server::Set& st = S->sets[si];
    if (st.finished.empty()) return false;
    std::vector<uint64_t> rids(st.finished.size());
    for (size_t i = 0; i < st.finished.size(); ++i) {
        const uint32_t sl = st.finished[i];
        tp_server::Req& r = *S->reqs.at(st.slot_req[sl]);
        if (!r.cancelled) std::memcpy(r.out + st.slot_idx[sl] * st.W, st.h_out + i * st.W, (size_t)st.W * 8);
        rids[i] = r.id;
    }
    std::vector<int> done(st.finished.size());
    tp_sched_release(S->sched, si, st.finished.data(), rids.data(), st.finished.size(), done.data());
    bool any = false;
    const Clock::time_point now = Clock::now();
    for (size_t i = 0; i < st.finished.size(); ++i) {
        const uint32_t sl = st.finished[i];
        tp_server::Req& r = *S->reqs.at(rids[i]);
        st.slot_req[sl] = 0;
        r.live_slots--;
        S->ciphertexts++;
        if (done[i] && !r.done) {
            r.done = true;
            r.t_done = now;
            any = true;
        }
        if (r.live_slots == 0 && (r.done || r.cancelled)) free_tv_row(st, r);
    }
    st.finished.clear();
    return any;
}

void fail_all(tp_server* S, int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->failed = code;
    S->fail_msg = msg;
    S->cv_done.notify_all();
}

// One worker per parameter set: each advances its own slots on its own stream, so a set with
// short steps is not held back by a slower one; the sets share the device and the HBM budget.
void server_main(tp_server* S, uint32_t si) {
    tp_server::Set& st = S->sets[si];
    if (tp::bind_device(st.ctx)) {
        fail_all(S, TP_E_DEVICE, "server thread could not select the device");
        return;
    }
    while (!S->stop.load()) {
        int rc = TP_OK;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->failed) return;
            rc = admit(S, si);
        }
        if (!rc) rc = advance(S, st);
        if (!rc && !st.active.empty() && hipStreamSynchronize(tp::context_stream(st.ctx)) != hipSuccess)
            rc = tp::set_error(TP_E_DEVICE, "server stream failed");
        if (rc) {
            fail_all(S, rc, tp_last_error());
            return;
        }
        std::unique_lock<std::mutex> lk(S->mu);
        if (deliver(S, si)) S->cv_done.notify_all();
        S->iterations++;
        if (st.active.empty()) S->cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
}

void destroy_sets(tp_server* S) {
    for (tp_server::Set& st : S->sets) {
        if (!st.ctx) continue;
        tp::bind_device(st.ctx);
        (void)hipStreamSynchronize(tp::context_stream(st.ctx));  // teardown: best effort
        for (void* p : {(void*)st.d_lwe, (void*)st.d_state, (void*)st.d_ext, (void*)st.d_out, (void*)st.d_tv,
                        (void*)st.d_ctrl})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)st.h_ctrl, (void*)st.h_in, (void*)st.h_out})
            if (p) (void)hipHostFree(p);
    }
}

}  // namespace

extern "C" {

int tp_server_create(tp_server** out, tp_context* const* ctxs, uint32_t nsets, const tp_server_config* cfg) {
    if (!out || !ctxs || nsets == 0 || !cfg) return tp::set_error(TP_E_INVALID, "bad server arguments");
    *out = nullptr;
    for (uint32_t si = 0; si < nsets; ++si) {
        if (!ctxs[si]) return tp::set_error(TP_E_INVALID, "null context");
        tp_params p;
        tp::context_params(ctxs[si], &p);
        if (!p.ks_out_dim) return tp::set_error(TP_E_INVALID, "every parameter set needs a key switch");
        if (!tp::context_ready(ctxs[si])) return tp::set_error(TP_E_STATE, "parameter set keys not imported");
    }
    int rc = tp::bind_device(ctxs[0]);
    if (rc) return rc;
    size_t budget = cfg->hbm_budget_bytes;
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        SRV_HIP(hipMemGetInfo(&free_b, &total_b));
        budget = free_b / 10 * 9;
    }
    auto S = std::make_unique<tp_server>();
    S->chunk_steps = cfg->chunk_steps ? cfg->chunk_steps : 64;
    std::vector<size_t> slot_bytes(nsets);
    std::vector<uint32_t> caps(nsets);
    S->sets.resize(nsets);
    for (uint32_t si = 0; si < nsets && !rc; ++si) {
        tp_server::Set& st = S->sets[si];
        tp_params p;
        tp::context_params(ctxs[si], &p);
        st.ctx = ctxs[si];
        st.N = p.ring_dim;
        st.n = p.lwe_dim;
        st.W = p.ks_out_dim + 1;
        st.total_steps = p.method == TP_METHOD_BMMP ? p.lwe_dim / 2 : p.lwe_dim;
        // HBM per slot: input, accumulator, extracted LWE, output, test-vector row, control words
        slot_bytes[si] = ((size_t)(st.n + 1) + 2 * st.N + (st.N + 1) + st.W + 2 * st.N) * 8 + 16;
        size_t cap = cfg->max_slots_per_set ? cfg->max_slots_per_set : 4096;
        cap = std::min(cap, budget / slot_bytes[si]);
        if (cap == 0) {
            rc = tp::set_error(TP_E_NOMEM, "HBM budget smaller than one slot");
            break;
        }
        st.capacity = caps[si] = (uint32_t)cap;
        if (tp::bind_device(st.ctx) || hipMalloc((void**)&st.d_lwe, cap * (st.n + 1) * 8) != hipSuccess ||
            hipMalloc((void**)&st.d_state, cap * 2 * st.N * 8) != hipSuccess ||
            hipMalloc((void**)&st.d_ext, cap * (st.N + 1) * 8) != hipSuccess ||
            hipMalloc((void**)&st.d_out, cap * st.W * 8) != hipSuccess ||
            hipMalloc((void**)&st.d_tv, cap * 2 * st.N * 8) != hipSuccess ||
            hipMalloc((void**)&st.d_ctrl, cap * 4 * 4) != hipSuccess) {
            rc = tp::set_error(TP_E_NOMEM, "slot pool allocation failed");
            break;
        }
        if (hipHostMalloc((void**)&st.h_ctrl, cap * 4 * 4, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&st.h_in, cap * (st.n + 1) * 8, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&st.h_out, cap * st.W * 8, hipHostMallocDefault) != hipSuccess) {
            rc = tp::set_error(TP_E_NOMEM, "pinned staging allocation failed");
            break;
        }
        std::memset(st.h_ctrl, 0, cap * 4 * 4);
        st.step.assign(cap, 0);
        st.slot_req.assign(cap, 0);
        st.slot_idx.assign(cap, 0);
        for (int r = (int)cap - 1; r >= 0; --r) st.tv_free.push_back(r);
    }
    if (!rc) rc = tp_sched_create(&S->sched, nsets, slot_bytes.data(), budget, caps.data());
    if (rc) {
        destroy_sets(S.get());
        return rc;
    }
    tp_server* raw = S.release();
    for (uint32_t si = 0; si < nsets; ++si) raw->workers.emplace_back(server_main, raw, si);
    *out = raw;
    return TP_OK;
}

int tp_server_destroy(tp_server* S) {
    if (!S) return TP_OK;
    S->stop.store(true);
    S->cv_work.notify_all();
    for (std::thread& w : S->workers)
        if (w.joinable()) w.join();
    destroy_sets(S);
    tp_sched_destroy(S->sched);
    delete S;
    return TP_OK;
}

int tp_server_submit(tp_server* S, uint32_t set, size_t count, const uint64_t* lwes, const uint64_t* tv,
                     uint64_t* out, uint64_t* id) {
    if (!S || !lwes || !tv || !out || !id) return tp::set_error(TP_E_INVALID, "null argument");
    if (set >= S->sets.size()) return tp::set_error(TP_E_INVALID, "unknown parameter set");
    const tp_server::Set& st = S->sets[set];
    for (size_t i = 0; i < count * (st.n + 1); ++i)
        if (lwes[i] >= 2 * (uint64_t)st.N) return tp::set_error(TP_E_INVALID, "lwe entries must lie in [0, 2N)");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->failed) return tp::set_error(S->failed, S->fail_msg.c_str());
    uint64_t rid = 0;
    int rc = tp_sched_submit(S->sched, set, count, &rid);
    if (rc) return rc;
    auto r = std::make_unique<tp_server::Req>();
    r->id = rid;
    r->set = set;
    r->count = count;
    r->lwes = lwes;
    r->out = out;
    r->tv.assign(tv, tv + 2 * st.N);
    r->t_submit = Clock::now();
    S->reqs[rid] = std::move(r);
    *id = rid;
    S->cv_work.notify_all();
    return TP_OK;
}

int tp_server_poll(tp_server* S, uint64_t id, int* done) {
    if (!S || !done) return tp::set_error(TP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    if (S->failed && !it->second->done) return tp::set_error(S->failed, S->fail_msg.c_str());
    *done = it->second->done;
    return TP_OK;
}

int tp_server_wait(tp_server* S, uint64_t id, double timeout_s) {
    if (!S) return tp::set_error(TP_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    tp_server::Req* r = it->second.get();
    const auto deadline =
        Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (!r->done) {
        if (S->failed) return tp::set_error(S->failed, S->fail_msg.c_str());
        if (S->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && !r->done)
            return S->failed ? tp::set_error(S->failed, S->fail_msg.c_str())
                             : tp::set_error(TP_E_TIMEOUT, "request not complete before the timeout");
    }
    return TP_OK;
}

int tp_server_cancel(tp_server* S, uint64_t id) {
    if (!S) return tp::set_error(TP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    tp_server::Req& r = *it->second;
    if (r.done) return TP_OK;
    tp_sched_cancel(S->sched, id);
    r.cancelled = true;
    if (r.live_slots == 0) {  // nothing of it on the device: complete now
        r.done = true;
        r.t_done = Clock::now();
        free_tv_row(S->sets[r.set], r);
        S->cv_done.notify_all();
    }
    return TP_OK;
}

int tp_server_release(tp_server* S, uint64_t id) {
    if (!S) return tp::set_error(TP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    if (!it->second->done) return tp::set_error(TP_E_STATE, "request still running");
    tp_sched_forget(S->sched, id);
    S->reqs.erase(it);
    return TP_OK;
}

int tp_server_times(tp_server* S, uint64_t id, double* submitted, double* completed) {
    if (!S) return tp::set_error(TP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return tp::set_error(TP_E_INVALID, "unknown request");
    if (submitted) *submitted = seconds(it->second->t_submit);
    if (completed) *completed = it->second->done ? seconds(it->second->t_done) : -1.0;
    return TP_OK;
}

int tp_server_stats(tp_server* S, uint64_t* iterations, uint64_t* launches, uint64_t* ciphertexts) {
    if (!S) return tp::set_error(TP_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (iterations) *iterations = S->iterations;
    if (launches) *launches = S->launches.load();
    if (ciphertexts) *ciphertexts = S->ciphertexts;
    return TP_OK;
}

}  // extern "C"
