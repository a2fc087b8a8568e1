// serve.cpp — continuous-batching request loop for the VIA-C and VIA-B
// answer paths.
//
// The reference has no server process (it is a library answering one query
// per call); this loop is the serving layer around it. Clients are admitted
// against an HBM budget for their evaluation keys and stay resident. Requests
// arrive at any time: VIA-C queries (Server::answer) and VIA-B batches of T
// queries (Server::answer_batch::<T>, against a second database packed at the
// VIA-B record degree). A worker thread takes work as soon as one of `slots`
// device contexts is free: a VIA-B request runs as one launch sequence of its
// own, VIA-C queries pending at that moment are gathered into one batch of up
// to max_batch. Work is launched asynchronously on the slot's stream and
// retired when its completion event fires, so launches overlap on the GPU.
// Every answer is bit-identical to answering its request alone.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_internal.h"

using namespace via;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    uint64_t client;
    uint32_t T = 0;  // 0: VIA-C query; else a VIA-B batch of T queries
    int pool = -1;   // VIA-C query: its entry of the pinned query pool, or -1 (then in lwes)
    std::vector<uint64_t> lwes;
    Clock::time_point t_submit;
};

struct Result {
    bool done = false;
    via_status status = VIA_OK;
    std::vector<uint64_t> mask, body;
    double latency_s = 0;
};

struct Slot {
    via_ctx *ctx = nullptr;
    uint64_t *h_q = nullptr;   // pinned query staging (q_cap words)
    uint64_t *h_a = nullptr;   // pinned answers, max_batch x 2 x n2
    uint64_t *d_q = nullptr;   // device queries
    uint64_t *d_a = nullptr;   // device answers
    hipEvent_t done = nullptr;      // the batch's answers are on the host
    hipEvent_t computed = nullptr;  // its kernels have run (the answers' copy may not have)
    std::vector<Request> batch;
    bool busy = false;
};

}  // namespace

struct via_server {
    via_params pp;       // VIA-C database parameters
    via_params kp;       // client key parameters (record degree of the VIA-B database when present)
    const via_db *db = nullptr;
    const via_db *db_b = nullptr;
    int device = 0;
    size_t key_budget = 0, key_bytes = 0;
    uint32_t max_batch = 1;
    size_t n_lwe = 0, lwe_words = 0;      // one VIA-C query
    size_t n_lwe_b = 0, lwe_words_b = 0;  // one query of a VIA-B batch
    size_t q_cap = 0;                     // staging words per slot
    // Pinned pool of VIA-C queries: submit copies (and checks) a query into a
    // free entry in one pass, and a batch's queries go to the device by DMA
    // from there, without a copy on the worker thread. A submit that finds
    // the pool empty keeps its query in the request (copied at launch).
    uint64_t *pool = nullptr;
    std::vector<int> pool_free;
    std::vector<Slot> slots;
    // The slot whose batch was launched last: the next batch's kernels wait
    // for its kernels (its query copy does not), so batches run back to back
    // on the device instead of contending for it, while copies overlap.
    Slot *last = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::unordered_map<uint64_t, Result> results;
    std::unordered_map<uint64_t, via_keys *> clients;
    std::unordered_map<uint64_t, int> client_inflight;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t answered = 0, batches = 0, max_seen_batch = 0, records = 0, answered_b = 0;
    std::vector<double> latencies, latencies_b;
};

namespace {

// Orders slot sl's kernels after those of the batch launched before it.
via_status chain_after_last(via_server *s, Slot &sl) {
    if (s->last && s->last != &sl && s->last->busy &&
        hipStreamWaitEvent(sl.ctx->stream, s->last->computed, 0) != hipSuccess)
        return VIA_E_DEVICE;
    s->last = &sl;
    return VIA_OK;
}

// Launch one batch on a free slot (caller holds no lock; the slot is owned by
// the worker thread).
via_status launch_vb(via_server *s, Slot &sl) {
    const Request &r = sl.batch[0];
    const via_keys *k;
    {
        std::lock_guard<std::mutex> g(s->mu);
        k = s->clients.at(r.client);
    }
    const size_t words = (size_t)r.T * s->lwe_words_b;
    memcpy(sl.h_q, r.lwes.data(), sizeof(uint64_t) * words);
    hipStream_t stream = sl.ctx->stream;
    if (hipMemcpyAsync(sl.d_q, sl.h_q, sizeof(uint64_t) * words, hipMemcpyHostToDevice, stream) != hipSuccess)
        return VIA_E_DEVICE;
    via_status st = chain_after_last(s, sl);
    if (st) return st;
    st = answer_vb_impl(sl.ctx, s->db_b, k, (int)r.T, sl.d_q, s->n_lwe_b, sl.d_a, sl.d_a + s->pp.n2, nullptr);
    if (st) return st;
    if (hipEventRecord(sl.computed, stream) != hipSuccess) return VIA_E_DEVICE;
    if (hipMemcpyAsync(sl.h_a, sl.d_a, sizeof(uint64_t) * 2 * s->pp.n2, hipMemcpyDeviceToHost, stream) != hipSuccess)
        return VIA_E_DEVICE;
    if (hipEventRecord(sl.done, stream) != hipSuccess) return VIA_E_DEVICE;
    return VIA_OK;
}

via_status launch_batch(via_server *s, Slot &sl) {
    if (sl.batch.size() == 1 && sl.batch[0].T > 0) return launch_vb(s, sl);
    const int Q = (int)sl.batch.size();
    via_status st = ensure_workspace(sl.ctx, &s->pp, (size_t)Q, s->n_lwe);
    if (st) return st;
    std::vector<const via_keys *> keys(Q);
    std::vector<const uint64_t *> dq(Q);
    std::vector<uint64_t *> dm(Q), dbd(Q);
    hipStream_t stream = sl.ctx->stream;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (int q = 0; q < Q; q++) keys[q] = s->clients.at(sl.batch[q].client);
    }
    const size_t qbytes = sizeof(uint64_t) * s->lwe_words;
    for (int q = 0; q < Q; q++) {
        const Request &r = sl.batch[q];
        dq[q] = sl.d_q + (size_t)q * s->lwe_words;
        dm[q] = sl.d_a + (size_t)q * 2 * s->pp.n2;
        dbd[q] = dm[q] + s->pp.n2;
        const uint64_t *src = r.pool >= 0 ? s->pool + (size_t)r.pool * s->lwe_words : sl.h_q + (size_t)q * s->lwe_words;
        if (r.pool < 0) memcpy(sl.h_q + (size_t)q * s->lwe_words, r.lwes.data(), qbytes);
        if (hipMemcpyAsync(sl.d_q + (size_t)q * s->lwe_words, src, qbytes, hipMemcpyHostToDevice, stream) !=
            hipSuccess)
            return VIA_E_DEVICE;
    }
    if ((st = chain_after_last(s, sl))) return st;
    st = answer_batch_impl(sl.ctx, s->db, Q, keys.data(), dq.data(), s->n_lwe, dm.data(), dbd.data(), nullptr);
    if (st) return st;
    if (hipEventRecord(sl.computed, stream) != hipSuccess) return VIA_E_DEVICE;
    if (hipMemcpyAsync(sl.h_a, sl.d_a, sizeof(uint64_t) * 2 * s->pp.n2 * Q, hipMemcpyDeviceToHost, stream) !=
        hipSuccess)
        return VIA_E_DEVICE;
    if (hipEventRecord(sl.done, stream) != hipSuccess) return VIA_E_DEVICE;
    return VIA_OK;
}

void finish_batch(via_server *s, Slot &sl, via_status st) {
    const auto now = Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    const int Q = (int)sl.batch.size();
    for (int q = 0; q < Q; q++) {
        const Request &r = sl.batch[q];
        Result &res = s->results[r.ticket];
        res.status = st;
        if (st == VIA_OK) {
            const uint64_t *a = sl.h_a + (size_t)q * 2 * s->pp.n2;
            res.mask.assign(a, a + s->pp.n2);
            res.body.assign(a + s->pp.n2, a + 2 * s->pp.n2);
        }
        res.latency_s = std::chrono::duration<double>(now - r.t_submit).count();
        res.done = true;
        s->latencies.push_back(res.latency_s);
        if (r.T) {
            s->latencies_b.push_back(res.latency_s);
            s->answered_b++;
        }
        s->records += r.T ? r.T : 1;
        s->client_inflight[r.client]--;
        if (r.pool >= 0) s->pool_free.push_back(r.pool);  // its DMA preceded the completion event
    }
    s->answered += Q;
    s->batches++;
    if (!(Q == 1 && sl.batch[0].T)) s->max_seen_batch = std::max<uint64_t>(s->max_seen_batch, (uint64_t)Q);
    sl.batch.clear();
    sl.busy = false;
    s->cv_done.notify_all();
}

void worker_loop(via_server *s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        // retire finished batches; any status but "not ready" is a device
        // failure that the batch's waiters receive as VIA_E_DEVICE
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) finish_batch(s, sl, VIA_OK);
            else if (q != hipErrorNotReady) finish_batch(s, sl, VIA_E_DEVICE);
        }
        // launch pending work on free slots
        bool launched = false;
        bool other_busy = false;
        for (Slot &sl : s->slots) other_busy |= sl.busy;
        for (Slot &sl : s->slots) {
            if (sl.busy) continue;
            {
                std::unique_lock<std::mutex> lk(s->mu);
                if (s->pending.empty()) break;
                // while a batch runs, VIA-C queries gather until a full batch
                // is pending (it then queues behind the running one); a VIA-B
                // request, or an idle device, launches at once
                if (other_busy && !s->pending.front().T) {
                    size_t c = 0;
                    for (const Request &r : s->pending) c += r.T ? 0 : 1;
                    if (c < s->max_batch) break;
                }
                if (s->pending.front().T) {
                    // a VIA-B request runs on its own
                    sl.batch.push_back(std::move(s->pending.front()));
                    s->pending.pop_front();
                } else {
                    // gather the pending VIA-C queries, oldest first
                    for (auto it = s->pending.begin(); it != s->pending.end() && sl.batch.size() < s->max_batch;) {
                        if (it->T) {
                            ++it;
                            continue;
                        }
                        sl.batch.push_back(std::move(*it));
                        it = s->pending.erase(it);
                    }
                }
            }
            sl.busy = true;
            via_status st = launch_batch(s, sl);
            if (st != VIA_OK) {
                (void)hipStreamSynchronize(sl.ctx->stream);
                finish_batch(s, sl, st);
            }
            launched = true;
            other_busy = true;
        }
        bool any_busy = false;
        for (Slot &sl : s->slots) any_busy |= sl.busy;
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stop && !any_busy && s->pending.empty()) return;
        if (launched) continue;
        if (any_busy) {
            // poll the in-flight batches; new arrivals wake us early
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        } else {
            s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        }
    }
}

}  // namespace

extern "C" {

via_status via_server_create(int device, const via_params *pp, const via_db *db, const via_db *db_b,
                             size_t key_budget_bytes, uint32_t max_batch, uint32_t slots, via_server **out) {
    if (!pp || !db || !out || max_batch < 1 || slots < 1) return VIA_E_INVALID;
    if (max_batch > (uint32_t)kMaxQ) max_batch = kMaxQ;
    if (!same_client_params(db->pp, *pp) || db->pp.n_rec != pp->n_rec) return VIA_E_DIM;
    // db_b: the same client parameters and grid at a finer record degree
    if (db_b && (!same_client_params(db_b->pp, *pp) || db_b->pp.n_rec >= pp->n_rec)) return VIA_E_DIM;
    if (hipSetDevice(device) != hipSuccess) return VIA_E_DEVICE;
    auto s = std::make_unique<via_server>();
    s->pp = *pp;
    s->kp = db_b ? db_b->pp : *pp;
    s->db = db;
    s->db_b = db_b;
    s->device = device;
    s->key_budget = key_budget_bytes;
    s->max_batch = max_batch;
    auto lg = [](uint32_t x) { return x <= 1 ? 0u : 32u - (uint32_t)__builtin_clz(x - 1); };
    const uint32_t bits = lg(pp->num_rows) + lg(pp->num_cols) + lg(pp->n1 / pp->n_rec);
    s->n_lwe = (size_t)bits * pp->l_query;
    s->lwe_words = s->n_lwe * pp->q1_nl * (pp->n1 + 1);
    s->q_cap = s->lwe_words * max_batch;
    const uint32_t t_b = db_b ? pp->n2 / db_b->pp.n_rec : 0;  // the VIA-B batch size of the record degree
    if (db_b) {
        const uint32_t bits_b = lg(pp->num_rows) + lg(pp->num_cols) + lg(pp->n1 / db_b->pp.n_rec);
        s->n_lwe_b = (size_t)bits_b * pp->l_query;
        s->lwe_words_b = s->n_lwe_b * pp->q1_nl * (pp->n1 + 1);
        s->q_cap = std::max(s->q_cap, s->lwe_words_b * std::min<uint32_t>(t_b, kServeMaxT));
    }
    {
        const int entries = (int)(2 * slots * max_batch + max_batch);
        if (hipHostMalloc((void **)&s->pool, sizeof(uint64_t) * s->lwe_words * entries) != hipSuccess)
            return VIA_E_NOMEM;
        for (int e = entries - 1; e >= 0; e--) s->pool_free.push_back(e);
    }
    s->slots.resize(slots);
    for (Slot &sl : s->slots) {
        via_status st = via_ctx_create(device, &sl.ctx);
        if (st) return st;
        // slots overlap on the device: a cooperative grid could not count on
        // having every CU, so the trees launch level by level
        sl.ctx->allow_persistent = false;
        sl.ctx->step_events = false;  // nothing reads a slot's step times
        if (hipHostMalloc((void **)&sl.h_q, sizeof(uint64_t) * s->q_cap) != hipSuccess ||
            hipHostMalloc((void **)&sl.h_a, sizeof(uint64_t) * 2 * pp->n2 * max_batch) != hipSuccess ||
            hipMalloc((void **)&sl.d_q, sizeof(uint64_t) * s->q_cap) != hipSuccess ||
            hipMalloc((void **)&sl.d_a, sizeof(uint64_t) * 2 * pp->n2 * max_batch) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.computed, hipEventDisableTiming) != hipSuccess)
            return VIA_E_NOMEM;
        st = ensure_workspace(sl.ctx, pp, max_batch, s->n_lwe);
        if (st) return st;
        if (db_b) {
            const uint32_t tmax = std::min<uint32_t>(t_b, kServeMaxT);
            st = ensure_workspace(sl.ctx, &db_b->pp, std::min<uint32_t>(tmax, kMaxQ), s->n_lwe_b);
            if (st) return st;
            st = ensure_repack_workspace(sl.ctx, &db_b->pp, (int)tmax, s->n_lwe_b);
            if (st) return st;
        }
        st = prepare_tables(sl.ctx, pp);
        if (st) return st;
    }
    via_server *raw = s.release();
    raw->worker = std::thread(worker_loop, raw);
    *out = raw;
    return VIA_OK;
}

void via_server_destroy(via_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto &kv : s->clients) via_keys_destroy(kv.second);
    for (Slot &sl : s->slots) {
        if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream);
        (void)hipHostFree(sl.h_q);
        (void)hipHostFree(sl.h_a);
        (void)hipFree(sl.d_q);
        (void)hipFree(sl.d_a);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.computed) (void)hipEventDestroy(sl.computed);
        via_ctx_destroy(sl.ctx);
    }
    (void)hipHostFree(s->pool);
    delete s;
}

via_status via_server_admit(via_server *s, uint64_t client_id, const uint64_t *cascade_key, const uint64_t *conv_key,
                            const uint64_t *rsk) {
    if (!s) return VIA_E_INVALID;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->clients.count(client_id)) return VIA_OK;
    }
    const size_t need = keys_footprint(&s->kp);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->key_bytes + need > s->key_budget) return VIA_E_NOMEM;
        s->key_bytes += need;  // reserve
    }
    via_keys *k = nullptr;
    via_status st = via_keys_create(s->slots[0].ctx, &s->kp, cascade_key, conv_key, rsk, &k);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (st) {
            s->key_bytes -= need;
            return st;
        }
        if (!s->clients.count(client_id)) {
            s->key_bytes += keys_device_bytes(k) - need;
            s->clients[client_id] = k;
            s->client_inflight[client_id] = 0;
            return VIA_OK;
        }
        // a concurrent admit of the same client won: keep its keys
        s->key_bytes -= need;
    }
    via_keys_destroy(k);
    return VIA_OK;
}

via_status via_server_evict(via_server *s, uint64_t client_id) {
    if (!s) return VIA_E_INVALID;
    via_keys *k = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->clients.find(client_id);
        if (it == s->clients.end()) return VIA_E_INVALID;
        // wait until none of its queries is pending or in flight
        s->cv_done.wait(lk, [&] {
            if (s->client_inflight[client_id] > 0) return false;
            for (const Request &r : s->pending)
                if (r.client == client_id) return false;
            return true;
        });
        k = it->second;
        s->key_bytes -= keys_device_bytes(k);
        s->clients.erase(it);
        s->client_inflight.erase(client_id);
    }
    via_keys_destroy(k);
    return VIA_OK;
}

static via_status enqueue(via_server *s, Request &&r, uint64_t client_id, uint64_t *ticket) {
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->clients.count(client_id)) return VIA_E_INVALID;
        r.ticket = s->next_ticket++;
        *ticket = r.ticket;
        s->results[r.ticket] = Result();
        s->client_inflight[client_id]++;
        s->pending.push_back(std::move(r));
    }
    s->cv_work.notify_one();
    return VIA_OK;
}

via_status via_server_submit(via_server *s, uint64_t client_id, const uint64_t *lwes, size_t n_lwe,
                             uint64_t *ticket) {
    if (!s || !lwes || !ticket) return VIA_E_INVALID;
    if (n_lwe != s->n_lwe) return VIA_E_QUERY_LEN;
    Request r;
    r.client = client_id;
    r.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->pool_free.empty()) {
            r.pool = s->pool_free.back();
            s->pool_free.pop_back();
        }
    }
    bool ok;
    if (r.pool >= 0) {
        ok = lwes_copy_canonical(&s->pp, s->pool + (size_t)r.pool * s->lwe_words, lwes, n_lwe);
    } else {
        r.lwes.resize(s->lwe_words);
        ok = lwes_copy_canonical(&s->pp, r.lwes.data(), lwes, n_lwe);
    }
    if (!ok) {
        if (r.pool >= 0) {
            std::lock_guard<std::mutex> g(s->mu);
            s->pool_free.push_back(r.pool);
        }
        return VIA_E_INVALID;
    }
    const int pool = r.pool;
    via_status st = enqueue(s, std::move(r), client_id, ticket);
    if (st && pool >= 0) {  // refused (client not admitted): the entry was not queued
        std::lock_guard<std::mutex> g(s->mu);
        s->pool_free.push_back(pool);
    }
    return st;
}

via_status via_server_submit_batch(via_server *s, uint64_t client_id, uint32_t T, const uint64_t *lwes,
                                   size_t n_lwe, uint64_t *ticket) {
    if (!s || !lwes || !ticket || !s->db_b) return VIA_E_INVALID;
    if ((size_t)T * s->lwe_words_b > s->q_cap || T == 0) return VIA_E_INVALID;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->clients.find(client_id);
        if (it == s->clients.end()) return VIA_E_INVALID;
        via_status st = check_vb(s->db_b, it->second, T);
        if (st) return st;
    }
    if (n_lwe != s->n_lwe_b) return VIA_E_QUERY_LEN;
    if (!lwes_canonical(&s->pp, lwes, (size_t)T * n_lwe)) return VIA_E_INVALID;
    Request r;
    r.client = client_id;
    r.T = T;
    r.lwes.assign(lwes, lwes + (size_t)T * s->lwe_words_b);
    r.t_submit = Clock::now();
    return enqueue(s, std::move(r), client_id, ticket);
}

via_status via_server_poll(via_server *s, uint64_t ticket, int *done) {
    if (!s || !done) return VIA_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return VIA_E_INVALID;
    *done = it->second.done ? 1 : 0;
    return VIA_OK;
}

via_status via_server_wait(via_server *s, uint64_t ticket, uint64_t *mask, uint64_t *body, double *latency_s) {
    if (!s || !mask || !body) return VIA_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return VIA_E_INVALID;
    s->cv_done.wait(lk, [&] { return s->results[ticket].done; });
    Result res = std::move(s->results[ticket]);
    s->results.erase(ticket);
    lk.unlock();
    if (res.status) return res.status;
    memcpy(mask, res.mask.data(), sizeof(uint64_t) * s->pp.n2);
    memcpy(body, res.body.data(), sizeof(uint64_t) * s->pp.n2);
    if (latency_s) *latency_s = res.latency_s;
    return VIA_OK;
}

via_status via_server_get_stats(via_server *s, via_server_stats *out) {
    if (!s || !out) return VIA_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    memset(out, 0, sizeof *out);
    out->answered = s->answered;
    out->batches = s->batches;
    out->max_batch_seen = s->max_seen_batch;
    out->clients = s->clients.size();
    out->key_bytes = s->key_bytes;
    out->pending = s->pending.size();
    out->records = s->records;
    out->answered_b = s->answered_b;
    auto pcts = [](std::vector<double> l, double *p50, double *p99) {
        if (l.empty()) return;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
        *p50 = pct(0.50);
        *p99 = pct(0.99);
    };
    pcts(s->latencies, &out->p50_s, &out->p99_s);
    pcts(s->latencies_b, &out->p50_b_s, &out->p99_b_s);
    return VIA_OK;
}

via_status via_server_reset_stats(via_server *s) {
    if (!s) return VIA_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    s->answered = s->batches = s->max_seen_batch = s->records = s->answered_b = 0;
    s->latencies.clear();
    s->latencies_b.clear();
    return VIA_OK;
}

}  // extern "C"
