// serve.cpp -- the request loop: continuous batching on one GPU.
//
// Slots: max_inflight worker threads, each owning a full pcf_ctx (its own
// default-priority HIP stream and resident device buffers: reduction
// partials, counters, the LSM path store and its captured graphs, cached
// basket factors).  Work issued by different slots overlaps on the device.
//
// Memory: every slot is charged for the request-sized device buffers it may
// hold -- path stores and per-path state, reduction partials, binomial
// tables -- and the charges of all slots never exceed mem_budget_bytes.  An
// idle slot is charged what it keeps from its last request (measured); a
// busy slot what it kept plus an upper bound of what the unit can add: for
// the American the growth of the kept path store and state it reuses (the
// buffers only grow) plus its other buffers; the tables for the binomial;
// partials otherwise.  A free slot takes
// the oldest queued request whose charge fits (smaller requests may pass one
// that does not fit yet; the oldest American request passes all others while
// fewer than two slots hold one), first by dropping its own kept buffers, then by
// freeing idle slots' kept buffers, least recently used first; it also takes
// every further queued request with the same batch key (single-launch kinds
// only: European, Asian, basket) up to kMaxBatch, and prices them in one
// batched launch.  Requests that arrive while others run join the next
// launch of a free slot.  The per-slot buffers that do not grow with a
// request (job table, arrival counters, results, basket factor slots) are
// outside the budget; pcf_server_stats reports them.
//
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pcf_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Item {
    uint64_t ticket;
    pcf_request req;
    uint64_t key;       // 0: runs alone
    int64_t footprint;  // device bytes while in flight
    double submit_ms, start_ms;
};

struct Finished {
    pcf_result res;
    int status;
    pcf_request_timing timing;
};

}  // namespace

struct pcf_server {
    pcf_ctx* front = nullptr;  // validation / error messages for submit
    std::vector<pcf_ctx*> slots;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::mutex front_mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Item> queue;
    std::unordered_map<uint64_t, Finished> done;
    std::unordered_map<uint64_t, int> live;  // submitted, not yet collected
    int64_t budget = 0;
    // Per slot: request-sized bytes it holds (measured when it went idle),
    // its charge against the budget, busy flag and last use (for eviction).
    std::vector<int64_t> held, charge;
    std::vector<char> busy;
    std::vector<char> busy_amer;  // the slot's unit is an American request
    std::vector<int64_t> kept_store, kept_state;  // American path store / state bytes among held
    std::vector<uint64_t> last_used;
    uint64_t use_clock = 0;
    int64_t charged = 0;  // sum of charge
    pcf_server_stats st{};
    bool stop = false;
    uint64_t next_ticket = 1;
    clk::time_point t0;
    int device = 0;
    std::string last_error;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
};

namespace {

// American units a free slot takes ahead of older requests (see take_unit).
constexpr int kAmerAhead = 2;

int64_t footprint_of(const pcf_request& q, int64_t nblk_hint) {
    switch (q.kind) {
        case PCF_KIND_AMER:
            return pcf_amer_footprint(q.N, q.M, q.opts.precision, q.opts.rng, q.opts.ref_exact);
        case PCF_KIND_BINOM:
            // chain: glibc log table + prefix table, (N+1) doubles each; double-
            // double: log k! table, (N+1) x 16 B; both: one partial per 256 terms
            return (q.opts.ref_exact ? 2 * (q.N + 1) * 8 : (q.N + 1) * 16) + ((q.N + 256) / 256 + 1) * 16 + 4096;
        default:
            return nblk_hint * 4 * 16 + 4096;
    }
}

// Request-sized device bytes a context holds (the buffers that grow with the
// requests it runs).
int64_t request_bytes(pcf_ctx* ctx) {
    int64_t b = 0;
    for (const DevBuf* d : {&ctx->partials, &ctx->normals, &ctx->logtab, &ctx->prefix, &ctx->logfact,
                            &ctx->lsm_store, &ctx->lsm_state, &ctx->lsm_coef, &ctx->scratch, &ctx->lattice})
        b += (int64_t)d->bytes;
    return b;
}

// Frees a context's request-sized buffers (the context must be idle).
void release_request_buffers(pcf_ctx* ctx) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pcf_lsm_release(ctx);
    for (DevBuf* d : {&ctx->partials, &ctx->normals, &ctx->logtab, &ctx->prefix, &ctx->logfact, &ctx->lsm_store,
                      &ctx->lsm_state, &ctx->lsm_coef, &ctx->scratch, &ctx->lattice})
        d->release();
    ctx->logtab_n = ctx->prefix_n = ctx->logfact_n = -1;
    ctx->lattice_key[0] = ctx->lattice_key[1] = ctx->lattice_key[2] = 0;
    ctx->lattice_ptr = nullptr;
}

void set_charge(pcf_server* srv, int slot, int64_t c) {
    srv->charged += c - srv->charge[slot];
    srv->charge[slot] = c;
    if (srv->charged > srv->st.peak_charged_bytes) srv->st.peak_charged_bytes = srv->charged;
}

// Runs one unit (items sharing a batch key, or one item) on a slot context.
void run_unit(pcf_server* srv, int slot, pcf_ctx* ctx, std::vector<Item>& unit, std::vector<Finished>& fin) {
    const int n = (int)unit.size();
    fin.assign(n, Finished{});
    std::vector<const pcf_request*> reqs(n);
    for (int i = 0; i < n; ++i) reqs[i] = &unit[i].req;
    std::vector<pcf_result> res(n);
    std::vector<int> status(n, PCF_OK);
    int st = PCF_OK;
    const int kind = unit[0].req.kind;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        if (kind == PCF_KIND_EUR) st = pcf_eur_run(ctx, n, reqs.data(), res.data(), status.data());
        else if (kind == PCF_KIND_ASIA) st = pcf_asia_run(ctx, n, reqs.data(), res.data(), status.data());
        else if (kind == PCF_KIND_BASKET) st = pcf_basket_run(ctx, n, reqs.data(), res.data(), status.data());
        // batched requests asking for the FP64 comparison get their FP64 run here
        // (an American request does it inside pcf_mc_amer)
        if (st == PCF_OK && kind != PCF_KIND_AMER && kind != PCF_KIND_BINOM)
            for (int i = 0; i < n; ++i)
                if (status[i] == PCF_OK) status[i] = pcf_apply_fp64_check(ctx, unit[i].req, &res[i]);
    }
    if (kind == PCF_KIND_AMER || kind == PCF_KIND_BINOM) {
        for (int i = 0; i < n; ++i) {
            const pcf_request& q = unit[i].req;
            status[i] = kind == PCF_KIND_AMER
                            ? pcf_mc_amer(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.M, q.payoff_fun, &q.opts, &res[i])
                            : pcf_binom_embar(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun, &q.opts, &res[i]);
        }
    }
    const double t = srv->now_ms();
    bool failed = st != PCF_OK;
    for (int i = 0; i < n; ++i) failed |= status[i] != PCF_OK;
    if (failed) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->last_error = ctx->last_error;
    }
    for (int i = 0; i < n; ++i) {
        fin[i].res = res[i];
        fin[i].status = st ? st : status[i];
        fin[i].timing.submit_ms = unit[i].submit_ms;
        fin[i].timing.start_ms = unit[i].start_ms;
        fin[i].timing.done_ms = t;
        fin[i].timing.latency_ms = t - unit[i].submit_ms;
    }
}

// Picks the next unit for `slot` under srv->mu and charges it; returns false
// if nothing is admissible.  *drop_own: the slot must free its own kept
// buffers before running the unit (its charge no longer includes them).
// Idle slots' kept buffers freed here to make room are freed at once.
bool take_unit(pcf_server* srv, int slot, std::vector<Item>& unit, bool* drop_own) {
    unit.clear();
    *drop_own = false;
    auto& q = srv->queue;
    const int64_t others = srv->charged - srv->charge[slot];
    const int64_t own = srv->held[slot];
    // Kept buffers of other idle slots that may be freed, least recently used first.
    std::vector<int> idle;
    int64_t evictable = 0;
    for (int s = 0; s < (int)srv->slots.size(); ++s)
        if (s != slot && !srv->busy[s] && srv->held[s] > 0) {
            idle.push_back(s);
            evictable += srv->held[s];
        }
    // Candidates in queue order, except that the oldest American request comes
    // first while fewer than kAmerAhead slots hold one.  American requests run
    // one after another on the device's persistent stream (their grids need
    // every CU); keeping the next one enqueued behind the running one keeps
    // that chain, the longest in a mixed load, from waiting on the host.
    std::vector<size_t> order;
    order.reserve(q.size());
    int amer_busy = 0;
    for (size_t s = 0; s < srv->slots.size(); ++s) amer_busy += srv->busy_amer[s] ? 1 : 0;
    size_t first_amer = q.size();
    if (amer_busy < kAmerAhead)
        for (size_t i = 0; i < q.size(); ++i)
            if (q[i].req.kind == PCF_KIND_AMER) {
                first_amer = i;
                order.push_back(i);
                break;
            }
    for (size_t i = 0; i < q.size(); ++i)
        if (i != first_amer) order.push_back(i);
    for (size_t i : order) {
        // An American request reuses the path store and state the slot keeps
        // (its buffers only grow): kept, it adds only the growth beyond them.
        const int64_t fp_drop = q[i].footprint;
        int64_t fp_keep = fp_drop;
        if (q[i].req.kind == PCF_KIND_AMER) {
            const pcf_request& r = q[i].req;
            const PcfAmerBytes b = pcf_amer_bytes(r.N, r.M, r.opts.precision, r.opts.rng, r.opts.ref_exact);
            fp_keep = std::max<int64_t>(0, b.store - srv->kept_store[slot]) +
                      std::max<int64_t>(0, b.state - srv->kept_state[slot]) + b.rest;
        }
        // Keep the slot's own buffers if that fits, after freeing idle slots'
        // if need be; else drop them; else the request waits.
        bool drop = false;
        int64_t fp = fp_keep;
        if (others + own + fp_keep > srv->budget && others - evictable + own + fp_keep > srv->budget) {
            if (others - evictable + fp_drop > srv->budget) continue;
            drop = own > 0;
            fp = fp_drop;
        }
        int64_t room = srv->budget - others - (drop ? 0 : own);
        if (fp > room) {
            std::sort(idle.begin(), idle.end(),
                      [&](int a, int b) { return srv->last_used[a] < srv->last_used[b]; });
            for (int s : idle) {
                if (fp <= room) break;
                release_request_buffers(srv->slots[s]);
                room += srv->held[s];
                srv->held[s] = 0;
                srv->kept_store[s] = srv->kept_state[s] = 0;
                set_charge(srv, s, 0);
                ++srv->st.evictions;
            }
        }
        Item head = q[i];
        q.erase(q.begin() + (long)i);
        int64_t total = fp;
        unit.push_back(head);
        if (head.key != 0) {
            for (size_t j = i; j < q.size() && (int)unit.size() < kMaxBatch;) {
                if (q[j].key == head.key && total + q[j].footprint <= room) {
                    total += q[j].footprint;
                    unit.push_back(q[j]);
                    q.erase(q.begin() + (long)j);
                } else {
                    ++j;
                }
            }
        }
        *drop_own = drop;
        if (drop) {
            srv->held[slot] = 0;
            srv->kept_store[slot] = srv->kept_state[slot] = 0;
            ++srv->st.evictions;
        }
        srv->busy[slot] = 1;
        srv->busy_amer[slot] = head.req.kind == PCF_KIND_AMER;
        set_charge(srv, slot, (drop ? 0 : own) + total);
        const double t = srv->now_ms();
        for (auto& it : unit) it.start_ms = t;
        return true;
    }
    return false;
}

void worker_main(pcf_server* srv, int slot) {
    pcf_ctx* ctx = srv->slots[slot];
    (void)hipSetDevice(srv->device);
    std::vector<Item> unit;
    std::vector<Finished> fin;
    for (;;) {
        bool drop_own = false;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            // Queued work is drained before a stop takes effect.
            srv->cv_work.wait(lk, [&] {
                return take_unit(srv, slot, unit, &drop_own) || (srv->stop && srv->queue.empty());
            });
            if (unit.empty()) return;
        }
        if (drop_own) release_request_buffers(ctx);
        run_unit(srv, slot, ctx, unit, fin);
        int64_t now_held, now_store, now_state;
        {
            std::lock_guard<std::recursive_mutex> lk(ctx->mu);
            now_held = request_bytes(ctx);
            now_store = (int64_t)ctx->lsm_store.bytes;
            now_state = (int64_t)ctx->lsm_state.bytes;
        }
        bool over = false;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            // The charge bounded what the unit could hold; anything above it
            // is an estimate error, counted and then freed.
            if (now_held > srv->charge[slot]) {
                ++srv->st.underestimates;
                over = true;
            }
            for (size_t i = 0; i < unit.size(); ++i) srv->done[unit[i].ticket] = fin[i];
            srv->st.completed += (int64_t)unit.size();
        }
        if (over) {
            release_request_buffers(ctx);
            now_held = now_store = now_state = 0;
        }
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->held[slot] = now_held;
            srv->kept_store[slot] = now_store;
            srv->kept_state[slot] = now_state;
            set_charge(srv, slot, now_held);
            srv->busy[slot] = 0;
            srv->busy_amer[slot] = 0;
            srv->last_used[slot] = ++srv->use_clock;
            int64_t sum = 0;
            for (int64_t h : srv->held) sum += h;
            // busy slots count what they held before their unit: a lower
            // bound of their current use; the charge is the upper bound
            if (sum > srv->st.peak_held_bytes) srv->st.peak_held_bytes = sum;
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();  // budget freed
        unit.clear();
    }
}

}  // namespace

extern "C" int pcf_server_create(pcf_ctx* ctx, const pcf_server_config* cfg, pcf_server** out) {
    if (!ctx || !out) return PCF_ERR_INVALID_ARG;
    *out = nullptr;
    pcf_server_config c{};
    if (cfg) c = *cfg;
    const int nslots = c.max_inflight > 0 ? std::min(c.max_inflight, 32) : 8;
    auto* srv = new pcf_server();
    srv->device = ctx->device;
    srv->budget = c.mem_budget_bytes > 0 ? c.mem_budget_bytes : ((int64_t)8 << 30);
    srv->held.assign(nslots, 0);
    srv->charge.assign(nslots, 0);
    srv->busy.assign(nslots, 0);
    srv->busy_amer.assign(nslots, 0);
    srv->kept_store.assign(nslots, 0);
    srv->kept_state.assign(nslots, 0);
    srv->last_used.assign(nslots, 0);
    srv->st.budget_bytes = srv->budget;
    srv->st.slots = nslots;
    srv->t0 = clk::now();
    int st = pcf_ctx_create(ctx->device, &srv->front);
    if (st != PCF_OK) {
        delete srv;
        return st;
    }
    for (int i = 0; i < nslots; ++i) {
        pcf_ctx* s = nullptr;
        st = pcf_ctx_create(ctx->device, &s);
        if (st != PCF_OK) {
            srv->stop = true;
            pcf_server_destroy(srv);
            return st;
        }
        srv->slots.push_back(s);
    }
    for (int i = 0; i < nslots; ++i) srv->workers.emplace_back(worker_main, srv, i);
    *out = srv;
    return PCF_OK;
}

extern "C" int pcf_server_submit(pcf_server* srv, const pcf_request* req, uint64_t* ticket) {
    if (!srv || !req || !ticket) return PCF_ERR_INVALID_ARG;
    Item it{};
    it.req = *req;
    if (it.req.opts.rng != PCF_RNG_PHILOX) {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->last_error = "served requests draw Philox normals (rng must be PCF_RNG_PHILOX)";
        return PCF_ERR_INVALID_ARG;
    }
    int status = PCF_OK;
    uint64_t key = 0;
    int64_t nblk_hint = 1;
    {
        // Validate and form the batch key up front: a bad request is refused here
        // and never reaches a batch.
        std::lock_guard<std::mutex> fl(srv->front_mu);
        pcf_ctx* f = srv->front;
        switch (it.req.kind) {
            case PCF_KIND_EUR: key = pcf_eur_batch_key(f, it.req, &status); break;
            case PCF_KIND_ASIA: key = pcf_asia_batch_key(f, it.req, &status); break;
            case PCF_KIND_BASKET: key = pcf_basket_batch_key(f, it.req, &status); break;
            case PCF_KIND_AMER:
                status = pcf_check_common(f, it.req.S0, it.req.E, it.req.r, it.req.sigma, it.req.T, it.req.N,
                                          it.req.payoff_fun);
                if (!status) status = pcf_check_opts(f, &it.req.opts);
                if (!status && (it.req.N % 2 || it.req.M < 1 || !(it.req.E > 0)))
                    status = pcf_fail(f, PCF_ERR_INVALID_ARG, "American: N even, M >= 1, E > 0 required");
                break;
            case PCF_KIND_BINOM:
                status = pcf_check_common(f, it.req.S0, it.req.E, it.req.r, it.req.sigma, it.req.T, it.req.N,
                                          it.req.payoff_fun);
                break;
            default:
                status = pcf_fail(f, PCF_ERR_INVALID_ARG, "unknown request kind %d", it.req.kind);
        }
        if (status) {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->last_error = f->last_error;
            return status;
        }
        nblk_hint = pcf_blocks_for(it.req.N, 64, 1);
    }
    it.key = key;
    it.footprint = footprint_of(it.req, nblk_hint);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return PCF_ERR_SHUTDOWN;
        if (it.footprint > srv->budget) {
            srv->last_error = "request needs more device memory than the server budget";
            return PCF_ERR_OUT_OF_MEMORY;
        }
        it.ticket = srv->next_ticket++;
        it.submit_ms = srv->now_ms();
        srv->live[it.ticket] = 1;
        srv->queue.push_back(it);
        *ticket = it.ticket;
    }
    srv->cv_work.notify_one();
    return PCF_OK;
}

extern "C" int pcf_server_wait(pcf_server* srv, uint64_t ticket, pcf_result* res, pcf_request_timing* timing) {
    if (!srv) return PCF_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (!srv->live.count(ticket)) return PCF_ERR_NOT_FOUND;
    srv->cv_done.wait(lk, [&] { return srv->done.count(ticket) > 0; });
    Finished f = srv->done[ticket];
    srv->done.erase(ticket);
    srv->live.erase(ticket);
    if (res) *res = f.res;
    if (timing) *timing = f.timing;
    return f.status;
}

extern "C" int pcf_server_destroy(pcf_server* srv) {
    if (!srv) return PCF_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& w : srv->workers)
        if (w.joinable()) w.join();
    for (pcf_ctx* c : srv->slots) pcf_ctx_destroy(c);
    if (srv->front) pcf_ctx_destroy(srv->front);
    delete srv;
    return PCF_OK;
}

extern "C" const char* pcf_server_last_error(pcf_server* srv) {
    if (!srv) return "null server";
    std::lock_guard<std::mutex> lk(srv->mu);
    return srv->last_error.c_str();
}

extern "C" int pcf_server_get_stats(pcf_server* srv, pcf_server_stats* out) {
    if (!srv || !out) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->st;
    out->charged_bytes = srv->charged;
    int64_t h = 0;
    for (int64_t v : srv->held) h += v;
    out->held_bytes = h;
    // arrival counters, result slots, job table, basket factor slots
    out->fixed_bytes_per_slot = (int64_t)(kCounterSlots * sizeof(unsigned int) + kOutDoubles * sizeof(double) +
                                          kMaxBatch * kJobBytes) + kBasketConstBytes;
    return PCF_OK;
}
