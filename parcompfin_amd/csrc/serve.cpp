// serve.cpp -- the request loop: continuous batching on one GPU.
//
// Slots: max_inflight worker threads, each owning a full pcf_ctx (its own
// default-priority HIP stream and resident device buffers: reduction
// partials, counters, the LSM path store and its captured graphs, cached
// basket factors).  Work issued by different slots overlaps on the device.
//
// Admission: a request's device footprint (the full path store for the
// American, reduction partials otherwise) is charged against mem_budget_bytes
// while it is in flight.  A free slot takes the oldest queued request that
// fits the remaining budget (smaller requests may pass one that does not fit
// yet) and, for the single-launch kinds (European, Asian, basket), every
// further queued request with the same batch key -- kernel variant and
// launch shape -- up to kMaxBatch, and prices them in one batched launch.
// Requests that arrive while others run join the next launch of a free slot.
//
// Results do not depend on the mix: a batched launch reduces each request
// over exactly the blocks and per-thread paths its lone launch would use
// (shapes come from the process-wide tuned table), and the Philox stream is
// keyed by (seed, path, step).  The American runs its captured per-class
// graph on the slot; the binomial sum runs alone on its slot.
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
    int64_t budget = 0, in_use = 0;
    // Path-store bytes that slots keep between American requests (per slot,
    // and their total), bounded by the budget as a whole.
    std::vector<int64_t> resident;
    int64_t resident_total = 0;
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

int64_t footprint_of(const pcf_request& q, int64_t nblk_hint) {
    switch (q.kind) {
        case PCF_KIND_AMER:
            return pcf_amer_footprint(q.N, q.M, q.opts.precision, q.opts.rng, q.opts.ref_exact);
        case PCF_KIND_BINOM:
            return (q.N + 1) * 8 + ((q.N / 2 + 256) / 256 + 1) * 16 + 4096;
        default:
            return nblk_hint * 4 * 16 + 4096;
    }
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
    }
    if (kind == PCF_KIND_AMER || kind == PCF_KIND_BINOM) {
        for (int i = 0; i < n; ++i) {
            const pcf_request& q = unit[i].req;
            status[i] = kind == PCF_KIND_AMER
                            ? pcf_mc_amer(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.M, q.payoff_fun, &q.opts, &res[i])
                            : pcf_binom_embar(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun, &q.opts, &res[i]);
        }
        // A slot keeps its path store for the next American request unless
        // the stores kept by all slots would exceed the budget; then this one
        // is freed.  A fixed share per slot (budget / slots) would instead
        // free and reallocate a store on every request once that share is
        // smaller than one store.
        if (kind == PCF_KIND_AMER) {
            const int64_t held = (int64_t)(ctx->lsm_store.bytes + ctx->lsm_state.bytes);
            bool release;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                srv->resident_total += held - srv->resident[slot];
                release = srv->resident_total > srv->budget;
                if (release) srv->resident_total -= held;
                srv->resident[slot] = release ? 0 : held;
            }
            if (release) {
                std::lock_guard<std::recursive_mutex> lk(ctx->mu);
                pcf_lsm_release(ctx);
                ctx->lsm_store.release();
                ctx->lsm_state.release();
            }
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

// Picks the next unit under srv->mu; returns false if nothing is admissible.
bool take_unit(pcf_server* srv, std::vector<Item>& unit) {
    unit.clear();
    auto& q = srv->queue;
    for (size_t i = 0; i < q.size(); ++i) {
        if (q[i].footprint > srv->budget - srv->in_use) continue;
        Item head = q[i];
        q.erase(q.begin() + (long)i);
        int64_t fp = head.footprint;
        unit.push_back(head);
        if (head.key != 0) {
            for (size_t j = i; j < q.size() && (int)unit.size() < kMaxBatch;) {
                if (q[j].key == head.key && fp + q[j].footprint <= srv->budget - srv->in_use) {
                    fp += q[j].footprint;
                    unit.push_back(q[j]);
                    q.erase(q.begin() + (long)j);
                } else {
                    ++j;
                }
            }
        }
        srv->in_use += fp;
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
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            // Queued work is drained before a stop takes effect.
            srv->cv_work.wait(lk, [&] {
                return take_unit(srv, unit) || (srv->stop && srv->queue.empty());
            });
            if (unit.empty()) return;
        }
        run_unit(srv, slot, ctx, unit, fin);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (size_t i = 0; i < unit.size(); ++i) {
                srv->in_use -= unit[i].footprint;
                srv->done[unit[i].ticket] = fin[i];
            }
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
    srv->resident.assign(nslots, 0);
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
