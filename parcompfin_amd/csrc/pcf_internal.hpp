// pcf_internal.hpp -- host-side internals of libpcf: the context (one HIP
// stream, timing events, resident device buffers), error plumbing and the
// launch-shape helpers shared by the product files.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "pcf.h"
#include "pcf_device.hpp"

// Grow-only device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    // Ensures at least n bytes; contents are not preserved.  zero=true clears
    // a newly allocated buffer (once, at allocation).
    hipError_t ensure(size_t n, bool zero = false) {
        if (n <= bytes && ptr) return hipSuccess;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            bytes = 0;
        }
        size_t want = n < 256 ? 256 : n;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = want;
        if (zero) {
            // hipMemset on device memory is asynchronous to the host and runs
            // on the null stream, which the library's non-blocking streams do
            // not wait for: complete it before the buffer is handed out.
            e = hipMemset(ptr, 0, want);
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Requests per batched launch; each owns the pcf::kCounterWords arrival
// counters of one grid reduction and kOutPerJob doubles of results.
constexpr int kMaxBatch = 64;
constexpr int kCounterSlots = kMaxBatch * pcf::kCounterWords;
constexpr int kOutPerJob = 8;
constexpr int kOutDoubles = kMaxBatch * kOutPerJob;
// Bytes of one job record (launch descriptor) staged per batched launch.
constexpr size_t kJobBytes = 512;
// Device-resident basket constants: kMaxBatch slots of 64 KB per context.
constexpr size_t kBasketConstSlot = 64 * 1024;
constexpr int64_t kBasketConstBytes = (int64_t)kBasketConstSlot * kMaxBatch;

// Device-resident basket constants (factor + per-asset terms), LRU slots.
struct BasketCache {
    std::unordered_map<uint64_t, int> where;
    std::vector<uint64_t> key, stamp;
    uint64_t clock = 0;
};

struct pcf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_kbeg = nullptr, ev_kend = nullptr, ev_end = nullptr;
    hipEvent_t ev_ready = nullptr;  // hand-off to another stream (no timing)
    std::string last_error;
    std::recursive_mutex mu;
    int cu_count = 0;
    int64_t total_mem = 0;
    char name[256] = {0};

    DevBuf partials;   // per-block compensated partial sums
    DevBuf counters;   // kCounterSlots uint32 arrival counters, zero at rest
    DevBuf out;        // kOutDoubles final values
    DevBuf normals;    // verification normals copied from the host
    DevBuf logtab;     // binom: glibc log(n), n = 0..N (resident per N)
    int64_t logtab_n = -1;
    DevBuf prefix;     // binom: prefix chain of the descending log sum
    int64_t prefix_n = -1;
    DevBuf logfact;    // binom production: log k! as double-double, k = 0..logfact_n
    int64_t logfact_n = -1;
    DevBuf lsm_store;  // mc_amer path store
    DevBuf lsm_state;  // mc_amer per-path state
    DevBuf lsm_coef;   // mc_amer per-date coefficients / flags
    DevBuf basket_L;   // basket Cholesky factor + per-asset constants
    DevBuf scratch;    // misc (normals dump)
    DevBuf jobs;       // job records of the current batched launch
    DevBuf lattice;    // binomial lattice: u^i, d^k tables + two level rows
    uint64_t lattice_key[3] = {0, 0, 0};  // (u, d, N) the tables hold
    void* lattice_ptr = nullptr;          // where they were written
    BasketCache basket_cache;
    double* h_out = nullptr;  // pinned mirror of `out`
    double* h_out_dev = nullptr;  // h_out as kernels address it (persistent LSM results)
    void* h_jobs = nullptr;   // pinned staging of `jobs`
};

// Releases the LSM graphs captured for a context (lsm.hip).
void pcf_lsm_release(pcf_ctx* ctx);

// Error helpers ------------------------------------------------------------
int pcf_fail(pcf_ctx* ctx, int code, const char* fmt, ...);
int pcf_hip_fail(pcf_ctx* ctx, hipError_t e, const char* what);

#define PCF_HIP_CHECK(ctx, expr)                                      \
    do {                                                              \
        hipError_t _e = (expr);                                       \
        if (_e != hipSuccess) return pcf_hip_fail((ctx), _e, #expr);  \
    } while (0)

// Common validation of the reference's scalar arguments.
int pcf_check_common(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                     int64_t N, double payoff_fun);
int pcf_check_opts(pcf_ctx* ctx, const pcf_opts* o);

// Copies `count` FP64 normals to ctx->normals.
int pcf_upload_normals(pcf_ctx* ctx, const pcf_opts* o, int64_t count);

// Waits for an event by polling it from this thread (yielding, then every
// 10 us).  A blocking wait sleeps in the driver until the completion
// interrupt and can wake a few hundred microseconds after the work ended.
hipError_t pcf_wait_event(hipEvent_t ev);

// Ordering around persistent LSM launches, which need all their blocks
// resident at once.  A persistent launch waits for the last kernel every
// other context on its device has enqueued (its kernel-end event), and a
// heavy kernel of any other kind waits for the last persistent launch, so
// neither kind holds CUs the other's blocks are waiting for.
//   pcf_wait_other_contexts: stream s (the persistent stream) waits for the
//     kernel-end event of every other live context on self's device.
//   pcf_after_persistent: called before a request's kernels are enqueued on
//     ctx's stream; when `work` (paths x steps of the launch, paths x assets
//     for a basket, times kPersistentGateFp64 for FP64 arithmetic; HUGE_VAL
//     for the lattices and the per-date LSM form) reaches
//     kPersistentGateWork, the stream waits for the last persistent launch on
//     its device.  2^26 path-steps is about 40 us of the FP32 Asian kernel;
//     its FP64 form takes about 6 times as long.  With `hold`, a heavy launch
//     keeps the device's persistent-launch order locked in *hold; the caller
//     releases it once the kernels' end event (ctx->ev_kend) is recorded, so
//     a persistent launch enqueued from another thread in between cannot miss
//     these kernels (it would wait for that context's previous end event).
constexpr double kPersistentGateWork = 67108864.0;  // 2^26
constexpr double kPersistentGateFp64 = 6.0;
hipError_t pcf_wait_other_contexts(const pcf_ctx* self, hipStream_t s);
hipError_t pcf_after_persistent(pcf_ctx* ctx, double work, std::unique_lock<std::mutex>* hold = nullptr);

// Records the begin event, returns status.  Timing: begin/kbeg/kend/end.
int pcf_timing_begin(pcf_ctx* ctx);
// Copies `nout` doubles of results to ctx->h_out, waits, fills timing fields.
int pcf_timing_end(pcf_ctx* ctx, int nout, pcf_result* res);

static inline pcf_opts pcf_opts_or_default(const pcf_opts* o) {
    pcf_opts d;
    pcf_opts_init(&d);
    return o ? *o : d;
}

static inline int pcf_div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Blocks for `items` work items at B threads x ipt items per thread (>= 1).
static inline int pcf_blocks_for(int64_t items, int B, int ipt) {
    const int64_t per = (int64_t)B * (ipt > 0 ? ipt : 1);
    int64_t b = (items + per - 1) / per;
    if (b < 1) b = 1;
    if (b > (1 << 20)) b = 1 << 20;
    return (int)b;
}

// Smallest power-of-two multiple of `ipt` (items per thread) that keeps a
// grid-stride launch over `items` at B threads per block within max_blocks
// blocks.  Beyond a few blocks per CU, extra blocks of a short kernel only
// add dispatch work and arrivals at its final reduction.
static inline int pcf_capped_ipt(int64_t items, int B, int ipt, int64_t max_blocks) {
    while (ipt < 1024 && (items + (int64_t)B * ipt - 1) / ((int64_t)B * ipt) > max_blocks) ipt <<= 1;
    return ipt;
}

// Size class of a request: ceil(log2 N) clamped to [10, 24].
static inline int pcf_size_class(int64_t N) {
    int c = 0;
    while (((int64_t)1 << c) < N && c < 62) ++c;
    return c < 10 ? 10 : (c > 24 ? 24 : c);
}

// Launch shape from the autotuner's table (autotune.cpp); leaves *B / *ipt
// untouched when the (kind, mode, size class) entry was never tuned.
void pcf_tuned_shape(int kind, int mode, int64_t N, int* B, int* ipt);

static inline uint64_t pcf_make_batch_key(int kind, int mode, int variant, int B) {
    return 1ull | ((uint64_t)(kind & 0xff) << 8) | ((uint64_t)(mode & 0xff) << 16) |
           ((uint64_t)(variant & 0xffff) << 24) | ((uint64_t)(B & 0xffff) << 40);
}

static inline pcf_request pcf_make_request(int kind, double S0, double E, double r, double sigma,
                                           double T, int64_t N, double pf, const pcf_opts* o) {
    pcf_request q{};
    q.kind = kind;
    q.S0 = S0;
    q.E = E;
    q.r = r;
    q.sigma = sigma;
    q.T = T;
    q.N = N;
    q.payoff_fun = pf;
    q.assets = 1;
    q.opts = pcf_opts_or_default(o);
    return q;
}

// Per-product batched runners: n requests sharing one batch key, one launch.
// Per-request validation failures land in status[i]; the return value is a
// launch-level failure.  *_batch_key returns 0 and sets *status on invalid input.
uint64_t pcf_eur_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status);
int pcf_eur_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status);
uint64_t pcf_asia_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status);
int pcf_asia_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status);
uint64_t pcf_basket_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status);
int pcf_basket_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status);
// Device bytes an American request holds while in flight: the path store and
// per-path state (grow-only buffers a context keeps between requests) and the
// rest (coefficients, partials, sync words); the footprint is their sum.
struct PcfAmerBytes {
    int64_t store = 0, state = 0, rest = 0;
};
PcfAmerBytes pcf_amer_bytes(int64_t N, int M, int precision, int rng, int ref_exact);
int64_t pcf_amer_footprint(int64_t N, int M, int precision, int rng, int ref_exact);

// opts.fp64_check on a PCF_FP32K Philox request of a Monte Carlo kind: prices
// q again in PCF_FP64 on the same Philox counters and stores |price - FP64
// price| in res->fp64_err (res->fp64_checked = 1).  A no-op otherwise.
int pcf_apply_fp64_check(pcf_ctx* ctx, const pcf_request& q, pcf_result* res);

// glibc pow, never folded by the compiler.  Verification constants follow
// the reference's -O0 build, which calls glibc pow(x, 2.0); compilers fold
// pow(x,2) into x*x, which differs in the last bit for ~0.1% of inputs.
static inline double libm_pow(double x, double y) {
    double (*volatile f)(double, double) = &::pow;
    return f(x, y);
}
