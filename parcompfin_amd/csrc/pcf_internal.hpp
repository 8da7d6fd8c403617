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
            e = hipMemset(ptr, 0, want);
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

// Number of reduction slots (last-block counters) a context keeps; each
// launch in a sequence that is in flight at once uses its own slot.
constexpr int kCounterSlots = 64;
// Max doubles of final results a kernel writes.
constexpr int kOutDoubles = 64;

struct pcf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_kbeg = nullptr, ev_kend = nullptr, ev_end = nullptr;
    std::string last_error;
    std::recursive_mutex mu;
    int cu_count = 0;
    int64_t total_mem = 0;
    char name[256] = {0};

    DevBuf partials;   // per-block compensated partial sums
    DevBuf counters;   // kCounterSlots uint32 last-block counters, zero at rest
    DevBuf out;        // kOutDoubles final values
    DevBuf normals;    // verification normals copied from the host
    DevBuf logtab;     // binom: glibc log(n), n = 0..N (resident per N)
    int64_t logtab_n = -1;
    DevBuf prefix;     // binom: prefix chain of the descending log sum
    int64_t prefix_n = -1;
    DevBuf lsm_store;  // mc_amer path store
    DevBuf lsm_state;  // mc_amer per-path state
    DevBuf lsm_coef;   // mc_amer per-date coefficients / flags
    DevBuf basket_L;   // basket Cholesky factor + per-asset constants
    DevBuf scratch;    // misc (normals dump)
    double* h_out = nullptr;  // pinned mirror of `out`
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

// glibc pow, never folded by the compiler.  Verification constants follow
// the reference's -O0 build, which calls glibc pow(x, 2.0); compilers fold
// pow(x,2) into x*x, which differs in the last bit for ~0.1% of inputs.
static inline double libm_pow(double x, double y) {
    double (*volatile f)(double, double) = &::pow;
    return f(x, y);
}
