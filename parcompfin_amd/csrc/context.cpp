// context.cpp -- pcf_ctx lifetime, errors, argument checks, timing, and the
// generator dump entry point.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "pcf_internal.hpp"

namespace {
// Every live context, so that a persistent launch can be ordered after the
// kernels that other contexts on its device have already enqueued.
std::mutex& ctx_registry_mu() {
    static std::mutex m;
    return m;
}
std::vector<pcf_ctx*>& ctx_registry() {
    static std::vector<pcf_ctx*> v;
    return v;
}
}  // namespace

hipError_t pcf_wait_other_contexts(const pcf_ctx* self, hipStream_t s) {
    std::lock_guard<std::mutex> lk(ctx_registry_mu());
    for (pcf_ctx* c : ctx_registry()) {
        if (c == self || c->device != self->device) continue;
        const hipError_t e = hipStreamWaitEvent(s, c->ev_kend, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int pcf_fail(pcf_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int pcf_hip_fail(pcf_ctx* ctx, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? PCF_ERR_OUT_OF_MEMORY
                         : PCF_ERR_HIP;
    return pcf_fail(ctx, code, "%s: %s (%s)", what, hipGetErrorString(e), hipGetErrorName(e));
}

extern "C" {

int pcf_abi_version(void) { return PCF_ABI_VERSION; }

void pcf_opts_init(pcf_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->seed = 0x5EEDull;
    o->precision = PCF_FP32K;
    o->rng = PCF_RNG_PHILOX;
}

int pcf_ctx_create(int device, pcf_ctx** out) {
    if (!out) return PCF_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PCF_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return PCF_ERR_NO_DEVICE;
    pcf_ctx* ctx = new pcf_ctx();
    ctx->device = device;
    auto bail = [&](hipError_t e, const char* what) {
        int code = pcf_hip_fail(ctx, e, what);
        std::fprintf(stderr, "pcf_ctx_create: %s\n", ctx->last_error.c_str());
        pcf_ctx_destroy(ctx);
        return code;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess)
        return bail(e, "hipGetDeviceProperties");
    ctx->cu_count = prop.multiProcessorCount;
    ctx->total_mem = (int64_t)prop.totalGlobalMem;
    std::snprintf(ctx->name, sizeof ctx->name, "%s (%s)", prop.name, prop.gcnArchName);
    // Non-blocking stream at the default priority (every stream here keeps it).
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    for (hipEvent_t* ev : {&ctx->ev_begin, &ctx->ev_kbeg, &ctx->ev_kend, &ctx->ev_end})
        if ((e = hipEventCreate(ev)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&ctx->ev_ready, hipEventDisableTiming)) != hipSuccess)
        return bail(e, "hipEventCreate");
    if ((e = ctx->counters.ensure(kCounterSlots * sizeof(unsigned int), true)) != hipSuccess)
        return bail(e, "alloc counters");
    if ((e = ctx->out.ensure(kOutDoubles * sizeof(double), true)) != hipSuccess)
        return bail(e, "alloc out");
    // Coherent: the persistent LSM kernel writes its results here directly,
    // and the host reads them once the launch has completed.
    if ((e = hipHostMalloc((void**)&ctx->h_out, kOutDoubles * sizeof(double), hipHostMallocCoherent)) !=
        hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = hipHostGetDevicePointer((void**)&ctx->h_out_dev, ctx->h_out, 0)) != hipSuccess)
        return bail(e, "hipHostGetDevicePointer");
    if ((e = hipHostMalloc(&ctx->h_jobs, kMaxBatch * kJobBytes, 0)) != hipSuccess)
        return bail(e, "hipHostMalloc jobs");
    if ((e = ctx->jobs.ensure(kMaxBatch * kJobBytes)) != hipSuccess) return bail(e, "alloc jobs");
    {
        std::lock_guard<std::mutex> lk(ctx_registry_mu());
        ctx_registry().push_back(ctx);
    }
    *out = ctx;
    return PCF_OK;
}

int pcf_ctx_destroy(pcf_ctx* ctx) {
    if (!ctx) return PCF_OK;
    {
        std::lock_guard<std::mutex> lk(ctx_registry_mu());
        auto& v = ctx_registry();
        v.erase(std::remove(v.begin(), v.end(), ctx), v.end());
    }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    pcf_lsm_release(ctx);
    for (DevBuf* b : {&ctx->partials, &ctx->counters, &ctx->out, &ctx->normals, &ctx->logtab,
                      &ctx->prefix, &ctx->logfact, &ctx->lsm_store, &ctx->lsm_state, &ctx->lsm_coef,
                      &ctx->basket_L, &ctx->scratch, &ctx->jobs, &ctx->lattice})
        b->release();
    if (ctx->h_out) (void)hipHostFree(ctx->h_out);
    if (ctx->h_jobs) (void)hipHostFree(ctx->h_jobs);
    for (hipEvent_t ev : {ctx->ev_begin, ctx->ev_kbeg, ctx->ev_kend, ctx->ev_end, ctx->ev_ready})
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PCF_OK;
}

const char* pcf_ctx_last_error(const pcf_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int pcf_ctx_synchronize(pcf_ctx* ctx) {
    if (!ctx) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    PCF_HIP_CHECK(ctx, hipDeviceSynchronize());
    return PCF_OK;
}

int pcf_device_info(pcf_ctx* ctx, char* name, int name_len, int* compute_units,
                    int64_t* total_mem) {
    if (!ctx) return PCF_ERR_INVALID_ARG;
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", ctx->name);
    if (compute_units) *compute_units = ctx->cu_count;
    if (total_mem) *total_mem = ctx->total_mem;
    return PCF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
int pcf_check_common(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                     int64_t N, double payoff_fun) {
    if (payoff_fun != 1.0 && payoff_fun != -1.0)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "Unknown payoff function (payoff_fun=%g)", payoff_fun);
    if (!(N > 0)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "N must be positive (N=%lld)", (long long)N);
    if (!std::isfinite(S0) || !std::isfinite(E) || !std::isfinite(r) || !std::isfinite(sigma) ||
        !std::isfinite(T))
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "non-finite parameter");
    if (!(T > 0)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "T must be positive");
    if (sigma < 0) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "sigma must be non-negative");
    return PCF_OK;
}

int pcf_check_opts(pcf_ctx* ctx, const pcf_opts* o) {
    if (o->precision != PCF_FP32K && o->precision != PCF_FP64)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "unknown precision %d", o->precision);
    if (o->rng != PCF_RNG_PHILOX && o->rng != PCF_RNG_HOST)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "unknown rng %d", o->rng);
    if (o->block_size != 0 &&
        (o->block_size < 64 || o->block_size > 1024 || (o->block_size % 64) != 0))
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "block_size must be a multiple of 64 in [64,1024]");
    if (o->items_per_thread < 0 || o->items_per_thread > 4096)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "items_per_thread out of range");
    return PCF_OK;
}

int pcf_upload_normals(pcf_ctx* ctx, const pcf_opts* o, int64_t count) {
    if (!o->host_normals)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "PCF_RNG_HOST needs host_normals");
    if (o->n_host_normals < count)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "need %lld host normals, got %lld",
                        (long long)count, (long long)o->n_host_normals);
    const size_t bytes = (size_t)count * sizeof(double);
    PCF_HIP_CHECK(ctx, ctx->normals.ensure(bytes));
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->normals.ptr, o->host_normals, bytes,
                                      hipMemcpyHostToDevice, ctx->stream));
    return PCF_OK;
}

int pcf_apply_fp64_check(pcf_ctx* ctx, const pcf_request& q, pcf_result* res) {
    const pcf_opts& o = q.opts;
    if (!o.fp64_check || o.precision != PCF_FP32K || o.rng != PCF_RNG_PHILOX) return PCF_OK;
    pcf_opts o64 = o;
    o64.precision = PCF_FP64;
    o64.fp64_check = 0;
    o64.when_out = nullptr;
    pcf_result r64{};
    int st;
    switch (q.kind) {
        case PCF_KIND_EUR:
            st = pcf_mc_eur(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun, &o64, &r64);
            break;
        case PCF_KIND_ASIA:
            st = pcf_mc_asia(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.M, q.payoff_fun, &o64, &r64);
            break;
        case PCF_KIND_AMER:
            st = pcf_mc_amer(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.M, q.payoff_fun, &o64, &r64);
            break;
        case PCF_KIND_BASKET:
            st = pcf_mc_eur_multi(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun, q.assets, q.rho, q.basket,
                                  &o64, &r64);
            break;
        default:
            return PCF_OK;
    }
    if (st) return st;
    res->fp64_err = std::fabs(res->price - r64.price);
    res->fp64_checked = 1;
    return PCF_OK;
}

int pcf_timing_begin(pcf_ctx* ctx) {
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_begin, ctx->stream));
    return PCF_OK;
}

hipError_t pcf_wait_event(hipEvent_t ev) {
    // Yield for the first few polls, then poll every 10 us: with many slots
    // waiting at once, tight polling took host time from the threads that
    // launch the next requests (19 American requests served back to back:
    // 7.07 ms per set with tight polling, 6.71-6.96 ms with this back-off).
    for (int spin = 0;; ++spin) {
        const hipError_t e = hipEventQuery(ev);
        if (e != hipErrorNotReady) return e;
        if (spin < 64) std::this_thread::yield();
        else std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
}

int pcf_timing_end(pcf_ctx* ctx, int nout, pcf_result* res) {
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_end, ctx->stream));
    if (nout > 0)
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_out, ctx->out.ptr, (size_t)nout * sizeof(double),
                                          hipMemcpyDeviceToHost, ctx->stream));
    // everything this call enqueued is done once ev_ready is
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_ready, ctx->stream));
    PCF_HIP_CHECK(ctx, pcf_wait_event(ctx->ev_ready));
    float kms = 0.f, tms = 0.f;
    PCF_HIP_CHECK(ctx, hipEventElapsedTime(&kms, ctx->ev_kbeg, ctx->ev_kend));
    PCF_HIP_CHECK(ctx, hipEventElapsedTime(&tms, ctx->ev_begin, ctx->ev_end));
    res->kernel_ms = kms;
    res->total_ms = tms;
    return PCF_OK;
}

// ---------------------------------------------------------------------------
// Generator dump: the exact normals the pricing kernels consume.
namespace {
template <typename Real>
__global__ void k_normals(uint64_t seed, uint32_t stream, uint64_t path0, int64_t n_paths,
                          int K, double* __restrict__ out) {
    const int nblk = (K + 3) / 4;
    const int64_t total = n_paths * nblk;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / nblk;
        const int b = (int)(t % nblk);
        Real z[4];
        pcf::normals4<Real>(seed, stream, path0 + (uint64_t)p, (uint32_t)b, z);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (4 * b + j < K) out[p * K + 4 * b + j] = (double)z[j];
    }
}
}  // namespace

extern "C" int pcf_normals(pcf_ctx* ctx, uint64_t seed, uint32_t stream, uint64_t path0,
                           int64_t n_paths, int32_t K, int32_t precision, double* out) {
    if (!ctx || !out || n_paths <= 0 || K <= 0) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "bad args");
    if (precision != PCF_FP32K && precision != PCF_FP64)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "unknown precision");
    if (n_paths * (int64_t)K > (int64_t)1 << 28)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "dump too large");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t bytes = (size_t)n_paths * K * sizeof(double);
    PCF_HIP_CHECK(ctx, ctx->scratch.ensure(bytes));
    const int64_t work = n_paths * ((K + 3) / 4);
    const int grid = (int)std::min<int64_t>((work + 255) / 256, 65535);
    if (precision == PCF_FP64)
        k_normals<double><<<grid, 256, 0, ctx->stream>>>(seed, stream, path0, n_paths, K,
                                                         ctx->scratch.as<double>());
    else
        k_normals<float><<<grid, 256, 0, ctx->stream>>>(seed, stream, path0, n_paths, K,
                                                        ctx->scratch.as<double>());
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(out, ctx->scratch.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PCF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return PCF_OK;
}

// sizeof of the ABI records, so bindings in other languages can check their
// mirrors of include/pcf.h without a device.
extern "C" int pcf_struct_size(const char* name) {
    if (!name) return -1;
    const std::string n(name);
    if (n == "pcf_opts") return (int)sizeof(pcf_opts);
    if (n == "pcf_result") return (int)sizeof(pcf_result);
    if (n == "pcf_basket_ext") return (int)sizeof(pcf_basket_ext);
    if (n == "pcf_request") return (int)sizeof(pcf_request);
    if (n == "pcf_server_config") return (int)sizeof(pcf_server_config);
    if (n == "pcf_request_timing") return (int)sizeof(pcf_request_timing);
    if (n == "pcf_tune_entry") return (int)sizeof(pcf_tune_entry);
    if (n == "pcf_server_stats") return (int)sizeof(pcf_server_stats);
    return -1;
}
