// asian.hip -- fixed-strike arithmetic-average Asian option.
// Hot loop of src/mc_asia.cpp:27-37, per path and step m:
//     I += S (1 + r dt/2 + sigma dB/2);   S *= exp((r - sigma^2/2) dt + sigma dB)
// then payoff(I/M).  (I is the reference's first-order trapezoid of the
// running integral: it includes S_0 and excludes S_M; reproduced as written.)
//
// One path per lane, paths contiguous across lanes; one Philox call feeds 4
// steps.  One kernel serves one request or a batch (see eur.hip: each job
// reduces over exactly the blocks a lone launch of it would use).  FP32 form: S and the exponent in FP32 on the hardware exp2 unit, I
// Kahan-compensated in FP32, per-path payoffs summed in FP64.  The optional
// geometric control averages log S over t_1..t_M on the same paths.
#include "pcf_internal.hpp"

namespace {

struct AsiaArgs {
    double S0, E, pf;
    double a1, a2;          // 1 + r dt/2, sigma sqrt(dt)/2  (I update, standard normal z)
    double mu, vol;         // (r - s^2/2) dt, sigma sqrt(dt)
    double h_c1, h_mu, sigma;  // verification: 1 + r*dt/2, (r - pow(s,2)/2)*dt, sigma
    float S0f, Ef, pff, a1f, a2f, mu2f, vol2f;  // FP32; exponent constants x log2(e)
    double invM, sdt;       // 1/M, sqrt(dt)
    int64_t N;
    int32_t M;
    uint64_t seed;
};

enum { kFp32k = 0, kFp64 = 1, kHost = 2, kHostProd = 3 };

template <bool CONTROL>
__device__ __forceinline__ void asia_path_f32(const AsiaArgs& a, uint64_t p, float& pay,
                                              float& geo) {
    float S = a.S0f, I = 0.f, Ic = 0.f, ls = 0.f, lsum = 0.f;
    auto step = [&](float z) {
        const float t = S * fmaf(a.a2f, z, a.a1f);
        // Kahan: I += t
        const float y = t - Ic;
        const float u = I + y;
        Ic = (u - I) - y;
        I = u;
        const float e = fmaf(a.vol2f, z, a.mu2f);  // log2 of the step factor
        S *= pcf::exp2_f32(e);
        if (CONTROL) {
            ls += e;
            lsum += ls;
        }
    };
    const int nfull = a.M >> 2;
    for (int b = 0; b < nfull; ++b) {
        float z[4];
        pcf::normals4<float>(a.seed, PCF_STREAM_ASIA, p, (uint32_t)b, z);
        step(z[0]);
        step(z[1]);
        step(z[2]);
        step(z[3]);
    }
    const int rem = a.M & 3;
    if (rem) {
        float z[4];
        pcf::normals4<float>(a.seed, PCF_STREAM_ASIA, p, (uint32_t)nfull, z);
        for (int j = 0; j < rem; ++j) step(z[j]);
    }
    const float avg = I * (float)a.invM;
    pay = pcf::payoff<float>(avg, a.Ef, a.pff);
    if (CONTROL) {
        const float G = a.S0f * pcf::exp2_f32(lsum * (float)a.invM);
        geo = pcf::payoff<float>(G, a.Ef, a.pff);
    }
}

template <bool CONTROL>
__device__ __forceinline__ void asia_path_f64(const AsiaArgs& a, uint64_t p, double& pay,
                                              double& geo) {
    double S = a.S0, I = 0.0, ls = 0.0, lsum = 0.0;
    auto step = [&](double z) {
        I += S * fma(a.a2, z, a.a1);
        const double e = fma(a.vol, z, a.mu);
        S *= exp(e);
        if (CONTROL) {
            ls += e;
            lsum += ls;
        }
    };
    const int nfull = a.M >> 2;
    for (int b = 0; b < nfull; ++b) {
        double z[4];
        pcf::normals4<double>(a.seed, PCF_STREAM_ASIA, p, (uint32_t)b, z);
        step(z[0]);
        step(z[1]);
        step(z[2]);
        step(z[3]);
    }
    const int rem = a.M & 3;
    if (rem) {
        double z[4];
        pcf::normals4<double>(a.seed, PCF_STREAM_ASIA, p, (uint32_t)nfull, z);
        for (int j = 0; j < rem; ++j) step(z[j]);
    }
    pay = pcf::payoff<double>(I / (double)a.M, a.E, a.pf);
    if (CONTROL) geo = pcf::payoff<double>(a.S0 * exp(lsum / (double)a.M), a.E, a.pf);
}

// The production FP64 step of asia_path_f64 on the reference's draws z = dB/sqrt(dt).
__device__ __forceinline__ double asia_path_prod_host(const AsiaArgs& a, const double* __restrict__ w) {
    double S = a.S0, I = 0.0;
    for (int m = 0; m < a.M; ++m) {
        const double z = w[m] / a.sdt;
        I += S * fma(a.a2, z, a.a1);
        S *= exp(fma(a.vol, z, a.mu));
    }
    return pcf::payoff<double>(I / (double)a.M, a.E, a.pf);
}

// Verification: the reference's expression order, FP64, no contraction.
__device__ __forceinline__ double asia_path_ref(const AsiaArgs& a, const double* __restrict__ z) {
#pragma clang fp contract(off)
    double St = a.S0, I = 0.0;
    for (int m = 0; m < a.M; ++m) {
        const double dBi = z[m];
        I += St * (a.h_c1 + a.sigma * dBi / 2);
        St *= exp(a.h_mu + a.sigma * dBi);
    }
    const double v = a.pf * (I / (double)a.M - a.E);
    return v > 0.0 ? v : 0.0;
}

struct AsiaJob {
    AsiaArgs a;
    const double* hz;
    pcf::CSum* partials;
    unsigned int* counter;
    double* out;
    unsigned int first_block, nblk;
};
static_assert(sizeof(AsiaJob) <= kJobBytes, "job record exceeds its staging slot");

template <int MODE, bool CONTROL>
__device__ __forceinline__ void asia_body(const AsiaArgs& a, const double* __restrict__ hz,
                                          pcf::CSum* __restrict__ partials,
                                          unsigned int* __restrict__ counter, double* __restrict__ out,
                                          unsigned int blk, unsigned int nblk, pcf::CSum* lds) {
    constexpr int K = CONTROL ? 4 : 2;
    double s = 0.0, s2 = 0.0, g = 0.0, g2 = 0.0;
    const int64_t stride = (int64_t)nblk * blockDim.x;
    for (int64_t p = (int64_t)blk * blockDim.x + threadIdx.x; p < a.N; p += stride) {
        double v, gv = 0.0;
        if (MODE == kFp32k) {
            float pv, pg = 0.f;
            asia_path_f32<CONTROL>(a, (uint64_t)p, pv, pg);
            v = pv;
            gv = pg;
        } else if (MODE == kFp64) {
            asia_path_f64<CONTROL>(a, (uint64_t)p, v, gv);
        } else if (MODE == kHostProd) {
            v = asia_path_prod_host(a, hz + p * (int64_t)a.M);
        } else {
            v = asia_path_ref(a, hz + p * (int64_t)a.M);
        }
        s += v;
        s2 += v * v;
        if (CONTROL) {
            g += gv;
            g2 += gv * gv;
        }
    }
    pcf::Acc<K> acc, tot;
    acc.v[0] = {s, 0.0};
    acc.v[1] = {s2, 0.0};
    if (CONTROL) {
        acc.v[2 % K] = {g, 0.0};
        acc.v[3 % K] = {g2, 0.0};
    }
    if (pcf::grid_reduce<K>(acc, partials, counter, lds, tot, blk, nblk) && threadIdx.x == 0)
        pcf::write_moments<K / 2>(tot, a.N, out);
}

template <int MODE, bool CONTROL>
__global__ __launch_bounds__(1024) void k_asia(const AsiaJob* __restrict__ jobs, int njobs) {
    __shared__ pcf::CSum lds[16 * (CONTROL ? 4 : 2)];
    const int j = pcf::find_job(jobs, njobs, blockIdx.x);
    const AsiaJob& J = jobs[j];
    asia_body<MODE, CONTROL>(J.a, J.hz, J.partials, J.counter, J.out, blockIdx.x - J.first_block,
                             J.nblk, lds);
}

struct AsiaPrep {
    AsiaArgs a;
    int mode, B, nblk;
    bool control;
    double disc;
};

int asia_prepare(pcf_ctx* ctx, const pcf_request& q, AsiaPrep& p) {
    const pcf_opts& o = q.opts;
    int st = pcf_check_common(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    const int M = q.M;
    if (M <= 0) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "M must be positive (M=%d)", M);
    const double dt = (double)q.T / (double)M;
    const double sdt = std::sqrt(dt);
    const double log2e = 1.4426950408889634;
    AsiaArgs& a = p.a;
    a = AsiaArgs{};
    a.S0 = q.S0;
    a.E = q.E;
    a.pf = q.payoff_fun;
    a.a1 = 1.0 + q.r * dt / 2;
    a.a2 = q.sigma * sdt / 2;
    a.mu = (q.r - 0.5 * q.sigma * q.sigma) * dt;
    a.vol = q.sigma * sdt;
    a.h_c1 = 1 + q.r * dt / 2;
    a.h_mu = (q.r - libm_pow(q.sigma, 2) / 2) * dt;
    a.sigma = q.sigma;
    a.S0f = (float)q.S0;
    a.Ef = (float)q.E;
    a.pff = (float)q.payoff_fun;
    a.a1f = (float)a.a1;
    a.a2f = (float)a.a2;
    a.mu2f = (float)(a.mu * log2e);
    a.vol2f = (float)(a.vol * log2e);
    a.invM = 1.0 / (double)M;
    a.sdt = sdt;
    a.N = q.N;
    a.M = M;
    a.seed = o.seed;
    if (o.rng == PCF_RNG_HOST && o.kernel_variant != 0 && o.kernel_variant != 1)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "mc_asia with host normals: kernel_variant 0 (reference "
                        "arithmetic) or 1 (production arithmetic), got %d", o.kernel_variant);
    p.mode = o.rng == PCF_RNG_HOST ? (o.kernel_variant == 1 ? kHostProd : kHost)
                                   : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    p.control = o.control != 0 && p.mode != kHost && p.mode != kHostProd;
    int B = 256, ipt = 1;
    pcf_tuned_shape(PCF_KIND_ASIA, p.mode, q.N, &B, &ipt);
    if (o.block_size) B = o.block_size;
    if (o.items_per_thread) ipt = o.items_per_thread;
    p.B = B;
    p.nblk = pcf_blocks_for(q.N, B, ipt);
    p.disc = std::exp(-q.r * q.T);
    return PCF_OK;
}

// Result record of one request from its 8 output doubles (write_moments:
// sum, sum of squares, mean, variance of the payoffs, then of the control).
void asia_result(const AsiaPrep& p, const double* h, int64_t N, int M, int B, int batch, const pcf_result& timing,
                 pcf_result& r) {
    // The reference returns exp(-r*T)*result/(double)N (src/mc_asia.cpp:39).
    r.price = p.disc * h[0] / (double)N;
    r.mean = h[2];
    r.variance = h[3];
    r.std_err = p.disc * std::sqrt(h[3] / (double)N);
    if (p.control) {
        r.control_price = p.disc * h[4] / (double)N;
        r.control_se = p.disc * std::sqrt(h[7] / (double)N);
    }
    r.n_paths = N;
    r.n_steps = M;
    r.kernel_ms = timing.kernel_ms;
    r.total_ms = timing.total_ms;
    // 14 algorithmic flops per path-step: normal 6, I update 2 FMAs (4),
    // S update FMA + exp + mul (4).  No HBM traffic in the Philox forms.
    r.flops = 14.0 * (double)N * (double)M;
    r.bytes = p.mode >= kHost ? 8.0 * (double)N * (double)M : 0.0;
    r.grid_blocks = p.nblk;
    r.block_size = B;
    r.batch_size = batch;
}

template <int MODE, bool CONTROL>
void launch_asia(pcf_ctx* ctx, unsigned int grid, int B, int njobs) {
    k_asia<MODE, CONTROL><<<grid, B, 0, ctx->stream>>>(ctx->jobs.as<AsiaJob>(), njobs);
}

}  // namespace

uint64_t pcf_asia_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status) {
    AsiaPrep p;
    *status = asia_prepare(ctx, q, p);
    if (*status) return 0;
    return pcf_make_batch_key(PCF_KIND_ASIA, p.mode, p.control ? 1 : 0, p.B);
}

int pcf_asia_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status) {
    if (n < 1 || n > kMaxBatch) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch size %d", n);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    AsiaPrep prep[kMaxBatch];
    int live[kMaxBatch], nl = 0;
    for (int i = 0; i < n; ++i) {
        res[i] = pcf_result{};
        status[i] = asia_prepare(ctx, *reqs[i], prep[i]);
        if (status[i] == PCF_OK) live[nl++] = i;
    }
    if (nl == 0) return PCF_OK;
    const int mode = prep[live[0]].mode, B = prep[live[0]].B;
    const bool control = prep[live[0]].control;
    const int K = control ? 4 : 2;
    int64_t nparts = 0, nb = 0, nnorm = 0;
    for (int k = 0; k < nl; ++k) {
        const AsiaPrep& p = prep[live[k]];
        if (p.mode != mode || p.B != B || p.control != control)
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "mixed batch keys");
        nparts += (int64_t)p.nblk * K;
        nb += p.nblk;
        if (p.mode >= kHost) nnorm += reqs[live[k]]->N * (int64_t)reqs[live[k]]->M;
    }
    if (nb > (1ll << 30)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch too large");
    int st = pcf_timing_begin(ctx);
    if (st) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)nparts * sizeof(pcf::CSum)));
    if (nnorm) PCF_HIP_CHECK(ctx, ctx->normals.ensure((size_t)nnorm * sizeof(double)));
    AsiaJob* jobs = (AsiaJob*)ctx->h_jobs;
    int64_t poff = 0, noff = 0;
    unsigned int boff = 0;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        const AsiaPrep& p = prep[i];
        AsiaJob& J = jobs[k];
        J.a = p.a;
        J.hz = nullptr;
        if (p.mode >= kHost) {
            const pcf_opts& o = reqs[i]->opts;
            const int64_t cnt = reqs[i]->N * (int64_t)reqs[i]->M;
            if (!o.host_normals || o.n_host_normals < cnt)
                return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "need %lld host normals", (long long)cnt);
            double* dst = ctx->normals.as<double>() + noff;
            PCF_HIP_CHECK(ctx, hipMemcpyAsync(dst, o.host_normals, (size_t)cnt * sizeof(double),
                                              hipMemcpyHostToDevice, ctx->stream));
            J.hz = dst;
            noff += cnt;
        }
        J.partials = ctx->partials.as<pcf::CSum>() + poff;
        J.counter = ctx->counters.as<unsigned int>() + (size_t)k * pcf::kCounterWords;
        J.out = ctx->out.as<double>() + (size_t)kOutPerJob * k;
        J.first_block = boff;
        J.nblk = (unsigned int)p.nblk;
        poff += (int64_t)p.nblk * K;
        boff += (unsigned int)p.nblk;
    }
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->jobs.ptr, jobs, sizeof(AsiaJob) * nl, hipMemcpyHostToDevice,
                                      ctx->stream));
    double work = 0.0;
    for (int k = 0; k < nl; ++k) work += (double)reqs[live[k]]->N * (double)reqs[live[k]]->M;
    std::unique_lock<std::mutex> order;  // held until the end event is recorded
    PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, mode == kFp32k ? work : kPersistentGateFp64 * work, &order));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    if (mode == kFp32k) {
        if (control) launch_asia<kFp32k, true>(ctx, boff, B, nl);
        else launch_asia<kFp32k, false>(ctx, boff, B, nl);
    } else if (mode == kFp64) {
        if (control) launch_asia<kFp64, true>(ctx, boff, B, nl);
        else launch_asia<kFp64, false>(ctx, boff, B, nl);
    } else if (mode == kHost) {
        launch_asia<kHost, false>(ctx, boff, B, nl);
    } else {
        launch_asia<kHostProd, false>(ctx, boff, B, nl);
    }
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if (order.owns_lock()) order.unlock();
    pcf_result timing{};
    if ((st = pcf_timing_end(ctx, kOutPerJob * nl, &timing))) return st;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        asia_result(prep[i], ctx->h_out + (size_t)kOutPerJob * k, reqs[i]->N, reqs[i]->M, B, nl, timing, res[i]);
    }
    return PCF_OK;
}

// ---- the no-allocation entry point -----------------------------------------
// pcf_mc_asia_noalloc runs the same kernels as pcf_mc_asia with every device
// buffer they need -- the job record, the per-block compensated partials, the
// arrival counters and the results -- in one file-scope device array, so a
// call allocates no device memory.  Kernel arguments and host code use the
// array's device address, obtained once per device with hipGetSymbolAddress.
namespace {

constexpr int kNoAllocMaxBlocks = 32768;  // grid blocks of one launch
constexpr size_t kNoAllocPartialsBytes = (size_t)kNoAllocMaxBlocks * 4 * sizeof(pcf::CSum);  // 2 MiB
constexpr size_t kNoAllocJobOff = kNoAllocPartialsBytes;
constexpr size_t kNoAllocCounterOff = kNoAllocJobOff + kJobBytes;
constexpr size_t kNoAllocOutOff =
    kNoAllocCounterOff + (pcf::kCounterWords * sizeof(unsigned int) + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocOutOff + 256;

__global__ void k_zero_words(unsigned int* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace

// The scratch array (2 MiB + 5 KiB per device, held until the process ends).
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Per device: the array's device address, and a lock that serialises the
// calls using it (each call returns only after its work is complete).
struct ScratchSlot {
    std::mutex mu;
    char* ptr = nullptr;
};
ScratchSlot& scratch_slot(int device) {
    static ScratchSlot slots[64];
    return slots[device & 63];
}

}  // namespace

extern "C" int pcf_mc_asia_noalloc(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                                   int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts,
                                   pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_request q = pcf_make_request(PCF_KIND_ASIA, S0, E, r, sigma, T, N, payoff_fun, opts);
    q.M = M;
    *res = pcf_result{};
    if (q.opts.rng != PCF_RNG_PHILOX)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "pcf_mc_asia_noalloc draws Philox normals (host normals would "
                        "need a device buffer: use pcf_mc_asia)");
    if (q.opts.fp64_check)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "pcf_mc_asia_noalloc does not make the FP64 comparison run");
    if (ctx->device < 0 || ctx->device >= 64) return pcf_fail(ctx, PCF_ERR_NO_DEVICE, "device index");
    AsiaPrep p;
    int st = asia_prepare(ctx, q, p);
    if (st) return st;
    const int K = p.control ? 4 : 2;
    if ((size_t)p.nblk * K * sizeof(pcf::CSum) > kNoAllocPartialsBytes)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG,
                        "pcf_mc_asia_noalloc holds the partials of at most %zu blocks with %d sums (this launch "
                        "has %d blocks of %d threads): use pcf_mc_asia",
                        kNoAllocPartialsBytes / (K * sizeof(pcf::CSum)), K, p.nblk, p.B);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ScratchSlot& slot = scratch_slot(ctx->device);
    std::lock_guard<std::mutex> sl(slot.mu);
    if (!slot.ptr) {
        void* dp = nullptr;
        PCF_HIP_CHECK(ctx, hipGetSymbolAddress(&dp, HIP_SYMBOL(scratch)));
        slot.ptr = static_cast<char*>(dp);
    }
    char* const base = slot.ptr;
    unsigned int* const counter = reinterpret_cast<unsigned int*>(base + kNoAllocCounterOff);
    AsiaJob* const djob = reinterpret_cast<AsiaJob*>(base + kNoAllocJobOff);
    double* const dout = reinterpret_cast<double*>(base + kNoAllocOutOff);
    AsiaJob* hj = static_cast<AsiaJob*>(ctx->h_jobs);
    *hj = AsiaJob{};
    hj->a = p.a;
    hj->hz = nullptr;
    hj->partials = reinterpret_cast<pcf::CSum*>(base);
    hj->counter = counter;
    hj->out = dout;
    hj->first_block = 0;
    hj->nblk = (unsigned int)p.nblk;
    if ((st = pcf_timing_begin(ctx))) return st;
    // the arrival counters start at zero (a fill kernel: no memset node)
    k_zero_words<<<1, 256, 0, ctx->stream>>>(counter, pcf::kCounterWords);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(djob, hj, sizeof(AsiaJob), hipMemcpyHostToDevice, ctx->stream));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    const unsigned int grid = (unsigned int)p.nblk;
    if (p.mode == kFp32k) {
        if (p.control) k_asia<kFp32k, true><<<grid, p.B, 0, ctx->stream>>>(djob, 1);
        else k_asia<kFp32k, false><<<grid, p.B, 0, ctx->stream>>>(djob, 1);
    } else {
        if (p.control) k_asia<kFp64, true><<<grid, p.B, 0, ctx->stream>>>(djob, 1);
        else k_asia<kFp64, false><<<grid, p.B, 0, ctx->stream>>>(djob, 1);
    }
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_end, ctx->stream));
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_out, dout, kOutPerJob * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->stream));
    PCF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    pcf_result timing{};
    float kms = 0.f, tms = 0.f;
    PCF_HIP_CHECK(ctx, hipEventElapsedTime(&kms, ctx->ev_kbeg, ctx->ev_kend));
    PCF_HIP_CHECK(ctx, hipEventElapsedTime(&tms, ctx->ev_begin, ctx->ev_end));
    timing.kernel_ms = kms;
    timing.total_ms = tms;
    asia_result(p, ctx->h_out, N, M, p.B, 1, timing, *res);
    return PCF_OK;
}

extern "C" int pcf_mc_asia(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                           int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts,
                           pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_request q = pcf_make_request(PCF_KIND_ASIA, S0, E, r, sigma, T, N, payoff_fun, opts);
    q.M = M;
    const pcf_request* qp = &q;
    int status = PCF_OK;
    int st = pcf_asia_run(ctx, 1, &qp, res, &status);
    if (st || status) return st ? st : status;
    return pcf_apply_fp64_check(ctx, q, res);
}
