// eur.hip -- European option on one GBM asset, single terminal step.
// Hot loop of src/mc_eur.cpp:23-25: sum_n payoff(S0 exp((r - s^2/2)T + s sqrt(T) z_n)).
//
// One kernel serves one request or a batch of requests: job j owns a range
// of the grid's blocks and reduces over exactly the blocks (and the per-thread
// path assignment) a lone launch of it would use, so its result does not
// depend on what it is batched with.  Within a job: grid-stride over groups of
// 4 paths (one Philox block feeds 4 paths), per-thread FP64 sums, compensated
// block / last-block reduction (pcf_device.hpp).
#include "pcf_internal.hpp"

namespace {

struct EurArgs {
    double S0, E, pf;
    double mu, vol;       // production: (r - s^2/2) T and s sqrt(T)
    double href_mu;       // verification: (r - pow(s,2)/2) * T evaluated on the host
    double sigma;
    float S0f, Ef, pff, mu2f, vol2f;  // FP32 form, exponent constants scaled by log2(e)
    int64_t N;
    uint64_t seed;
};

struct EurJob {
    EurArgs a;
    const double* hz;
    pcf::CSum* partials;
    unsigned int* counter;
    double* out;
    unsigned int first_block, nblk;
};
static_assert(sizeof(EurJob) <= kJobBytes, "job record exceeds its staging slot");

enum { kFp32k = 0, kFp64 = 1, kHost = 2 };

// Verification arithmetic in the reference's order, no contraction:
//   S0*exp((r-pow(sigma,2)/2)*T + sigma*z)
__device__ __forceinline__ double eur_ref_path(const EurArgs& a, double z) {
#pragma clang fp contract(off)
    const double x = a.href_mu + a.sigma * z;
    const double S = a.S0 * exp(x);
    const double v = a.pf * (S - a.E);
    return v > 0.0 ? v : 0.0;
}

template <int MODE>
__device__ __forceinline__ void eur_body(const EurArgs& a, const double* __restrict__ hz,
                                         pcf::CSum* __restrict__ partials,
                                         unsigned int* __restrict__ counter, double* __restrict__ out,
                                         unsigned int blk, unsigned int nblk, pcf::CSum* lds) {
    double s = 0.0, s2 = 0.0;
    const int64_t ngroups = (a.N + 3) >> 2;
    const int64_t stride = (int64_t)nblk * blockDim.x;
    for (int64_t g = (int64_t)blk * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
        const int64_t n0 = g << 2;
        if (MODE == kFp32k) {
            float z[4];
            pcf::normals4<float>(a.seed, PCF_STREAM_EUR, (uint64_t)g, 0u, z);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float S = a.S0f * pcf::exp2_f32(fmaf(a.vol2f, z[j], a.mu2f));
                const float v = pcf::payoff<float>(S, a.Ef, a.pff);
                const double dv = (n0 + j < a.N) ? (double)v : 0.0;
                s += dv;
                s2 += dv * dv;
            }
        } else if (MODE == kFp64) {
            double z[4];
            pcf::normals4<double>(a.seed, PCF_STREAM_EUR, (uint64_t)g, 0u, z);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double S = a.S0 * exp(fma(a.vol, z[j], a.mu));
                const double v = pcf::payoff<double>(S, a.E, a.pf);
                const double dv = (n0 + j < a.N) ? v : 0.0;
                s += dv;
                s2 += dv * dv;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (n0 + j < a.N) {
                    const double v = eur_ref_path(a, hz[n0 + j]);
                    s += v;
                    s2 += v * v;
                }
            }
        }
    }
    pcf::Acc<2> acc, tot;
    acc.v[0] = {s, 0.0};
    acc.v[1] = {s2, 0.0};
    if (pcf::grid_reduce<2>(acc, partials, counter, lds, tot, blk, nblk) && threadIdx.x == 0)
        pcf::write_moments<1>(tot, a.N, out);
}

template <int MODE>
__global__ __launch_bounds__(1024) void k_eur(const EurJob* __restrict__ jobs, int njobs) {
    __shared__ pcf::CSum lds[16 * 2];
    const int j = pcf::find_job(jobs, njobs, blockIdx.x);
    const EurJob& J = jobs[j];
    eur_body<MODE>(J.a, J.hz, J.partials, J.counter, J.out, blockIdx.x - J.first_block, J.nblk, lds);
}

struct EurPrep {
    EurArgs a;
    int mode, B, nblk;
    double disc;
};

int eur_prepare(pcf_ctx* ctx, const pcf_request& q, EurPrep& p) {
    const pcf_opts& o = q.opts;
    int st = pcf_check_common(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    EurArgs& a = p.a;
    a = EurArgs{};
    a.S0 = q.S0;
    a.E = q.E;
    a.pf = q.payoff_fun;
    a.sigma = q.sigma;
    a.mu = (q.r - 0.5 * q.sigma * q.sigma) * q.T;
    a.vol = q.sigma * std::sqrt(q.T);
    a.href_mu = (q.r - libm_pow(q.sigma, 2) / 2) * q.T;
    const double log2e = 1.4426950408889634;
    a.S0f = (float)q.S0;
    a.Ef = (float)q.E;
    a.pff = (float)q.payoff_fun;
    a.mu2f = (float)(a.mu * log2e);
    a.vol2f = (float)(a.vol * log2e);
    a.N = q.N;
    a.seed = o.seed;
    p.mode = o.rng == PCF_RNG_HOST ? kHost : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    int B = 256, ipt = 4;
    pcf_tuned_shape(PCF_KIND_EUR, p.mode, q.N, &B, &ipt);
    if (o.block_size) B = o.block_size;
    if (o.items_per_thread) ipt = o.items_per_thread;
    p.B = B;
    p.nblk = pcf_blocks_for((q.N + 3) / 4, B, ipt);
    p.disc = std::exp(-q.r * q.T);
    return PCF_OK;
}

}  // namespace

uint64_t pcf_eur_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status) {
    EurPrep p;
    *status = eur_prepare(ctx, q, p);
    if (*status) return 0;
    return pcf_make_batch_key(PCF_KIND_EUR, p.mode, 0, p.B);
}

// Runs n requests of one batch key in one launch; statuses per request.
int pcf_eur_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status) {
    if (n < 1 || n > kMaxBatch) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch size %d", n);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    EurPrep prep[kMaxBatch];
    int live[kMaxBatch], nl = 0;
    for (int i = 0; i < n; ++i) {
        res[i] = pcf_result{};
        status[i] = eur_prepare(ctx, *reqs[i], prep[i]);
        if (status[i] == PCF_OK) live[nl++] = i;
    }
    if (nl == 0) return PCF_OK;
    const int mode = prep[live[0]].mode, B = prep[live[0]].B;
    int64_t nparts = 0, nb = 0, nnorm = 0;
    for (int k = 0; k < nl; ++k) {
        const EurPrep& p = prep[live[k]];
        if (p.mode != mode || p.B != B) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "mixed batch keys");
        nparts += (int64_t)p.nblk * 2;
        nb += p.nblk;
        if (p.mode == kHost) nnorm += reqs[live[k]]->N;
    }
    if (nb > (1ll << 30)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch too large");
    int st = pcf_timing_begin(ctx);
    if (st) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)nparts * sizeof(pcf::CSum)));
    if (nnorm) PCF_HIP_CHECK(ctx, ctx->normals.ensure((size_t)nnorm * sizeof(double)));
    EurJob* jobs = (EurJob*)ctx->h_jobs;
    int64_t poff = 0, noff = 0;
    unsigned int boff = 0;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        const EurPrep& p = prep[i];
        EurJob& J = jobs[k];
        J.a = p.a;
        J.hz = nullptr;
        if (p.mode == kHost) {
            const pcf_opts& o = reqs[i]->opts;
            if (!o.host_normals || o.n_host_normals < reqs[i]->N)
                return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "need %lld host normals", (long long)reqs[i]->N);
            double* dst = ctx->normals.as<double>() + noff;
            PCF_HIP_CHECK(ctx, hipMemcpyAsync(dst, o.host_normals, (size_t)reqs[i]->N * sizeof(double),
                                              hipMemcpyHostToDevice, ctx->stream));
            J.hz = dst;
            noff += reqs[i]->N;
        }
        J.partials = ctx->partials.as<pcf::CSum>() + poff;
        J.counter = ctx->counters.as<unsigned int>() + (size_t)k * pcf::kCounterWords;
        J.out = ctx->out.as<double>() + (size_t)kOutPerJob * k;
        J.first_block = boff;
        J.nblk = (unsigned int)p.nblk;
        poff += (int64_t)p.nblk * 2;
        boff += (unsigned int)p.nblk;
    }
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->jobs.ptr, jobs, sizeof(EurJob) * nl, hipMemcpyHostToDevice, ctx->stream));
    double work = 0.0;
    for (int k = 0; k < nl; ++k) work += (double)reqs[live[k]]->N;
    std::unique_lock<std::mutex> order;  // held until the end event is recorded
    PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, mode == kFp32k ? work : kPersistentGateFp64 * work, &order));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    const EurJob* dj = ctx->jobs.as<EurJob>();
    if (mode == kFp32k) k_eur<kFp32k><<<boff, B, 0, ctx->stream>>>(dj, nl);
    else if (mode == kFp64) k_eur<kFp64><<<boff, B, 0, ctx->stream>>>(dj, nl);
    else k_eur<kHost><<<boff, B, 0, ctx->stream>>>(dj, nl);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if (order.owns_lock()) order.unlock();
    pcf_result timing{};
    if ((st = pcf_timing_end(ctx, kOutPerJob * nl, &timing))) return st;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        const EurPrep& p = prep[i];
        const double* h = ctx->h_out + (size_t)kOutPerJob * k;
        const int64_t N = reqs[i]->N;
        pcf_result& r = res[i];
        // The reference returns (exp(-r*T)*result)/(double)N (src/mc_eur.cpp:26).
        r.price = (p.disc * h[0]) / (double)N;
        r.mean = h[2];
        r.variance = h[3];
        r.std_err = p.disc * std::sqrt(h[3] / (double)N);
        r.n_paths = N;
        r.n_steps = 1;
        r.kernel_ms = timing.kernel_ms;
        r.total_ms = timing.total_ms;
        // 13 algorithmic flops per path (normal 6, drift FMA 2, exp 1, x S0 1,
        // payoff 2, accumulate 1); bytes: host normals only in verification.
        r.flops = 13.0 * (double)N;
        r.bytes = p.mode == kHost ? 8.0 * (double)N : 0.0;
        r.grid_blocks = p.nblk;
        r.block_size = B;
        r.batch_size = nl;
    }
    return PCF_OK;
}

extern "C" int pcf_mc_eur(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                          int64_t N, double payoff_fun, const pcf_opts* opts, pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_request q = pcf_make_request(PCF_KIND_EUR, S0, E, r, sigma, T, N, payoff_fun, opts);
    const pcf_request* qp = &q;
    int status = PCF_OK;
    int st = pcf_eur_run(ctx, 1, &qp, res, &status);
    if (st || status) return st ? st : status;
    return pcf_apply_fp64_check(ctx, q, res);
}
