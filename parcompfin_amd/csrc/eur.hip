// eur.hip -- European option on one GBM asset, single terminal step.
// Hot loop of src/mc_eur.cpp:23-25: sum_n payoff(S0 exp((r - s^2/2)T + s sqrt(T) z_n)).
//
// Launch: grid-stride over groups of 4 paths (one Philox block feeds 4 paths),
// per-thread FP64 sums, compensated block/grid reduction (pcf_device.hpp).
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
__global__ __launch_bounds__(1024) void k_eur(EurArgs a, const double* __restrict__ hz,
                                              pcf::CSum* __restrict__ partials,
                                              unsigned int* __restrict__ counter,
                                              double* __restrict__ out) {
    __shared__ pcf::CSum lds[16 * 2];
    double s = 0.0, s2 = 0.0;
    const int64_t ngroups = (a.N + 3) >> 2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
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
    if (pcf::grid_reduce<2>(acc, partials, counter, lds, tot) && threadIdx.x == 0) {
        const double sum = tot.v[0].s + tot.v[0].c;
        const double sq = tot.v[1].s + tot.v[1].c;
        const double n = (double)a.N;
        const double mean = sum / n;
        out[0] = sum;
        out[1] = sq;
        out[2] = mean;
        out[3] = a.N > 1 ? (sq - sum * mean) / (n - 1.0) : 0.0;
    }
}

}  // namespace

extern "C" int pcf_mc_eur(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                          int64_t N, double payoff_fun, const pcf_opts* opts_in,
                          pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const pcf_opts o = pcf_opts_or_default(opts_in);
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    EurArgs a{};
    a.S0 = S0;
    a.E = E;
    a.pf = payoff_fun;
    a.sigma = sigma;
    a.mu = (r - 0.5 * sigma * sigma) * T;
    a.vol = sigma * std::sqrt(T);
    a.href_mu = (r - libm_pow(sigma, 2) / 2) * T;
    const double log2e = 1.4426950408889634;
    a.S0f = (float)S0;
    a.Ef = (float)E;
    a.pff = (float)payoff_fun;
    a.mu2f = (float)(a.mu * log2e);
    a.vol2f = (float)(a.vol * log2e);
    a.N = N;
    a.seed = o.seed;

    const int mode = o.rng == PCF_RNG_HOST ? kHost : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    const int B = o.block_size ? o.block_size : 256;
    const int ipt = o.items_per_thread ? o.items_per_thread : 4;
    const int64_t ngroups = (N + 3) / 4;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((ngroups + (int64_t)B * ipt - 1) /
                                                                    ((int64_t)B * ipt),
                                                                1 << 20));
    if ((st = pcf_timing_begin(ctx))) return st;
    if (mode == kHost && (st = pcf_upload_normals(ctx, &o, N))) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)grid * 2 * sizeof(pcf::CSum)));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    auto* part = ctx->partials.as<pcf::CSum>();
    auto* cnt = ctx->counters.as<unsigned int>();
    auto* out = ctx->out.as<double>();
    const double* hz = ctx->normals.as<double>();
    if (mode == kFp32k)
        k_eur<kFp32k><<<grid, B, 0, ctx->stream>>>(a, hz, part, cnt, out);
    else if (mode == kFp64)
        k_eur<kFp64><<<grid, B, 0, ctx->stream>>>(a, hz, part, cnt, out);
    else
        k_eur<kHost><<<grid, B, 0, ctx->stream>>>(a, hz, part, cnt, out);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if ((st = pcf_timing_end(ctx, 4, res))) return st;

    const double* h = ctx->h_out;
    const double disc = std::exp(-r * T);
    // The reference returns (exp(-r*T)*result)/(double)N (src/mc_eur.cpp:26).
    res->price = (disc * h[0]) / (double)N;
    res->mean = h[2];
    res->variance = h[3];
    res->std_err = disc * std::sqrt(h[3] / (double)N);
    res->n_paths = N;
    res->n_steps = 1;
    // 13 algorithmic flops per path (normal 6, drift FMA 2, exp 1, x S0 1,
    // payoff 2, accumulate 1); bytes: host normals only in verification.
    res->flops = 13.0 * (double)N;
    res->bytes = mode == kHost ? 8.0 * (double)N : 0.0;
    res->grid_blocks = grid;
    res->block_size = B;
    return PCF_OK;
}
