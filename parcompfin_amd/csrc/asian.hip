// asian.hip -- fixed-strike arithmetic-average Asian option.
// Hot loop of src/mc_asia.cpp:27-37, per path and step m:
//     I += S (1 + r dt/2 + sigma dB/2);   S *= exp((r - sigma^2/2) dt + sigma dB)
// then payoff(I/M).  (I is the reference's first-order trapezoid of the
// running integral: it includes S_0 and excludes S_M; reproduced as written.)
//
// One path per lane, paths contiguous across lanes; one Philox call feeds 4
// steps.  FP32 form: S and the exponent in FP32 on the hardware exp2 unit, I
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
    double invM;
    int64_t N;
    int32_t M;
    uint64_t seed;
};

enum { kFp32k = 0, kFp64 = 1, kHost = 2 };

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

template <int MODE, bool CONTROL>
__global__ __launch_bounds__(1024) void k_asia(AsiaArgs a, const double* __restrict__ hz,
                                               pcf::CSum* __restrict__ partials,
                                               unsigned int* __restrict__ counter,
                                               double* __restrict__ out) {
    constexpr int K = CONTROL ? 4 : 2;
    __shared__ pcf::CSum lds[16 * K];
    double s = 0.0, s2 = 0.0, g = 0.0, g2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.N; p += stride) {
        double v, gv = 0.0;
        if (MODE == kFp32k) {
            float pv, pg = 0.f;
            asia_path_f32<CONTROL>(a, (uint64_t)p, pv, pg);
            v = pv;
            gv = pg;
        } else if (MODE == kFp64) {
            asia_path_f64<CONTROL>(a, (uint64_t)p, v, gv);
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
    if (pcf::grid_reduce<K>(acc, partials, counter, lds, tot) && threadIdx.x == 0) {
        const double n = (double)a.N;
#pragma unroll
        for (int k = 0; k < K; k += 2) {
            const double sum = tot.v[k].s + tot.v[k].c;
            const double sq = tot.v[k + 1].s + tot.v[k + 1].c;
            const double mean = sum / n;
            out[2 * k + 0] = sum;
            out[2 * k + 1] = sq;
            out[2 * k + 2] = mean;
            out[2 * k + 3] = a.N > 1 ? (sq - sum * mean) / (n - 1.0) : 0.0;
        }
    }
}

template <int MODE, bool CONTROL>
void launch_asia(pcf_ctx* ctx, int grid, int B, const AsiaArgs& a) {
    k_asia<MODE, CONTROL><<<grid, B, 0, ctx->stream>>>(
        a, ctx->normals.as<double>(), ctx->partials.as<pcf::CSum>(),
        ctx->counters.as<unsigned int>(), ctx->out.as<double>());
}

}  // namespace

extern "C" int pcf_mc_asia(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                           int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts_in,
                           pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const pcf_opts o = pcf_opts_or_default(opts_in);
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    if (M <= 0) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "M must be positive (M=%d)", M);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    const double dt = (double)T / (double)M;
    const double sdt = std::sqrt(dt);
    const double log2e = 1.4426950408889634;
    AsiaArgs a{};
    a.S0 = S0;
    a.E = E;
    a.pf = payoff_fun;
    a.a1 = 1.0 + r * dt / 2;
    a.a2 = sigma * sdt / 2;
    a.mu = (r - 0.5 * sigma * sigma) * dt;
    a.vol = sigma * sdt;
    a.h_c1 = 1 + r * dt / 2;
    a.h_mu = (r - libm_pow(sigma, 2) / 2) * dt;
    a.sigma = sigma;
    a.S0f = (float)S0;
    a.Ef = (float)E;
    a.pff = (float)payoff_fun;
    a.a1f = (float)a.a1;
    a.a2f = (float)a.a2;
    a.mu2f = (float)(a.mu * log2e);
    a.vol2f = (float)(a.vol * log2e);
    a.invM = 1.0 / (double)M;
    a.N = N;
    a.M = M;
    a.seed = o.seed;

    const int mode = o.rng == PCF_RNG_HOST ? kHost : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    const bool control = o.control != 0 && mode != kHost;
    const int B = o.block_size ? o.block_size : 256;
    const int ipt = o.items_per_thread ? o.items_per_thread : 1;
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((N + (int64_t)B * ipt - 1) / ((int64_t)B * ipt), 1 << 20));
    const int K = control ? 4 : 2;

    if ((st = pcf_timing_begin(ctx))) return st;
    if (mode == kHost && (st = pcf_upload_normals(ctx, &o, N * (int64_t)M))) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)grid * K * sizeof(pcf::CSum)));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    if (mode == kFp32k) {
        if (control) launch_asia<kFp32k, true>(ctx, grid, B, a);
        else launch_asia<kFp32k, false>(ctx, grid, B, a);
    } else if (mode == kFp64) {
        if (control) launch_asia<kFp64, true>(ctx, grid, B, a);
        else launch_asia<kFp64, false>(ctx, grid, B, a);
    } else {
        launch_asia<kHost, false>(ctx, grid, B, a);
    }
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if ((st = pcf_timing_end(ctx, control ? 8 : 4, res))) return st;

    const double* h = ctx->h_out;
    const double disc = std::exp(-r * T);
    // The reference returns exp(-r*T)*result/(double)N (src/mc_asia.cpp:39).
    res->price = disc * h[0] / (double)N;
    res->mean = h[2];
    res->variance = h[3];
    res->std_err = disc * std::sqrt(h[3] / (double)N);
    if (control) {
        res->control_price = disc * h[4] / (double)N;
        res->control_se = disc * std::sqrt(h[7] / (double)N);
    }
    res->n_paths = N;
    res->n_steps = M;
    // 14 algorithmic flops per path-step: normal 6, I update 2 FMAs (4),
    // S update FMA + exp + mul (4).  No HBM traffic in the Philox forms.
    res->flops = 14.0 * (double)N * (double)M;
    res->bytes = mode == kHost ? 8.0 * (double)N * (double)M : 0.0;
    res->grid_blocks = grid;
    res->block_size = B;
    return PCF_OK;
}
