// lattice.hip -- backward induction on the reference's binomial lattice.
// Reference: src/binom_vanilla_eur.cpp:5-39 (European) and
// src/binom_vanilla_amer.cpp:5-40 (American), on the beta tree of
// src/binom_embar.cpp:16-24.
//
// The reference sweeps level n = N-1 .. 0 over nodes i = 0 .. n:
//   European  v[i] = (p v[i+1] + q v[i]) / R
//   American  v[i] = max((p v[i+1] + q v[i]) / R, payoff(S0 u^i d^(n-i), E))
// from the leaves v[i] = max((S0 u^i d^(N-i) - E) pf, 0).  Every node depends
// only on its two children, so any schedule that forms each node from its
// children with the same operations reproduces the reference's bits: FP64,
// no contraction, IEEE division, std::max's "a < b ? b : a", and u^i, d^k
// from tables made on the host with glibc pow (the call the reference makes).
//
// Schedule: one launch advances kLatLevels levels.  Block b owns output nodes
// [b W, b W + W) of the last of them; it loads the W + kLatLevels nodes of the
// first level it needs into LDS and steps down level by level (ping-pong LDS
// rows, one barrier per level).  The kLatLevels-wide right edge it forms is
// also formed by its neighbour, with identical values.  Device data: two
// level rows of N + 1 doubles and the tables u^i, d^k (i, k = 0..N), kept per
// context for the last (u, d, N).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pcf_internal.hpp"

namespace {

constexpr int kLatThreads = 256;
constexpr int kLatW = 512;        // output nodes per block and launch
constexpr int kLatLevels = 128;   // levels per launch
constexpr int64_t kLatMaxN = (int64_t)1 << 21;

struct LatParams {
    double S0, E, pf, p, q, R;
    int64_t N;
};

// std::max(a, b) as the reference calls it.
__device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }

// payoff(St, E, pf) of include/common.h:56-60: max(pf * (St - E), 0).
__device__ __forceinline__ double ref_payoff(double St, double E, double pf) {
#pragma clang fp contract(off)
    return ref_max(pf * (St - E), 0.0);
}

template <bool AMER>
__global__ __launch_bounds__(kLatThreads) void k_lat_leaves(LatParams P, const double* __restrict__ up,
                                                            const double* __restrict__ dn, double* __restrict__ v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= P.N;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma clang fp contract(off)
        const double S = P.S0 * up[i] * dn[P.N - i];
        // European: std::max(((S0*pow(u,i)*pow(d,N-i))-E)*payoff_fun, 0.0)
        // (src/binom_vanilla_eur.cpp:29); American: payoff(...) (src/binom_vanilla_amer.cpp:29).
        v[i] = AMER ? ref_payoff(S, P.E, P.pf) : ref_max((S - P.E) * P.pf, 0.0);
    }
}

// Levels n0-1 .. n0-K from level n0 (vin) into vout.  The grid covers the
// n0-K+1 nodes of the last level; the host checks K <= kLatLevels.
template <bool AMER>
__global__ __launch_bounds__(kLatThreads) void k_lat_steps(LatParams P, const double* __restrict__ up,
                                                           const double* __restrict__ dn,
                                                           const double* __restrict__ vin, double* __restrict__ vout,
                                                           int64_t n0, int K) {
    __shared__ double ra[kLatW + kLatLevels], rb[kLatW + kLatLevels];
    __shared__ double su[AMER ? kLatW + kLatLevels : 1];       // u^j, j = j0 + l
    __shared__ double sd[AMER ? kLatW + 2 * kLatLevels : 1];   // d^k, k = kmin + t
    const int64_t j0 = (int64_t)blockIdx.x * kLatW;
    // Level n0 nodes this block needs: j0 .. j0 + W + K - 1 (clipped to n0).
    const int nin = (int)(std::min<int64_t>(j0 + kLatW + K, n0 + 1) - j0);
    for (int l = threadIdx.x; l < nin; l += blockDim.x) ra[l] = vin[j0 + l];
    int64_t kmin = 0;
    if constexpr (AMER) {
        for (int l = threadIdx.x; l < nin; l += blockDim.x) su[l] = up[j0 + l];
        // d^(n - j) for n in [n0-K, n0-1], j in [j0, j0 + W + K - 2], n - j >= 0.
        kmin = std::max<int64_t>(0, n0 - j0 - kLatW - 2 * K + 2);
        const int nd = (int)(n0 - 1 - j0 - kmin + 1);
        for (int t = threadIdx.x; t < nd; t += blockDim.x) sd[t] = dn[kmin + t];
    }
    __syncthreads();
    double* cur = ra;
    double* nxt = rb;
    for (int s = 1; s <= K; ++s) {
        const int64_t n = n0 - s;
        const int cnt = (int)(std::min<int64_t>(j0 + kLatW + K - s, n + 1) - j0);
        for (int l = threadIdx.x; l < cnt; l += blockDim.x) {
#pragma clang fp contract(off)
            double v = (P.p * cur[l + 1] + P.q * cur[l]) / P.R;
            if constexpr (AMER) {
                const double S = P.S0 * su[l] * sd[(n - (j0 + l)) - kmin];
                v = ref_max(v, ref_payoff(S, P.E, P.pf));
            }
            nxt[l] = v;
        }
        __syncthreads();
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    const int nout = (int)(std::min<int64_t>(j0 + kLatW, n0 - K + 1) - j0);
    for (int l = threadIdx.x; l < nout; l += blockDim.x) vout[j0 + l] = cur[l];
}

}  // namespace

extern "C" int pcf_binom_vanilla(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T, int64_t N,
                                 double payoff_fun, int32_t american, const pcf_opts* opts_in, pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_opts o = pcf_opts_or_default(opts_in);
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    if (N > kLatMaxN)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "lattice N=%lld exceeds %lld", (long long)N, (long long)kLatMaxN);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    // Tree parameters exactly as src/binom_vanilla_eur.cpp:16-24 (host, glibc).
    const double dt = (double)T / (double)N;
    const double beta = 0.5 * (std::exp(-r * dt) + std::exp((r + libm_pow(sigma, 2)) * dt));
    const double u = beta + std::sqrt(libm_pow(beta, 2) - 1);
    const double d = beta - std::sqrt(libm_pow(beta, 2) - 1);
    const double R = std::exp(r * dt);
    const double p = (R - d) / (u - d);
    const double q = 1 - p;
    const LatParams P{S0, E, payoff_fun, p, q, R, N};

    // Tables u^i, d^k (i, k = 0..N) -- resident for the last (u, d, N).
    const size_t row = (size_t)N + 1;
    uint64_t key[3];
    std::memcpy(&key[0], &u, 8);
    std::memcpy(&key[1], &d, 8);
    key[2] = (uint64_t)N;
    if ((st = pcf_timing_begin(ctx))) return st;
    PCF_HIP_CHECK(ctx, ctx->lattice.ensure(4 * row * sizeof(double)));
    double* up = ctx->lattice.as<double>();
    double* dn = up + row;
    double* A = dn + row;
    double* B = A + row;
    if (std::memcmp(key, ctx->lattice_key, sizeof key) != 0 || ctx->lattice_ptr != (void*)up) {
        std::vector<double> tab(2 * row);
        for (size_t i = 0; i < row; ++i) {
            tab[i] = libm_pow(u, (double)i);        // pow(u,i), src/binom_vanilla_eur.cpp:29
            tab[row + i] = libm_pow(d, (double)i);  // pow(d,N-i)
        }
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(up, tab.data(), 2 * row * sizeof(double), hipMemcpyHostToDevice,
                                          ctx->stream));
        PCF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        std::memcpy(ctx->lattice_key, key, sizeof key);
        ctx->lattice_ptr = (void*)up;
    }

    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    const int lgrid = (int)std::min<int64_t>(4096, ((int64_t)row + kLatThreads - 1) / kLatThreads);
    if (american) k_lat_leaves<true><<<lgrid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A);
    else k_lat_leaves<false><<<lgrid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    int launches = 1;
    for (int64_t n0 = N; n0 > 0;) {
        const int K = (int)std::min<int64_t>(kLatLevels, n0);
        const int64_t outs = n0 - K + 1;
        const int grid = (int)((outs + kLatW - 1) / kLatW);
        if (american) k_lat_steps<true><<<grid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A, B, n0, K);
        else k_lat_steps<false><<<grid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A, B, n0, K);
        PCF_HIP_CHECK(ctx, hipGetLastError());
        std::swap(A, B);
        n0 -= K;
        ++launches;
    }
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->out.ptr, A, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((st = pcf_timing_end(ctx, 1, res))) return st;

    res->price = ctx->h_out[0];  // v_ij[0], returned as is (src/binom_vanilla_eur.cpp:38)
    res->mean = res->price;
    res->n_paths = 1;
    res->n_steps = N;
    // Node updates: (N+1)(N+2)/2 - 1 over the sweep; 4 flops each European
    // (2 mul, add, div), 9 American (+ 2 mul for S, sub, mul, 2 max).
    const double nodes = 0.5 * (double)(N + 1) * (double)(N + 2) - 1.0;
    res->flops = nodes * (american ? 9.0 : 4.0);
    res->bytes = 0.0;
    res->grid_blocks = launches;
    res->block_size = kLatThreads;
    res->batch_size = 1;
    return PCF_OK;
}
