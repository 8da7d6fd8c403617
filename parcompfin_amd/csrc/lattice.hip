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
// Schedule: one launch advances K levels (LatTile).  Each wavefront owns
// Wout = 64 R - K output nodes of the last of them and holds the 64 R nodes
// of the first level it needs in registers, R consecutive nodes per lane.
// A level forms every node from itself and its right neighbour; the right
// neighbour of a lane's last node is the next lane's first node, one DPP
// wave shift away.  The span of valid nodes shrinks by one per level, so
// after K levels the wave's first Wout nodes are exact and the waves need no
// barrier or LDS at all.  The K-wide right edge a wave forms is also formed
// by its neighbour, with identical values.  Device data: two level rows of
// N + 1 doubles and the tables u^i, d^k (i, k = 0..N), kept per context for
// the last (u, d, N).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pcf_internal.hpp"

namespace {

constexpr int kLatThreads = 256;
constexpr int64_t kLatMaxN = (int64_t)1 << 21;

// Tile shape per exercise style: R nodes per lane, K levels per launch, Wout =
// 64 R - K output nodes per wave and launch.  A level costs a wave about R
// node updates in sequence, and a large lattice holds only ~n/Wout waves for
// the 1024 SIMDs, so small tiles run faster per level than they lose to
// redundant edges and extra launches.  Measured over 46 lattices (N = 513 ..
// 10^5) against R = 4, K = 128: American R = 1, K = 32 171.5 -> 104.9 ms in
// total; European R = 2, K = 64 69.3 -> 51.5 ms; bits unchanged.
template <bool AMER>
struct LatTile {
    static constexpr int R = AMER ? 1 : 2;
    static constexpr int K = AMER ? 32 : 64;
    static constexpr int Wout = 64 * R - K;
    static_assert(Wout > 0, "a wave must keep some exact output nodes");
};

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

// Right neighbour across lanes: lane l receives lane l+1's value (DPP
// wave_shl:1; lane 63 receives 0, which only ever reaches nodes past the
// wave's valid span).  FP64 as two 32-bit halves.
__device__ __forceinline__ double from_next_lane(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x130, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x130, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Levels n0-1 .. n0-K from level n0 (vin) into vout.  The grid covers the
// n0-K+1 nodes of the last level with waves of LatTile::Wout outputs; the
// host keeps K <= LatTile::K.  Every lane runs every level (no early exit), so the
// cross-lane reads always see active lanes.
template <bool AMER>
__global__ __launch_bounds__(kLatThreads) void k_lat_steps(LatParams P, const double* __restrict__ up,
                                                           const double* __restrict__ dn,
                                                           const double* __restrict__ vin, double* __restrict__ vout,
                                                           int64_t n0, int K) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    constexpr int kR = LatTile<AMER>::R, kWout = LatTile<AMER>::Wout;
    const int64_t j0 = wave * kWout + (int64_t)lane * kR;  // this lane's first node
    // v: node values; AMER: u = u^j (fixed per node) and d = d^(n-j) for the
    // level n being formed.  d^(n-j) of node j at level n-1 is d^(n-(j+1)) of
    // node j+1 at level n, so d shifts one node left per level (the last node
    // of a lane takes the next lane's first); only lane 63's last node takes a
    // value it does not own, and that node is past the valid span already.
    double v[kR], u[kR], d[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const int64_t j = j0 + r;
        v[r] = j <= n0 ? vin[j] : 0.0;
        if constexpr (AMER) {
            u[r] = j <= n0 ? up[j] : 0.0;
            d[r] = j <= n0 - 1 ? dn[n0 - 1 - j] : 0.0;
        }
    }
    for (int s = 1; s <= K; ++s) {
        const int64_t n = n0 - s;
        const double nb = from_next_lane(v[0]);
        double w[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
#pragma clang fp contract(off)
            const double right = r + 1 < kR ? v[r + 1 < kR ? r + 1 : 0] : nb;
            w[r] = (P.p * right + P.q * v[r]) / P.R;
            if constexpr (AMER) {
                const int64_t j = j0 + r;
                if (j <= n) {
                    const double S = P.S0 * u[r] * d[r];
                    w[r] = ref_max(w[r], ref_payoff(S, P.E, P.pf));
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) v[r] = w[r];
        if constexpr (AMER) {
            const double dnext = from_next_lane(d[0]);
#pragma unroll
            for (int r = 0; r + 1 < kR; ++r) d[r] = d[r + 1];
            d[kR - 1] = dnext;
        }
    }
    // The wave's first kWout nodes of level n0-K (and nothing past it).
    const int64_t wend = wave * kWout + kWout;
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const int64_t j = j0 + r;
        if (j < wend && j <= n0 - K) vout[j] = v[r];
    }
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

    PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, HUGE_VAL));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    const int lgrid = (int)std::min<int64_t>(4096, ((int64_t)row + kLatThreads - 1) / kLatThreads);
    if (american) k_lat_leaves<true><<<lgrid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A);
    else k_lat_leaves<false><<<lgrid, kLatThreads, 0, ctx->stream>>>(P, up, dn, A);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    int launches = 1;
    const int kmax = american ? LatTile<true>::K : LatTile<false>::K;
    const int wout = american ? LatTile<true>::Wout : LatTile<false>::Wout;
    for (int64_t n0 = N; n0 > 0;) {
        const int K = (int)std::min<int64_t>(kmax, n0);
        const int64_t outs = n0 - K + 1;
        const int64_t waves = (outs + wout - 1) / wout;
        const int grid = (int)((waves * 64 + kLatThreads - 1) / kLatThreads);
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
