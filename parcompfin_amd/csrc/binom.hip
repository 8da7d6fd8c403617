// binom.hip -- closed binomial sum over the beta tree (src/binom_embar.cpp:5-50).
//
//   price = e^{-rT} sum_i C(N,i) p^i q^{N-i} payoff(S0 u^i d^{N-i})
//
// with the reference's mirrored pairing: term index i < ceil(N/2) contributes
// the terms i and N-i sharing one log C(N,i), and for even N the middle term
// i = N/2 once.  Here lane i (0 <= i < ceil(N/2), plus lane N/2 for even N)
// owns term i and, if i < ceil(N/2), its mirror N-i; the middle term's
// expression comb(N,N/2) + N/2 log p + N/2 log q is exactly term 1 of lane N/2.
//
// Two coefficient forms:
//  * chain (reference-exact): log C(N,i) replayed bit for bit as the
//    reference's sequential chain (include/common.h:63-72):
//        acc = sum_{n=N..i+1} log n  (descending),  then acc -= log j, j=2..N-i
//    with glibc log values from a host table.  The descending part of every
//    term is a prefix of one chain (log N + log(N-1) + ... in that order), so
//    it comes from a host-made prefix table, rounding for rounding; only the
//    second part runs per lane.  Its table index is uniform across a
//    wavefront: the block stages the table through LDS and every wave reads
//    its values as LDS broadcasts; lanes that have reached their own end are
//    masked, never reordered.
//  * double-double (production): one lane per term i = 0..N, weight
//        t = lf[N] - lf[i] - lf[N-i] + i log p + (N-i) log q,   w = exp(t)
//    in double-double arithmetic from a resident table lf[k] = log k! (hi,
//    lo).  The FP64 forms of these logs round at ulp(log N!) = 1.9e-9 for
//    N = 1e6; the reference's own chain carries that rounding into a price
//    1.5e-8 from the exact sum, while this form stays within 1e-15 of it.
#include <cmath>
#include <mutex>
#include <vector>

#include "pcf_dd.hpp"
#include "pcf_internal.hpp"

namespace {

struct BinomArgs {
    double S0, E, pf, u, d, logp, logq;
    int64_t N;
    int64_t lanes;   // number of term lanes
    int64_t until;   // ceil(N/2): lanes below this also own the mirror term
};

__device__ __forceinline__ double term_value(const BinomArgs& a, double comb_val,
                                             int64_t k /*exponent of u*/) {
#pragma clang fp contract(off)
    // binom = comb_val + k*log(p) + (N-k)*log(q)  (src/binom_embar.cpp:36-37)
    const double b = comb_val + (double)k * a.logp + (double)(a.N - k) * a.logq;
    const double S = a.S0 * pow(a.u, (double)k) * pow(a.d, (double)(a.N - k));
    const double v = a.pf * (S - a.E);
    return exp(b) * (v > 0.0 ? v : 0.0);
}

// Chain form.  One lane per term index; all 64 lanes of a wave run the same
// trip counts (wave-uniform loop bounds from the wave's first index).
//
// The block copies the table through LDS, kStageChunk entries at a time
// (vector loads of the next chunk in flight while the current one is
// subtracted, one barrier per chunk, the same chunk count for every wave of
// the block), and each wave reads its values as LDS broadcasts.  Scalar loads
// of the table can return out of order, so a use waits for every scalar load
// outstanding and their prefetch cannot run deeper than one batch; LDS reads
// complete in order.
constexpr int kStageChunk = 1024;

// acc - v of lane 16 r + K, for every lane of row r (r = 0..3), as one
// v_fmac_f64 whose first source is read through the DPP row broadcast:
// acc + v * -1 rounds exactly as acc - v.  Only VALU operands: the value
// comes from an LDS load, which the compiler's wait counts cover.
#define PCF_SUB_BCAST(K)                                                                   \
    asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #K " row_mask:0xf bank_mask:0xf" \
                 : "+v"(acc)                                                                \
                 : "v"(v), "v"(neg1))
// The 16 values of each row, subtracted in lane order.  A DPP read of a VGPR
// needs two wait states after a VALU write of it (v may have just been moved
// into place), which the compiler does not see inside the assembly text: the
// s_nop 1 provides them.
__device__ __forceinline__ void sub_row16(double& acc, double v, double neg1) {
    asm volatile("s_nop 1");
    PCF_SUB_BCAST(0);
    PCF_SUB_BCAST(1);
    PCF_SUB_BCAST(2);
    PCF_SUB_BCAST(3);
    PCF_SUB_BCAST(4);
    PCF_SUB_BCAST(5);
    PCF_SUB_BCAST(6);
    PCF_SUB_BCAST(7);
    PCF_SUB_BCAST(8);
    PCF_SUB_BCAST(9);
    PCF_SUB_BCAST(10);
    PCF_SUB_BCAST(11);
    PCF_SUB_BCAST(12);
    PCF_SUB_BCAST(13);
    PCF_SUB_BCAST(14);
    PCF_SUB_BCAST(15);
}
#undef PCF_SUB_BCAST
__global__ __launch_bounds__(256) void k_binom_chain(BinomArgs a, const double* __restrict__ tab,
                                                     const double* __restrict__ pre,
                                                     pcf::CSum* __restrict__ partials,
                                                     unsigned int* __restrict__ counter,
                                                     double* __restrict__ out) {
    __shared__ pcf::CSum lds[4];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // First index of this wave (uniform).
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(pcf::kWave - 1));
    const int N = (int)a.N;
    const int w0 = __builtin_amdgcn_readfirstlane((int)i0);
    // Part 1: n = N down to i+1, i.e. the first N-i steps of the descending
    // chain (pre[k], k = 0..N; lanes past N/2 + 1 carry no term).
    const int64_t k1 = a.N - i;
    double acc = pre[k1 < 0 ? 0 : (k1 > a.N ? a.N : k1)];
    // Part 2: j = 2 .. N-i.  Unmasked while j <= N-(i0+63).
    {
        const int lo = (int)(threadIdx.x & 15);
        const double neg1 = -1.0;
        __shared__ double stage[2][kStageChunk];
        const int b0 = __builtin_amdgcn_readfirstlane((int)((int64_t)blockIdx.x * blockDim.x));
        const int blk_hi = N - b0;                   // largest j of any lane of the block
        const int hi_all = N - (w0 + pcf::kWave - 1);  // this wave: unmasked up to here
        const int hi_any = N - w0;                   // masked up to here (lane i stops at N - i)
        const int nchunks = blk_hi >= 2 ? (blk_hi - 2) / kStageChunk + 1 : 0;
        constexpr int kPer = kStageChunk / 256;      // entries per thread per chunk
        double v[kPer];
        auto fetch = [&](int c) {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int e = 2 + c * kStageChunk + u * 256 + (int)threadIdx.x;
                v[u] = tab[e < N ? e : N];
            }
        };
        if (nchunks > 0) {
            fetch(0);
#pragma unroll
            for (int u = 0; u < kPer; ++u) stage[0][u * 256 + threadIdx.x] = v[u];
        }
        __syncthreads();
        for (int c = 0; c < nchunks; ++c) {
            const bool more = c + 1 < nchunks;
            if (more) fetch(c + 1);
            const double* S = stage[c & 1];
            const int base = 2 + c * kStageChunk;
            const int last = base + kStageChunk - 1;
            const int e_all = last < hi_all ? last : hi_all;
            int j = base;
            // Groups of 16: lane l loads entry j + l%16 (each 16-lane row holds
            // the same 16 values) and subtracts the row's 16 values in order,
            // one broadcast DPP instruction each; the next group loads meanwhile.
            if (j + 15 <= e_all) {
                double vcur = S[j - base + lo];
                for (; j + 15 <= e_all; j += 16) {
                    // unconditional (clamped into the chunk; unused past e_all)
                    const int nx = j + 16 - base + lo;
                    const double vnext = S[nx < kStageChunk ? nx : kStageChunk - 1];
                    sub_row16(acc, vcur, neg1);
                    vcur = vnext;
                }
            }
            for (; j <= e_all; ++j) acc -= S[j - base];
            const int e_any = last < hi_any ? last : hi_any;
            for (j = j > hi_all + 1 ? j : hi_all + 1; j <= e_any; ++j) {
                const double t = S[j - base];
                acc = (j <= N - i) ? acc - t : acc;
            }
            if (more) {
                // buffer (c+1)&1 was last read in chunk c-1, which every wave
                // finished before the previous barrier
#pragma unroll
                for (int u = 0; u < kPer; ++u) stage[(c + 1) & 1][u * 256 + threadIdx.x] = v[u];
            }
            __syncthreads();
        }
    }
    double comb_val = acc;
    if (i == 0) comb_val = 0.0;                // i==0 || i==N
    else if (i == 1) comb_val = tab[N];        // i==1 || i==N-1: log(N)
    const double s = (i < a.lanes) ? term_value(a, comb_val, i) : 0.0;
    const double s_b = (i < a.until) ? term_value(a, comb_val, a.N - i) : 0.0;
    pcf::Acc<1> ac, tot;
    ac.v[0] = {s, 0.0};
    pcf::csum_add(ac.v[0], s_b);
    if (pcf::grid_reduce<1>(ac, partials, counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0) {
        out[0] = tot.v[0].s + tot.v[0].c;
    }
}

struct BinomDDArgs {
    double S0, E, pf, u, d;
    pcf::DD lp, lq;  // log p, log q
    int64_t N;
};

// Double-double form: lane i prices term i (0 <= i <= N).  lf holds log k! as
// (hi, lo) pairs; the weight's logs are combined without rounding error
// beyond ~1e-32 relative, then w = e^{t.hi} (1 + t.lo).
__global__ __launch_bounds__(256) void k_binom_dd(BinomDDArgs a, const double2* __restrict__ lf,
                                                  pcf::CSum* __restrict__ partials,
                                                  unsigned int* __restrict__ counter,
                                                  double* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ pcf::CSum lds[4];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i <= a.N) {
        const double2 fN = lf[a.N], fi = lf[i], fk = lf[a.N - i];
        pcf::DD t = pcf::dd_add(pcf::DD{fN.x, fN.y}, pcf::DD{-fi.x, -fi.y});
        t = pcf::dd_add(t, pcf::DD{-fk.x, -fk.y});
        t = pcf::dd_add(t, pcf::dd_mul_d(a.lp, (double)i));
        t = pcf::dd_add(t, pcf::dd_mul_d(a.lq, (double)(a.N - i)));
        const double e = exp(t.h);
        const double w = e + e * t.l;
        if (w != 0.0) {
            const double S = a.S0 * pow(a.u, (double)i) * pow(a.d, (double)(a.N - i));
            const double v = a.pf * (S - a.E);
            s = w * (v > 0.0 ? v : 0.0);
        }
    }
    pcf::Acc<1> ac, tot;
    ac.v[0] = {s, 0.0};
    if (pcf::grid_reduce<1>(ac, partials, counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0)
        out[0] = tot.v[0].s + tot.v[0].c;
}

// Process-wide host table lf[k] = log k! (hi, lo), k = 0..n, grown on demand
// and never shrunk.  Built from IEEE double operations only:
//   log 2 as a double-double constant; log k = log(k-1) + 2 atanh(1/(2k-1))
//   for k >= 3 (series in double-double); lf[k] = lf[k-1] + log k.
// Each step rounds at ~1e-32 relative, so lf[10^6] ~ 1.3e7 is within ~1e-20
// of log 10^6!.
struct LogFactTable {
    std::mutex mu;
    std::vector<double> v;           // 2 (n + 1) doubles
    pcf::DD Lk{0.0, 0.0}, lf{0.0, 0.0};  // log k and log k! of the last entry
    int64_t n = -1;
};
LogFactTable& logfact_table() {
    static LogFactTable t;
    return t;
}

// Makes ctx->logfact hold lf[0..N] (at least).  Resident: a context keeps the
// largest table it was asked for.
int ensure_logfact(pcf_ctx* ctx, int64_t N) {
    if (ctx->logfact_n >= N) return PCF_OK;
    LogFactTable& T = logfact_table();
    std::lock_guard<std::mutex> lk(T.mu);
    if (T.n < N) {
        T.v.resize(2 * ((size_t)N + 1));
        for (int64_t k = T.n + 1; k <= N; ++k) {
            if (k >= 2) {
                if (k == 2) T.Lk = pcf::kDdLn2;
                else T.Lk = pcf::dd_add(T.Lk, pcf::dd_2atanh(pcf::dd_div_d(pcf::DD{1.0, 0.0}, (double)(2 * k - 1))));
                T.lf = pcf::dd_add(T.lf, T.Lk);
            }
            T.v[2 * (size_t)k] = T.lf.h;
            T.v[2 * (size_t)k + 1] = T.lf.l;
        }
        T.n = N;
    }
    const size_t bytes = 2 * ((size_t)T.n + 1) * sizeof(double);
    PCF_HIP_CHECK(ctx, ctx->logfact.ensure(bytes));
    PCF_HIP_CHECK(ctx, hipMemcpy(ctx->logfact.ptr, T.v.data(), bytes, hipMemcpyHostToDevice));
    ctx->logfact_n = T.n;
    return PCF_OK;
}

// glibc log(n) for n = 0..N (entry 0 unused) and the descending chain's
// prefixes pre[k] = (..((0 + log N) + log(N-1)) + ..) + log(N-k+1), k = 0..N,
// made in that order on the host (plain IEEE double adds), resident per N.
int ensure_logtab(pcf_ctx* ctx, int64_t N) {
    if (ctx->logtab_n == N && ctx->prefix_n == N) return PCF_OK;
    std::vector<double> t((size_t)N + 1), pre((size_t)N + 1);
    t[0] = 0.0;
    for (int64_t n = 1; n <= N; ++n) t[(size_t)n] = std::log((double)n);
    volatile double acc = 0.0;  // one rounding per add, in the reference's order
    pre[0] = 0.0;
    for (int64_t k = 1; k <= N; ++k) {
        acc = acc + t[(size_t)(N - k + 1)];
        pre[(size_t)k] = acc;
    }
    PCF_HIP_CHECK(ctx, ctx->logtab.ensure(t.size() * sizeof(double)));
    PCF_HIP_CHECK(ctx, ctx->prefix.ensure(pre.size() * sizeof(double)));
    PCF_HIP_CHECK(ctx, hipMemcpy(ctx->logtab.ptr, t.data(), t.size() * sizeof(double),
                                 hipMemcpyHostToDevice));
    PCF_HIP_CHECK(ctx, hipMemcpy(ctx->prefix.ptr, pre.data(), pre.size() * sizeof(double),
                                 hipMemcpyHostToDevice));
    ctx->logtab_n = N;
    ctx->prefix_n = N;
    return PCF_OK;
}

}  // namespace

extern "C" int pcf_binom_embar(pcf_ctx* ctx, double S0, double E, double r, double sigma,
                               double T, int64_t N, double payoff_fun, const pcf_opts* opts_in,
                               pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_opts o = pcf_opts_or_default(opts_in);
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    if (N > (int64_t)2147483647 - 64)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "N exceeds the reference's int range");
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    // Tree parameters exactly as src/binom_embar.cpp:16-24 (host, glibc).
    const double dt = (double)T / (double)N;
    const double beta = 0.5 * (std::exp(-r * dt) + std::exp((r + libm_pow(sigma, 2)) * dt));
    const double u = beta + std::sqrt(libm_pow(beta, 2) - 1);
    const double d = beta - std::sqrt(libm_pow(beta, 2) - 1);
    const double R = std::exp(r * dt);
    const double p = (R - d) / (u - d);
    const double q = 1 - p;
    BinomArgs a{};
    a.S0 = S0;
    a.E = E;
    a.pf = payoff_fun;
    a.u = u;
    a.d = d;
    a.logp = std::log(p);
    a.logq = std::log(q);
    a.N = N;
    a.until = (N % 2 != 0) ? (N + 1) / 2 : N / 2;
    a.lanes = a.until + ((N % 2 == 0) ? 1 : 0);

    const int B = 256;
    const bool chain = o.ref_exact != 0;
    const int grid = (int)(((chain ? a.lanes : N + 1) + B - 1) / B);
    BinomDDArgs ad{};
    ad.S0 = S0;
    ad.E = E;
    ad.pf = payoff_fun;
    ad.u = u;
    ad.d = d;
    ad.lp = pcf::dd_log(p);
    ad.lq = pcf::dd_log(q);
    ad.N = N;
    if ((st = pcf_timing_begin(ctx))) return st;
    if (chain && (st = ensure_logtab(ctx, N))) return st;
    if (!chain && (st = ensure_logfact(ctx, N))) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)grid * sizeof(pcf::CSum)));
    std::unique_lock<std::mutex> order;  // held until the end event is recorded
    PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, HUGE_VAL, &order));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    if (chain)
        k_binom_chain<<<grid, B, 0, ctx->stream>>>(a, ctx->logtab.as<double>(), ctx->prefix.as<double>(),
                                                  ctx->partials.as<pcf::CSum>(),
                                                  ctx->counters.as<unsigned int>(),
                                                  ctx->out.as<double>());
    else
        k_binom_dd<<<grid, B, 0, ctx->stream>>>(ad, ctx->logfact.as<double2>(), ctx->partials.as<pcf::CSum>(),
                                               ctx->counters.as<unsigned int>(), ctx->out.as<double>());
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if (order.owns_lock()) order.unlock();
    if ((st = pcf_timing_end(ctx, 1, res))) return st;
    // The reference returns exp(-r*T)*result (src/binom_embar.cpp:49).
    res->mean = ctx->h_out[0];
    res->price = std::exp(-r * T) * ctx->h_out[0];
    res->n_paths = chain ? a.lanes : N + 1;
    res->n_steps = N;
    // chain: the reference's 0.75 N^2 log-adds become table adds
    // (sum over lanes of (N-i) + (N-i-1)); double-double: ~20 algorithmic
    // flops per term (the reference's term arithmetic, counted once).
    double adds = 0.0;
    if (chain) {
        const double L = (double)a.until;
        adds = 2.0 * (double)N * L - L * (L - 1.0) - L;  // sum_{i<until} (2N - 2i - 1)
        if (N % 2 == 0) adds += (double)N - 1.0;         // middle lane
    }
    res->flops = chain ? adds + 24.0 * (double)a.lanes : 20.0 * (double)(N + 1);
    res->bytes = chain ? 8.0 * adds : 48.0 * (double)(N + 1);  // three 16-byte table entries per term
    res->grid_blocks = grid;
    res->block_size = B;
    res->batch_size = 1;
    return PCF_OK;
}
