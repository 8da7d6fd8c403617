// basket.hip -- arithmetic basket on d correlated GBM assets, one terminal step.
// Reference: src/mc_eur_multi.cpp:26-33 with mvnorm (include/mvn.h:42-82):
//     B = L Z (L = Cholesky factor of the correlation matrix, Z ~ N(0, I))
//     basket_n = sum_a w_a S0_a exp((r - s_a^2/2) T + s_a sqrt(T) B_a)
// (the reference has one scalar sigma, S0, weights 1/d, equicorrelation rho,
// and omits sqrt(T): ref_exact=1 reproduces that; all configs have T = 1).
//
// Kernels:
//  * k_basket_fma<Real, DMAX, MODE, FULL>: one path per lane; d normals in
//    registers, B = L z with scalar-broadcast L, per-asset exp and the weighted
//    sum fused; FP32/Kahan, FP64 and host-normal verification forms.
//  * k_basket_mfma: FP32, 16 <= d <= 32.  One wavefront prices 32 paths per
//    tile: B^T (32 assets x 32 paths) = L (32 x 32, zero padded) * Z^T on
//    v_mfma_f32_32x32x2f32, 16 issues per tile.  Lane l feeds path l%32 and
//    generates exactly the normals its half (l/32) contributes, so Z and B
//    never touch memory; the epilogue (exp, weights, basket sum) runs on the
//    16 accumulator registers, halves combined with one lane exchange.
#include <cmath>
#include <cstring>
#include <vector>

#include "pcf_internal.hpp"

namespace {

constexpr int kMaxAssets = 64;
enum { kFp32k = 0, kFp64 = 1, kHost = 2 };

struct BasketArgs {
    double E, pf;
    float Ef, pff;
    int64_t N;
    int32_t d;
    uint64_t seed;
};

// Device constants, one buffer: FP64 block then FP32 block.
//   FP64: L[d*d] (row-major), c0[d] = (r - s^2/2)T, c1[d] = s sqrt(T) (or s),
//         ws[d] = w S0
//   FP32: L32[32*32] zero padded, c0f[32], c1f[32] (x log2 e), wsf[32]; for
//         the FMA kernel the first d*d of Lf hold the d x d factor.
struct BasketLayout {
    size_t off_L64, off_c064, off_c164, off_ws64, off_Lf, off_Lpad, off_c0f, off_c1f, off_wsf, bytes;
    static BasketLayout make(int d) {
        BasketLayout l;
        size_t o = 0;
        l.off_L64 = o; o += sizeof(double) * d * d;
        l.off_c064 = o; o += sizeof(double) * d;
        l.off_c164 = o; o += sizeof(double) * d;
        l.off_ws64 = o; o += sizeof(double) * d;
        l.off_Lf = o; o += sizeof(float) * d * d;
        l.off_Lpad = o; o += sizeof(float) * 32 * 32;
        l.off_c0f = o; o += sizeof(float) * kMaxAssets;
        l.off_c1f = o; o += sizeof(float) * kMaxAssets;
        l.off_wsf = o; o += sizeof(float) * kMaxAssets;
        l.bytes = o;
        return l;
    }
};

template <typename Real, int DMAX, int MODE, bool FULL>
__global__ __launch_bounds__(256) void k_basket_fma(BasketArgs a, const Real* __restrict__ Lm,
                                                    const Real* __restrict__ c0,
                                                    const Real* __restrict__ c1,
                                                    const Real* __restrict__ ws,
                                                    const double* __restrict__ hz,
                                                    pcf::CSum* __restrict__ partials,
                                                    unsigned int* __restrict__ counter,
                                                    double* __restrict__ out) {
    __shared__ pcf::CSum lds[4 * 2];
    const int d = a.d;
    double s = 0.0, s2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.N; p += stride) {
        Real z[DMAX];
        if (MODE == kHost) {
#pragma unroll
            for (int k = 0; k < DMAX; ++k) {
                if (k < d) z[k] = (Real)hz[p * d + k];
                else z[k] = Real(0);
            }
        } else {
#pragma unroll
            for (int b = 0; b < DMAX / 4; ++b) {
                if (4 * b < d) {
                    Real t[4];
                    pcf::normals4<Real>(a.seed, PCF_STREAM_BASKET, (uint64_t)p, (uint32_t)b, t);
#pragma unroll
                    for (int q = 0; q < 4; ++q) z[4 * b + q] = t[q];
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) z[4 * b + q] = Real(0);
                }
            }
        }
        Real basket = 0;
#pragma unroll
        for (int i = 0; i < DMAX; ++i) {
            if (i < d) {
                Real B = 0;
                if (MODE == kHost) {
#pragma clang fp contract(off)
#pragma unroll
                    for (int k = 0; k < DMAX; ++k)
                        if ((FULL ? true : k <= i) && k < d) B += Lm[i * d + k] * z[k];
                    // (w S0) * exp((r - pow(s,2)/2) T + s B)  (src/mc_eur_multi.cpp:30)
                    basket += ws[i] * exp(c0[i] + c1[i] * B);
                } else {
#pragma unroll
                    for (int k = 0; k < DMAX; ++k)
                        if ((FULL ? true : k <= i) && k < d) B = fma(Lm[i * d + k], z[k], B);
                    if constexpr (sizeof(Real) == 4)
                        basket = fmaf(ws[i], pcf::exp2_f32(fmaf(c1[i], B, c0[i])), basket);
                    else
                        basket = fma(ws[i], exp(fma(c1[i], B, c0[i])), basket);
                }
            }
        }
        Real v = pcf::payoff<Real>(basket, (Real)a.E, (Real)a.pf);
        const double dv = (double)v;
        s += dv;
        s2 += dv * dv;
    }
    pcf::Acc<2> acc, tot;
    acc.v[0] = {s, 0.0};
    acc.v[1] = {s2, 0.0};
    if (pcf::grid_reduce<2>(acc, partials, counter, lds, tot) && threadIdx.x == 0) {
        const double sum = tot.v[0].s + tot.v[0].c, sq = tot.v[1].s + tot.v[1].c, n = (double)a.N;
        out[0] = sum;
        out[1] = sq;
        out[2] = sum / n;
        out[3] = a.N > 1 ? (sq - sum * (sum / n)) / (n - 1.0) : 0.0;
    }
}

typedef float v16f __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void k_basket_mfma(BasketArgs a, const float* __restrict__ Lpad,
                                                     const float* __restrict__ c0,
                                                     const float* __restrict__ c1,
                                                     const float* __restrict__ ws,
                                                     pcf::CSum* __restrict__ partials,
                                                     unsigned int* __restrict__ counter,
                                                     double* __restrict__ out) {
    __shared__ pcf::CSum lds[4 * 2];
    const int lane = threadIdx.x & 63;
    const int j = lane & 31;  // path within the tile; also the A-operand row (asset)
    const int h = lane >> 5;  // k half fed by this lane
    // A operand: instruction s uses k = 8*(s/4) + 4*h + (s%4), i.e. element s%4 of
    // Philox block 2*(s/4) + h; this lane supplies L[j][k].
    float Areg[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) Areg[s] = Lpad[j * 32 + 8 * (s >> 2) + 4 * h + (s & 3)];
    // Accumulator register r of this lane holds row i = 8*(r/4) + 4*h + (r%4).
    float e0[16], e1[16], ew[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r >> 2) + 4 * h + (r & 3);
        e0[r] = c0[i];
        e1[r] = c1[i];
        ew[r] = ws[i];
    }
    const int nblk = (a.d + 7) >> 3;  // k blocks of 8 actually used
    double s = 0.0, s2 = 0.0;
    const int64_t ntiles = (a.N + 31) >> 5;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t t = wave; t < ntiles; t += nwaves) {
        const int64_t p = (t << 5) + j;
        v16f acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            if (cc < nblk) {
                float z[4];
                pcf::normals4<float>(a.seed, PCF_STREAM_BASKET, (uint64_t)p, (uint32_t)(2 * cc + h), z);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Areg[4 * cc + q], z[q], acc, 0, 0, 0);
            }
        }
        float part = 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) part = fmaf(ew[r], pcf::exp2_f32(fmaf(e1[r], acc[r], e0[r])), part);
        const float basket = part + __shfl_xor(part, 32, 64);
        const float v = pcf::payoff<float>(basket, a.Ef, a.pff);
        if (h == 0 && p < a.N) {
            const double dv = (double)v;
            s += dv;
            s2 += dv * dv;
        }
    }
    pcf::Acc<2> acc2, tot;
    acc2.v[0] = {s, 0.0};
    acc2.v[1] = {s2, 0.0};
    if (pcf::grid_reduce<2>(acc2, partials, counter, lds, tot) && threadIdx.x == 0) {
        const double sum = tot.v[0].s + tot.v[0].c, sq = tot.v[1].s + tot.v[1].c, n = (double)a.N;
        out[0] = sum;
        out[1] = sq;
        out[2] = sum / n;
        out[3] = a.N > 1 ? (sq - sum * (sum / n)) / (n - 1.0) : 0.0;
    }
}

// ---- host linear algebra (FP64) ---------------------------------------------
// Lower Cholesky factor, textbook column order; false if not positive definite.
bool cholesky(int d, const std::vector<double>& C, std::vector<double>& L) {
    L.assign((size_t)d * d, 0.0);
    for (int j = 0; j < d; ++j) {
        double s = C[j * d + j];
        for (int k = 0; k < j; ++k) s -= L[j * d + k] * L[j * d + k];
        if (!(s > 0)) return false;
        const double ljj = std::sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; ++i) {
            double t = C[i * d + j];
            for (int k = 0; k < j; ++k) t -= L[i * d + k] * L[j * d + k];
            L[i * d + j] = t / ljj;
        }
    }
    return true;
}

// Symmetric eigen-decomposition by cyclic Jacobi rotations; the factor
// V sqrt(Lambda) used by mvnorm when the Cholesky factorisation fails
// (include/mvn.h:73-76).  Returns false if an eigenvalue is clearly negative.
bool eigen_factor(int d, const std::vector<double>& C, std::vector<double>& F) {
    std::vector<double> A = C, V((size_t)d * d, 0.0);
    for (int i = 0; i < d; ++i) V[i * d + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < d; ++p)
            for (int q = p + 1; q < d; ++q) off += A[p * d + q] * A[p * d + q];
        if (off < 1e-30) break;
        for (int p = 0; p < d; ++p)
            for (int q = p + 1; q < d; ++q) {
                const double apq = A[p * d + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (A[q * d + q] - A[p * d + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < d; ++k) {
                    const double akp = A[k * d + p], akq = A[k * d + q];
                    A[k * d + p] = c * akp - sn * akq;
                    A[k * d + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < d; ++k) {
                    const double apk = A[p * d + k], aqk = A[q * d + k];
                    A[p * d + k] = c * apk - sn * aqk;
                    A[q * d + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < d; ++k) {
                    const double vkp = V[k * d + p], vkq = V[k * d + q];
                    V[k * d + p] = c * vkp - sn * vkq;
                    V[k * d + q] = sn * vkp + c * vkq;
                }
            }
    }
    F.assign((size_t)d * d, 0.0);
    for (int k = 0; k < d; ++k) {
        double lam = A[k * d + k];
        if (lam < -1e-10) return false;
        lam = lam > 0 ? std::sqrt(lam) : 0.0;
        for (int i = 0; i < d; ++i) F[i * d + k] = V[i * d + k] * lam;
    }
    return true;
}

template <typename Real, int DMAX, int MODE>
void launch_fma(pcf_ctx* ctx, int grid, int B, const BasketArgs& a, bool full,
                const BasketLayout& lay, int d) {
    char* base = ctx->basket_L.as<char>();
    const bool f64 = sizeof(Real) == 8;
    const Real* Lm = (const Real*)(base + (f64 ? lay.off_L64 : lay.off_Lf));
    const Real* c0 = (const Real*)(base + (f64 ? lay.off_c064 : lay.off_c0f));
    const Real* c1 = (const Real*)(base + (f64 ? lay.off_c164 : lay.off_c1f));
    const Real* ws = (const Real*)(base + (f64 ? lay.off_ws64 : lay.off_wsf));
    (void)d;
    auto* part = ctx->partials.as<pcf::CSum>();
    auto* cnt = ctx->counters.as<unsigned int>();
    auto* out = ctx->out.as<double>();
    const double* hz = ctx->normals.as<double>();
    if (full)
        k_basket_fma<Real, DMAX, MODE, true><<<grid, B, 0, ctx->stream>>>(a, Lm, c0, c1, ws, hz, part, cnt, out);
    else
        k_basket_fma<Real, DMAX, MODE, false><<<grid, B, 0, ctx->stream>>>(a, Lm, c0, c1, ws, hz, part, cnt, out);
}

template <typename Real, int MODE>
void launch_fma_d(pcf_ctx* ctx, int grid, int B, const BasketArgs& a, bool full,
                  const BasketLayout& lay, int d) {
    if (d <= 4) launch_fma<Real, 4, MODE>(ctx, grid, B, a, full, lay, d);
    else if (d <= 8) launch_fma<Real, 8, MODE>(ctx, grid, B, a, full, lay, d);
    else if (d <= 16) launch_fma<Real, 16, MODE>(ctx, grid, B, a, full, lay, d);
    else if (d <= 32) launch_fma<Real, 32, MODE>(ctx, grid, B, a, full, lay, d);
    else launch_fma<Real, 64, MODE>(ctx, grid, B, a, full, lay, d);
}

}  // namespace

extern "C" int pcf_mc_eur_multi(pcf_ctx* ctx, double S0, double E, double r, double sigma,
                                double T, int64_t N, double payoff_fun, int32_t assets, double rho,
                                const pcf_basket_ext* ext, const pcf_opts* opts_in,
                                pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const pcf_opts o = pcf_opts_or_default(opts_in);
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    const int d = assets;
    if (d < 1 || d > kMaxAssets)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "assets must be in [1, %d] (got %d)", kMaxAssets, d);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    // Correlation matrix: ext->corr or equicorrelation rho (include/mvn.h:55-60).
    std::vector<double> C((size_t)d * d);
    for (int i = 0; i < d; ++i)
        for (int k = 0; k < d; ++k) {
            double v = (ext && ext->corr) ? ext->corr[i * d + k] : (i != k ? rho : 1.0);
            if (!std::isfinite(v)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "non-finite correlation");
            C[(size_t)i * d + k] = v;
        }
    for (int i = 0; i < d; ++i)
        for (int k = 0; k < i; ++k)
            if (std::fabs(C[i * d + k] - C[k * d + i]) > 1e-12)
                return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "correlation matrix is not symmetric");
    std::vector<double> L;
    bool full = false;
    if (!cholesky(d, C, L)) {
        if (!eigen_factor(d, C, L))
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "correlation matrix is not positive semi-definite");
        full = true;
    }
    // Per-asset constants.
    const BasketLayout lay = BasketLayout::make(d);
    std::vector<char> hb(lay.bytes, 0);
    double* L64 = (double*)(hb.data() + lay.off_L64);
    double* c064 = (double*)(hb.data() + lay.off_c064);
    double* c164 = (double*)(hb.data() + lay.off_c164);
    double* ws64 = (double*)(hb.data() + lay.off_ws64);
    float* Lf = (float*)(hb.data() + lay.off_Lf);
    float* Lpad = (float*)(hb.data() + lay.off_Lpad);
    float* c0f = (float*)(hb.data() + lay.off_c0f);
    float* c1f = (float*)(hb.data() + lay.off_c1f);
    float* wsf = (float*)(hb.data() + lay.off_wsf);
    const double log2e = 1.4426950408889634;
    const double sqT = o.ref_exact ? 1.0 : std::sqrt(T);
    const double w_i = 1.0 / (double)d;
    for (int i = 0; i < d * d; ++i) {
        L64[i] = L[i];
        Lf[i] = (float)L[i];
    }
    if (d <= 32)
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k) Lpad[i * 32 + k] = (float)L[i * d + k];
    for (int i = 0; i < d; ++i) {
        const double sg = (ext && ext->sigma) ? ext->sigma[i] : sigma;
        const double s0 = (ext && ext->S0) ? ext->S0[i] : S0;
        const double wt = (ext && ext->weights) ? ext->weights[i] : w_i;
        if (!(sg >= 0) || !std::isfinite(s0) || !std::isfinite(wt))
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "bad per-asset parameter at %d", i);
        c064[i] = (r - libm_pow(sg, 2) / 2) * T;
        c164[i] = sg * sqT;
        ws64[i] = wt * s0;
        c0f[i] = (float)(c064[i] * log2e);
        c1f[i] = (float)(c164[i] * log2e);
        wsf[i] = (float)ws64[i];
    }

    const int mode = o.rng == PCF_RNG_HOST ? kHost : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    if (o.kernel_variant == 2 && !(mode == kFp32k && d >= 16 && d <= 32))
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "MFMA basket kernel needs FP32/Philox and 16 <= assets <= 32");
    const bool use_mfma = mode == kFp32k && d >= 16 && d <= 32 && o.kernel_variant != 1;
    BasketArgs a{};
    a.E = E;
    a.pf = payoff_fun;
    a.Ef = (float)E;
    a.pff = (float)payoff_fun;
    a.N = N;
    a.d = d;
    a.seed = o.seed;

    const int B = o.block_size ? o.block_size : 256;
    int grid;
    if (use_mfma) {
        const int64_t ntiles = (N + 31) / 32;
        const int waves_per_block = B / 64;
        const int ipt = o.items_per_thread ? o.items_per_thread : 2;  // tiles per wave
        grid = (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + (int64_t)waves_per_block * ipt - 1) /
                                                               ((int64_t)waves_per_block * ipt),
                                                           1 << 20));
    } else {
        const int ipt = o.items_per_thread ? o.items_per_thread : 1;
        grid = (int)std::max<int64_t>(
            1, std::min<int64_t>((N + (int64_t)B * ipt - 1) / ((int64_t)B * ipt), 1 << 20));
    }

    if ((st = pcf_timing_begin(ctx))) return st;
    PCF_HIP_CHECK(ctx, ctx->basket_L.ensure(lay.bytes));
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->basket_L.ptr, hb.data(), lay.bytes, hipMemcpyHostToDevice,
                                      ctx->stream));
    if (mode == kHost && (st = pcf_upload_normals(ctx, &o, N * (int64_t)d))) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)grid * 2 * sizeof(pcf::CSum)));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    if (use_mfma) {
        char* base = ctx->basket_L.as<char>();
        k_basket_mfma<<<grid, B, 0, ctx->stream>>>(
            a, (const float*)(base + lay.off_Lpad), (const float*)(base + lay.off_c0f),
            (const float*)(base + lay.off_c1f), (const float*)(base + lay.off_wsf),
            ctx->partials.as<pcf::CSum>(), ctx->counters.as<unsigned int>(), ctx->out.as<double>());
    } else if (mode == kFp32k) {
        launch_fma_d<float, kFp32k>(ctx, grid, B, a, full, lay, d);
    } else if (mode == kFp64) {
        launch_fma_d<double, kFp64>(ctx, grid, B, a, full, lay, d);
    } else {
        launch_fma_d<double, kHost>(ctx, grid, B, a, full, lay, d);
    }
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if ((st = pcf_timing_end(ctx, 4, res))) return st;

    const double* h = ctx->h_out;
    const double disc = std::exp(-r * T);
    // The reference returns (exp(-r*T)*result)/(double)N (src/mc_eur_multi.cpp:34).
    res->price = (disc * h[0]) / (double)N;
    res->mean = h[2];
    res->variance = h[3];
    res->std_err = disc * std::sqrt(h[3] / (double)N);
    res->n_paths = N;
    res->n_steps = 1;
    // d^2 + 12 d + 3 algorithmic flops per path: normals 6d, triangular Z L^T
    // d(d+1), per asset FMA + exp + FMA 5d, payoff 3.
    res->flops = ((double)d * d + 12.0 * d + 3.0) * (double)N;
    res->bytes = mode == kHost ? 8.0 * d * (double)N : 0.0;
    res->grid_blocks = grid;
    res->block_size = B;
    res->lsm_fallback = full ? 1 : 0;
    return PCF_OK;
}
