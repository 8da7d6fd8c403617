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
//  * k_basket_mfma: FP32, 16 <= d <= 32.  One wavefront prices 16 paths per
//    tile: the exponents X^T (32 assets x 16 paths) = c0 + (c1 L) Z^T, with
//    L zero padded to 32 x 32 and each row scaled by c1 log2 e, on
//    v_mfma_f32_16x16x4f32 in two blocks of 16 assets, skipping the factor's
//    zero upper-right block, the accumulation starting from c0 log2 e.  Each
//    lane generates exactly the normals it feeds, so Z and X never touch
//    memory; the epilogue (exp2, weights, basket sum) runs on the 8
//    accumulator registers, the four lane groups combined by lane exchange.
#include <cmath>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "pcf_internal.hpp"

namespace {

constexpr int kMaxAssets = 64;
// Threads per block of both basket kernels; their LDS reduction buffers are
// sized for it and the host refuses larger blocks.
constexpr int kBasketThreads = 256;
enum { kFp32k = 0, kFp64 = 1, kHost = 2 };

struct BasketArgs {
    double E, pf;
    float Ef, pff;
    int64_t N;
    int32_t d;
    int32_t full;  // the factor is not lower triangular (eigen fallback)
    uint64_t seed;
};

// Device constants, one buffer: FP64 block then FP32 block.
//   FP64: L[d*d] (row-major), c0[d] = (r - s^2/2)T, c1[d] = s sqrt(T) (or s),
//         ws[d] = w S0
//   FP32: Lf[d*d] (the FMA kernel's factor), L32[32*32] zero padded with row
//         i scaled by c1[i] log2 e (the MFMA kernel's), c0f[32], c1f[32]
//         (x log2 e), wsf[32].
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

struct BasketJob {
    BasketArgs a;
    const void* Lm;   // FMA: d x d factor (Real); MFMA: 32 x 32 zero-padded FP32 factor
    const void* c0;
    const void* c1;
    const void* ws;
    const double* hz;
    pcf::CSum* partials;
    unsigned int* counter;
    double* out;
    unsigned int first_block, nblk;
};
static_assert(sizeof(BasketJob) <= kJobBytes, "job record exceeds its staging slot");

// The factor and the per-asset constants are read-only and the same for every
// lane, so the FMA kernel reads them through the constant address space: the
// compiler then issues scalar loads whose results feed the FMAs as SGPR
// operands.  Through generic pointers every asset's row was a flat vector load
// followed by a wait on the vector memory counter inside the path loop.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

template <typename Real, int DMAX, int MODE, bool FULL>
__device__ __forceinline__ void basket_fma_body(const BasketArgs& a, cptr<Real> Lm, cptr<Real> c0,
                                                cptr<Real> c1, cptr<Real> ws,
                                                const double* __restrict__ hz,
                                                pcf::CSum* __restrict__ partials,
                                                unsigned int* __restrict__ counter,
                                                double* __restrict__ out, unsigned int blk,
                                                unsigned int nblk, pcf::CSum* lds) {
    const int d = a.d;
    double s = 0.0, s2 = 0.0;
    const int64_t stride = (int64_t)nblk * blockDim.x;
    for (int64_t p = (int64_t)blk * blockDim.x + threadIdx.x; p < a.N; p += stride) {
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
    if (pcf::grid_reduce<2>(acc, partials, counter, lds, tot, blk, nblk) && threadIdx.x == 0)
        pcf::write_moments<1>(tot, a.N, out);
}

template <typename Real, int DMAX, int MODE, bool FULL>
__global__ __launch_bounds__(kBasketThreads) void k_basket_fma(const BasketJob* __restrict__ jobs, int njobs) {
    __shared__ pcf::CSum lds[(kBasketThreads / pcf::kWave) * 2];
    const int j = pcf::find_job(jobs, njobs, blockIdx.x);
    const BasketJob& J = jobs[j];
    basket_fma_body<Real, DMAX, MODE, FULL>(J.a, (cptr<Real>)J.Lm, (cptr<Real>)J.c0, (cptr<Real>)J.c1,
                                            (cptr<Real>)J.ws, J.hz, J.partials, J.counter, J.out,
                                            blockIdx.x - J.first_block, J.nblk, lds);
}

typedef float v4f __attribute__((ext_vector_type(4)));

// One wavefront prices a tile of 16 paths with v_mfma_f32_16x16x4f32 (A: 16
// rows x 4 k, lane l supplies row l%16, k l/16; B: 4 k x 16 columns, lane l
// supplies k l/16, column l%16; D: lane l holds rows 4(l/16)..+3 of column
// l%16).  Rows are assets in two blocks of 16 (a = 0, 1), columns the tile's
// paths.  Step s (0..7) of a block sums over k = 16(s/4) + 4g + s%4 for the
// four lane groups g = l/16: lane (g, j) feeds the 8 normals of Philox blocks
// g and 4 + g of path j, whole blocks, in that order.  A Cholesky factor is
// lower triangular, so asset block 0 needs only k < 16, steps 0..3: 12 MFMAs
// of 32 cycles per tile instead of the 16 of 64 a padded 32 x 32 x 2 form
// takes.  A full factor (the eigen fallback) runs block 0's steps 4..7 too.
__device__ __forceinline__ void basket_mfma_body(const BasketArgs& a, const float* __restrict__ Lpad,
                                                 const float* __restrict__ c0,
                                                 const float* __restrict__ ws,
                                                 pcf::CSum* __restrict__ partials,
                                                 unsigned int* __restrict__ counter,
                                                 double* __restrict__ out, unsigned int blk,
                                                 unsigned int nblk, pcf::CSum* lds) {
    const int lane = threadIdx.x & 63;
    const int j = lane & 15;  // path column of the tile; also the A-operand row
    const int g = lane >> 4;  // k group fed by this lane
    const bool hi = a.d > 16;  // asset block 1 and the k >= 16 steps carry terms
    const bool full = a.full != 0;
    // A operands of steps 0..7; this lane supplies the scaled factor's
    // [16 a + j][16 (s/4) + 4 g + s%4].
    float A0[8], A1[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        A0[s] = Lpad[j * 32 + 16 * (s >> 2) + 4 * g + (s & 3)];
        A1[s] = Lpad[(16 + j) * 32 + 16 * (s >> 2) + 4 * g + (s & 3)];
    }
    // Accumulator register r of block a holds asset 16 a + 4 g + r; every
    // tile's accumulation starts from c0 log2 e of that asset.
    v4f base0, base1;
    float ew0[4], ew1[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        base0[r] = c0[4 * g + r];
        base1[r] = c0[16 + 4 * g + r];
        ew0[r] = ws[4 * g + r];
        ew1[r] = ws[16 + 4 * g + r];
    }
    double s = 0.0, s2 = 0.0;
    const int64_t ntiles = (a.N + 15) >> 4;
    const int64_t wave = ((int64_t)blk * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)nblk * blockDim.x) >> 6;
    for (int64_t t = wave; t < ntiles; t += nwaves) {
        const int64_t p = (t << 4) + j;
        float z[4];
        pcf::normals4<float>(a.seed, PCF_STREAM_BASKET, (uint64_t)p, (uint32_t)g, z);
        v4f acc0 = base0, acc1 = base1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(A0[q], z[q], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A1[q], z[q], acc1, 0, 0, 0);
        }
        if (hi) {
            pcf::normals4<float>(a.seed, PCF_STREAM_BASKET, (uint64_t)p, (uint32_t)(4 + g), z);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A1[4 + q], z[q], acc1, 0, 0, 0);
            if (full)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(A0[4 + q], z[q], acc0, 0, 0, 0);
        }
        float part = 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) part = fmaf(ew0[r], pcf::exp2_f32(acc0[r]), part);
        if (hi)
#pragma unroll
            for (int r = 0; r < 4; ++r) part = fmaf(ew1[r], pcf::exp2_f32(acc1[r]), part);
        // the 32 assets of path j sit in the four lane groups
        part += __shfl_xor(part, 16, 64);
        const float basket = part + __shfl_xor(part, 32, 64);
        const float v = pcf::payoff<float>(basket, a.Ef, a.pff);
        if (g == 0 && p < a.N) {
            const double dv = (double)v;
            s += dv;
            s2 += dv * dv;
        }
    }
    pcf::Acc<2> acc2, tot;
    acc2.v[0] = {s, 0.0};
    acc2.v[1] = {s2, 0.0};
    if (pcf::grid_reduce<2>(acc2, partials, counter, lds, tot, blk, nblk) && threadIdx.x == 0)
        pcf::write_moments<1>(tot, a.N, out);
}

// At least four waves per SIMD: the operands and accumulators then stay in
// VGPRs (86, no AGPRs, no spills: five waves).
__global__ __launch_bounds__(kBasketThreads) __attribute__((amdgpu_waves_per_eu(4)))
void k_basket_mfma(const BasketJob* __restrict__ jobs, int njobs) {
    __shared__ pcf::CSum lds[(kBasketThreads / pcf::kWave) * 2];
    const int j = pcf::find_job(jobs, njobs, blockIdx.x);
    const BasketJob& J = jobs[j];
    basket_mfma_body(J.a, (const float*)J.Lm, (const float*)J.c0, (const float*)J.ws,
                     J.partials, J.counter, J.out, blockIdx.x - J.first_block, J.nblk, lds);
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

// ---- host: factor cache, preparation, batched launch ------------------------

// Process-wide cache of factorised baskets (Cholesky or eigen factor plus the
// per-asset constants), found by a hash of everything they depend on; an
// entry keeps that full key and is used only when it equals the request's
// (a hash collision is a miss).  id is unique per entry and never reused, so
// the contexts' device slots are keyed by it.
struct BasketConsts {
    std::vector<char> blob;
    std::vector<double> key;
    uint64_t id = 0;
    BasketLayout lay;
    bool full;
};

std::atomic<uint64_t>& consts_next_id() {
    static std::atomic<uint64_t> n{1};
    return n;
}
std::atomic<uint64_t>& consts_hash_mask() {
    static std::atomic<uint64_t> m{~0ull};
    return m;
}

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= c[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::mutex& consts_mu() {
    static std::mutex m;
    return m;
}
std::map<uint64_t, std::shared_ptr<BasketConsts>>& consts_cache() {
    static std::map<uint64_t, std::shared_ptr<BasketConsts>> c;
    return c;
}

struct BasketPrep {
    BasketArgs a;
    int mode, B, nblk, dmax;
    bool mfma, full;
    uint64_t key;  // the constants' entry id (device slots are keyed by it)
    std::shared_ptr<BasketConsts> consts;
    double disc;
};

int basket_prepare(pcf_ctx* ctx, const pcf_request& q, BasketPrep& p) {
    const pcf_opts& o = q.opts;
    int st = pcf_check_common(ctx, q.S0, q.E, q.r, q.sigma, q.T, q.N, q.payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    const int d = q.assets;
    if (d < 1 || d > kMaxAssets)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "assets must be in [1, %d] (got %d)", kMaxAssets, d);
    const pcf_basket_ext* ext = q.basket;
    // Key of the constants: d, rho, r, T, sqrt(T) convention, scalar sigma/S0
    // and every extension array (absent arrays marked by a NaN-free flag).
    const int refx = o.ref_exact ? 1 : 0;
    const double* arrs[4] = {ext ? ext->sigma : nullptr, ext ? ext->S0 : nullptr,
                             ext ? ext->weights : nullptr, ext ? ext->corr : nullptr};
    std::vector<double> key = {(double)d, q.rho, q.r, q.T, q.sigma, q.S0, (double)refx};
    for (int k = 0; k < 4; ++k) {
        key.push_back(arrs[k] ? 1.0 : 0.0);
        if (arrs[k]) key.insert(key.end(), arrs[k], arrs[k] + (k == 3 ? (size_t)d * d : (size_t)d));
    }
    const uint64_t h = fnv(1469598103934665603ull, key.data(), key.size() * sizeof(double)) &
                       consts_hash_mask().load(std::memory_order_relaxed);
    std::shared_ptr<BasketConsts> c;
    {
        std::lock_guard<std::mutex> lk(consts_mu());
        auto it = consts_cache().find(h);
        // bitwise comparison of the full key: a hash collision is a miss
        if (it != consts_cache().end() && it->second->key.size() == key.size() &&
            std::memcmp(it->second->key.data(), key.data(), key.size() * sizeof(double)) == 0)
            c = it->second;
    }
    if (!c) {
        // Correlation matrix: ext->corr or equicorrelation rho (include/mvn.h:55-60).
        std::vector<double> C((size_t)d * d);
        for (int i = 0; i < d; ++i)
            for (int k = 0; k < d; ++k) {
                const double v = (ext && ext->corr) ? ext->corr[i * d + k] : (i != k ? q.rho : 1.0);
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
        c = std::make_shared<BasketConsts>();
        c->full = full;
        c->key = key;
        c->id = consts_next_id().fetch_add(1);
        c->lay = BasketLayout::make(d);
        c->blob.assign(c->lay.bytes, 0);
        char* hb = c->blob.data();
        const BasketLayout& lay = c->lay;
        double* L64 = (double*)(hb + lay.off_L64);
        double* c064 = (double*)(hb + lay.off_c064);
        double* c164 = (double*)(hb + lay.off_c164);
        double* ws64 = (double*)(hb + lay.off_ws64);
        float* Lf = (float*)(hb + lay.off_Lf);
        float* Lpad = (float*)(hb + lay.off_Lpad);
        float* c0f = (float*)(hb + lay.off_c0f);
        float* c1f = (float*)(hb + lay.off_c1f);
        float* wsf = (float*)(hb + lay.off_wsf);
        const double log2e = 1.4426950408889634;
        const double sqT = refx ? 1.0 : std::sqrt(q.T);
        const double w_i = 1.0 / (double)d;
        for (int i = 0; i < d * d; ++i) {
            L64[i] = L[i];
            Lf[i] = (float)L[i];
        }
        for (int i = 0; i < d; ++i) {
            const double sg = (ext && ext->sigma) ? ext->sigma[i] : q.sigma;
            const double s0 = (ext && ext->S0) ? ext->S0[i] : q.S0;
            const double wt = (ext && ext->weights) ? ext->weights[i] : w_i;
            if (!(sg >= 0) || !std::isfinite(sg) || !std::isfinite(s0) || !std::isfinite(wt))
                return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "bad per-asset parameter at %d", i);
            // (w_i*S0)*exp((r-pow(sigma,2)/2)*T + sigma*B)  (src/mc_eur_multi.cpp:30)
            c064[i] = (q.r - libm_pow(sg, 2) / 2) * q.T;
            c164[i] = sg * sqT;
            ws64[i] = wt * s0;
            c0f[i] = (float)(c064[i] * log2e);
            c1f[i] = (float)(c164[i] * log2e);
            wsf[i] = (float)ws64[i];
            // The MFMA kernel's factor carries each row's c1 log2 e, so its
            // accumulator holds the exponent itself (rounded once from FP64).
            if (d <= 32)
                for (int k = 0; k < d; ++k) Lpad[i * 32 + k] = (float)(L[i * d + k] * (c164[i] * log2e));
        }
        std::lock_guard<std::mutex> lk(consts_mu());
        if (consts_cache().size() > 1024) consts_cache().clear();
        consts_cache()[h] = c;
    }
    p.consts = c;
    p.key = c->id;
    p.full = c->full;
    p.mode = o.rng == PCF_RNG_HOST ? kHost : (o.precision == PCF_FP64 ? kFp64 : kFp32k);
    if (o.kernel_variant == 2 && !(p.mode == kFp32k && d >= 16 && d <= 32))
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "MFMA basket kernel needs FP32/Philox and 16 <= assets <= 32");
    p.mfma = p.mode == kFp32k && d >= 16 && d <= 32 && o.kernel_variant != 1;
    p.dmax = d <= 4 ? 4 : d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 64;
    BasketArgs& a = p.a;
    a = BasketArgs{};
    a.E = q.E;
    a.pf = q.payoff_fun;
    a.Ef = (float)q.E;
    a.pff = (float)q.payoff_fun;
    a.N = q.N;
    a.d = d;
    a.full = p.full ? 1 : 0;
    a.seed = o.seed;
    // Built-in shape: 256 threads, at most 16 blocks per CU (the MFMA kernel's
    // items are 16-path tiles, one wavefront per tile at a time).
    const int64_t ntiles = (q.N + 15) / 16;
    const int64_t max_blocks = 16LL * (ctx->cu_count > 0 ? ctx->cu_count : 256);
    int B = 256;
    int ipt = p.mfma ? pcf_capped_ipt(ntiles, B / 64, 2, max_blocks) : pcf_capped_ipt(q.N, B, 1, max_blocks);
    pcf_tuned_shape(PCF_KIND_BASKET, p.mode * 2 + (p.mfma ? 1 : 0), q.N, &B, &ipt);
    if (o.block_size) B = o.block_size;
    if (o.items_per_thread) ipt = o.items_per_thread;
    if (B > kBasketThreads)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "mc_eur_multi takes at most %d threads per block (block_size %d)",
                        kBasketThreads, B);
    p.B = B;
    p.nblk = p.mfma ? pcf_blocks_for(ntiles, B / 64, ipt) : pcf_blocks_for(q.N, B, ipt);
    p.disc = std::exp(-q.r * q.T);
    return PCF_OK;
}

// Device-resident constants: fixed-size slots in ctx->basket_L, LRU by stamp.
constexpr size_t kConstSlot = kBasketConstSlot;
constexpr int kConstSlots = kMaxBatch;

int resident_consts(pcf_ctx* ctx, const BasketPrep& p, uint64_t batch_stamp, char** dev) {
    auto& bc = ctx->basket_cache;
    if (bc.key.empty()) {
        bc.key.assign(kConstSlots, 0);
        bc.stamp.assign(kConstSlots, 0);
        PCF_HIP_CHECK(ctx, ctx->basket_L.ensure(kConstSlot * kConstSlots));
    }
    auto it = bc.where.find(p.key);
    int slot;
    if (it != bc.where.end()) {
        slot = it->second;
    } else {
        slot = -1;
        for (int s = 0; s < kConstSlots; ++s)
            if (bc.stamp[s] < batch_stamp && (slot < 0 || bc.stamp[s] < bc.stamp[slot])) slot = s;
        if (slot < 0) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "too many distinct baskets in one batch");
        if (bc.key[slot]) bc.where.erase(bc.key[slot]);
        bc.key[slot] = p.key;
        bc.where[p.key] = slot;
        if (p.consts->blob.size() > kConstSlot) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "basket too large");
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->basket_L.as<char>() + kConstSlot * slot, p.consts->blob.data(),
                                          p.consts->blob.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    bc.stamp[slot] = ++bc.clock;
    *dev = ctx->basket_L.as<char>() + kConstSlot * slot;
    return PCF_OK;
}

template <typename Real, int DMAX, int MODE>
void launch_fma(pcf_ctx* ctx, unsigned int grid, int B, int nj, bool full) {
    const BasketJob* dj = ctx->jobs.as<BasketJob>();
    if (full) k_basket_fma<Real, DMAX, MODE, true><<<grid, B, 0, ctx->stream>>>(dj, nj);
    else k_basket_fma<Real, DMAX, MODE, false><<<grid, B, 0, ctx->stream>>>(dj, nj);
}

template <typename Real, int MODE>
void launch_fma_d(pcf_ctx* ctx, unsigned int grid, int B, int nj, bool full, int dmax) {
    if (dmax == 4) launch_fma<Real, 4, MODE>(ctx, grid, B, nj, full);
    else if (dmax == 8) launch_fma<Real, 8, MODE>(ctx, grid, B, nj, full);
    else if (dmax == 16) launch_fma<Real, 16, MODE>(ctx, grid, B, nj, full);
    else if (dmax == 32) launch_fma<Real, 32, MODE>(ctx, grid, B, nj, full);
    else launch_fma<Real, 64, MODE>(ctx, grid, B, nj, full);
}

}  // namespace

uint64_t pcf_basket_batch_key(pcf_ctx* ctx, const pcf_request& q, int* status) {
    BasketPrep p;
    *status = basket_prepare(ctx, q, p);
    if (*status) return 0;
    const int variant = p.mfma ? 1 : (2 + (p.full ? 1 : 0) + 2 * p.dmax);
    return pcf_make_batch_key(PCF_KIND_BASKET, p.mode, variant, p.B);
}

int pcf_basket_run(pcf_ctx* ctx, int n, const pcf_request* const* reqs, pcf_result* res, int* status) {
    if (n < 1 || n > kMaxBatch) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch size %d", n);
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    static thread_local BasketPrep prep[kMaxBatch];
    int live[kMaxBatch], nl = 0;
    for (int i = 0; i < n; ++i) {
        res[i] = pcf_result{};
        status[i] = basket_prepare(ctx, *reqs[i], prep[i]);
        if (status[i] == PCF_OK) live[nl++] = i;
    }
    if (nl == 0) return PCF_OK;
    const BasketPrep& p0 = prep[live[0]];
    int64_t nparts = 0, nb = 0, nnorm = 0;
    for (int k = 0; k < nl; ++k) {
        const BasketPrep& p = prep[live[k]];
        if (p.mode != p0.mode || p.B != p0.B || p.mfma != p0.mfma || p.dmax != p0.dmax || p.full != p0.full)
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "mixed batch keys");
        nparts += (int64_t)p.nblk * 2;
        nb += p.nblk;
        if (p.mode == kHost) nnorm += reqs[live[k]]->N * (int64_t)reqs[live[k]]->assets;
    }
    if (nb > (1ll << 30)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "batch too large");
    int st = pcf_timing_begin(ctx);
    if (st) return st;
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)nparts * sizeof(pcf::CSum)));
    if (nnorm) PCF_HIP_CHECK(ctx, ctx->normals.ensure((size_t)nnorm * sizeof(double)));
    BasketJob* jobs = (BasketJob*)ctx->h_jobs;
    const uint64_t batch_stamp = ctx->basket_cache.clock + 1;
    int64_t poff = 0, noff = 0;
    unsigned int boff = 0;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        const BasketPrep& p = prep[i];
        char* dev = nullptr;
        if ((st = resident_consts(ctx, p, batch_stamp, &dev))) return st;
        const BasketLayout& lay = p.consts->lay;
        const bool f64 = p.mode != kFp32k;
        BasketJob& J = jobs[k];
        J.a = p.a;
        J.Lm = dev + (p.mfma ? lay.off_Lpad : (f64 ? lay.off_L64 : lay.off_Lf));
        J.c0 = dev + (f64 ? lay.off_c064 : lay.off_c0f);
        J.c1 = dev + (f64 ? lay.off_c164 : lay.off_c1f);
        J.ws = dev + (f64 ? lay.off_ws64 : lay.off_wsf);
        J.hz = nullptr;
        if (p.mode == kHost) {
            const pcf_opts& o = reqs[i]->opts;
            const int64_t cnt = reqs[i]->N * (int64_t)reqs[i]->assets;
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
        poff += (int64_t)p.nblk * 2;
        boff += (unsigned int)p.nblk;
    }
    PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->jobs.ptr, jobs, sizeof(BasketJob) * nl, hipMemcpyHostToDevice,
                                      ctx->stream));
    double work = 0.0;
    for (int k = 0; k < nl; ++k) work += (double)reqs[live[k]]->N * (double)reqs[live[k]]->assets;
    std::unique_lock<std::mutex> order;  // held until the end event is recorded
    PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, p0.mode == kFp32k ? work : kPersistentGateFp64 * work, &order));
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
    if (p0.mfma)
        k_basket_mfma<<<boff, p0.B, 0, ctx->stream>>>(ctx->jobs.as<BasketJob>(), nl);
    else if (p0.mode == kFp32k)
        launch_fma_d<float, kFp32k>(ctx, boff, p0.B, nl, p0.full, p0.dmax);
    else if (p0.mode == kFp64)
        launch_fma_d<double, kFp64>(ctx, boff, p0.B, nl, p0.full, p0.dmax);
    else
        launch_fma_d<double, kHost>(ctx, boff, p0.B, nl, p0.full, p0.dmax);
    PCF_HIP_CHECK(ctx, hipGetLastError());
    PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
    if (order.owns_lock()) order.unlock();
    pcf_result timing{};
    if ((st = pcf_timing_end(ctx, kOutPerJob * nl, &timing))) return st;
    for (int k = 0; k < nl; ++k) {
        const int i = live[k];
        const BasketPrep& p = prep[i];
        const double* h = ctx->h_out + (size_t)kOutPerJob * k;
        const int64_t N = reqs[i]->N;
        const double d = reqs[i]->assets;
        pcf_result& r = res[i];
        // The reference returns (exp(-r*T)*result)/(double)N (src/mc_eur_multi.cpp:34).
        r.price = (p.disc * h[0]) / (double)N;
        r.mean = h[2];
        r.variance = h[3];
        r.std_err = p.disc * std::sqrt(h[3] / (double)N);
        r.n_paths = N;
        r.n_steps = 1;
        r.kernel_ms = timing.kernel_ms;
        r.total_ms = timing.total_ms;
        // d^2 + 12 d + 3 algorithmic flops per path: normals 6d, triangular
        // Z L^T d(d+1), per asset FMA + exp + FMA 5d, payoff 3.
        r.flops = (d * d + 12.0 * d + 3.0) * (double)N;
        r.bytes = p.mode == kHost ? 8.0 * d * (double)N : 0.0;
        r.grid_blocks = p.nblk;
        r.block_size = p.B;
        r.batch_size = nl;
        r.lsm_fallback = p.full ? 1 : 0;
    }
    for (int k = 0; k < nl; ++k) prep[live[k]].consts.reset();
    return PCF_OK;
}

extern "C" int pcf_mc_eur_multi(pcf_ctx* ctx, double S0, double E, double r, double sigma,
                                double T, int64_t N, double payoff_fun, int32_t assets, double rho,
                                const pcf_basket_ext* ext, const pcf_opts* opts, pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_request q = pcf_make_request(PCF_KIND_BASKET, S0, E, r, sigma, T, N, payoff_fun, opts);
    q.assets = assets;
    q.rho = rho;
    q.basket = ext;
    const pcf_request* qp = &q;
    int status = PCF_OK;
    int st = pcf_basket_run(ctx, 1, &qp, res, &status);
    if (st || status) return st ? st : status;
    return pcf_apply_fp64_check(ctx, q, res);
}

extern "C" void pcf_debug_basket_hash_mask(uint64_t mask) {
    consts_hash_mask().store(mask, std::memory_order_relaxed);
}
