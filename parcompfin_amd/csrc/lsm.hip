// lsm.hip -- American option by Longstaff-Schwartz on an antithetic path store.
// Reference: src/mc_amer.cpp:5-114, paths from pathsfinder
// (include/common.h:168-208), 3x3 solve include/common.h:98-141.
//
// Device data (HBM, structure of arrays, coalesced along the path index n):
//   store[(m-1)*N + n], m = 1..M  path values (time-major like the reference's
//                                 paths[m][n]; row 0 = S0 is not stored)
//   when[n] (int32), cf[n], st[n]  per-path exercise date, true payoff at that
//                                 date (the reference's y gathers it from
//                                 paths[when][n]; here it is carried), and the
//                                 booked cash flow (differs from cf only under
//                                 the reference's payoff(S-E,E) exercise test)
//   coef[m] = {flag, c0, c1, c2}  regression result per date
// Launch sequence (one stream; built once per shape as a hipGraph and replayed):
//   k_lsm_paths        store + state init + European control + sums of date M-1
//   k_lsm_step(m)      exercise decision at m, then sums of date m-1   (m = M-1..2)
//   k_lsm_final        decision at date 1, then sum of discounted cash flows
// Each sums pass reduces 8 compensated FP64 sums (n, x, x^2, x^3, x^4, y,
// xy, x^2 y) through the last-arriving block, which solves the normal
// equations in FP64 and writes coef for the next launch.
//
// Arithmetic forms:
//  * REF (ref_exact): FP64, basis {1, S-E, (S-E)^2}, cofactor inverse with the
//    reference's operation order, exercise test payoff(S-E, E)
//    (src/mc_amer.cpp:100; wrong for puts, reproduced for parity), det <= 0
//    flagged (the reference throws).
//  * production: basis {1, x, x^2} with x = S/E (same span as {1, S, S^2}),
//    FP64 sums and a 3x3 LDL^T solve with fallback to {1, x} then {1},
//    exercise test payoff(S, E); FP32 or FP64 path store.
// Dates with 1-2 in-the-money paths compare payoff with the discounted cash
// flow directly (src/mc_amer.cpp:75-83), in both forms.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "pcf_internal.hpp"

namespace {

enum { kPhilox = 0, kHost = 1 };
enum { kFlagSkip = 0, kFlagSmall = 1, kFlagRegress = 2, kFlagSingular = 3, kFlagReduced = 4 };

struct LsmParams {
    double S0, E, pf, invE;
    double nrdt;        // (-r)*dt            reference discount of y
    double nr, dt;      // -r, dt             reference final discount exp((-r*when)*dt)
    double h_mu, sigma; // (r-0.5*sigma*sigma)*dt, sigma   reference store recurrence
    double mu, vol;     // (r - s^2/2) dt, s sqrt(dt)      production FP64
    float mu2f, vol2f, S0f, Ef, pff, pad;  // production FP32 (x log2 e)
    int64_t N, half;
    int32_t M, pad2;
    uint64_t seed;
};

struct LsmBufs {
    void* store;
    int32_t* when;
    void* cf;
    void* st;
    double* coef;  // per date 0..M: flag + coefficients (stride 4 for the 3-term basis, else 8)
    double* disc;  // exp(-r dt k), k = 0..M
    const double* hw;  // host normals (verification)
    pcf::CSum* partials;
    unsigned int* counter;
    double* out;
    const LsmParams* P;
};

template <typename Real>
__device__ __forceinline__ Real pay(Real S, Real E, Real pf) {
    return pcf::payoff<Real>(S, E, pf);
}

// Regression sums of one path at date m, in two halves: the x-moments
// acc[0..4] = {n, x, x^2, x^3, x^4} depend only on the path value S at date m;
// the y-terms acc[5..7] = {y, yx, yx^2} also on the exercise state after date
// m+1.  Each sum receives the same terms in the same path order whichever
// half runs first, so the persistent kernel can form the x-moments of the
// next date while it waits for the other blocks.
template <typename Real, bool REF>
__device__ __forceinline__ void lsm_accumulate_x(const LsmParams& P, Real S, double acc[5]) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    if (pv > 0.0) {
        if (REF) {
#pragma clang fp contract(off)
            const double x = (double)S - P.E;
            acc[0] += 1.0;
            acc[1] += x;
            acc[2] += x * x;
            acc[3] += x * x * x;
            acc[4] += x * x * x * x;
        } else {
            const double x = (double)S * P.invE;
            const double x2 = x * x;
            acc[0] += 1.0;
            acc[1] += x;
            acc[2] += x2;
            acc[3] += x2 * x;
            acc[4] += x2 * x2;
        }
    }
}

template <typename Real, bool REF>
__device__ __forceinline__ void lsm_accumulate_y(const LsmParams& P, const double* __restrict__ disc, int m,
                                                 Real S, int32_t when, Real cf, double acc[8]) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    if (pv > 0.0) {
        if (REF) {
#pragma clang fp contract(off)
            const double x = (double)S - P.E;
            const double y = exp(P.nrdt * ((double)when - (double)m)) * (double)cf;
            acc[5] += y;
            acc[6] += y * x;
            acc[7] += y * x * x;
        } else {
            const double x = (double)S * P.invE;
            const double y = disc[when - m] * (double)cf;
            const double x2 = x * x;
            acc[5] += y;
            acc[6] += y * x;
            acc[7] += y * x2;
        }
    }
}

template <typename Real, bool REF>
__device__ __forceinline__ void lsm_accumulate(const LsmParams& P, const double* __restrict__ disc,
                                               int m, Real S, int32_t when, Real cf, double acc[8]) {
    lsm_accumulate_x<Real, REF>(P, S, acc);
    lsm_accumulate_y<Real, REF>(P, disc, m, S, when, cf, acc);
}

// The production quadratic's three passes over a date's path value -- the
// x-moments (formed a date early), the y-terms and the decision -- share x =
// S/E and the in-the-money test.  The persistent kernel keeps x from the
// x-moments pass (-1 out of the money; x >= 0 in the money) and the other two
// passes read it instead of repeating the payoff, the conversion and the
// product.  Same operations on the same values, so the same bits.
template <typename Real>
__device__ __forceinline__ void lsm_accumulate_x_keep(const LsmParams& P, Real S, double acc[5], double& xo) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    xo = -1.0;
    if (pv > 0.0) {
        const double x = (double)S * P.invE;
        const double x2 = x * x;
        acc[0] += 1.0;
        acc[1] += x;
        acc[2] += x2;
        acc[3] += x2 * x;
        acc[4] += x2 * x2;
        xo = x;
    }
}

template <typename Real>
__device__ __forceinline__ void lsm_accumulate_y_kept(const double* __restrict__ disc, int m, double xo,
                                                      int32_t when, Real cf, double acc[8]) {
    if (xo >= 0.0) {
        const double y = disc[when - m] * (double)cf;
        const double x2 = xo * xo;
        acc[5] += y;
        acc[6] += y * xo;
        acc[7] += y * x2;
    }
}

// Normal equations from the 8 totals g = {n, sx, sx2, sx3, sx4, sy, sxy,
// sx2y} -> c = {flag, c0, c1, c2}.  Pure function of g: every thread that
// calls it with the same totals gets the same coefficients.
template <bool REF>
__device__ __forceinline__ void lsm_solve_g(const double g[8], double c[4]) {
    const double n = g[0];
    c[1] = c[2] = c[3] = 0.0;
    if (n == 0.0) {
        c[0] = kFlagSkip;
        return;
    }
    if (n <= 2.0) {
        c[0] = kFlagSmall;
        return;
    }
    if (REF) {
#pragma clang fp contract(off)
        // include/common.h:98-124 then :128-141, same order.
        const double x[3][3] = {{n, g[1], g[2]}, {g[1], g[2], g[3]}, {g[2], g[3], g[4]}};
        const double y[3] = {g[5], g[6], g[7]};
        double det = 0;
        for (int i = 0; i < 3; ++i)
            det += (x[0][i] * (x[1][(i + 1) % 3] * x[2][(i + 2) % 3] - x[1][(i + 2) % 3] * x[2][(i + 1) % 3]));
        if (det <= 0) {
            c[0] = kFlagSingular;
            return;
        }
        double inv[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                inv[j][i] = ((x[(j + 1) % 3][(i + 1) % 3] * x[(j + 2) % 3][(i + 2) % 3]) -
                             (x[(j + 1) % 3][(i + 2) % 3] * x[(j + 2) % 3][(i + 1) % 3])) / det;
        for (int i = 0; i < 3; ++i) {
            double mm = 0;
            for (int j = 0; j < 3; ++j) mm += inv[i][j] * y[j];
            c[1 + i] = mm;
        }
        c[0] = kFlagRegress;
        return;
    }
    // Production: FP64 square-root-free Cholesky, G = L D L^T, of
    // G = [[n,s1,s2],[s1,s2,s3],[s2,s3,s4]].  Its pivots d1, d2 are the
    // squares of the Cholesky factor's diagonal, so the fallback rule is the
    // Cholesky one: basis column j is kept while d_j > 1e-13 G_jj.  Without
    // square roots the dependent chain is three divisions long; every block
    // runs it between the date's all-reduce and its decisions.
    const double eps = 1e-13;
    const double s1 = g[1], s2 = g[2], s3 = g[3], s4 = g[4];
    const double b0 = g[5], b1 = g[6], b2 = g[7];
    const double i0 = 1.0 / n;
    const double l10 = s1 * i0, l20 = s2 * i0;
    const double d1 = fma(-l10, s1, s2);  // s2 - s1^2 / n
    if (d1 > eps * s2) {
        const double i1 = 1.0 / d1;
        const double t21 = fma(-l20, s1, s3);  // l21 d1
        const double l21 = t21 * i1;
        const double d2 = fma(-l21, t21, fma(-l20, s2, s4));
        const double z1 = fma(-l10, b0, b1);  // forward substitution (z0 = b0)
        if (d2 > eps * s4) {
            const double z2 = fma(-l21, z1, fma(-l20, b0, b2));
            const double x2 = z2 / d2;
            const double x1 = fma(-l21, x2, z1 * i1);
            const double x0 = fma(-l20, x2, fma(-l10, x1, b0 * i0));
            c[1] = x0;
            c[2] = x1;
            c[3] = x2;
            c[0] = kFlagRegress;
            return;
        }
        // {1, x}
        const double x1 = z1 * i1;
        c[1] = fma(-l10, x1, b0 * i0);
        c[2] = x1;
        c[0] = kFlagReduced;
        return;
    }
    c[1] = b0 / n;  // {1}: the mean continuation value
    c[0] = kFlagReduced;
}

// Normal equations -> coef[date] in global memory.  Runs in one thread.
template <bool REF>
__device__ void lsm_solve(const pcf::Acc<8>& tot, double* __restrict__ coef, int date) {
    double g[8], c[4];
#pragma unroll
    for (int k = 0; k < 8; ++k) g[k] = tot.v[k].s + tot.v[k].c;
    lsm_solve_g<REF>(g, c);
#pragma unroll
    for (int k = 0; k < 4; ++k) coef[4 * date + k] = c[k];
}

// Exercise decision at date m for one path.  Returns true if state changed.
template <typename Real, bool REF>
__device__ __forceinline__ bool lsm_decide(const LsmParams& P, const double* __restrict__ disc,
                                           const double* __restrict__ cm, int m, Real S,
                                           int32_t& when, Real& cf, Real& st) {
    const int flag = (int)cm[0];
    if (flag != kFlagSmall && flag != kFlagRegress && flag != kFlagReduced) return false;
    const Real pvr = pay<Real>(S, (Real)P.E, (Real)P.pf);
    const double pv = (double)pvr;
    if (!(pv > 0.0)) return false;
    if (REF) {
#pragma clang fp contract(off)
        const double x = (double)S - P.E;
        if (flag == kFlagSmall) {
            const double y = exp(P.nrdt * ((double)when - (double)m)) * (double)cf;
            if (pv > y) {
                when = m;
                cf = pvr;
                st = pvr;
                return true;
            }
            return false;
        }
        // The regression pass skips x == -1, the reference's "not in the money"
        // marker (src/mc_amer.cpp:98).
        if (x == -1.0) return false;
        const double yhat = cm[1] + cm[2] * x + cm[3] * (x * x);
        const double test = pay<double>(x, P.E, P.pf);  // payoff(x[i],E,payoff_fun)
        if (test > yhat) {
            when = m;
            cf = pvr;
            st = (Real)test;
            return true;
        }
        return false;
    }
    if (flag == kFlagSmall) {
        const double y = disc[when - m] * (double)cf;
        if (pv > y) {
            when = m;
            cf = pvr;
            st = pvr;
            return true;
        }
        return false;
    }
    const double x = (double)S * P.invE;
    const double yhat = cm[1] + x * (cm[2] + x * cm[3]);
    if (pv > yhat) {
        when = m;
        cf = pvr;
        st = pvr;
        return true;
    }
    return false;
}

// lsm_decide's production branch with x kept from the x-moments pass.
template <typename Real>
__device__ __forceinline__ void lsm_decide_kept(const LsmParams& P, const double* __restrict__ disc,
                                                const double* __restrict__ cm, int m, Real S, double xo,
                                                int32_t& when, Real& cf) {
    const int flag = (int)cm[0];
    if (flag != kFlagSmall && flag != kFlagRegress && flag != kFlagReduced) return;
    if (!(xo >= 0.0)) return;
    const Real pvr = pay<Real>(S, (Real)P.E, (Real)P.pf);
    const double pv = (double)pvr;
    if (flag == kFlagSmall) {
        const double y = disc[when - m] * (double)cf;
        if (pv > y) {
            when = m;
            cf = pvr;
        }
        return;
    }
    const double yhat = cm[1] + xo * (cm[2] + xo * cm[3]);
    if (pv > yhat) {
        when = m;
        cf = pvr;
    }
}

// The three kept-x passes without a branch per path (used up to 2 pairs per
// thread), so that a thread's paths interleave -- their LDS reads issued
// together -- instead of running one exec-masked region after another.  A
// path out of the money, or not owned (its staged value may be stale),
// contributes exactly zero: its S (x-moments) or discount factor (y-terms) is
// replaced by 0 before any product, so each fused term becomes fma(a, 0, acc)
// with a finite, i.e. acc + 0, which leaves the non-negative sums unchanged.  The fused forms are those the branchy passes above
// compile to (acc1 = fma(1/E, S, acc1), acc2 = fma(x, x, acc2), ...,
// acc5 = fma(disc, cf, acc5), acc6 = fma(x, y, acc6)), so every sum receives
// the same roundings in the same path order: the same bits.
template <typename Real>
__device__ __forceinline__ void lsm_x_keep_flat(const LsmParams& P, bool own, Real S, double acc[5], double& xo) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    const bool itm = own && pv > 0.0;
    const double Sd = itm ? (double)S : 0.0;  // then x = x2 = 0 and every term adds 0
    const double x = Sd * P.invE;
    const double x2 = x * x;
    acc[0] += (double)itm;
    acc[1] = fma(P.invE, Sd, acc[1]);
    acc[2] = fma(x, x, acc[2]);
    acc[3] = fma(x, x2, acc[3]);
    acc[4] = fma(x2, x2, acc[4]);
    xo = itm ? x : -1.0;
}
template <typename Real>
__device__ __forceinline__ void lsm_y_kept_flat(const double* __restrict__ disc, int m, double xo, int32_t when,
                                                Real cf, double acc[8]) {
    const bool itm = xo >= 0.0;
    const double d = itm ? disc[when - m] : 0.0;  // (cf is finite: 0 or a payoff)
    const double cfd = (double)cf;
    const double y = d * cfd;
    acc[5] = fma(d, cfd, acc[5]);
    acc[6] = fma(xo, y, acc[6]);
    acc[7] = fma(xo * xo, y, acc[7]);
}
// lsm_decide_kept for a date whose flag is small, regress or reduced (the
// caller tests the flag, which is the same in every thread).
template <typename Real>
__device__ __forceinline__ void lsm_decide_kept_flat(const LsmParams& P, const double* __restrict__ disc,
                                                     const double* __restrict__ cm, bool small, int m, Real S,
                                                     double xo, int32_t& when, Real& cf) {
    const Real pvr = pay<Real>(S, (Real)P.E, (Real)P.pf);
    const double pv = (double)pvr;
    const double cont = small ? disc[when - m] * (double)cf : cm[1] + xo * (cm[2] + xo * cm[3]);
    const bool ex = xo >= 0.0 && pv > cont;
    when = ex ? m : when;
    cf = ex ? pvr : cf;
}

// ---- K-term basis {1, x, ..., x^(K-1)}, x = S/E - 1 (K = 2, 4, 5, 6) --------
// The quadratic regression above generalised to K terms (the reference
// regresses on {1, x, x^2}, src/mc_amer.cpp:38-108; its raw-S experiments
// depr/mc_amer/v3/mc_amer.cpp:70-148 keep three terms).  x is centred at the
// strike so that the Hankel Gram stays usable in FP64 up to K = 6 (condition
// ~5e9 for in-the-money puts, against ~4e11 on S/E).  Per date the sums are
// the x-moments g[j] = sum x^j (j = 0..2K-2, g[0] = count) and the y-terms
// g[2K-1+i] = sum y x^i (i = 0..K-1); they travel as chunks of 8 through the
// same wave and grid reductions as the quadratic's 8 sums.  Production
// arithmetic only (the reference-exact mode keeps the reference's basis).
template <int K>
struct LsmBasis {
    static constexpr int kMom = 2 * K - 1;            // x-moments
    static constexpr int kSums = 3 * K - 1;           // + y-terms
    static constexpr int kChunks = (kSums + 7) / 8;   // reductions of 8
    static constexpr int kCoef = K == 3 ? 4 : 8;      // per-date stride: flag + K coefficients
};

template <typename Real, int K>
__device__ __forceinline__ void lsm_poly_x(const LsmParams& P, Real S, double* __restrict__ acc) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    if (pv > 0.0) {
        const double x = (double)S * P.invE - 1.0;
        double p = x;
        acc[0] += 1.0;
        acc[1] += p;
#pragma unroll
        for (int j = 2; j < 2 * K - 1; ++j) {
            p = p * x;
            acc[j] += p;
        }
    }
}

// y-terms into acc[0..K-1] (the caller passes &sums[2K-1]).
template <typename Real, int K>
__device__ __forceinline__ void lsm_poly_y(const LsmParams& P, const double* __restrict__ disc, int m, Real S,
                                           int32_t when, Real cf, double* __restrict__ acc) {
    const double pv = (double)pay<Real>(S, (Real)P.E, (Real)P.pf);
    if (pv > 0.0) {
        const double x = (double)S * P.invE - 1.0;
        double q = disc[when - m] * (double)cf;
        acc[0] += q;
#pragma unroll
        for (int i = 1; i < K; ++i) {
            q = q * x;
            acc[i] += q;
        }
    }
}

// Normal equations of the K-term basis -> c = {flag, c_0 .. c_(K-1)}.
// Cholesky of G[i][j] = g[i+j]; column j is accepted while its pivot exceeds
// 1e-13 G[j][j], and at the first rejected column the leading kk x kk system
// is solved (flag "reduced", the dropped coefficients 0).  A pure function of
// g, evaluated redundantly by every thread.
template <int K>
__device__ __forceinline__ void lsm_solve_poly(const double* g, double c[K + 1]) {
    const double n = g[0];
#pragma unroll
    for (int i = 1; i <= K; ++i) c[i] = 0.0;
    if (n == 0.0) {
        c[0] = kFlagSkip;
        return;
    }
    if (n <= 2.0) {
        c[0] = kFlagSmall;
        return;
    }
    double L[K][K] = {}, id[K] = {}, y[K], x[K];
    int kk = K;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double p = g[2 * j];
#pragma unroll
        for (int t = 0; t < j; ++t) p -= L[j][t] * L[j][t];
        if (kk == K && j > 0 && !(p > 1e-13 * g[2 * j])) kk = j;
        if (j < kk) {
            L[j][j] = sqrt(p);
            id[j] = 1.0 / L[j][j];
#pragma unroll
            for (int i = j + 1; i < K; ++i) {
                double s = g[i + j];
#pragma unroll
                for (int t = 0; t < j; ++t) s -= L[i][t] * L[j][t];
                L[i][j] = s * id[j];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double s = g[2 * K - 1 + i];
#pragma unroll
        for (int t = 0; t < i; ++t) s -= L[i][t] * y[t];
        y[i] = i < kk ? s * id[i] : 0.0;
    }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int t = i + 1; t < K; ++t) s -= L[t][i] * x[t];  // x[t] = 0 past kk
        x[i] = i < kk ? s * id[i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < K; ++i) c[1 + i] = x[i];
    c[0] = kk == K ? kFlagRegress : kFlagReduced;
}

// Exercise decision at date m with the K-term fit (Horner in x).
template <typename Real, int K>
__device__ __forceinline__ void lsm_poly_decide(const LsmParams& P, const double* __restrict__ disc,
                                                const double* __restrict__ cm, int m, Real S, int32_t& when,
                                                Real& cf) {
    const int flag = (int)cm[0];
    if (flag != kFlagSmall && flag != kFlagRegress && flag != kFlagReduced) return;
    const Real pvr = pay<Real>(S, (Real)P.E, (Real)P.pf);
    const double pv = (double)pvr;
    if (!(pv > 0.0)) return;
    double v;
    if (flag == kFlagSmall) {
        v = disc[when - m] * (double)cf;
    } else {
        const double x = (double)S * P.invE - 1.0;
        v = cm[K];
#pragma unroll
        for (int i = K - 1; i >= 1; --i) v = cm[i] + x * v;
    }
    if (pv > v) {
        when = m;
        cf = pvr;
    }
}

// Normal equations from grid totals tot.v[off..] -> coef[date]: the quadratic
// (lsm_solve) for K = 3, the K-term basis otherwise.  Runs in one thread.
template <bool REF, int K, int NT>
__device__ void lsm_solve_tot(const pcf::Acc<NT>& tot, int off, double* __restrict__ coef, int date) {
    if constexpr (K == 3) {
        pcf::Acc<8> g;
#pragma unroll
        for (int k = 0; k < 8; ++k) g.v[k] = tot.v[off + k];
        lsm_solve<REF>(g, coef, date);
    } else {
        double g[LsmBasis<K>::kSums], c[K + 1];
#pragma unroll
        for (int k = 0; k < LsmBasis<K>::kSums; ++k) g[k] = tot.v[off + k].s + tot.v[off + k].c;
        lsm_solve_poly<K>(g, c);
#pragma unroll
        for (int k = 0; k <= K; ++k) coef[LsmBasis<K>::kCoef * date + k] = c[k];
    }
}

// --- k_lsm_paths ------------------------------------------------------------
// Verification store: host normals w[n*M + (m-1)] (pathsfinder's draw order),
// the reference's recurrence in FP64 without contraction.
// Threads per block of the per-date (graph) form's kernels; their LDS
// reduction buffers are sized for it and the host refuses larger blocks.
constexpr int kLsmGraphThreads = 256;
// Paths whose loads a thread of the per-date kernels issues together.
constexpr int kLsmBatch = 4;

template <bool REF, int K = 3>
__global__ __launch_bounds__(kLsmGraphThreads) void k_lsm_paths_host(LsmBufs b) {
    using Real = double;
    constexpr int NS = K == 3 ? 8 : LsmBasis<K>::kSums;
    __shared__ pcf::CSum lds[(kLsmGraphThreads / pcf::kWave) * (NS + 2)];
    const LsmParams& P = *b.P;
    Real* __restrict__ store = (Real*)b.store;
    Real* __restrict__ cfv = (Real*)b.cf;
    Real* __restrict__ stv = (Real*)b.st;
    const int64_t N = P.N, half = P.half;
    const int M = P.M;
    double acc[NS] = {};
    double e0 = 0.0, e1 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < half; n += stride) {
        Real Sp = (Real)P.S0, Sm = (Real)P.S0, Sp_prev = Sp, Sm_prev = Sm;
        for (int m = 1; m <= M; ++m) {
            Sp_prev = Sp;
            Sm_prev = Sm;
            {
#pragma clang fp contract(off)
                const double w = b.hw[n * M + (m - 1)];
                // paths[m][n] = paths[m-1][n]*exp((r-0.5*sigma*sigma)*dt +/- sigma*w)
                Sp = Sp * exp(P.h_mu + P.sigma * w);
                Sm = Sm * exp(P.h_mu - P.sigma * w);
            }
            store[(int64_t)(m - 1) * N + n] = Sp;
            store[(int64_t)(m - 1) * N + n + half] = Sm;
        }
        const Real cp = pay<Real>(Sp, (Real)P.E, (Real)P.pf);
        const Real cm_ = pay<Real>(Sm, (Real)P.E, (Real)P.pf);
        b.when[n] = M;
        b.when[n + half] = M;
        cfv[n] = cp;
        cfv[n + half] = cm_;
        if (REF) {
            stv[n] = cp;
            stv[n + half] = cm_;
        }
        e0 += (double)cp + (double)cm_;
        e1 += (double)cp * (double)cp + (double)cm_ * (double)cm_;
        if (M >= 2) {
            if constexpr (K == 3) {
                lsm_accumulate<Real, REF>(P, b.disc, M - 1, Sp_prev, M, cp, acc);
                lsm_accumulate<Real, REF>(P, b.disc, M - 1, Sm_prev, M, cm_, acc);
            } else {
                lsm_poly_x<Real, K>(P, Sp_prev, acc);
                lsm_poly_y<Real, K>(P, b.disc, M - 1, Sp_prev, M, cp, acc + LsmBasis<K>::kMom);
                lsm_poly_x<Real, K>(P, Sm_prev, acc);
                lsm_poly_y<Real, K>(P, b.disc, M - 1, Sm_prev, M, cm_, acc + LsmBasis<K>::kMom);
            }
        }
    }
    pcf::Acc<NS + 2> a, tot;
    a.v[0] = {e0, 0.0};
    a.v[1] = {e1, 0.0};
#pragma unroll
    for (int k = 0; k < NS; ++k) a.v[2 + k] = {acc[k], 0.0};
    if (pcf::grid_reduce<NS + 2>(a, b.partials, b.counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0) {
        b.out[0] = tot.v[0].s + tot.v[0].c;
        b.out[1] = tot.v[1].s + tot.v[1].c;
        if (M >= 2) lsm_solve_tot<REF, K>(tot, 2, b.coef, M - 1);
    }
}

// Production store: one Philox block (pair n, block (m-1)/4) feeds 4 dates.
template <typename Real, bool REF, int K = 3>
__global__ __launch_bounds__(kLsmGraphThreads) void k_lsm_paths_philox(LsmBufs b) {
    constexpr int NS = K == 3 ? 8 : LsmBasis<K>::kSums;
    __shared__ pcf::CSum lds[(kLsmGraphThreads / pcf::kWave) * (NS + 2)];
    const LsmParams& P = *b.P;
    Real* __restrict__ store = (Real*)b.store;
    Real* __restrict__ cfv = (Real*)b.cf;
    Real* __restrict__ stv = (Real*)b.st;
    const int64_t N = P.N, half = P.half;
    const int M = P.M;
    double acc[NS] = {};
    double e0 = 0.0, e1 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < half; n += stride) {
        Real Sp = (Real)P.S0, Sm = (Real)P.S0, Sp_prev = Sp, Sm_prev = Sm;
        for (int m0 = 1; m0 <= M; m0 += 4) {
            Real z[4];
            pcf::normals4<Real>(P.seed, PCF_STREAM_AMER, (uint64_t)n, (uint32_t)((m0 - 1) >> 2), z);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + q;
                if (m <= M) {
                    Sp_prev = Sp;
                    Sm_prev = Sm;
                    if constexpr (sizeof(Real) == 4) {
                        Sp *= pcf::exp2_f32(fmaf(P.vol2f, z[q], P.mu2f));
                        Sm *= pcf::exp2_f32(fmaf(-P.vol2f, z[q], P.mu2f));
                    } else {
                        Sp *= exp(fma(P.vol, (double)z[q], P.mu));
                        Sm *= exp(fma(-P.vol, (double)z[q], P.mu));
                    }
                    store[(int64_t)(m - 1) * N + n] = Sp;
                    store[(int64_t)(m - 1) * N + n + half] = Sm;
                }
            }
        }
        const Real cp = pay<Real>(Sp, (Real)P.E, (Real)P.pf);
        const Real cm_ = pay<Real>(Sm, (Real)P.E, (Real)P.pf);
        b.when[n] = M;
        b.when[n + half] = M;
        cfv[n] = cp;
        cfv[n + half] = cm_;
        if (REF) {
            stv[n] = cp;
            stv[n + half] = cm_;
        }
        e0 += (double)cp + (double)cm_;
        e1 += (double)cp * (double)cp + (double)cm_ * (double)cm_;
        if (M >= 2) {
            if constexpr (K == 3) {
                lsm_accumulate<Real, REF>(P, b.disc, M - 1, Sp_prev, M, cp, acc);
                lsm_accumulate<Real, REF>(P, b.disc, M - 1, Sm_prev, M, cm_, acc);
            } else {
                lsm_poly_x<Real, K>(P, Sp_prev, acc);
                lsm_poly_y<Real, K>(P, b.disc, M - 1, Sp_prev, M, cp, acc + LsmBasis<K>::kMom);
                lsm_poly_x<Real, K>(P, Sm_prev, acc);
                lsm_poly_y<Real, K>(P, b.disc, M - 1, Sm_prev, M, cm_, acc + LsmBasis<K>::kMom);
            }
        }
    }
    pcf::Acc<NS + 2> a, tot;
    a.v[0] = {e0, 0.0};
    a.v[1] = {e1, 0.0};
#pragma unroll
    for (int k = 0; k < NS; ++k) a.v[2 + k] = {acc[k], 0.0};
    if (pcf::grid_reduce<NS + 2>(a, b.partials, b.counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0) {
        b.out[0] = tot.v[0].s + tot.v[0].c;
        b.out[1] = tot.v[1].s + tot.v[1].c;
        if (M >= 2) lsm_solve_tot<REF, K>(tot, 2, b.coef, M - 1);
    }
}

// --- k_lsm_step(m): decide at m, sums at m-1 --------------------------------
template <typename Real, bool REF, int K = 3>
__global__ __launch_bounds__(kLsmGraphThreads) void k_lsm_step(LsmBufs b, int m) {
    constexpr int NS = K == 3 ? 8 : LsmBasis<K>::kSums;
    __shared__ pcf::CSum lds[(kLsmGraphThreads / pcf::kWave) * NS];
    const LsmParams& P = *b.P;
    const Real* __restrict__ store = (const Real*)b.store;
    Real* __restrict__ cfv = (Real*)b.cf;
    Real* __restrict__ stv = (Real*)b.st;
    const int64_t N = P.N;
    const double* __restrict__ cm = b.coef + LsmBasis<K>::kCoef * m;
    double acc[NS] = {};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // A thread's paths are taken kLsmBatch at a time: every load of the batch
    // is issued before the first path is worked on, so the loads' latency is
    // paid once per batch rather than once per path.  Paths are still summed
    // in the order n, n + stride, n + 2 stride, ...
    for (int64_t n0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n0 < N; n0 += kLsmBatch * stride) {
        int32_t wv[kLsmBatch];
        Real cv[kLsmBatch], Sv[kLsmBatch], Spv[kLsmBatch], sv[kLsmBatch];
#pragma unroll
        for (int u = 0; u < kLsmBatch; ++u) {
            const int64_t n = n0 + u * stride;
            const bool in = n < N;
            wv[u] = in ? b.when[n] : 0;
            cv[u] = in ? cfv[n] : (Real)0;
            Sv[u] = in ? store[(int64_t)(m - 1) * N + n] : (Real)0;
            Spv[u] = in ? store[(int64_t)(m - 2) * N + n] : (Real)0;
            sv[u] = REF && in ? stv[n] : (Real)0;
        }
#pragma unroll
        for (int u = 0; u < kLsmBatch; ++u) {
            const int64_t n = n0 + u * stride;
            if (n >= N) break;
            int32_t when = wv[u];
            Real cf = cv[u];
            const Real S = Sv[u];
            const Real Sprev = Spv[u];
            if constexpr (K == 3) {
                Real st = REF ? sv[u] : cf;
                if (lsm_decide<Real, REF>(P, b.disc, cm, m, S, when, cf, st)) {
                    b.when[n] = when;
                    cfv[n] = cf;
                    if (REF) stv[n] = st;
                }
                lsm_accumulate<Real, REF>(P, b.disc, m - 1, Sprev, when, cf, acc);
            } else {
                const int32_t when0 = when;
                lsm_poly_decide<Real, K>(P, b.disc, cm, m, S, when, cf);
                if (when != when0) {
                    b.when[n] = when;
                    cfv[n] = cf;
                }
                lsm_poly_x<Real, K>(P, Sprev, acc);
                lsm_poly_y<Real, K>(P, b.disc, m - 1, Sprev, when, cf, acc + LsmBasis<K>::kMom);
            }
        }
    }
    pcf::Acc<NS> a, tot;
#pragma unroll
    for (int k = 0; k < NS; ++k) a.v[k] = {acc[k], 0.0};
    if (pcf::grid_reduce<NS>(a, b.partials, b.counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0)
        lsm_solve_tot<REF, K>(tot, 0, b.coef, m - 1);
}

// --- k_lsm_final: decide at date 1 (if M >= 2), sum discounted cash flows ----
template <typename Real, bool REF, int K = 3>
__global__ __launch_bounds__(kLsmGraphThreads) void k_lsm_final(LsmBufs b) {
    __shared__ pcf::CSum lds[(kLsmGraphThreads / pcf::kWave) * 2];
    const LsmParams& P = *b.P;
    const Real* __restrict__ store = (const Real*)b.store;
    Real* __restrict__ cfv = (Real*)b.cf;
    Real* __restrict__ stv = (Real*)b.st;
    const int64_t N = P.N;
    const int M = P.M;
    const double* __restrict__ cm = b.coef + LsmBasis<K>::kCoef;
    double s = 0.0, s2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t n0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n0 < N; n0 += kLsmBatch * stride) {
        int32_t wv[kLsmBatch];
        Real cv[kLsmBatch], Sv[kLsmBatch], sv[kLsmBatch];
#pragma unroll
        for (int u = 0; u < kLsmBatch; ++u) {
            const int64_t n = n0 + u * stride;
            const bool in = n < N;
            wv[u] = in ? b.when[n] : 0;
            cv[u] = in ? cfv[n] : (Real)0;
            Sv[u] = in && M >= 2 ? store[n] : (Real)0;  // row of date 1
            sv[u] = REF && in ? stv[n] : (Real)0;
        }
#pragma unroll
        for (int u = 0; u < kLsmBatch; ++u) {
            const int64_t n = n0 + u * stride;
            if (n >= N) break;
            int32_t when = wv[u];
            Real cf = cv[u];
            Real st = REF ? sv[u] : cf;
            if (M >= 2) {
                const Real S = Sv[u];
                if constexpr (K == 3) {
                    if (lsm_decide<Real, REF>(P, b.disc, cm, 1, S, when, cf, st)) {
                        b.when[n] = when;
                        cfv[n] = cf;
                        if (REF) stv[n] = st;
                    }
                } else {
                    const int32_t when0 = when;
                    lsm_poly_decide<Real, K>(P, b.disc, cm, 1, S, when, cf);
                    st = cf;
                    if (when != when0) {
                        b.when[n] = when;
                        cfv[n] = cf;
                    }
                }
            }
            double v;
            if (REF) {
#pragma clang fp contract(off)
                // if(exercise_st[n]!=0) result+=exp(-r*exercise_when[n]*dt)*exercise_st[n];
                v = ((double)st != 0.0) ? exp(P.nr * (double)when * P.dt) * (double)st : 0.0;
            } else {
                v = b.disc[when] * (double)st;
            }
            s += v;
            s2 += v * v;
        }
    }
    pcf::Acc<2> a, tot;
    a.v[0] = {s, 0.0};
    a.v[1] = {s2, 0.0};
    if (pcf::grid_reduce<2>(a, b.partials, b.counter, lds, tot, blockIdx.x, gridDim.x) && threadIdx.x == 0) {
        b.out[2] = tot.v[0].s + tot.v[0].c;
        b.out[3] = tot.v[1].s + tot.v[1].c;
    }
}

// --- persistent form ----------------------------------------------------------
// One launch prices the whole request (grid co-resident).  Thread t owns antithetic
// pairs n = t, t + T, ... (T threads in the grid, at most PMAX pairs each) for
// the forward path generation and the whole backward sweep, so the per-path
// state (exercise date, true and booked cash flows) stays in registers and
// each thread reads back only the store entries it wrote itself.  Per date,
// one all-reduce: every block publishes its 8 sums and counts its arrival on
// a counter that only grows (date k of the sweep is complete at k * grid
// arrivals); once all blocks have arrived, every block sums the partials in
// the same fixed order and solves the same normal equations itself, so no
// block waits for another's solve.  Partials alternate between two slot sets
// by date parity: a block can write date m only after every block arrived at
// date m+1, that is after every block finished reading date m+2's slots.
// Every wait is bounded: past the limit a block sets the error word and stops
// waiting (the host reports PCF_ERR_TIMEOUT), so every wave always reaches
// the end of the kernel.
// Per-date arrivals are spread over kArrivalCounters counters, each on its
// own 128-byte line (block b counts on counter b % kArrivalCounters): device-
// scope atomics on one address serialise at the memory side (~25 ns each,
// measured), so one shared counter costs ~grid x 25 ns per date.
constexpr int kArrivalCounters = 32;
constexpr int kCounterStride = 32;  // u32 words = 128 bytes
constexpr int kSyncWords = kArrivalCounters * kCounterStride + 32;  // + fin, err

struct LsmSync {
    unsigned int* arrive;  // kArrivalCounters counters, kCounterStride apart; each grows
                           // by its block count per date
    unsigned int* fin;     // arrivals at the final sums
    unsigned int* err;     // nonzero once any wait gave up
    unsigned long long* sv;  // self-validating date slots (lsm_publish_sv), zero at rest
};

constexpr unsigned int kSpinLimit = 1u << 20;

// Publish the block total (held by thread 0), count arrivals; true in the
// last block, whose thread 0 then holds the fixed-order total in `tot`.
// Coherent stores/loads + vmcnt ordering, no fences (see pcf_device.hpp).
template <int K>
__device__ __forceinline__ bool lsm_arrive(pcf::Acc<K> a, pcf::CSum* __restrict__ partials,
                                           unsigned int* arrive, pcf::CSum* lds, pcf::Acc<K>& tot) {
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            pcf::publish(&partials[(size_t)blockIdx.x * K + k].s, a.v[k].s);
            pcf::publish(&partials[(size_t)blockIdx.x * K + k].c, a.v[k].c);
        }
        pcf::stores_complete();
        const unsigned int prev = atomicAdd(arrive, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    pcf::Acc<K> r;
    r.zero();
    for (unsigned int bb = threadIdx.x; bb < gridDim.x; bb += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            pcf::CSum p;
            p.s = pcf::fetch(&partials[(size_t)bb * K + k].s);
            p.c = pcf::fetch(&partials[(size_t)bb * K + k].c);
            r.v[k] = pcf::csum_merge(r.v[k], p);
        }
    }
    __syncthreads();
    r = pcf::block_reduce<K>(r, lds);
    if (threadIdx.x == 0) tot = r;
    return true;
}

// Plain-FP64 block reduction of K sums in a fixed order (wave shuffles, then
// waves in order); the result is valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum_plain(double v[K], double* lds) {
#pragma unroll
    for (int off = pcf::kWave / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, pcf::kWave);
    const int lane = threadIdx.x & (pcf::kWave - 1), wave = threadIdx.x / pcf::kWave;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + pcf::kWave - 1) / pcf::kWave;
        for (int w = 1; w < nw; ++w)
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] += lds[w * K + k];
    }
}

// Sum of 8 values over the 64 lanes of a wave by transposition: at each of
// the xor-32/16/8 steps a lane keeps half of its remaining sums and adds its
// partner's copy of them; xor-4/2/1 then finish a plain tree.  Afterwards
// every lane whose bits 3..5 encode k (k = b3 + 2 b4 + 4 b5) holds sum k.
// 10 FP64 exchanges instead of 48; IEEE addition commutes, so both partners
// of an exchange form the same sum and the order is fixed.
// Lane exchanges without the LDS crossbar (gfx950): DPP within 16-lane rows
// (row_ror:n gives lane l the value of lane (l - n) mod 16; quad_perm 0x4E and
// 0xB1 are xor 2 and xor 1), and v_permlane32_swap / v_permlane16_swap, which
// swap the upper 32 lanes of x with the lower 32 of y (and the odd 16-lane
// rows of x with the even rows of y).  FP64 travels as two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <bool ROWS16>
__device__ __forceinline__ void permlane_swap_f64(double& x, double& y) {
    const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
    const uint32_t xl = (uint32_t)ux, xh = (uint32_t)(ux >> 32), yl = (uint32_t)uy, yh = (uint32_t)(uy >> 32);
    uint32_t rl0, rl1, rh0, rh1;
    if constexpr (ROWS16) {
        const auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        rl0 = l[0]; rl1 = l[1]; rh0 = h[0]; rh1 = h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        rl0 = l[0]; rl1 = l[1]; rh0 = h[0]; rh1 = h[1];
    }
    x = __builtin_bit_cast(double, ((uint64_t)rh0 << 32) | rl0);
    y = __builtin_bit_cast(double, ((uint64_t)rh1 << 32) | rl1);
}

// (The exchanges below pair exactly the lanes of the xor-32/16/8/4/2/1
// shuffles described above and add the same two operands, so the sums are
// the shuffles' sums bit for bit: after a swap, lane l holds its own kept
// value and its partner's given value in x and y, and x + y == y + x.)
__device__ __forceinline__ double wave_sum8(const double v[8]) {
    const int lane = threadIdx.x & (pcf::kWave - 1);
    double a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // xor 32: lanes < 32 keep v[j], lanes >= 32 keep v[j + 4]
        double x = v[j], y = v[j + 4];
        permlane_swap_f64<false>(x, y);
        a[j] = x + y;
    }
    double c[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {  // xor 16: even rows keep a[j], odd rows keep a[j + 2]
        double x = a[j], y = a[j + 2];
        permlane_swap_f64<true>(x, y);
        c[j] = x + y;
    }
    const bool b3 = lane & 8, b2 = lane & 4;
    const double keep = b3 ? c[1] : c[0];
    const double give = b3 ? c[0] : c[1];
    double d = keep + dpp_f64<0x128>(give);                      // xor 8: row_ror:8
    // xor 4: row_ror:4 for lanes with bit 2 set, row_ror:12 for the others.  Both
    // run in every lane before the select (a DPP in a divergent branch would
    // read inactive source lanes as 0).
    const double r4 = dpp_f64<0x124>(d), r12 = dpp_f64<0x12C>(d);
    d += b2 ? r4 : r12;
    d += dpp_f64<0x4E>(d);                                       // xor 2
    d += dpp_f64<0xB1>(d);                                       // xor 1
    return d;
}

// Block sum of 8 values: the wave sums above, then the waves in order, one
// lane per sum.  Returns sum k in thread k < 8 (0 elsewhere).  `lds` holds
// 16 waves x 8 doubles.
__device__ __forceinline__ double block_sum8(const double v[8], double* lds) {
    const double d = wave_sum8(v);
    const int lane = threadIdx.x & (pcf::kWave - 1), wave = threadIdx.x / pcf::kWave;
    if ((lane & 7) == 0) lds[wave * 8 + ((lane >> 3) & 1) + 2 * ((lane >> 4) & 1) + 4 * ((lane >> 5) & 1)] = d;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < 8) {
        const int nw = (blockDim.x + pcf::kWave - 1) / pcf::kWave;
        t = lds[threadIdx.x];
        for (int w = 1; w < nw; ++w) t += lds[w * 8 + threadIdx.x];
    }
    return t;
}

// Per-date all-reduce of the 8 regression sums, in three steps so that the
// kernel can work between the arrival and the wait:
//   lsm_publish: the block's 8 sums (in threads 0..7) go to this date's slots
//     with coherent stores that complete before the block's arrival is
//     counted: sum-major (slots[k * grid + block], so a wave's load of sum k
//     over consecutive blocks is one contiguous run), or block-major
//     (slots[8 * block + k]) where the gather lands them in LDS (LAND);
//   lsm_wait_all: wave 0 polls every arrival counter until each has counted
//     its blocks `gen` times (dates completed including this one);
//   lsm_gather_solve: every block reads all partials coherently and sums
//     them in the same fixed order, so the grid totals are bit-identical in
//     every block, and solves the normal equations from them.
template <bool LAND>
__device__ __forceinline__ void lsm_publish(double mine, double* __restrict__ slots, const LsmSync& s) {
    if (threadIdx.x < pcf::kWave) {  // wave 0, uniformly
        const int lane = threadIdx.x;
        if (lane < 8)
            pcf::publish(LAND ? slots + (size_t)blockIdx.x * 8 + lane : slots + (size_t)lane * gridDim.x + blockIdx.x,
                         mine);
        pcf::stores_complete();
        if (lane == 0)
            __hip_atomic_fetch_add(s.arrive + (blockIdx.x % kArrivalCounters) * kCounterStride, 1u,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Self-validating date slots, the all-reduce of small grids of the quadratic
// (at most kSvMaxGrid blocks, 1-2 pairs per thread).  A slot word is zero at
// rest and holds the bitwise complement of a published sum, which is never
// zero (the complement of an all-ones NaN), so a reader polls the sums
// themselves: the writer neither waits for its stores nor counts an arrival,
// and a date costs one trip to the coherence point (stores, then the readers'
// loads) instead of three (stores and their acknowledgement, the arrival and
// the counter poll, the loads of the sums).
// Three parities of slots, sum-major (slot k * grid + block): block b clears
// its slot of date m+1 once its gather of date m is complete (every block has
// then read date m+1), and that clear is complete before b publishes date m-1
// (b waits for its stores first).  A reader of date m-2, the next user of the
// parity, first saw b's date m-1 sums, so it finds b's slot cleared or
// rewritten, never holding date m+1.  The kernel's last block clears every
// parity, so each launch starts from zero.  Word-sized atomic stores and loads
// (single-copy atomic), as for every other published value.
// Every waiting lane polls its own block's sums, so the polling grows with
// the square of the grid: with 24-64 blocks a date took 0.5-2 us less than
// with arrival counters (same bits), with 256 blocks 0.7-3 us more (the polls
// crowd the lines being written and the row copies of the next date).
constexpr int kSvMaxGrid = 64;
constexpr int kSvParities = 3;
constexpr int kSvWords = kSvParities * kSvMaxGrid * 8;
__device__ __forceinline__ unsigned long long* lsm_sv_slots(unsigned long long* sv, int m) {
    return sv + (size_t)(m % kSvParities) * gridDim.x * 8;
}
__device__ __forceinline__ void lsm_publish_sv(double mine, unsigned long long* __restrict__ slots) {
    if (threadIdx.x < pcf::kWave) {  // wave 0, uniformly
        pcf::stores_complete();      // this wave's clear of date m+2's slot (a date ago) is complete
        if (threadIdx.x < 8)
            __hip_atomic_store(slots + (size_t)threadIdx.x * gridDim.x + blockIdx.x,
                               ~__builtin_bit_cast(unsigned long long, mine), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}
// Clear this block's slot of date m+1 (threads 0..7, the publishing lanes),
// after its gather of date m.
__device__ __forceinline__ void lsm_clear_sv(unsigned long long* __restrict__ slots) {
    if (threadIdx.x < 8)
        __hip_atomic_store(slots + (size_t)threadIdx.x * gridDim.x + blockIdx.x, 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
// A poll (of the slots, or of the arrival counters in lsm_wait_all) re-reads
// the error word once per kSvErrPoll rounds, not on every round: each read is
// a trip to the coherence point.
constexpr unsigned int kSvErrPoll = 64;

__device__ __forceinline__ void lsm_wait_all(const LsmSync& s, unsigned int gen) {
    if (threadIdx.x < pcf::kWave) {  // wave 0, uniformly
        const int lane = threadIdx.x;
        // Lane c < kArrivalCounters waits for counter c to reach (its blocks) x gen.
        const unsigned int blocks_c =
            lane < kArrivalCounters && (unsigned)lane < gridDim.x
                ? (gridDim.x - 1 - (unsigned)lane) / kArrivalCounters + 1 : 0u;
        const unsigned int target = blocks_c * gen;
        unsigned int it = 0;
        while (true) {
            const unsigned int v =
                lane < kArrivalCounters ? pcf::fetch_u32(s.arrive + lane * kCounterStride) : 0u;
            if (__all(v >= target)) break;
            __builtin_amdgcn_s_sleep(1);
            // Past the bound, or once any block has given up, stop waiting.
            if (++it > kSpinLimit || (it % kSvErrPoll == 0 && pcf::fetch_u32(s.err) != 0u)) {
                if (lane == 0) atomicOr(s.err, 1u);
                break;
            }
        }
    }
    __syncthreads();
}

// 64-bit readlane: lane l's value of v, uniform over the wave.
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// lsm_gather then the solve of the normal equations, once per block.  Wave 0
// forms the 8 totals (threads 0..7, waves in order), takes all of them from
// those lanes, solves, and hands the coefficients to the block through LDS;
// the other waves wait at the barrier.  Had every thread solved, the waves
// sharing a SIMD (4 at 1024 threads) would each issue the same dependent
// FP64 chain (reciprocals) in turn.  The solve is a pure
// function of the totals, which are the same bits in every block, so every
// block still derives the same coefficients.
// With LAND (kernels of 3, 4, 6 or 8 pairs per thread), each lane's block published
// 64 contiguous bytes, and four 16-byte coherent loads per lane land them in
// LDS (global_load_lds: no registers), all in flight at once.  Through
// registers, these kernels sit at the 128-VGPR limit and the compiler issued
// the 8 loads one after another into the same register pair.  Segment s (64
// blocks) takes land[512 s ..]: load j of lane l writes sums 2j, 2j+1 at
// land[512 s + 128 j + 2 l].  Grids of the quadratic form hold at most
// kLandBlocks blocks (checked by the host).
constexpr int kLandBlocks = 256;
// Cache-policy operand of those loads: SC1, the bit pcf::fetch's agent-scope
// atomic loads carry on gfx950, so both read the published sums the same way.
constexpr int kLoadCoherent = 16;
// Gather forms: coherent loads into registers after the arrival wait, loads
// landed in LDS after that wait, or polled self-validating slots.
enum { kGatherFetch = 0, kGatherLand = 1, kGatherSv = 2 };
template <bool REF, int FORM>
__device__ __forceinline__ void lsm_gather_solve(double cm[4], const void* __restrict__ slots_v, double* red_b,
                                                 double* tot_sh, double* land, unsigned int* err) {
    const double* __restrict__ slots = (const double*)slots_v;
    // Only the first min(grid, block) threads fetch; the waves holding them
    // sum in the fixed order of block_sum8 (wave trees, then waves in order),
    // the others stay out.  Thread t adds blocks t, t + blockDim, ... in that
    // order on either path.
    const unsigned int active = gridDim.x < blockDim.x ? gridDim.x : blockDim.x;
    const int nw = (int)((active + pcf::kWave - 1) / pcf::kWave);
    const int lane = threadIdx.x & (pcf::kWave - 1), wave = threadIdx.x / pcf::kWave;
    if (wave < nw) {
        double t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = 0.0;
        if constexpr (FORM == kGatherLand) {
            for (unsigned int b0 = (unsigned)wave * pcf::kWave; b0 < gridDim.x; b0 += blockDim.x) {
                if (b0 + lane < gridDim.x) {
                    const double* src = slots + (size_t)(b0 + lane) * 8;
                    double* dst = land + (size_t)(b0 / pcf::kWave) * 512;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        __builtin_amdgcn_global_load_lds((const void*)(src + 2 * j),
                                                         (__attribute__((address_space(3))) void*)(dst + 128 * j),
                                                         16, 0, kLoadCoherent);
                }
            }
            pcf::stores_complete();  // this wave's loads have landed
            for (unsigned int b0 = (unsigned)wave * pcf::kWave; b0 < gridDim.x; b0 += blockDim.x)
                if (b0 + lane < gridDim.x) {
                    const double* row = land + (size_t)(b0 / pcf::kWave) * 512 + 2 * lane;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        t[2 * j] += row[128 * j];
                        t[2 * j + 1] += row[128 * j + 1];
                    }
                }
        } else if constexpr (FORM == kGatherSv) {
            // A lane whose blocks' sums are not all published yet sums them
            // again from zero (the same order, so the same bits), until every
            // lane of the wave has all of them; past the bound, or once any
            // block has given up, it stops waiting.
            const unsigned long long* __restrict__ sv = (const unsigned long long*)slots_v;
            bool miss = true;
            for (unsigned int it = 1;; ++it) {
                if (miss) {
                    miss = false;
#pragma unroll
                    for (int k = 0; k < 8; ++k) t[k] = 0.0;
                    for (unsigned int bb = threadIdx.x; bb < gridDim.x; bb += blockDim.x)
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const unsigned long long u = pcf::fetch_u64(sv + (size_t)k * gridDim.x + bb);
                            miss |= u == 0ull;
                            t[k] += __builtin_bit_cast(double, ~u);
                        }
                }
                if (!__any(miss)) break;
                __builtin_amdgcn_s_sleep(1);
                if (it > kSpinLimit || (it % kSvErrPoll == 0 && pcf::fetch_u32(err) != 0u)) {
                    if (miss) atomicOr(err, 1u);
                    break;
                }
            }
        } else {
            for (unsigned int bb = threadIdx.x; bb < gridDim.x; bb += blockDim.x)
#pragma unroll
                for (int k = 0; k < 8; ++k) t[k] += pcf::fetch(slots + (size_t)k * gridDim.x + bb);
        }
        const double d = wave_sum8(t);
        if ((lane & 7) == 0) red_b[wave * 8 + ((lane >> 3) & 1) + 2 * ((lane >> 4) & 1) + 4 * ((lane >> 5) & 1)] = d;
    }
    __syncthreads();
    if (wave == 0) {  // the whole wave (block sizes are multiples of 64)
        double tot = 0.0;
        if (lane < 8) {
            tot = red_b[lane];
            for (int w = 1; w < nw; ++w) tot += red_b[w * 8 + lane];
        }
        double g[8], c[4];
#pragma unroll
        for (int k = 0; k < 8; ++k) g[k] = readlane_f64(tot, k);
        lsm_solve_g<REF>(g, c);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 4; ++k) tot_sh[k] = c[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) cm[k] = tot_sh[k];
}

// The same three steps for 8C sums (C chunks of 8), used by the K-term basis:
// chunk c of the block's sums is reduced by block_sum8's tree into its own LDS
// rows (lds + 128 c), threads 0..7 then hold sum 8c + k in t[c]; the slots
// hold 8C sums per block, sum-major.
template <int C>
__device__ __forceinline__ void block_sum_chunks(const double* v, double* lds, double t[C]) {
    const int lane = threadIdx.x & (pcf::kWave - 1), wave = threadIdx.x / pcf::kWave;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const double d = wave_sum8(v + 8 * c);
        if ((lane & 7) == 0)
            lds[128 * c + wave * 8 + ((lane >> 3) & 1) + 2 * ((lane >> 4) & 1) + 4 * ((lane >> 5) & 1)] = d;
    }
    __syncthreads();
    const int nw = (blockDim.x + pcf::kWave - 1) / pcf::kWave;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        t[c] = 0.0;
        if (threadIdx.x < 8) {
            t[c] = lds[128 * c + threadIdx.x];
            for (int w = 1; w < nw; ++w) t[c] += lds[128 * c + w * 8 + threadIdx.x];
        }
    }
}

template <int C>
__device__ __forceinline__ void lsm_publish_chunks(const double t[C], double* __restrict__ slots, const LsmSync& s) {
    if (threadIdx.x < pcf::kWave) {  // wave 0, uniformly
        const int lane = threadIdx.x;
        if (lane < 8)
#pragma unroll
            for (int c = 0; c < C; ++c) pcf::publish(slots + (size_t)(8 * c + lane) * gridDim.x + blockIdx.x, t[c]);
        pcf::stores_complete();
        if (lane == 0)
            __hip_atomic_fetch_add(s.arrive + (blockIdx.x % kArrivalCounters) * kCounterStride, 1u,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The chunked gather and the K-term solve, once per block as in
// lsm_gather_solve; the coefficients (K + 1 <= 8C values) go through tot_sh.
template <int C, int K>
__device__ __forceinline__ void lsm_gather_solve_chunks(double cm[K + 1], const double* __restrict__ slots,
                                                        double* red_b, double* tot_sh) {
    const unsigned int active = gridDim.x < blockDim.x ? gridDim.x : blockDim.x;
    const int nw = (int)((active + pcf::kWave - 1) / pcf::kWave);
    const int lane = threadIdx.x & (pcf::kWave - 1), wave = threadIdx.x / pcf::kWave;
    if (wave < nw) {
        double t[8 * C];
#pragma unroll
        for (int k = 0; k < 8 * C; ++k) t[k] = 0.0;
        for (unsigned int bb = threadIdx.x; bb < gridDim.x; bb += blockDim.x)
#pragma unroll
            for (int k = 0; k < 8 * C; ++k) t[k] += pcf::fetch(slots + (size_t)k * gridDim.x + bb);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const double d = wave_sum8(t + 8 * c);
            if ((lane & 7) == 0)
                red_b[128 * c + wave * 8 + ((lane >> 3) & 1) + 2 * ((lane >> 4) & 1) + 4 * ((lane >> 5) & 1)] = d;
        }
    }
    __syncthreads();
    static_assert(K + 1 <= 8 * C, "coefficients pass through tot_sh");
    if (wave == 0) {
        double g[8 * C], c[K + 1];
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            double tot = 0.0;
            if (lane < 8) {
                tot = red_b[128 * cc + lane];
                for (int w = 1; w < nw; ++w) tot += red_b[128 * cc + w * 8 + lane];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) g[8 * cc + k] = readlane_f64(tot, k);
        }
        lsm_solve_poly<K>(g, c);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k <= K; ++k) tot_sh[k] = c[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k <= K; ++k) cm[k] = tot_sh[k];
}

// Row staging in LDS.  With kLdsRow, the path values of two dates live in the
// block's dynamic LDS, [2 date parities][2 * PMAX slots][threads], and the
// next date's row is copied there straight from HBM (global_load_lds, no
// registers) while the current date is processed.  Each wave copies exactly
// the 64 entries of every slot its own threads read (4-byte units: 1 copy per
// slot for FP32, 2 for FP64), so a wave only waits for its own copies.
template <typename Real, int PMAX>
constexpr bool lsm_lds_row() {
    return 2 * PMAX * sizeof(Real) <= 64;  // two parities x 1024 threads <= 128 KB
}
template <typename Real, int PMAX>
constexpr size_t lsm_lds_row_bytes(int threads) {
    return lsm_lds_row<Real, PMAX>() ? (size_t)2 * 2 * PMAX * threads * sizeof(Real) : 0;
}

// Copies the owned values of date m (store row m-1) into rowbuf parity m & 1.
// Lanes whose pair index is past N/2 copy nothing (their entries are never
// read), so every load stays inside the row.  The copies are non-temporal
// (each row is read once): A/B over two runs, bits unchanged, 2^21-2^22 paths
// 2-3% faster (FP32 2^22 0.776 -> 0.758 ms), smaller sizes unchanged.
template <typename Real, int PMAX>
__device__ __forceinline__ void lsm_stage_row(Real* rowbuf, const Real* __restrict__ store, int m, int64_t N,
                                              int64_t half, int64_t T) {
    constexpr int kUnits = (int)(sizeof(Real) / 4);  // 4-byte copies per entry
    const unsigned B = blockDim.x;
    const int wave = threadIdx.x / pcf::kWave, lane = threadIdx.x & (pcf::kWave - 1);
    Real* const buf = rowbuf + (size_t)(m & 1) * 2 * PMAX * B;
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
        const int64_t nw0 = (int64_t)blockIdx.x * B + (int64_t)k * T + (int64_t)wave * pcf::kWave;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const char* g = (const char*)(store + (int64_t)(m - 1) * N + nw0 + h * half);
            char* l = (char*)(buf + (size_t)(2 * k + h) * B + (size_t)wave * pcf::kWave);
#pragma unroll
            for (int j = 0; j < kUnits; ++j) {
                const int unit = j * pcf::kWave + lane;  // 4-byte unit of this wave's 64 entries
                if (nw0 + unit / kUnits < half)
                    __builtin_amdgcn_global_load_lds((const void*)(g + 4 * unit),
                                                     (__attribute__((address_space(3))) void*)(l + 4 * j * pcf::kWave),
                                                     4, 0, 2);
            }
        }
    }
}

// Discount factors exp(-r dt k), k = 0..M, live in LDS (the persistent form
// needs M + 1 <= kDiscLds; longer schedules run as the graph of launches).
constexpr int kDiscLds = 1024;

// Everything a persistent launch reads besides its buffers travels in its
// kernel arguments: the parameter block and, for M + 1 <= kDiscArgs, the
// discount table.  A request then needs no copy before its launch.
constexpr int kDiscArgs = 400;
struct LsmPersistArgs {
    LsmBufs b;
    LsmSync sy;
    LsmParams P;
    double disc[kDiscArgs];  // exp(-r dt k), k = 0..M; longer tables are read from b.disc
};
static_assert(sizeof(LsmPersistArgs) <= 4096, "kernel arguments are limited to 4 KB");

// Results of a persistent launch, written by the last block to arrive at the
// final sums: out[0..3] = European control sum and sum of squares, discounted
// cash-flow sum and sum of squares; out[4..6] = dates with 1-2 in-the-money
// paths, with a reduced basis, with a singular reference Gram matrix;
// out[7] = 1 if a date barrier wait gave up.  The same block then clears the
// sync words, so the next launch needs no fill.
constexpr int kPersistOut = 8;
static_assert(kPersistOut <= kOutDoubles, "the persistent results fit the context's pinned result buffer");

// K: regression basis size (3: the quadratic of lsm_solve_g, else the K-term
// basis of lsm_solve_poly).  SV: the quadratic at 1-2 pairs per thread
// all-reduces through self-validating slots (grids of at most kSvMaxGrid
// blocks, chosen by the host).
template <typename Real, int SRC, bool REF, int PMAX, int K = 3, bool SV = false>
__global__ __launch_bounds__(1024) void k_lsm_persistent(const LsmPersistArgs args) {
    using Bs = LsmBasis<K>;
    static_assert(K == 3 || !REF, "the reference-exact mode keeps the reference's basis");
    const LsmBufs& b = args.b;
    const LsmSync& sy = args.sy;
    int n_small = 0, n_reduced = 0, n_singular = 0;  // date flags (the same in every thread)
    __shared__ pcf::CSum lds[16 * 8];
    __shared__ double ldsp[16 * 8];
    __shared__ double red_a[16 * 8 * Bs::kChunks], red_b[16 * 8 * Bs::kChunks], tot_sh[8 * Bs::kChunks];
    // The per-date gather lands in LDS for the quadratic at 3-8 pairs per
    // thread (lsm_gather_solve).
    constexpr bool kLand = K == 3 && PMAX >= 3;
    __shared__ __attribute__((aligned(16))) double land[kLand ? 8 * kLandBlocks : 2];
    // Self-validating slots (lsm_publish_sv), with no arrival count.
    static_assert(!SV || (K == 3 && PMAX <= 2), "self-validating slots: the quadratic at 1-2 pairs per thread");
    constexpr bool kSv = SV;
    constexpr int kGather = kLand ? kGatherLand : (kSv ? kGatherSv : kGatherFetch);
    __shared__ double disc[kDiscLds];
    extern __shared__ __align__(16) unsigned char lsm_dyn_lds[];  // row staging (kLdsRow)
    // A private copy: the row copies into LDS count as memory writes, and
    // through a reference every use would reload (and wait for) the block.
    const LsmParams P = args.P;
    Real* __restrict__ store = (Real*)b.store;
    const int64_t N = P.N, half = P.half;
    const int M = P.M;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (M + 1 <= kDiscArgs)
        for (int k = threadIdx.x; k <= M; k += blockDim.x) disc[k] = args.disc[k];
    else
        for (int k = threadIdx.x; k <= M; k += blockDim.x) disc[k] = b.disc[k];
    // Date slots: 2 parities x grid x 8 C doubles, then the final sums' slots.
    double* __restrict__ slots = (double*)b.partials;
    __syncthreads();
    // Forward: the owned pairs' paths into the store (runtime loop: no state yet).
    for (int64_t n = t0; n < half && n < t0 + (int64_t)PMAX * T; n += T) {
        Real Sp = (Real)P.S0, Sm = (Real)P.S0;
        if constexpr (SRC == kHost) {
            for (int m = 1; m <= M; ++m) {
#pragma clang fp contract(off)
                const double w = b.hw[n * M + (m - 1)];
                Sp = Sp * exp(P.h_mu + P.sigma * w);
                Sm = Sm * exp(P.h_mu - P.sigma * w);
                store[(int64_t)(m - 1) * N + n] = Sp;
                store[(int64_t)(m - 1) * N + n + half] = Sm;
            }
        } else {
            for (int m0 = 1; m0 <= M; m0 += 4) {
                Real z[4];
                pcf::normals4<Real>(P.seed, PCF_STREAM_AMER, (uint64_t)n, (uint32_t)((m0 - 1) >> 2), z);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = m0 + q;
                    if (m <= M) {
                        if constexpr (sizeof(Real) == 4) {
                            Sp *= pcf::exp2_f32(fmaf(P.vol2f, z[q], P.mu2f));
                            Sm *= pcf::exp2_f32(fmaf(-P.vol2f, z[q], P.mu2f));
                        } else {
                            Sp *= exp(fma(P.vol, (double)z[q], P.mu));
                            Sm *= exp(fma(-P.vol, (double)z[q], P.mu));
                        }
                        store[(int64_t)(m - 1) * N + n] = Sp;
                        store[(int64_t)(m - 1) * N + n + half] = Sm;
                    }
                }
            }
        }
    }
    // State of the owned paths (slot 2k: path n, 2k+1: path n + N/2), from the
    // entries this thread just wrote.
    int32_t when[2 * PMAX];
    Real cf[2 * PMAX];
    Real stb[REF ? 2 * PMAX : 1];
    double e0 = 0.0, e1 = 0.0;
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
        const int64_t n = t0 + (int64_t)k * T;
        when[2 * k] = when[2 * k + 1] = M;
        cf[2 * k] = cf[2 * k + 1] = Real(0);
        if (n < half) {
            cf[2 * k] = pay<Real>(store[(int64_t)(M - 1) * N + n], (Real)P.E, (Real)P.pf);
            cf[2 * k + 1] = pay<Real>(store[(int64_t)(M - 1) * N + n + half], (Real)P.E, (Real)P.pf);
            e0 += (double)cf[2 * k] + (double)cf[2 * k + 1];
            e1 += (double)cf[2 * k] * (double)cf[2 * k] + (double)cf[2 * k + 1] * (double)cf[2 * k + 1];
        }
        if constexpr (REF) {
            stb[2 * k] = cf[2 * k];
            stb[2 * k + 1] = cf[2 * k + 1];
        }
    }
    // Backward sweep: one grid all-reduce per exercise date.  While a block
    // waits for the others at date m it forms the x-moments xm of date m-1,
    // which do not depend on the decisions at m.  The path values of a date
    // come from the LDS row staging (kLdsRow: the row of date m-1 is copied in
    // at the top of date m) or, where two staged rows do not fit, from
    // registers: Sc holds date m and date m-1 is loaded into Sn after the
    // block's arrival (at 5-8 pairs per thread the row of m-1 is loaded into
    // transient registers instead, Sc being reloaded at the next date).
    constexpr bool kLdsRow = lsm_lds_row<Real, PMAX>();
    constexpr bool kPrefetch = !kLdsRow && PMAX <= 4;
    Real* const rowbuf = reinterpret_cast<Real*>(lsm_dyn_lds);
    Real Sc[kLdsRow ? 1 : 2 * PMAX], Sn[kPrefetch ? 2 * PMAX : 1];
    auto load_row = [&](int m, Real* dst) {
#pragma unroll
        for (int k = 0; k < PMAX; ++k) {
            const int64_t n = t0 + (int64_t)k * T;
            const bool own = n < half;
            dst[2 * k] = own ? store[(int64_t)(m - 1) * N + n] : Real(0);
            dst[2 * k + 1] = own ? store[(int64_t)(m - 1) * N + n + half] : Real(0);
        }
    };
    // Value of slot i = 2k + h at date m from the staging or from `row`.
    auto sval = [&](const Real* row, int m, int i) -> Real {
        if constexpr (kLdsRow)
            return rowbuf[((size_t)(m & 1) * 2 * PMAX + i) * blockDim.x + threadIdx.x];
        return row[i];
    };
    // x kept between the passes (production quadratic, up to 4 pairs per
    // thread: at 8 the extra registers would spill): xn for the date whose
    // x-moments were formed last, xc for the date being processed.
    constexpr bool kKeepX = K == 3 && !REF && PMAX <= 4;
    double xc[kKeepX ? 2 * PMAX : 1], xn[kKeepX ? 2 * PMAX : 1];
    constexpr bool kFlat = kKeepX && PMAX <= 2;  // branch-free passes (lsm_x_keep_flat)
    // At 3 and 4 pairs only the x-moments pass is branch-free: the other two
    // would spill.
    constexpr bool kFlatX4 = kKeepX && PMAX >= 3;
    // At 5-8 pairs (x not kept) the FP32 kernel's x-moments pass is
    // branch-free too; its other passes, and the FP64 kernel's, would spill.
    constexpr bool kFlatX8 = K == 3 && !REF && PMAX >= 5 && sizeof(Real) == 4;
    auto x_moments = [&](const Real* row, int m, double xm[Bs::kMom]) {
#pragma unroll
        for (int k = 0; k < Bs::kMom; ++k) xm[k] = 0.0;
        if constexpr (kKeepX) {
#pragma unroll
            for (int i = 0; i < 2 * PMAX; ++i) xn[i] = -1.0;
        }
#pragma unroll
        for (int k = 0; k < PMAX; ++k)
            if constexpr (kFlat || kFlatX4) {
                const bool own = t0 + (int64_t)k * T < half;
                lsm_x_keep_flat<Real>(P, own, sval(row, m, 2 * k), xm, xn[2 * k]);
                lsm_x_keep_flat<Real>(P, own, sval(row, m, 2 * k + 1), xm, xn[2 * k + 1]);
            } else if constexpr (kFlatX8) {
                const bool own = t0 + (int64_t)k * T < half;
                double xo;  // (not kept)
                lsm_x_keep_flat<Real>(P, own, sval(row, m, 2 * k), xm, xo);
                lsm_x_keep_flat<Real>(P, own, sval(row, m, 2 * k + 1), xm, xo);
            } else if (t0 + (int64_t)k * T < half) {
                if constexpr (kKeepX) {
                    lsm_accumulate_x_keep<Real>(P, sval(row, m, 2 * k), xm, xn[2 * k]);
                    lsm_accumulate_x_keep<Real>(P, sval(row, m, 2 * k + 1), xm, xn[2 * k + 1]);
                } else if constexpr (K == 3) {
                    lsm_accumulate_x<Real, REF>(P, sval(row, m, 2 * k), xm);
                    lsm_accumulate_x<Real, REF>(P, sval(row, m, 2 * k + 1), xm);
                } else {
                    lsm_poly_x<Real, K>(P, sval(row, m, 2 * k), xm);
                    lsm_poly_x<Real, K>(P, sval(row, m, 2 * k + 1), xm);
                }
            }
    };
    double xm[Bs::kMom] = {};
    if (M >= 2) {
        if constexpr (kLdsRow) {
            lsm_stage_row<Real, PMAX>(rowbuf, store, M - 1, N, half, T);
            pcf::stores_complete();  // this wave's copies have landed
        } else {
            load_row(M - 1, Sc);
        }
        x_moments(Sc, M - 1, xm);
    }
    unsigned int gen = 0;
    for (int m = M - 1; m >= 1; --m) {
        ++gen;
        if constexpr (kKeepX) {
#pragma unroll
            for (int i = 0; i < 2 * PMAX; ++i) xc[i] = xn[i];
        }
        if constexpr (kLdsRow) {
            if (m > 1) lsm_stage_row<Real, PMAX>(rowbuf, store, m - 1, N, half, T);
        } else if constexpr (!kPrefetch) {
            if (m != M - 1) load_row(m, Sc);
        }
        double* const slots_m = slots + (size_t)(m & 1) * gridDim.x * (8 * Bs::kChunks);
        if constexpr (K == 3) {
            double acc[8] = {xm[0], xm[1], xm[2], xm[3], xm[4], 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < PMAX; ++k) {
                const int64_t n = t0 + (int64_t)k * T;
                if constexpr (kFlat) {
                    lsm_y_kept_flat<Real>(disc, m, xc[2 * k], when[2 * k], cf[2 * k], acc);
                    lsm_y_kept_flat<Real>(disc, m, xc[2 * k + 1], when[2 * k + 1], cf[2 * k + 1], acc);
                } else if (n < half) {
                    if constexpr (kKeepX) {
                        lsm_accumulate_y_kept<Real>(disc, m, xc[2 * k], when[2 * k], cf[2 * k], acc);
                        lsm_accumulate_y_kept<Real>(disc, m, xc[2 * k + 1], when[2 * k + 1], cf[2 * k + 1], acc);
                    } else {
                        lsm_accumulate_y<Real, REF>(P, disc, m, sval(Sc, m, 2 * k), when[2 * k], cf[2 * k], acc);
                        lsm_accumulate_y<Real, REF>(P, disc, m, sval(Sc, m, 2 * k + 1), when[2 * k + 1],
                                                    cf[2 * k + 1], acc);
                    }
                }
            }
            const double bs = block_sum8(acc, red_a);
            if constexpr (kSv)
                lsm_publish_sv(bs, lsm_sv_slots(sy.sv, m));
            else
                lsm_publish<kLand>(bs, slots_m, sy);
        } else {
            double acc[8 * Bs::kChunks];
#pragma unroll
            for (int j = 0; j < 8 * Bs::kChunks; ++j) acc[j] = j < Bs::kMom ? xm[j] : 0.0;
#pragma unroll
            for (int k = 0; k < PMAX; ++k) {
                const int64_t n = t0 + (int64_t)k * T;
                if (n < half) {
                    lsm_poly_y<Real, K>(P, disc, m, sval(Sc, m, 2 * k), when[2 * k], cf[2 * k], acc + Bs::kMom);
                    lsm_poly_y<Real, K>(P, disc, m, sval(Sc, m, 2 * k + 1), when[2 * k + 1], cf[2 * k + 1],
                                        acc + Bs::kMom);
                }
            }
            double t[Bs::kChunks];
            block_sum_chunks<Bs::kChunks>(acc, red_a, t);
            lsm_publish_chunks<Bs::kChunks>(t, slots_m, sy);
        }
        if (m > 1) {
            if constexpr (kLdsRow) {
                pcf::stores_complete();  // this wave's copies of date m-1 have landed
                x_moments(Sc, m - 1, xm);
            } else if constexpr (kPrefetch) {
                load_row(m - 1, Sn);
                x_moments(Sn, m - 1, xm);
            } else {
                Real Sx[2 * PMAX];
                load_row(m - 1, Sx);
                x_moments(Sx, m - 1, xm);
            }
        }
        if constexpr (!kSv) lsm_wait_all(sy, gen);
        if constexpr (K == 3) {
            double cm[4];
            if constexpr (kSv) {
                lsm_gather_solve<REF, kGather>(cm, lsm_sv_slots(sy.sv, m), red_b, tot_sh, land, sy.err);
                if (m + 1 <= M - 1) lsm_clear_sv(lsm_sv_slots(sy.sv, m + 1));
            } else {
                lsm_gather_solve<REF, kGather>(cm, slots_m, red_b, tot_sh, land, sy.err);
            }
            if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
                for (int k = 0; k < 4; ++k) b.coef[4 * m + k] = cm[k];
            n_small += (int)cm[0] == kFlagSmall;
            n_reduced += (int)cm[0] == kFlagReduced;
            n_singular += (int)cm[0] == kFlagSingular;
            if ((int)cm[0] == kFlagSmall || (int)cm[0] == kFlagRegress || (int)cm[0] == kFlagReduced) {
#pragma unroll
                for (int k = 0; k < PMAX; ++k) {
                    const int64_t n = t0 + (int64_t)k * T;
                    if constexpr (kFlat) {
#pragma unroll
                        for (int h = 0; h < 2; ++h)
                            lsm_decide_kept_flat<Real>(P, disc, cm, (int)cm[0] == kFlagSmall, m, sval(Sc, m, 2 * k + h),
                                                       xc[2 * k + h], when[2 * k + h], cf[2 * k + h]);
                    } else if (n < half) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const int i = 2 * k + h;
                            if constexpr (kKeepX) {
                                lsm_decide_kept<Real>(P, disc, cm, m, sval(Sc, m, i), xc[i], when[i], cf[i]);
                            } else {
                                Real stv = REF ? stb[REF ? i : 0] : cf[i];
                                (void)lsm_decide<Real, REF>(P, disc, cm, m, sval(Sc, m, i), when[i], cf[i], stv);
                                if constexpr (REF) stb[i] = stv;
                            }
                        }
                    }
                }
            }
        } else {
            double cm[K + 1];
            lsm_gather_solve_chunks<Bs::kChunks, K>(cm, slots_m, red_b, tot_sh);
            if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
                for (int k = 0; k <= K; ++k) b.coef[Bs::kCoef * m + k] = cm[k];
            n_small += (int)cm[0] == kFlagSmall;
            n_reduced += (int)cm[0] == kFlagReduced;
#pragma unroll
            for (int k = 0; k < PMAX; ++k) {
                const int64_t n = t0 + (int64_t)k * T;
                if (n < half) {
                    lsm_poly_decide<Real, K>(P, disc, cm, m, sval(Sc, m, 2 * k), when[2 * k], cf[2 * k]);
                    lsm_poly_decide<Real, K>(P, disc, cm, m, sval(Sc, m, 2 * k + 1), when[2 * k + 1], cf[2 * k + 1]);
                }
            }
        }
        if constexpr (kPrefetch)
#pragma unroll
            for (int i = 0; i < 2 * PMAX; ++i) Sc[i] = Sn[i];
        // (tot_sh and the reduction buffers are rewritten only after the
        // barriers of the next date, which every thread reaches after this
        // date's reads; a staging parity is rewritten one date after its last
        // read, by the same wave)
    }
    // Final: discounted booked cash flows + European control.
    double s = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < PMAX; ++k) {
        const int64_t n = t0 + (int64_t)k * T;
        if (n < half) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = 2 * k + h;
                double v;
                if constexpr (REF) {
#pragma clang fp contract(off)
                    v = ((double)stb[i] != 0.0) ? exp(P.nr * (double)when[i] * P.dt) * (double)stb[i] : 0.0;
                } else {
                    v = disc[when[i]] * (double)cf[i];
                }
                s += v;
                s2 += v * v;
                b.when[n + h * half] = when[i];
            }
        }
    }
    double fin[4] = {e0, e1, s, s2};
    block_sum_plain<4>(fin, ldsp);
    pcf::Acc<4> a, tot;
#pragma unroll
    for (int k = 0; k < 4; ++k) a.v[k] = {threadIdx.x == 0 ? fin[k] : 0.0, 0.0};
    // Final sums: last-arriving block, on its own counter and slots.  Every
    // block has passed its last date barrier before it arrives here, so the
    // last one may read the error word and return all sync words to zero.
    if (lsm_arrive<4>(a, (pcf::CSum*)(slots + (size_t)2 * gridDim.x * (8 * Bs::kChunks)), sy.fin, lds, tot)) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) b.out[k] = tot.v[k].s + tot.v[k].c;
            b.out[4] = (double)n_small;
            b.out[5] = (double)n_reduced;
            b.out[6] = (double)n_singular;
            b.out[7] = pcf::fetch_u32(sy.err) != 0u ? 1.0 : 0.0;
            pcf::publish_u32(sy.err, 0u);
            pcf::publish_u32(sy.fin, 0u);
        }
        if (threadIdx.x < kArrivalCounters) pcf::publish_u32(sy.arrive + threadIdx.x * kCounterStride, 0u);
        if constexpr (kSv)
            for (unsigned int i = threadIdx.x; i < kSvParities * gridDim.x * 8; i += blockDim.x)
                __hip_atomic_store(sy.sv + i, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Dynamic LDS of a persistent launch: the row staging only (kLdsRow), so that
// a block can still be placed on a CU that runs other kernels' blocks.
template <typename Real, int SRC, bool REF, int PMAX>
size_t lsm_dyn_lds_bytes(int threads) {
    return lsm_lds_row_bytes<Real, PMAX>(threads);
}

template <typename Real, int SRC, bool REF, int PMAX, int K = 3, bool SV = false>
hipError_t launch_persistent_t(const LsmPersistArgs& pa, int grid, int B, hipStream_t s, hipEvent_t ev0,
                               hipEvent_t ev1) {
    LsmPersistArgs a = pa;
    void* args[] = {&a};
    // A plain launch: the grid fits the device's co-resident capacity (checked
    // by the caller), persistent launches run one after another on one stream
    // per device, each after the kernels other contexts had enqueued, and
    // every other kernel of the library finishes on its own, so all blocks
    // become resident -- at worst after a co-running kernel ends.  Every wait
    // is bounded all the same.  ev0 / ev1 take the kernel's
    // own start and end times (no marker packets between consecutive launches).
    return hipExtLaunchKernel((const void*)k_lsm_persistent<Real, SRC, REF, PMAX, K, SV>, dim3((unsigned)grid),
                              dim3((unsigned)B), args, lsm_dyn_lds_bytes<Real, SRC, REF, PMAX>(B), s, ev0, ev1, 0);
}

// The quadratic at 1-2 pairs per thread: self-validating slots for grids of
// at most kSvMaxGrid blocks, arrival counters above.  (Co-residency and LDS
// are checked on the counted kernel: the two have the same LDS, and each holds
// one 1024-thread block per CU at most.)
template <typename Real, int SRC, bool REF, int PMAX>
hipError_t launch_persistent_q(const LsmPersistArgs& pa, int grid, int B, hipStream_t s, hipEvent_t ev0,
                               hipEvent_t ev1) {
    if constexpr (PMAX <= 2)
        if (grid <= kSvMaxGrid) return launch_persistent_t<Real, SRC, REF, PMAX, 3, true>(pa, grid, B, s, ev0, ev1);
    return launch_persistent_t<Real, SRC, REF, PMAX>(pa, grid, B, s, ev0, ev1);
}

template <typename Real, int SRC, bool REF, int PMAX, int K = 3>
int max_coresident_t(int B, int cus, int device) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_lsm_persistent<Real, SRC, REF, PMAX, K>, B,
                                                     lsm_dyn_lds_bytes<Real, SRC, REF, PMAX>(B)) != hipSuccess)
        return 0;
    return nb * cus;
}

// True when the kernel's static LDS plus its row staging fit one block.
template <typename Real, int SRC, bool REF, int PMAX, int K = 3>
bool lds_fits_t(int B, int device) {
    static int lds_block[64];
    static size_t lds_static[64];
    static bool have[64];
    static std::mutex mu;
    if (device < 0 || device >= 64) return false;
    std::lock_guard<std::mutex> lk(mu);
    if (!have[device]) {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, (const void*)k_lsm_persistent<Real, SRC, REF, PMAX, K>) != hipSuccess ||
            hipDeviceGetAttribute(&lds_block[device], hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
                hipSuccess)
            return false;
        lds_static[device] = fa.sharedSizeBytes;
        have[device] = true;
    }
    return lds_static[device] + lsm_dyn_lds_bytes<Real, SRC, REF, PMAX>(B) <= (size_t)lds_block[device];
}

// Dispatch over (Real, SRC, REF), the register budget PMAX in {1,2,4,8} and
// the basis size K: 3 (every mode), or 2, 4, 5, 6 (production arithmetic,
// ref == false, checked by the caller).
template <int PMAX, int K>
hipError_t launch_persistent_k(bool f64, int src, bool ref, const LsmPersistArgs& pa, int grid, int B,
                               hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if constexpr (K == 3) {
        if (!f64) return launch_persistent_q<float, kPhilox, false, PMAX>(pa, grid, B, s, e0, e1);
        if (src == kHost)
            return ref ? launch_persistent_q<double, kHost, true, PMAX>(pa, grid, B, s, e0, e1)
                       : launch_persistent_q<double, kHost, false, PMAX>(pa, grid, B, s, e0, e1);
        return ref ? launch_persistent_q<double, kPhilox, true, PMAX>(pa, grid, B, s, e0, e1)
                   : launch_persistent_q<double, kPhilox, false, PMAX>(pa, grid, B, s, e0, e1);
    } else {
        if (!f64) return launch_persistent_t<float, kPhilox, false, PMAX, K>(pa, grid, B, s, e0, e1);
        (void)ref;
        return src == kHost ? launch_persistent_t<double, kHost, false, PMAX, K>(pa, grid, B, s, e0, e1)
                            : launch_persistent_t<double, kPhilox, false, PMAX, K>(pa, grid, B, s, e0, e1);
    }
}
template <int PMAX>
hipError_t launch_persistent(int K, bool f64, int src, bool ref, const LsmPersistArgs& pa, int grid,
                             int B, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if constexpr (PMAX == 3 || PMAX == 5 || PMAX == 6) return launch_persistent_k<PMAX, 3>(f64, src, ref, pa, grid, B, s, e0, e1);
    switch (K) {
        case 2: return launch_persistent_k<PMAX, 2>(f64, src, ref, pa, grid, B, s, e0, e1);
        case 4: return launch_persistent_k<PMAX, 4>(f64, src, ref, pa, grid, B, s, e0, e1);
        case 5: return launch_persistent_k<PMAX, 5>(f64, src, ref, pa, grid, B, s, e0, e1);
        case 6: return launch_persistent_k<PMAX, 6>(f64, src, ref, pa, grid, B, s, e0, e1);
        default: return launch_persistent_k<PMAX, 3>(f64, src, ref, pa, grid, B, s, e0, e1);
    }
}

template <int PMAX, int K>
bool lds_fits_k(bool f64, int src, bool ref, int B, int dev) {
    if (!f64) return lds_fits_t<float, kPhilox, false, PMAX, K>(B, dev);
    if constexpr (K == 3) {
        if (src == kHost)
            return ref ? lds_fits_t<double, kHost, true, PMAX>(B, dev)
                       : lds_fits_t<double, kHost, false, PMAX>(B, dev);
        return ref ? lds_fits_t<double, kPhilox, true, PMAX>(B, dev)
                   : lds_fits_t<double, kPhilox, false, PMAX>(B, dev);
    } else {
        (void)ref;
        return src == kHost ? lds_fits_t<double, kHost, false, PMAX, K>(B, dev)
                            : lds_fits_t<double, kPhilox, false, PMAX, K>(B, dev);
    }
}
template <int PMAX>
bool lds_fits(int K, bool f64, int src, bool ref, int B, int dev) {
    if constexpr (PMAX == 3 || PMAX == 5 || PMAX == 6) return lds_fits_k<PMAX, 3>(f64, src, ref, B, dev);
    switch (K) {
        case 2: return lds_fits_k<PMAX, 2>(f64, src, ref, B, dev);
        case 4: return lds_fits_k<PMAX, 4>(f64, src, ref, B, dev);
        case 5: return lds_fits_k<PMAX, 5>(f64, src, ref, B, dev);
        case 6: return lds_fits_k<PMAX, 6>(f64, src, ref, B, dev);
        default: return lds_fits_k<PMAX, 3>(f64, src, ref, B, dev);
    }
}

template <int PMAX, int K>
int max_coresident_k(bool f64, int src, bool ref, int B, int cus, int dev) {
    if (!f64) return max_coresident_t<float, kPhilox, false, PMAX, K>(B, cus, dev);
    if constexpr (K == 3) {
        if (src == kHost)
            return ref ? max_coresident_t<double, kHost, true, PMAX>(B, cus, dev)
                       : max_coresident_t<double, kHost, false, PMAX>(B, cus, dev);
        return ref ? max_coresident_t<double, kPhilox, true, PMAX>(B, cus, dev)
                   : max_coresident_t<double, kPhilox, false, PMAX>(B, cus, dev);
    } else {
        (void)ref;
        return src == kHost ? max_coresident_t<double, kHost, false, PMAX, K>(B, cus, dev)
                            : max_coresident_t<double, kPhilox, false, PMAX, K>(B, cus, dev);
    }
}
template <int PMAX>
int max_coresident(int K, bool f64, int src, bool ref, int B, int cus, int dev) {
    if constexpr (PMAX == 3 || PMAX == 5 || PMAX == 6) return max_coresident_k<PMAX, 3>(f64, src, ref, B, cus, dev);
    switch (K) {
        case 2: return max_coresident_k<PMAX, 2>(f64, src, ref, B, cus, dev);
        case 4: return max_coresident_k<PMAX, 4>(f64, src, ref, B, cus, dev);
        case 5: return max_coresident_k<PMAX, 5>(f64, src, ref, B, cus, dev);
        case 6: return max_coresident_k<PMAX, 6>(f64, src, ref, B, cus, dev);
        default: return max_coresident_k<PMAX, 3>(f64, src, ref, B, cus, dev);
    }
}

// Persistent LSM kernels run one at a time per device (two of them could each
// hold part of the CUs while waiting for their remaining blocks): every
// context launches them on one process-wide stream per device, in order,
// handing over with events, so no host thread waits on another's kernel.
hipError_t coop_stream(int device, hipStream_t* out) {
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    std::lock_guard<std::mutex> lk(mu);
    auto it = streams.find(device);
    if (it != streams.end()) {
        *out = it->second;
        return hipSuccess;
    }
    hipStream_t st = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);  // default priority
    if (e != hipSuccess) return e;
    streams[device] = st;
    *out = st;
    return hipSuccess;
}
// Serialises enqueueing onto that stream (launch + its bracketing events).
std::mutex& persistent_mu() {
    static std::mutex m;
    return m;
}
// Per device: an event recorded after every persistent launch (under
// persistent_mu), and whether it has been recorded yet.  Heavy kernels of
// other kinds wait for it (pcf_after_persistent): started while a persistent
// grid waits for CUs, a long Asian launch kept taking the CUs its blocks
// freed, and the grid's resident blocks spun for up to its whole duration.
struct PersistEnd {
    hipEvent_t ev = nullptr;
    bool recorded = false;
};
std::map<int, PersistEnd>& persist_end() {
    static std::map<int, PersistEnd> m;
    return m;
}

// ---- host: buffers, graph cache --------------------------------------------
// Graph shapes depend on the size class of N (capacity 2^ceil(log2 N)), not
// on N itself: the kernels read N from the parameter block, so any N of a
// class replays the same captured graph with the same launch shape.
struct GraphKey {
    int64_t cap;
    int M, real, src, ref, grid_paths, grid_step, B, K;
    bool operator<(const GraphKey& o) const {
        return std::tie(cap, M, real, src, ref, grid_paths, grid_step, B, K) <
               std::tie(o.cap, o.M, o.real, o.src, o.ref, o.grid_paths, o.grid_step, o.B, o.K);
    }
};

struct LsmShape {
    int64_t cap;
    int B, ipt, grid_paths, grid_step;
};

// Built-in shape of the per-date (graph) form: 256 threads, at least 4 paths
// per thread, and more until the grid is at most 2 blocks per CU (4 above 2^21
// paths).  A date's kernel ends in a grid reduction of 8-10 sums whose last
// block reads every block's partials, so its cost grows with the block count;
// with too few blocks each thread's chain of dependent loads (state, then the
// discount of its exercise date) is exposed instead.  Measured per request at
// M = 50, FP32: 2^20 1.56 -> 1.25 ms (1024 -> 512 blocks), 2^22 4.66 -> 2.81
// ms (4096 -> 1024), 2^23 8.28 -> 4.39 ms (8192 -> 1024).  The paths kernel,
// one thread per antithetic pair up to the cap and a grid reduction of 10 sums
// at its end, is capped at 4 blocks per CU: 2^23 3.76 -> 3.45 ms (16384 ->
// 1024 blocks).
LsmShape lsm_shape(int64_t N, int mode, const pcf_opts& o, int cus = 256) {
    LsmShape sh;
    sh.cap = (int64_t)1 << pcf_size_class(N);
    if (sh.cap < N) sh.cap = N;  // beyond the largest class
    (void)mode;
    int B = 256;  // the tuned table holds the persistent form's shape
    int ipt = pcf_capped_ipt(sh.cap, B, 4, (int64_t)(sh.cap <= ((int64_t)1 << 21) ? 2 : 4) * (cus > 0 ? cus : 256));
    if (o.block_size) B = o.block_size;
    if (o.items_per_thread) ipt = o.items_per_thread;
    sh.B = B;
    sh.ipt = ipt;
    sh.grid_paths = pcf_blocks_for(sh.cap / 2, B, pcf_capped_ipt(sh.cap / 2, B, 1, (int64_t)4 * (cus > 0 ? cus : 256)));
    sh.grid_step = pcf_blocks_for(sh.cap, B, ipt);
    return sh;
}

// Default shape of the persistent form.  Per date every block reads every
// block's partials (grid^2 coherent loads) and each CU works through its own
// pairs with latency exposed at few waves, so the grid targets 64 blocks (the
// self-validating all-reduce) for up to 2^15 antithetic pairs in FP64 and 2^17
// in FP32, and one per CU beyond, and the pairs of a block are spread over as
// many threads as a block holds (1024) before a thread takes more than one
// pair (FP32 above 2^15 pairs: 2 pairs per thread from the start): the fewest
// pairs per thread (1..8) that fit, not only powers of two (the kernel's
// register budget is the next power of two, or exactly 3, 5 or 6 pairs; the
// pair slots past the count hold no path).
// (At 2^15 and 2^16 paths, 64 blocks of 256 and 512 threads took 0.121 and
// 0.138 ms in FP32, 0.149 and 0.171 ms in FP64, against 0.161, 0.167, 0.182
// and 0.190 ms with 128 blocks and arrival counters.)  In FP32 the 64 blocks
// at 2 pairs per thread also took 76500-178714 paths 10-18% faster than the
// 200-256 counted blocks (2^17 paths 0.158 against 0.182 ms, and 0.167 at
// 1024 threads x 1 pair; 163840 paths 0.173 against 0.192) and 250052 paths
// as fast; in FP64 256 blocks stay ahead at 2^17 paths (0.212 against 0.227).
void lsm_persistent_shape(int64_t half, bool f64, int cus, int* bs, int* pm) {
    const int64_t sv_pairs = f64 ? (1 << 15) : (1 << 17);
    const int64_t target = half <= sv_pairs ? kSvMaxGrid : std::max(cus, 64);
    const int64_t per_block = (half + target - 1) / target;
    int p = !f64 && half > (1 << 15) && half <= sv_pairs ? 2 : 1;
    while (p < 8 && (per_block + p - 1) / p > 1024) ++p;
    const int64_t b = ((per_block + p - 1) / p + 63) / 64 * 64;
    *pm = p;
    *bs = (int)std::min<int64_t>(1024, std::max<int64_t>(64, b));
}

struct LsmGraph {
    hipGraphExec_t exec = nullptr;
    void* store_ptr = nullptr;  // buffers the graph was captured with
    void* state_ptr = nullptr;
    void* coef_ptr = nullptr;
    void* part_ptr = nullptr;
    void* norm_ptr = nullptr;
};

std::map<pcf_ctx*, std::map<GraphKey, LsmGraph>>& graph_cache() {
    static std::map<pcf_ctx*, std::map<GraphKey, LsmGraph>> c;
    return c;
}
std::mutex& graph_mu() {
    static std::mutex m;
    return m;
}

// Builds the date sweep as an explicit graph (kernel nodes in a chain).  No
// stream capture is involved, so other threads may use the device freely
// while a graph is being built.
template <typename Real, int SRC, bool REF, int K = 3>
hipError_t build_lsm_graph(const LsmBufs& b, int M, int grid_paths, int grid_step, int B,
                           hipGraph_t* out) {
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    if (e != hipSuccess) return e;
    hipGraphNode_t prev = nullptr;
    // kernel arguments are copied into the node at creation
    LsmBufs bb = b;
    std::vector<int> dates;
    for (int m = M - 1; m >= 2; --m) dates.push_back(m);
    auto add = [&](const void* fn, int grid, void** args) -> hipError_t {
        hipKernelNodeParams p{};
        p.func = const_cast<void*>(fn);
        p.gridDim = dim3((unsigned)grid);
        p.blockDim = dim3((unsigned)B);
        p.sharedMemBytes = 0;
        p.kernelParams = args;
        p.extra = nullptr;
        hipGraphNode_t n = nullptr;
        hipError_t r = hipGraphAddKernelNode(&n, g, prev ? &prev : nullptr, prev ? 1 : 0, &p);
        if (r == hipSuccess) prev = n;
        return r;
    };
    void* a0[] = {&bb};
    if constexpr (SRC == kHost)
        e = add((const void*)k_lsm_paths_host<REF, K>, grid_paths, a0);
    else
        e = add((const void*)k_lsm_paths_philox<Real, REF, K>, grid_paths, a0);
    for (size_t i = 0; e == hipSuccess && i < dates.size(); ++i) {
        void* as[] = {&bb, &dates[i]};
        e = add((const void*)k_lsm_step<Real, REF, K>, grid_step, as);
    }
    if (e == hipSuccess) {
        void* af[] = {&bb};
        e = add((const void*)k_lsm_final<Real, REF, K>, grid_step, af);
    }
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    *out = g;
    return hipSuccess;
}

// The K-term graphs (production arithmetic: FP32 or FP64 on Philox, FP64 on
// host normals).
template <int K>
hipError_t build_lsm_graph_k(bool f64, int src, const LsmBufs& b, int M, int grid_paths, int grid_step, int B,
                             hipGraph_t* out) {
    if (!f64) return build_lsm_graph<float, kPhilox, false, K>(b, M, grid_paths, grid_step, B, out);
    return src == kHost ? build_lsm_graph<double, kHost, false, K>(b, M, grid_paths, grid_step, B, out)
                        : build_lsm_graph<double, kPhilox, false, K>(b, M, grid_paths, grid_step, B, out);
}

}  // namespace

PcfAmerBytes pcf_amer_bytes(int64_t N, int M, int precision, int rng, int ref_exact) {
    const bool f64 = ref_exact || rng == PCF_RNG_HOST || precision == PCF_FP64;
    const int64_t rs = f64 ? 8 : 4;
    pcf_opts o;
    pcf_opts_init(&o);
    const LsmShape sh = lsm_shape(N, f64 ? 1 : 0, o);
    const int64_t grid_max = std::max(sh.grid_paths, sh.grid_step);
    PcfAmerBytes b;
    b.store = (int64_t)M * sh.cap * rs;
    b.state = sh.cap * (4 + 2 * rs);
    // coefficients, discounts, parameters, sync words and slots + the persistent
    // form's partials: at most 4096 co-resident blocks x 448 B (6-term basis)
    b.rest = grid_max * 10 * 16 + 8 * (9 * (int64_t)(M + 1)) + 8192 + 8 * kSvWords + 4096 * 448;
    return b;
}

int64_t pcf_amer_footprint(int64_t N, int M, int precision, int rng, int ref_exact) {
    const PcfAmerBytes b = pcf_amer_bytes(N, M, precision, rng, ref_exact);
    return b.store + b.state + b.rest;
}

// Releases a context's captured LSM graphs (called from pcf_ctx_destroy).
void pcf_lsm_release(pcf_ctx* ctx) {
    std::lock_guard<std::mutex> lk(graph_mu());
    auto it = graph_cache().find(ctx);
    if (it == graph_cache().end()) return;
    for (auto& kv : it->second)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    graph_cache().erase(it);
}

namespace {
int mc_amer_run(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T, int64_t N, int32_t M,
                double payoff_fun, const pcf_opts& o, pcf_result* res);
}  // namespace

extern "C" int pcf_mc_amer(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T,
                           int64_t N, int32_t M, double payoff_fun, const pcf_opts* opts_in,
                           pcf_result* res) {
    if (!ctx || !res) return PCF_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    pcf_request q = pcf_make_request(PCF_KIND_AMER, S0, E, r, sigma, T, N, payoff_fun, opts_in);
    q.M = M;
    const int st = mc_amer_run(ctx, S0, E, r, sigma, T, N, M, payoff_fun, q.opts, res);
    if (st) return st;
    return pcf_apply_fp64_check(ctx, q, res);
}

namespace {
int mc_amer_run(pcf_ctx* ctx, double S0, double E, double r, double sigma, double T, int64_t N, int32_t M,
                double payoff_fun, const pcf_opts& o, pcf_result* res) {
    int st = pcf_check_common(ctx, S0, E, r, sigma, T, N, payoff_fun);
    if (st) return st;
    if ((st = pcf_check_opts(ctx, &o))) return st;
    if (N % 2 != 0)  // include/common.h:180
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "N needs to be divisible by 2 for finding paths");
    if (M < 1 || M > 100000) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "M must be in [1, 100000]");
    if (!(E > 0)) return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "E must be positive");
    PCF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    *res = pcf_result{};

    const int src = o.rng == PCF_RNG_HOST ? kHost : kPhilox;
    const bool ref = o.ref_exact != 0;
    const bool f64 = ref || src == kHost || o.precision == PCF_FP64;
    const size_t rs = f64 ? sizeof(double) : sizeof(float);
    // Regression basis: K = 3 ({1, S/E, (S/E)^2}; the reference's own basis in
    // the reference-exact mode) or the K-term basis in x = S/E - 1.
    const int K = o.lsm_basis == 0 ? 3 : o.lsm_basis;
    if (K < 2 || K > 6)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "lsm_basis must be 0 (3 terms) or 2..6, got %d", o.lsm_basis);
    if (K != 3 && ref)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG,
                        "the reference-exact mode regresses on the reference's 3-term basis (lsm_basis %d)", K);
    const int CS = K == 3 ? 4 : 8;  // per-date coefficient stride: flag + K coefficients

    // Parameter block (host copy -> device), discount table, buffers.
    const double dt = T / M;
    LsmParams P{};
    P.S0 = S0;
    P.E = E;
    P.pf = payoff_fun;
    P.invE = 1.0 / E;
    P.nrdt = -r * dt;
    P.nr = -r;
    P.dt = dt;
    P.h_mu = (r - 0.5 * sigma * sigma) * dt;
    P.sigma = sigma;
    P.mu = (r - 0.5 * sigma * sigma) * dt;
    P.vol = sigma * std::sqrt(dt);
    const double log2e = 1.4426950408889634;
    P.mu2f = (float)(P.mu * log2e);
    P.vol2f = (float)(P.vol * log2e);
    P.S0f = (float)S0;
    P.Ef = (float)E;
    P.pff = (float)payoff_fun;
    P.N = N;
    P.half = N / 2;
    P.M = M;
    P.seed = o.seed;

    const int mode = f64 ? 1 : 0;
    const LsmShape sh = lsm_shape(N, mode, o, ctx->cu_count);
    const size_t store_bytes = (size_t)M * (size_t)sh.cap * rs;
    const size_t state_bytes = (size_t)sh.cap * (sizeof(int32_t) + 2 * rs);
    // persistent-form sync words and self-validating slots (at a fixed place,
    // zero at rest: zeroed when allocated, and every persistent launch leaves
    // them zero) | coef (CS per date) | discount table | parameter block
    const size_t words_bytes = (sizeof(unsigned int) * kSyncWords + 255) / 256 * 256;
    const size_t sync_bytes = words_bytes + sizeof(unsigned long long) * kSvWords;
    const size_t coef_bytes =
        sync_bytes + sizeof(double) * (CS * (size_t)(M + 1) + (size_t)(M + 1)) + sizeof(LsmParams);
    PCF_HIP_CHECK(ctx, ctx->lsm_store.ensure(store_bytes));
    PCF_HIP_CHECK(ctx, ctx->lsm_state.ensure(state_bytes));
    PCF_HIP_CHECK(ctx, ctx->lsm_coef.ensure(coef_bytes, true));

    const int B = sh.B;
    const int grid_paths = sh.grid_paths;
    const int grid_step = sh.grid_step;
    const int grid_max = std::max(grid_paths, grid_step);
    // graph form: (2 + sums) compensated partials per block (10 for the quadratic)
    PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)grid_max * std::max(10, 3 * K + 1) * sizeof(pcf::CSum)));

    LsmBufs b{};
    b.store = ctx->lsm_store.ptr;
    char* sp = ctx->lsm_state.as<char>();
    b.when = (int32_t*)sp;
    b.cf = sp + (size_t)sh.cap * sizeof(int32_t);
    b.st = sp + (size_t)sh.cap * (sizeof(int32_t) + rs);
    unsigned int* const sync_words = ctx->lsm_coef.as<unsigned int>();
    double* cbase = (double*)(ctx->lsm_coef.as<char>() + sync_bytes);
    b.coef = cbase;
    b.disc = cbase + CS * (size_t)(M + 1);
    LsmParams* Pdev = (LsmParams*)(cbase + (CS + 1) * (size_t)(M + 1));
    b.P = Pdev;
    b.partials = ctx->partials.as<pcf::CSum>();
    b.counter = ctx->counters.as<unsigned int>();
    b.out = ctx->out.as<double>();

    // Persistent form when the request fits a co-resident grid.  Shape: the
    // tuned (block size, pairs per thread) for this size class, else explicit
    // opts, else lsm_persistent_shape.
    bool persistent = false;
    int pgrid = 0, pmax = 0, PB = 0;
    if (o.kernel_variant == 2 && M + 1 > kDiscLds)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG, "persistent LSM form holds at most %d dates (M=%d)", kDiscLds - 1,
                        M);
    if (o.kernel_variant != 1 && M + 1 <= kDiscLds) {
        int bs = 0, pm = 0;
        const bool forced = o.block_size > 0 || o.items_per_thread > 0;
        if (forced) {
            lsm_persistent_shape(P.half, f64, std::min(ctx->cu_count, kLandBlocks), &bs, &pm);
            if (o.block_size > 0) bs = o.block_size;
            if (o.items_per_thread > 0) pm = o.items_per_thread;
        } else {
            if (K == 3) pcf_tuned_shape(PCF_KIND_AMER, f64 ? 1 : 0, N, &bs, &pm);  // tuned for the quadratic
            if (bs <= 0 || pm <= 0) lsm_persistent_shape(P.half, f64, std::min(ctx->cu_count, kLandBlocks), &bs, &pm);
        }
        if (bs % 64 || bs < 64 || bs > 1024 || pm < 1 || pm > 8)
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG,
                            "persistent LSM shape %d x %d: need a block size in 64..1024 (multiple of 64) and "
                            "1..8 pairs per thread", bs, pm);
        // Register budget: pair slots per thread, a power of two, or exactly
        // 3, 5 or 6 pairs of the quadratic.
        int p2 = 1;
        while (p2 < pm) p2 <<= 1;
        if (K == 3 && (pm == 3 || pm == 5 || pm == 6)) p2 = pm;
        // pm pairs per thread: slots pm..p2-1 hold no path anywhere
        const int64_t g = (P.half + (int64_t)pm * bs - 1) / ((int64_t)pm * bs);
        int gmax;
        if (p2 == 1) gmax = max_coresident<1>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else if (p2 == 2) gmax = max_coresident<2>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else if (p2 == 3) gmax = max_coresident<3>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else if (p2 == 4) gmax = max_coresident<4>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else if (p2 == 5) gmax = max_coresident<5>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else if (p2 == 6) gmax = max_coresident<6>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        else gmax = max_coresident<8>(K, f64, src, ref, bs, ctx->cu_count, ctx->device);
        bool lds_ok;
        if (p2 == 1) lds_ok = lds_fits<1>(K, f64, src, ref, bs, ctx->device);
        else if (p2 == 2) lds_ok = lds_fits<2>(K, f64, src, ref, bs, ctx->device);
        else if (p2 == 3) lds_ok = lds_fits<3>(K, f64, src, ref, bs, ctx->device);
        else if (p2 == 4) lds_ok = lds_fits<4>(K, f64, src, ref, bs, ctx->device);
        else if (p2 == 5) lds_ok = lds_fits<5>(K, f64, src, ref, bs, ctx->device);
        else if (p2 == 6) lds_ok = lds_fits<6>(K, f64, src, ref, bs, ctx->device);
        else lds_ok = lds_fits<8>(K, f64, src, ref, bs, ctx->device);
        if (!lds_ok) gmax = 0;  // the row staging would not fit: graph form
        // The quadratic form gathers at most kLandBlocks blocks' sums (through
        // LDS, or from the self-validating slots).
        if (K == 3 && gmax > kLandBlocks) gmax = kLandBlocks;
        if (g >= 1 && g <= gmax) {
            persistent = true;
            pgrid = (int)g;
            pmax = p2;
            PB = bs;
        }
        if (o.kernel_variant == 2 && !persistent)
            return pcf_fail(ctx, PCF_ERR_INVALID_ARG,
                            "persistent LSM form does not fit N=%lld (%lld blocks of %d, at most %d co-resident)",
                            (long long)N, (long long)g, bs, gmax);
    }
    std::vector<double> host_blk((CS + 1) * (size_t)(M + 1), 0.0);
    for (int k = 0; k <= M; ++k) host_blk[CS * (size_t)(M + 1) + k] = std::exp(-r * dt * k);

    // Nothing left on the context's stream from earlier calls (checked before
    // this call enqueues its own begin event).
    const bool ctx_idle = hipStreamQuery(ctx->stream) == hipSuccess;
    if ((st = pcf_timing_begin(ctx))) return st;
    if (src == kHost && (st = pcf_upload_normals(ctx, &o, (N / 2) * (int64_t)M))) return st;
    b.hw = src == kHost ? ctx->normals.as<double>() : nullptr;
    // coef | disc | params are contiguous on the device: one copy, from the
    // context's pinned staging buffer when it fits (the call synchronises
    // before returning, so the staging is free again for the next call).  A
    // persistent launch with M + 1 <= kDiscArgs takes them as kernel
    // arguments instead and needs no copy.
    const size_t blk_bytes = host_blk.size() * sizeof(double);
    if (persistent && M + 1 <= kDiscArgs) {
        // nothing to copy
    } else if (blk_bytes + sizeof P <= kMaxBatch * kJobBytes) {
        std::memcpy(ctx->h_jobs, host_blk.data(), blk_bytes);
        std::memcpy((char*)ctx->h_jobs + blk_bytes, &P, sizeof P);
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(cbase, ctx->h_jobs, blk_bytes + sizeof P, hipMemcpyHostToDevice,
                                          ctx->stream));
    } else {
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(cbase, host_blk.data(), blk_bytes, hipMemcpyHostToDevice, ctx->stream));
        PCF_HIP_CHECK(ctx, hipMemcpyAsync(Pdev, &P, sizeof P, hipMemcpyHostToDevice, ctx->stream));
    }


    // Results.  The persistent form writes out[0..7] (kPersistOut) straight
    // into the context's pinned buffer and clears its own sync words, so it
    // needs no read-back copy.  The graph form's per-date flags are read back
    // beside out[0..3] into the pinned buffer, so the one synchronisation in
    // pcf_timing_end covers every copy.
    const size_t ncoef = CS * (size_t)(M + 1);
    const bool coef_pinned = 8 + ncoef <= (size_t)kOutDoubles;
    auto stage_readback = [&]() -> int {
        if (coef_pinned)
            PCF_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_out + 8, cbase, ncoef * sizeof(double), hipMemcpyDeviceToHost,
                                              ctx->stream));
        return PCF_OK;
    };
    if (persistent) {
        // 2 date parities x 8C sums, then the final sums' 4 compensated pairs
        const size_t chunks = K == 6 ? 3 : (K >= 4 ? 2 : 1);
        PCF_HIP_CHECK(ctx, ctx->partials.ensure((size_t)pgrid *
                                                (2 * 8 * chunks * sizeof(double) + 4 * sizeof(pcf::CSum))));
        b.partials = ctx->partials.as<pcf::CSum>();
        b.out = ctx->h_out_dev;
        const int fin_w = kArrivalCounters * kCounterStride;
        LsmPersistArgs pa;
        pa.b = b;
        pa.sy = LsmSync{sync_words, sync_words + fin_w, sync_words + fin_w + 1,
                        (unsigned long long*)(ctx->lsm_coef.as<char>() + words_bytes)};
        pa.P = P;
        for (int k = 0; k <= M && k < kDiscArgs; ++k) pa.disc[k] = host_blk[CS * (size_t)(M + 1) + k];
        hipStream_t cs = nullptr;
        PCF_HIP_CHECK(ctx, coop_stream(ctx->device, &cs));
        // Inputs enqueued on the context's stream (host normals, the
        // coefficient block) are handed over with an event; with none, the
        // launch needs no wait.  The host then waits for the kernel's end
        // event itself: every later use of this context's buffers is enqueued
        // after that wait.
        const bool handover = !ctx_idle || src == kHost || M + 1 > kDiscArgs;
        if (handover) PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_ready, ctx->stream));
        {
            std::lock_guard<std::mutex> pl(persistent_mu());
            if (handover) PCF_HIP_CHECK(ctx, hipStreamWaitEvent(cs, ctx->ev_ready, 0));
            PCF_HIP_CHECK(ctx, pcf_wait_other_contexts(ctx, cs));
            hipEvent_t e0 = ctx->ev_kbeg, e1 = ctx->ev_kend;
            hipError_t le;
            if (pmax == 1) le = launch_persistent<1>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else if (pmax == 2) le = launch_persistent<2>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else if (pmax == 3) le = launch_persistent<3>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else if (pmax == 4) le = launch_persistent<4>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else if (pmax == 5) le = launch_persistent<5>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else if (pmax == 6) le = launch_persistent<6>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            else le = launch_persistent<8>(K, f64, src, ref, pa, pgrid, PB, cs, e0, e1);
            if (le != hipSuccess) return pcf_hip_fail(ctx, le, "launching the persistent LSM kernel");
            PersistEnd& pe = persist_end()[ctx->device];
            if (!pe.ev) PCF_HIP_CHECK(ctx, hipEventCreateWithFlags(&pe.ev, hipEventDisableTiming));
            PCF_HIP_CHECK(ctx, hipEventRecord(pe.ev, cs));
            pe.recorded = true;
        }
        PCF_HIP_CHECK(ctx, pcf_wait_event(ctx->ev_kend));
        PCF_HIP_CHECK(ctx, pcf_wait_event(ctx->ev_begin));
        float kms = 0.f, tms = 0.f;
        PCF_HIP_CHECK(ctx, hipEventElapsedTime(&kms, ctx->ev_kbeg, ctx->ev_kend));
        PCF_HIP_CHECK(ctx, hipEventElapsedTime(&tms, ctx->ev_begin, ctx->ev_kend));
        res->kernel_ms = kms;
        res->total_ms = tms;
        res->grid_blocks = pgrid;
        res->block_size = PB;
    }

    if (!persistent && B > kLsmGraphThreads)
        return pcf_fail(ctx, PCF_ERR_INVALID_ARG,
                        "the per-date LSM form takes at most %d threads per block (block_size %d)", kLsmGraphThreads,
                        B);

    // Graph per shape; rebuilt if a buffer moved.
    GraphKey key{sh.cap, M, f64 ? 1 : 0, src, ref ? 1 : 0, grid_paths, grid_step, B, K};
    hipGraphExec_t exec = nullptr;
    if (!persistent) {
        std::lock_guard<std::mutex> gl(graph_mu());
        auto& g = graph_cache()[ctx][key];
        const bool fresh = g.exec && g.store_ptr == b.store && g.state_ptr == (void*)b.when &&
                           g.coef_ptr == (void*)cbase && g.part_ptr == (void*)b.partials &&
                           g.norm_ptr == (void*)b.hw;
        if (!fresh) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = LsmGraph{};
            hipGraph_t graph = nullptr;
            hipError_t ce;
            if (K == 2) ce = build_lsm_graph_k<2>(f64, src, b, M, grid_paths, grid_step, B, &graph);
            else if (K == 4) ce = build_lsm_graph_k<4>(f64, src, b, M, grid_paths, grid_step, B, &graph);
            else if (K == 5) ce = build_lsm_graph_k<5>(f64, src, b, M, grid_paths, grid_step, B, &graph);
            else if (K == 6) ce = build_lsm_graph_k<6>(f64, src, b, M, grid_paths, grid_step, B, &graph);
            else if (f64) {
                if (src == kHost)
                    ce = ref ? build_lsm_graph<double, kHost, true>(b, M, grid_paths, grid_step, B, &graph)
                             : build_lsm_graph<double, kHost, false>(b, M, grid_paths, grid_step, B, &graph);
                else
                    ce = ref ? build_lsm_graph<double, kPhilox, true>(b, M, grid_paths, grid_step, B, &graph)
                             : build_lsm_graph<double, kPhilox, false>(b, M, grid_paths, grid_step, B, &graph);
            } else {
                ce = build_lsm_graph<float, kPhilox, false>(b, M, grid_paths, grid_step, B, &graph);
            }
            if (ce != hipSuccess) return pcf_hip_fail(ctx, ce, "building the LSM graph");
            ce = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ce != hipSuccess) {
                g.exec = nullptr;
                return pcf_hip_fail(ctx, ce, "hipGraphInstantiate");
            }
            g.store_ptr = b.store;
            g.state_ptr = (void*)b.when;
            g.coef_ptr = (void*)cbase;
            g.part_ptr = (void*)b.partials;
            g.norm_ptr = (void*)b.hw;
        }
        exec = g.exec;
    }
    if (!persistent) {
        std::unique_lock<std::mutex> order;  // held until the end event is recorded
        PCF_HIP_CHECK(ctx, pcf_after_persistent(ctx, HUGE_VAL, &order));
        PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kbeg, ctx->stream));
        PCF_HIP_CHECK(ctx, hipGraphLaunch(exec, ctx->stream));
        PCF_HIP_CHECK(ctx, hipEventRecord(ctx->ev_kend, ctx->stream));
        if (order.owns_lock()) order.unlock();
        if ((st = stage_readback())) return st;
        if ((st = pcf_timing_end(ctx, 4, res))) return st;
        res->grid_blocks = grid_step;
        res->block_size = B;
    }
    const double h0 = ctx->h_out[0], h1 = ctx->h_out[1], h2 = ctx->h_out[2], h3 = ctx->h_out[3];
    int small = 0, fallback = 0, singular = 0;
    if (persistent) {
        small = (int)ctx->h_out[4];
        fallback = (int)ctx->h_out[5];
        singular = (int)ctx->h_out[6];
        if (ctx->h_out[7] != 0.0) return pcf_fail(ctx, PCF_ERR_TIMEOUT, "LSM date barrier wait exceeded its bound");
    }

    // Per-date flags of the graph form beyond the pinned staging area, and the
    // optional exercise dates: one more copy each.
    std::vector<double> coefs;
    const double* cf_host = ctx->h_out + 8;
    if ((!persistent && !coef_pinned) || o.when_out) {
        if (!persistent && !coef_pinned) {
            coefs.resize(ncoef);
            PCF_HIP_CHECK(ctx, hipMemcpyAsync(coefs.data(), cbase, ncoef * sizeof(double), hipMemcpyDeviceToHost,
                                              ctx->stream));
            cf_host = coefs.data();
        }
        if (o.when_out)
            PCF_HIP_CHECK(ctx, hipMemcpyAsync(o.when_out, b.when, (size_t)N * sizeof(int32_t),
                                              hipMemcpyDeviceToHost, ctx->stream));
        PCF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (!persistent)
        for (int m = 1; m <= M - 1; ++m) {
            const int f = (int)cf_host[CS * m];
            small += f == kFlagSmall;
            fallback += f == kFlagReduced;
            singular += f == kFlagSingular;
        }

    const double n = (double)N;
    const double pay0 = std::max(payoff_fun * (S0 - E), 0.0);
    // return std::max(payoff(S0,E,payoff_fun),result/(double)N)  (src/mc_amer.cpp:113)
    res->price = std::max(pay0, h2 / n);
    res->mean = h2 / n;
    res->variance = N > 1 ? (h3 - h2 * (h2 / n)) / (n - 1.0) : 0.0;
    res->std_err = std::sqrt(res->variance / n);
    const double discT = std::exp(-r * T);
    res->control_price = discT * h0 / n;
    const double evar = N > 1 ? (h1 - h0 * (h0 / n)) / (n - 1.0) : 0.0;
    res->control_se = discT * std::sqrt(evar / n);
    res->n_paths = N;
    res->n_steps = M;
    res->lsm_small_itm = small;
    res->lsm_fallback = fallback;
    res->batch_size = 1;
    // Algorithmic HBM bytes: store written once and read back once, per-path
    // state {cash flow, date} read + written per backward date.
    res->bytes = n * ((double)M * rs + (double)(M - 1) * rs + 2.0 * (M - 1) * (4.0 + rs));
    res->flops = 0.0;
    if (singular)
        return pcf_fail(ctx, PCF_ERR_SINGULAR, "Detereminant is not > 0 (%d dates)", singular);
    return PCF_OK;
}
}  // namespace

hipError_t pcf_after_persistent(pcf_ctx* ctx, double work, std::unique_lock<std::mutex>* hold) {
    if (work < kPersistentGateWork) return hipSuccess;
    std::unique_lock<std::mutex> pl(persistent_mu());
    auto it = persist_end().find(ctx->device);
    hipError_t e = hipSuccess;
    if (it != persist_end().end() && it->second.recorded) e = hipStreamWaitEvent(ctx->stream, it->second.ev, 0);
    if (hold) *hold = std::move(pl);
    return e;
}
