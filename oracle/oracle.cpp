// oracle.cpp -- CPU restatement of the reference pricing arithmetic.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (parcompfin_amd/, the CLIs,
// libpcf.so) links, loads or calls this file.  It is used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker, never as the thing measured or shipped.
//
// Two families of functions live here:
//
//  (1) "ref" restatements: the serial reference programs' arithmetic in the
//      reference's own operation order, fed either by the reference's own
//      normal stream (libstdc++ std::mt19937 + std::normal_distribution, the
//      same library the reference links) or by an explicit array of normals in
//      the reference's draw order.  These are pinned bit-for-bit against
//      drivers built from the reference sources (oracle/_ref, see Makefile) and
//      against committed 17-digit fixtures (tests/golden/).
//
//  (2) "philox" restatements: the production estimators (counter-based
//      Philox4x32-10 normals, Box-Muller, FP64) that the GPU FP64 mode must
//      reproduce to ~1e-12, plus closed forms (Black-Scholes, discrete
//      geometric Asian) and the reference's binomial lattices.
//
// Parity status per function is stated in the comment above each one.
// Every function is compiled with -O2 -fno-builtin and no -ffast-math: on
// x86-64 without -mfma GCC emits no fused multiply-adds, and -fno-builtin keeps
// every pow/exp/log a call into glibc as in the reference's -O0 build (GCC
// otherwise folds pow(x,2) into x*x, which differs from glibc's pow(x,2.0) in
// the last bit for ~0.1% of inputs).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------------------
// payoff -- include/common.h:56-60
//   max(pf*(S-E), 0) with pf = +1 (call) / -1 (put).  The reference takes an
//   int and its callers pass a double that is converted; pf is exactly +-1.
// Pinned: closed form (trivial), and transitively by every fixture.
static inline double payoff(double St, double E, int pf) {
    return std::max(pf * (St - E), 0.0);
}

// ---------------------------------------------------------------------------
// Reference normal stream: libstdc++ mt19937 seeded with `seed` (the value the
// reference takes from time(), src/mc_eur.cpp:16-20, src/mc_asia.cpp:20-24,
// include/common.h:188-192), std::normal_distribution<>{0, sd}.
// Fills out[0..count).
// Pinned: fixtures made by the reference binaries with time() fixed.
void orc_normals_mt(uint64_t seed, double sd, int64_t count, double* out) {
    std::mt19937 gen;
    gen.seed(seed);
    std::normal_distribution<> norm{0, sd};
    for (int64_t k = 0; k < count; ++k) out[k] = norm(gen);
}

// ---------------------------------------------------------------------------
// mc_eur -- src/mc_eur.cpp:5-27.  z[n] are the reference draws N(0, sqrt(T)).
// Pinned: tests/golden/ref_kernels.csv (17 digits, reference driver) and
// Black-Scholes within a few SE.
double orc_mc_eur_z(double S0, double E, double r, double sigma, double T,
                    int64_t N, double payoff_fun, const double* z) {
    const int pf = (int)payoff_fun;
    double result = 0;
    for (int64_t n = 0; n < N; ++n)
        result += payoff(S0 * exp((r - pow(sigma, 2) / 2) * T + sigma * z[n]), E, pf);
    return (exp(-r * T) * result) / (double)N;
}

double orc_mc_eur_seed(double S0, double E, double r, double sigma, double T,
                       int64_t N, double payoff_fun, uint64_t seed) {
    std::vector<double> z(N);
    orc_normals_mt(seed, sqrt(T), N, z.data());
    return orc_mc_eur_z(S0, E, r, sigma, T, N, payoff_fun, z.data());
}

// ---------------------------------------------------------------------------
// mc_asia -- src/mc_asia.cpp:5-40.  z[n*M+m] are the reference draws
// N(0, sqrt(dt)) in the reference's path-major order.
// Pinned: ref driver fixtures (17 digits); geometric closed form on the same
// stream is only a statistical control.
double orc_mc_asia_z(double S0, double E, double r, double sigma, double T,
                     int64_t N, int M, double payoff_fun, const double* z) {
    const int pf = (int)payoff_fun;
    double dt = (double)T / (double)M;
    double result = 0;
    for (int64_t n = 0; n < N; ++n) {
        double St = S0;
        double I = 0;
        const double* zp = z + n * (int64_t)M;
        for (int m = 0; m < M; ++m) {
            double dBi = zp[m];
            I += St * (1 + r * dt / 2 + sigma * dBi / 2);
            St *= exp((r - pow(sigma, 2) / 2) * dt + sigma * dBi);
        }
        result += payoff(I / (double)M, E, pf);
    }
    return exp(-r * T) * result / (double)N;
}

double orc_mc_asia_seed(double S0, double E, double r, double sigma, double T,
                        int64_t N, int M, double payoff_fun, uint64_t seed) {
    std::vector<double> z((size_t)N * M);
    orc_normals_mt(seed, sqrt((double)T / (double)M), N * (int64_t)M, z.data());
    return orc_mc_asia_z(S0, E, r, sigma, T, N, M, payoff_fun, z.data());
}

// ---------------------------------------------------------------------------
// comb -- include/common.h:63-72.  log C(N,i) as a sequential chain of FP64
// adds of glibc log(n).  The rounding of this chain IS the reference result,
// so it is restated verbatim in order.
// Pinned: 12 published rows of results/results_binom_embar.csv (10 digits)
// and the reference driver (17 digits).
static double comb(int N, int i) {
    if (i == 0 || i == N) return 0;
    if (i == 1 || i == (N - 1)) return log((double)N);
    double result = 0;
    for (int n = N; n > i; --n) result += log((double)n);
    for (int j = 2; j <= (N - i); ++j) result -= log((double)j);
    return result;
}

double orc_comb(int N, int i) { return comb(N, i); }

// Tree parameters of src/binom_embar.cpp:16-24 (the beta tree, not CRR).
void orc_beta_tree(double r, double sigma, double T, int N, double* out4) {
    double dt = (double)T / (double)N;
    double beta = 0.5 * (exp(-r * dt) + exp((r + pow(sigma, 2)) * dt));
    double u = beta + sqrt(pow(beta, 2) - 1);
    double d = beta - sqrt(pow(beta, 2) - 1);
    double R = exp(r * dt);
    double p = (R - d) / (u - d);
    out4[0] = u; out4[1] = d; out4[2] = p; out4[3] = R;
}

// binom (embarrassingly parallel closed sum) -- src/binom_embar.cpp:5-50.
// omp_threads > 1 runs the same arithmetic under the reference's OpenMP
// schedule (src/binom_embar_omp.cpp:37): identical terms, reduction order may
// differ in the last bits.  Used as a CPU timing baseline only.
double orc_binom_embar(double S0, double E, double r, double sigma, double T,
                       int N, double payoff_fun, int omp_threads) {
    const int pf = (int)payoff_fun;
    double result = 0;
    double dt = (double)T / (double)N;
    double beta = 0.5 * (exp(-r * dt) + exp((r + pow(sigma, 2)) * dt));
    double u = beta + sqrt(pow(beta, 2) - 1);
    double d = beta - sqrt(pow(beta, 2) - 1);
    double R = exp(r * dt);
    double p = (R - d) / (u - d);
    double q = 1 - p;
    int until = (N % 2 != 0) ? (N + 1) / 2 : N / 2;
    (void)omp_threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1000) reduction(+ : result) num_threads(omp_threads > 0 ? omp_threads : 1)
#endif
    for (int i = 0; i < until; ++i) {
        double comb_val = comb(N, i);
        double binom1 = comb_val + i * log(p) + (N - i) * log(q);
        double binom2 = comb_val + (N - i) * log(p) + i * log(q);
        result += exp(binom1) * payoff(S0 * pow(u, i) * pow(d, N - i), E, pf);
        result += exp(binom2) * payoff(S0 * pow(u, N - i) * pow(d, i), E, pf);
        if (i == 0 && N % 2 == 0) {
            double binom_mid = comb(N, N / 2) + N / 2 * log(p) + N / 2 * log(q);
            result += exp(binom_mid) * payoff(S0 * pow(u, N / 2) * pow(d, N / 2), E, pf);
        }
    }
    return exp(-r * T) * result;
}

// Same result as orc_binom_embar(..., 1) bit for bit, in O(N^2/2) adds
// instead of O(N^2) logs, for large N:
//  * comb(N,i)'s first loop, sum_{n=N..i+1} log n descending from 0, is the
//    (N-i)-th prefix of ONE running sum, so every prefix is computed once;
//  * the second loop (subtract log 2..log(N-i)) starts from that prefix and is
//    replayed per term, in order, with the same glibc log values;
//  * terms are added to `result` serially in the reference's order.
// Pinned: equality with orc_binom_embar at every N the tests run.
double orc_binom_embar_prefix(double S0, double E, double r, double sigma, double T,
                              int N, double payoff_fun, int threads) {
    const int pf = (int)payoff_fun;
    double dt = (double)T / (double)N;
    double beta = 0.5 * (exp(-r * dt) + exp((r + pow(sigma, 2)) * dt));
    double u = beta + sqrt(pow(beta, 2) - 1);
    double d = beta - sqrt(pow(beta, 2) - 1);
    double R = exp(r * dt);
    double p = (R - d) / (u - d);
    double q = 1 - p;
    int until = (N % 2 != 0) ? (N + 1) / 2 : N / 2;
    std::vector<double> tab((size_t)N + 1), pre((size_t)N + 1);
    for (int n = 1; n <= N; ++n) tab[n] = log((double)n);
    pre[0] = 0;  // pre[k] = (((0 + log N) + log(N-1)) + ...) over k terms
    for (int k = 1; k <= N; ++k) pre[k] = pre[k - 1] + tab[N - k + 1];
    auto cmb = [&](int i) -> double {
        if (i == 0 || i == N) return 0;
        if (i == 1 || i == (N - 1)) return tab[N];
        double acc = pre[N - i];
        for (int j = 2; j <= (N - i); ++j) acc -= tab[j];
        return acc;
    };
    std::vector<double> t1(until), t2(until);
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
    for (int i = 0; i < until; ++i) {
        double comb_val = cmb(i);
        double binom1 = comb_val + i * log(p) + (N - i) * log(q);
        double binom2 = comb_val + (N - i) * log(p) + i * log(q);
        t1[i] = exp(binom1) * payoff(S0 * pow(u, i) * pow(d, N - i), E, pf);
        t2[i] = exp(binom2) * payoff(S0 * pow(u, N - i) * pow(d, i), E, pf);
    }
    double mid = 0;
    if (N % 2 == 0) {
        double binom_mid = cmb(N / 2) + N / 2 * log(p) + N / 2 * log(q);
        mid = exp(binom_mid) * payoff(S0 * pow(u, N / 2) * pow(d, N / 2), E, pf);
    }
    double result = 0;
    for (int i = 0; i < until; ++i) {
        result += t1[i];
        result += t2[i];
        if (i == 0 && N % 2 == 0) result += mid;
    }
    return exp(-r * T) * result;
}

// Production binomial sum: the same beta tree and terms as orc_binom_embar,
// with each weight C(N,i) p^i q^(N-i) formed in log space in double-double
// arithmetic, so the weights carry no rounding of the large intermediate
// logs (log N! ~ 1.3e7 at N = 1e6, whose FP64 ulp is 1.9e-9).
//   lf[k] = log k! : double-double table (orc_logfact_dd);
//   lp, lq         : log p, log q as double-double (dd_log, no libm);
//   t_i = lf[N] - lf[i] - lf[N-i] + i lp + (N-i) lq   (double-double),
//   w_i = e + e t_i.lo with e = exp(t_i.hi);
//   term_i = w_i payoff(S0 pow(u,i) pow(d,N-i))       (terms with w_i = 0 skipped)
// and exp(-rT) times the sum.  The double-double steps are error-free
// transformations (TwoSum; Dekker's split for TwoProduct, exact without FMA),
// so a device restatement with FMA forms the same t_i bit for bit.
// Pinned: the exact sum (mpmath, 40 digits, the same double u, d, p, q) at
// N = 10^3, 10^5, 10^6, call and put, to 1e-13 (tests/golden/binom_exact.csv).
// The reference's own comb() chain is 1.2e-10 (N = 1e5) and 1.5e-8 (N = 1e6)
// away from that exact sum.
struct DD {
    double h, l;
};
static DD dd_two_sum(double a, double b) {
    const double s = a + b, bp = s - a;
    return {s, (a - (s - bp)) + (b - bp)};
}
static DD dd_quick(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
static void dd_split(double a, double& hi, double& lo) {
    const double c = 134217729.0 * a;  // 2^27 + 1
    hi = c - (c - a);
    lo = a - hi;
}
static DD dd_two_prod(double a, double b) {
    const double p = a * b;
    double ah, al, bh, bl;
    dd_split(a, ah, al);
    dd_split(b, bh, bl);
    return {p, ((ah * bh - p) + ah * bl + al * bh) + al * bl};
}
static DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.h, b.h);
    s.l += a.l + b.l;
    return dd_quick(s.h, s.l);
}
static DD dd_mul_d(DD a, double b) {
    DD p = dd_two_prod(a.h, b);
    p.l += a.l * b;
    return dd_quick(p.h, p.l);
}

static DD dd_mul(DD a, DD b) {
    DD p = dd_two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return dd_quick(p.h, p.l);
}
// a / b for a double-double a and a double b: q = a.h / b, then the exact
// remainder a - q b (TwoProduct) divided once more.
static DD dd_div_d(DD a, double b) {
    const double q = a.h / b;
    const DD qb = dd_two_prod(q, b);
    const double rem = ((a.h - qb.h) - qb.l) + a.l;
    return dd_quick(q, rem / b);
}

static DD dd_div(DD a, DD b) {
    const double q1 = a.h / b.h;
    const DD r = dd_add(a, dd_mul_d(DD{-b.h, -b.l}, q1));
    const double q2 = r.h / b.h;
    return dd_quick(q1, q2);
}
// 2 atanh(t) = sum 2 t^(2j+1)/(2j+1), summed until a term is below 1e-34 of the sum.
static DD dd_2atanh(DD t) {
    const DD u = dd_mul(t, t);
    DD sum = t, pw = t;
    for (int j = 1; j < 200; ++j) {
        pw = dd_mul(pw, u);
        const DD term = dd_div_d(pw, (double)(2 * j + 1));
        sum = dd_add(sum, term);
        if (fabs(term.h) < 1e-34 * fabs(sum.h)) break;
    }
    return DD{2.0 * sum.h, 2.0 * sum.l};
}
static const DD kLn2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
// log x (x > 0 finite) as double-double without libm: x = m 2^e with m in
// [0.75, 1.5), log x = e log 2 + 2 atanh((m - 1)/(m + 1)).
static DD dd_log(double x) {
    int e = 0;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.75) {
        m *= 2.0;
        e -= 1;
    }
    const DD t = dd_div(DD{m - 1.0, 0.0}, dd_two_sum(m, 1.0));
    return dd_add(dd_mul_d(kLn2, (double)e), dd_2atanh(t));
}

// lf[k] = log k!, k = 0..N, as double-double (hi, lo interleaved: out[2k],
// out[2k+1]), from IEEE double operations only (no libm), so the table is the
// same on every host:
//   log 2 = the double-double constant (0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56);
//   log k = log(k-1) + 2 atanh(1/(2k-1))  (k >= 3), atanh(t) = sum t^(2j+1)/(2j+1)
//           summed until a term falls below 1e-34 of the sum;
//   lf[k] = lf[k-1] + log k.
// Each step rounds at ~1e-32 relative, so after 10^6 steps lf[N] ~ 1.3e7 is
// within ~1e-20 of log N!.
void orc_logfact_dd(int N, double* out) {
    out[0] = out[1] = 0;
    if (N >= 1) out[2] = out[3] = 0;
    DD Lk = kLn2, lf = {0.0, 0.0};
    for (int k = 2; k <= N; ++k) {
        if (k > 2) Lk = dd_add(Lk, dd_2atanh(dd_div_d(DD{1.0, 0.0}, (double)(2 * k - 1))));
        lf = dd_add(lf, Lk);
        out[2 * k] = lf.h;
        out[2 * k + 1] = lf.l;
    }
}

double orc_binom_dd(double S0, double E, double r, double sigma, double T, int N, double payoff_fun) {
    const int pf = (int)payoff_fun;
    double tree[4];
    orc_beta_tree(r, sigma, T, N, tree);
    const double u = tree[0], d = tree[1], p = tree[2];
    const double q = 1 - p;
    std::vector<double> lf(2 * ((size_t)N + 1));
    orc_logfact_dd(N, lf.data());
    const DD lp = dd_log(p), lq = dd_log(q);
    const DD lfN = {lf[2 * (size_t)N], lf[2 * (size_t)N + 1]};
    long double acc = 0;
    for (int i = 0; i <= N; ++i) {
        DD t = dd_add(lfN, DD{-lf[2 * (size_t)i], -lf[2 * (size_t)i + 1]});
        t = dd_add(t, DD{-lf[2 * (size_t)(N - i)], -lf[2 * (size_t)(N - i) + 1]});
        t = dd_add(t, dd_mul_d(lp, (double)i));
        t = dd_add(t, dd_mul_d(lq, (double)(N - i)));
        const double e = exp(t.h);
        const double w = e + e * t.l;
        if (w == 0.0) continue;
        acc += (long double)(w * payoff(S0 * pow(u, i) * pow(d, N - i), E, pf));
    }
    return exp(-r * T) * (double)acc;
}

// Vanilla lattices -- src/binom_vanilla_eur.cpp:5-39 and
// src/binom_vanilla_amer.cpp:5-40 (backward induction on the same beta tree).
// Pinned: reference driver (17 digits); the European one also against the
// closed sum above (the published rows differ by <= 1.1e-10 at N <= 1e4).
double orc_binom_vanilla(double S0, double E, double r, double sigma, double T,
                         int N, double payoff_fun, int american) {
    const int pf = (int)payoff_fun;
    double dt = (double)T / (double)N;
    double beta = 0.5 * (exp(-r * dt) + exp((r + pow(sigma, 2)) * dt));
    double u = beta + sqrt(pow(beta, 2) - 1);
    double d = beta - sqrt(pow(beta, 2) - 1);
    double R = exp(r * dt);
    double p = (R - d) / (u - d);
    double q = 1 - p;
    std::vector<double> v((size_t)N + 1);
    if (american) {
        for (int i = 0; i < N + 1; ++i) v[i] = payoff(S0 * pow(u, i) * pow(d, N - i), E, pf);
        for (int n = N - 1; n >= 0; --n)
            for (int i = 0; i < n + 1; ++i) {
                double sij = payoff(S0 * pow(u, i) * pow(d, n - i), E, pf);
                double jatk = (p * v[i + 1] + q * v[i]) / R;
                v[i] = std::max(jatk, sij);
            }
    } else {
        for (int i = 0; i < N + 1; ++i)
            v[i] = std::max(((S0 * pow(u, i) * pow(d, N - i)) - E) * payoff_fun, (double)0.0);
        for (int n = N - 1; n >= 0; --n)
            for (int i = 0; i < n + 1; ++i) v[i] = (p * v[i + 1] + q * v[i]) / R;
    }
    return v[0];
}

// ---------------------------------------------------------------------------
// 3x3 inverse by cofactors -- include/common.h:98-124, and mat_vec_mul
// include/common.h:128-141.  Returns 0 when det <= 0 (the reference throws
// "Detereminant is not > 0").
// Pinned: against an LU solve on random SPD matrices and the det<=0 case
// (tests/test_oracle.py).
int orc_inverse3_solve(const double* A /*3x3 row-major*/, const double* y, double* coef) {
    double det = 0;
    for (int i = 0; i < 3; ++i)
        det += (A[0 * 3 + i] * (A[1 * 3 + (i + 1) % 3] * A[2 * 3 + (i + 2) % 3] -
                                A[1 * 3 + (i + 2) % 3] * A[2 * 3 + (i + 1) % 3]));
    if (det <= 0) return 0;
    double inv[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            inv[j * 3 + i] = ((A[((j + 1) % 3) * 3 + (i + 1) % 3] * A[((j + 2) % 3) * 3 + (i + 2) % 3]) -
                              (A[((j + 1) % 3) * 3 + (i + 2) % 3] * A[((j + 2) % 3) * 3 + (i + 1) % 3])) / det;
    for (int i = 0; i < 3; ++i) {
        double m = 0;
        for (int j = 0; j < 3; ++j) m += inv[i * 3 + j] * y[j];
        coef[i] = m;
    }
    return 1;
}

// ---------------------------------------------------------------------------
// mc_amer (Longstaff-Schwartz) -- src/mc_amer.cpp:5-114 with the path store of
// pathsfinder, include/common.h:168-208.
//   w[n*M + (m-1)], n < N/2, m = 1..M: the reference draws N(0, sqrt(dt)) in
//   pathsfinder's order (antithetic: path n uses +w, path n+N/2 uses -w).
//   quirk = 1 reproduces the reference exactly, including the exercise test
//   payoff(S-E, E) at src/mc_amer.cpp:100 (wrong for puts, harmless for calls).
//   quirk = 0 is the corrected estimator (payoff(S, E)), still with the
//   reference's basis {1, S-E, (S-E)^2} and cofactor solve.
//   when_out (optional) receives exercise_when per path.
// Returns NAN when the reference would throw (singular Gram, odd N).
// Pinned: reference driver (17 digits, call and put; the put pins the quirk);
// call == European call (no early exercise without dividends); a brute-force
// N=8 case is checked in tests.
double orc_mc_amer_z(double S0, double E, double r, double sigma, double T,
                     int64_t N, int M, double payoff_fun, const double* w,
                     int quirk, int32_t* when_out) {
    const int pf = (int)payoff_fun;
    if (N % 2 != 0) return NAN;
    double dt = T / M;
    double result = 0;
    // pathsfinder: paths[m][n], (M+1) x N, time-major.
    std::vector<std::vector<double>> paths(M + 1);
    for (int i = 0; i < M + 1; ++i) paths[i].resize(N);
    for (int64_t n = 0; n < N / 2; ++n) {
        paths[0][n] = S0;
        paths[0][n + N / 2] = S0;
        for (int m = 1; m < M + 1; ++m) {
            double ww = w[n * (int64_t)M + (m - 1)];
            paths[m][n] = paths[m - 1][n] * exp((r - 0.5 * sigma * sigma) * dt + sigma * ww);
            paths[m][n + N / 2] = paths[m - 1][n + N / 2] * exp((r - 0.5 * sigma * sigma) * dt - sigma * ww);
        }
    }
    std::vector<double> exercise_when(N, M);
    std::vector<double> exercise_st(N);
    for (int64_t n = 0; n < N; ++n) exercise_st[n] = payoff(paths[M][n], E, pf);
    std::vector<double> x(N), y(N);
    for (int m = M - 1; m > 0; --m) {
        std::fill(x.begin(), x.end(), -1.0);
        std::fill(y.begin(), y.end(), -1.0);
        double sum_x = 0, sum_x2 = 0, sum_x3 = 0, sum_x4 = 0, sum_y = 0, sum_yx = 0, sum_yx2 = 0;
        double x_length = 0;
        std::vector<double> fst_po, fst_y;
        std::vector<int64_t> fst_n;
        for (int64_t n = 0; n < N; ++n) {
            double payoff_val = payoff(paths[m][n], E, pf);
            if (payoff_val > 0) {
                ++x_length;
                double exer = paths[m][n] - E;
                x[n] = exer;
                double cont = exp(-r * dt * (exercise_when[n] - m)) *
                              payoff(paths[(int)exercise_when[n]][n], E, pf);
                y[n] = cont;
                sum_x += exer;
                sum_x2 += exer * exer;
                sum_x3 += exer * exer * exer;
                sum_x4 += exer * exer * exer * exer;
                sum_y += cont;
                sum_yx += cont * exer;
                sum_yx2 += cont * exer * exer;
                if (x_length <= 2) {
                    fst_po.push_back(payoff_val);
                    fst_y.push_back(cont);
                    fst_n.push_back(n);
                }
            }
        }
        if (x_length == 0) {
            continue;
        } else if (x_length <= 2) {
            for (size_t i = 0; i < fst_n.size(); ++i)
                if (fst_po[i] > fst_y[i]) {
                    exercise_when[fst_n[i]] = m;
                    exercise_st[fst_n[i]] = fst_po[i];
                }
            continue;
        }
        double A[9] = {x_length, sum_x, sum_x2, sum_x, sum_x2, sum_x3, sum_x2, sum_x3, sum_x4};
        double rhs[3] = {sum_y, sum_yx, sum_yx2};
        double coef[3];
        if (!orc_inverse3_solve(A, rhs, coef)) return NAN;
        for (int64_t i = 0; i < N; ++i) {
            if (x[i] != -1) {
                double Y_hat = coef[0] + coef[1] * x[i] + coef[2] * pow(x[i], 2);
                double payoff_val = quirk ? payoff(x[i], E, pf) : payoff(paths[m][i], E, pf);
                if (payoff_val > Y_hat) {
                    exercise_when[i] = m;
                    exercise_st[i] = payoff_val;
                }
            }
        }
    }
    for (int64_t n = 0; n < N; ++n)
        if (exercise_st[n] != 0) result += exp(-r * exercise_when[n] * dt) * exercise_st[n];
    if (when_out)
        for (int64_t n = 0; n < N; ++n) when_out[n] = (int32_t)exercise_when[n];
    return std::max(payoff(S0, E, pf), result / (double)N);
}

double orc_mc_amer_seed(double S0, double E, double r, double sigma, double T,
                        int64_t N, int M, double payoff_fun, uint64_t seed, int quirk) {
    std::vector<double> w((size_t)(N / 2) * M);
    orc_normals_mt(seed, sqrt(T / M), (N / 2) * (int64_t)M, w.data());
    return orc_mc_amer_z(S0, E, r, sigma, T, N, M, payoff_fun, w.data(), quirk, nullptr);
}

// ---------------------------------------------------------------------------
// mc_amer with a K-term basis {1, x, ..., x^(K-1)}, x = S/E - 1 (K = 2..6): the
// generalisation of the reference's regression (src/mc_amer.cpp:38-108, basis
// {1, x, x^2}; the raw-S experiments in depr/mc_amer/v3/mc_amer.cpp:70-148)
// with the production decision rule (corrected exercise test payoff(S, E),
// continuation y = exp(-r dt (when - m)) * cash flow from a table).
// Per date, over the in-the-money paths in path order:
//   g[j]        = sum x^j,          j = 0..2K-2   (g[0] = count)
//   g[2K-1 + i] = sum y x^i,        i = 0..K-1
// then orc_lsm_solve_poly.  Fewer than 3 in-the-money paths: compare the
// payoff with y directly.  Store: the reference recurrence on w (as
// orc_mc_amer_z).
// Pinned: K = 3 agrees with orc_mc_amer_z(quirk = 0) (the same fitted values
// in exact arithmetic); the solve against numpy on random and exactly
// polynomial systems; the whole estimator against an independent numpy
// least-squares LSM (tests/test_oracle.py).
//
// Cholesky of the Hankel Gram G[i][j] = g[i+j]: column j is accepted while its
// pivot p_j = G[j][j] - sum_t L[j][t]^2 exceeds 1e-13 G[j][j]; at the first
// rejected column the leading kk x kk system is solved (kk < K: flag 4,
// "reduced").  c = {flag, coefficients of x^0..x^(K-1)} (unused ones 0).
// Flags: 0 no path in the money, 1 at most two, 2 full basis, 4 reduced.
int orc_lsm_solve_poly(const double* g, int K, double* c) {
    for (int i = 0; i <= K; ++i) c[i] = 0.0;
    const double n = g[0];
    if (n == 0.0) return (int)(c[0] = 0);
    if (n <= 2.0) return (int)(c[0] = 1);
    double L[8][8] = {{0}}, id[8] = {0}, y[8] = {0}, x[8] = {0};
    int kk = K;
    for (int j = 0; j < K; ++j) {
        double p = g[2 * j];
        for (int t = 0; t < j; ++t) p -= L[j][t] * L[j][t];
        if (j > 0 && !(p > 1e-13 * g[2 * j])) {
            kk = j;
            break;
        }
        L[j][j] = sqrt(p);
        id[j] = 1.0 / L[j][j];
        for (int i = j + 1; i < K; ++i) {
            double s = g[i + j];
            for (int t = 0; t < j; ++t) s -= L[i][t] * L[j][t];
            L[i][j] = s * id[j];
        }
    }
    for (int i = 0; i < kk; ++i) {
        double s = g[2 * K - 1 + i];
        for (int t = 0; t < i; ++t) s -= L[i][t] * y[t];
        y[i] = s * id[i];
    }
    for (int i = kk - 1; i >= 0; --i) {
        double s = y[i];
        for (int t = i + 1; t < kk; ++t) s -= L[t][i] * x[t];
        x[i] = s * id[i];
    }
    for (int i = 0; i < kk; ++i) c[1 + i] = x[i];
    c[0] = kk == K ? 2 : 4;
    return (int)c[0];
}

double orc_mc_amer_poly(double S0, double E, double r, double sigma, double T, int64_t N, int M,
                        double payoff_fun, const double* w, int K, int32_t* when_out, int32_t* flags_out) {
    const int pf = (int)payoff_fun;
    if (N % 2 != 0 || K < 2 || K > 6) return NAN;
    const double dt = T / M, invE = 1.0 / E;
    const int64_t half = N / 2;
    std::vector<std::vector<double>> S(M + 1, std::vector<double>(N));
    for (int64_t n = 0; n < half; ++n) {
        S[0][n] = S0;
        S[0][n + half] = S0;
        for (int m = 1; m <= M; ++m) {
            const double ww = w[n * (int64_t)M + (m - 1)];
            S[m][n] = S[m - 1][n] * exp((r - 0.5 * sigma * sigma) * dt + sigma * ww);
            S[m][n + half] = S[m - 1][n + half] * exp((r - 0.5 * sigma * sigma) * dt - sigma * ww);
        }
    }
    std::vector<double> disc(M + 1);
    for (int k = 0; k <= M; ++k) disc[k] = exp(-r * dt * k);
    std::vector<int32_t> when(N, M);
    std::vector<double> cf(N);
    for (int64_t n = 0; n < N; ++n) cf[n] = payoff(S[M][n], E, pf);
    for (int m = M - 1; m >= 1; --m) {
        double g[20] = {0};
        for (int64_t n = 0; n < N; ++n) {
            if (!(payoff(S[m][n], E, pf) > 0.0)) continue;
            const double x = S[m][n] * invE - 1.0;
            const double y = disc[when[n] - m] * cf[n];
            double p = x;
            g[0] += 1.0;
            g[1] += p;
            for (int j = 2; j <= 2 * K - 2; ++j) {
                p = p * x;
                g[j] += p;
            }
            double q = y;
            g[2 * K - 1] += q;
            for (int i = 1; i < K; ++i) {
                q = q * x;
                g[2 * K - 1 + i] += q;
            }
        }
        double c[8];
        const int flag = orc_lsm_solve_poly(g, K, c);
        if (flags_out) flags_out[m] = flag;
        if (flag == 0) continue;
        for (int64_t n = 0; n < N; ++n) {
            const double pv = payoff(S[m][n], E, pf);
            if (!(pv > 0.0)) continue;
            double v;
            if (flag == 1) {
                v = disc[when[n] - m] * cf[n];
            } else {
                const double x = S[m][n] * invE - 1.0;
                v = c[K];
                for (int i = K - 1; i >= 1; --i) v = c[i] + x * v;
            }
            if (pv > v) {
                when[n] = m;
                cf[n] = pv;
            }
        }
    }
    double s = 0.0;
    for (int64_t n = 0; n < N; ++n) s += disc[when[n]] * cf[n];
    if (when_out)
        for (int64_t n = 0; n < N; ++n) when_out[n] = when[n];
    return std::max(payoff(S0, E, pf), s / (double)N);
}

// ---------------------------------------------------------------------------
// Basket -- src/mc_eur_multi.cpp:6-35 with mvnorm, include/mvn.h:42-82.
//   Z: standard normals, column-major d x N (asset fastest), the order in
//   which mvnorm's NullaryExpr is filled.  L: lower Cholesky factor (d x d,
//   row-major) of the correlation matrix.  B(a,n) = sum_{k<=a} L(a,k) Z(k,n).
//   No sqrt(T) on B (the reference omits it; exact only at T=1).
//   sig/s0/wt may be NULL: then the reference's scalar sigma, S0 and 1/d.
// Parity unpinned at the bit level: Eigen and Boost are not installed, so the
// reference basket cannot be built; pinned instead by d=1 == mc_eur on the
// same normals, rho=1 == single-asset Black-Scholes, and L*L^T == C.
double orc_basket_z(double S0, double E, double r, double sigma, double T,
                    int64_t N, double payoff_fun, int d, const double* L,
                    const double* Z, const double* sig, const double* s0,
                    const double* wt) {
    const int pf = (int)payoff_fun;
    double w_i = 1.0 / (double)d;
    double result = 0;
    std::vector<double> B(d);
    for (int64_t n = 0; n < N; ++n) {
        const double* z = Z + n * (int64_t)d;
        for (int a = 0; a < d; ++a) {
            double acc = 0;
            for (int k = 0; k <= a; ++k) acc += L[a * d + k] * z[k];
            B[a] = acc;
        }
        double result_n = 0;
        for (int a = 0; a < d; ++a) {
            double sg = sig ? sig[a] : sigma;
            double S = s0 ? s0[a] : S0;
            double wa = wt ? wt[a] : w_i;
            result_n += wa * S * exp((r - pow(sg, 2) / 2) * T + sg * B[a]);
        }
        result += payoff(result_n, E, pf);
    }
    return (exp(-r * T) * result) / (double)N;
}

// Production basket (FP64): the reference's estimator with the sqrt(T) that
// mvnorm leaves out (include/mvn.h:78-80 samples L Z with Z ~ N(0, I) and
// src/mc_eur_multi.cpp:30 multiplies by sigma only), i.e. the terminal value
// of each asset of a GBM basket:
//     S_a = S0_a exp((r - s_a^2/2) T + (s_a sqrt(T)) B_a),  B = L z.
// Z column-major d x N (asset fastest) as in orc_basket_z; sig/s0/wt may be
// NULL (scalar sigma, S0, weights 1/d).  se_out (optional) receives the
// standard error.
// Pinned: d = 1 equals orc_mc_eur_z on the draws z sqrt(T) (the reference's
// mc_eur, src/mc_eur.cpp:15-27); a rank-one factor (rho = 1) prices the
// single-asset Black-Scholes value within a few standard errors at T != 1.
double orc_basket_prod_z(double S0, double E, double r, double sigma, double T, int64_t N, double payoff_fun,
                         int d, const double* L, const double* Z, const double* sig, const double* s0,
                         const double* wt, double* se_out) {
    const int pf = (int)payoff_fun;
    const double sqT = sqrt(T);
    std::vector<double> c0(d), c1(d), ws(d), B(d);
    for (int a = 0; a < d; ++a) {
        const double sg = sig ? sig[a] : sigma;
        c0[a] = (r - pow(sg, 2) / 2) * T;
        c1[a] = sg * sqT;
        ws[a] = (wt ? wt[a] : 1.0 / (double)d) * (s0 ? s0[a] : S0);
    }
    double s = 0, s2 = 0;
    for (int64_t n = 0; n < N; ++n) {
        const double* z = Z + n * (int64_t)d;
        double basket = 0;
        for (int a = 0; a < d; ++a) {
            double acc = 0;
            for (int k = 0; k < d; ++k) acc += L[a * d + k] * z[k];
            basket += ws[a] * exp(c0[a] + c1[a] * acc);
        }
        const double v = payoff(basket, E, pf);
        s += v;
        s2 += v * v;
    }
    const double disc = exp(-r * T);
    const double mean = s / N;
    const double var = (s2 - s * mean) / (double)(N > 1 ? N - 1 : 1);
    if (se_out) *se_out = disc * sqrt(var / N);
    return (disc * s) / (double)N;
}

// Lower Cholesky factor of a d x d SPD matrix (row-major), FP64, textbook
// column-by-column order.  Returns 0 if not positive definite.
// Pinned: L L^T == C residual and the equicorrelation closed form in tests.
int orc_cholesky(int d, const double* C, double* L) {
    for (int i = 0; i < d * d; ++i) L[i] = 0;
    for (int j = 0; j < d; ++j) {
        double s = C[j * d + j];
        for (int k = 0; k < j; ++k) s -= L[j * d + k] * L[j * d + k];
        if (!(s > 0)) return 0;
        double ljj = sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; ++i) {
            double t = C[i * d + j];
            for (int k = 0; k < j; ++k) t -= L[i * d + k] * L[j * d + k];
            L[i * d + j] = t / ljj;
        }
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC11).  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85, 10 rounds.
// Pinned: Random123 known-answer vectors and rocRAND's philox4x32_10 engine
// (tests/test_oracle_philox.py).
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Build normals of the production generator.  Counter layout (shared with the
// GPU kernels, documented in include/pcf.h):
//   key = (seed lo, seed hi); ctr = (block, path lo, path hi, stream);
//   uniform u_j = (x_j + 0.5) * 2^-32; Box-Muller on (u0,u1) and (u2,u3):
//   z0 = R cos(2 pi u1), z1 = R sin(2 pi u1), R = sqrt(-2 ln u0); same for z2, z3.
// out[0..3] are the 4 standard normals of block `block` of `path`.
void orc_philox_normals4(uint64_t seed, uint32_t stream, uint64_t path, uint32_t block, double out[4]) {
    uint32_t ctr[4] = {block, (uint32_t)path, (uint32_t)(path >> 32), stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    orc_philox4x32_10(ctr, key, x);
    const double two_pi = 6.283185307179586476925286766559;
    for (int h = 0; h < 2; ++h) {
        double u0 = ((double)x[2 * h] + 0.5) * 0x1p-32;
        double u1 = ((double)x[2 * h + 1] + 0.5) * 0x1p-32;
        double R = sqrt(-2.0 * log(u0));
        out[2 * h] = R * cos(two_pi * u1);
        out[2 * h + 1] = R * sin(two_pi * u1);
    }
}

// Stream ids of the products (include/pcf.h PCF_STREAM_*).
enum { ST_EUR = 1, ST_ASIA = 2, ST_AMER = 3, ST_BASKET = 4 };

// Normal k of the production stream for (path, step): block k/4, lane k%4.
static inline double pnorm(uint64_t seed, uint32_t stream, uint64_t path, uint32_t k) {
    double o[4];
    orc_philox_normals4(seed, stream, path, k >> 2, o);
    return o[k & 3];
}

// mc_eur production estimator (FP64): path p uses normal p%4 of block 0 of
// Philox path-group p/4.  Returns price; se_out receives the standard error.
double orc_philox_mc_eur(double S0, double E, double r, double sigma, double T,
                         int64_t N, double payoff_fun, uint64_t seed, double* se_out) {
    const int pf = (int)payoff_fun;
    double s = 0, s2 = 0;
    double mu = (r - 0.5 * sigma * sigma) * T, vol = sigma * sqrt(T);
    for (int64_t p = 0; p < N; p += 4) {
        double z[4];
        orc_philox_normals4(seed, ST_EUR, (uint64_t)(p >> 2), 0, z);
        for (int j = 0; j < 4 && p + j < N; ++j) {
            double v = payoff(S0 * exp(mu + vol * z[j]), E, pf);
            s += v; s2 += v * v;
        }
    }
    double disc = exp(-r * T);
    double mean = s / N;
    double var = (s2 - s * mean) / (double)(N > 1 ? N - 1 : 1);
    if (se_out) *se_out = disc * sqrt(var / N);
    return disc * mean;
}

// mc_asia production arithmetic (FP64): the recurrence of src/mc_asia.cpp:31-35
// written on standard normals z,
//     I += S (a1 + a2 z),  S *= exp(mu + vol z),
//     a1 = 1 + r dt/2, a2 = sigma sqrt(dt)/2, mu = (r - sigma^2/2) dt, vol = sigma sqrt(dt),
// which the reference writes on its draws dB = sqrt(dt) z.  The geometric
// control averages log S over t_1..t_M on the same path.
struct AsiaProd {
    double a1, a2, mu, vol, sdt;
};
static AsiaProd asia_prod(double r, double sigma, double T, int M) {
    AsiaProd c;
    const double dt = T / M;
    c.sdt = sqrt(dt);
    c.a1 = 1 + r * dt / 2;
    c.a2 = sigma * c.sdt / 2;
    c.mu = (r - 0.5 * sigma * sigma) * dt;
    c.vol = sigma * c.sdt;
    return c;
}
extern "C++" {
template <class ZF>
static void asia_prod_path(const AsiaProd& c, double S0, int M, ZF z_of, double* avg, double* geo_log) {
    double S = S0, I = 0, ls = log(S0), lsum = 0;
    for (int m = 0; m < M; ++m) {
        const double z = z_of(m);
        I += S * (c.a1 + c.a2 * z);
        S *= exp(c.mu + c.vol * z);
        ls += c.mu + c.vol * z;
        lsum += ls;
    }
    *avg = I / M;
    *geo_log = lsum / M;
}
}  // extern "C++"

// The production arithmetic fed by the reference's draws: w[n*M+m] are
// N(0, sqrt(dt)) as in orc_mc_asia_z, and z = w / sqrt(dt).  This is the
// production form on the reference's own stream, so it must agree with
// orc_mc_asia_z to rounding: a wrong a1, a2, mu or vol does not.
// Pinned: orc_mc_asia_z on the same draws to 1e-12 (tests/test_oracle.py).
double orc_mc_asia_prod_z(double S0, double E, double r, double sigma, double T, int64_t N, int M,
                          double payoff_fun, const double* w, double* se_out) {
    const int pf = (int)payoff_fun;
    const AsiaProd c = asia_prod(r, sigma, T, M);
    double s = 0, s2 = 0;
    for (int64_t p = 0; p < N; ++p) {
        const double* wp = w + p * (int64_t)M;
        double avg, gl;
        asia_prod_path(c, S0, M, [&](int m) { return wp[m] / c.sdt; }, &avg, &gl);
        const double v = payoff(avg, E, pf);
        s += v;
        s2 += v * v;
    }
    const double disc = exp(-r * T);
    const double mean = s / N;
    const double var = (s2 - s * mean) / (double)(N > 1 ? N - 1 : 1);
    if (se_out) *se_out = disc * sqrt(var / N);
    return disc * mean;
}

// mc_asia production estimator (FP64): the production arithmetic above driven
// by the Philox stream, step m of path p uses normal m.  geo_out (optional)
// receives the discounted price of the geometric-average control on the same
// paths, whose closed form is orc_geo_asian().
// Pinned: the arithmetic through orc_mc_asia_prod_z; the normals through
// orc_philox_normals4's known answers; the estimator statistically against
// orc_geo_asian.
double orc_philox_mc_asia(double S0, double E, double r, double sigma, double T,
                          int64_t N, int M, double payoff_fun, uint64_t seed,
                          double* se_out, double* geo_out) {
    const int pf = (int)payoff_fun;
    const AsiaProd c = asia_prod(r, sigma, T, M);
    double s = 0, s2 = 0, g = 0;
    for (int64_t p = 0; p < N; ++p) {
        double avg, gl;
        asia_prod_path(c, S0, M, [&](int m) { return pnorm(seed, ST_ASIA, (uint64_t)p, (uint32_t)m); }, &avg,
                       &gl);
        const double v = payoff(avg, E, pf);
        s += v;
        s2 += v * v;
        g += payoff(exp(gl), E, pf);
    }
    double disc = exp(-r * T);
    double mean = s / N;
    double var = (s2 - s * mean) / (double)(N > 1 ? N - 1 : 1);
    if (se_out) *se_out = disc * sqrt(var / N);
    if (geo_out) *geo_out = disc * g / N;
    return disc * mean;
}

// Production normals for the American path store: pair n (0..N/2-1), date m
// (1..M) uses normal m-1 of Philox path n, stream ST_AMER; path n takes +w,
// path n+N/2 takes -w (the antithetic layout of include/common.h:195-206).
void orc_philox_amer_normals(uint64_t seed, int64_t N, int M, double T, double* w) {
    double sdt = sqrt(T / M);
    for (int64_t n = 0; n < N / 2; ++n)
        for (int m = 1; m <= M; ++m)
            w[n * (int64_t)M + (m - 1)] = sdt * pnorm(seed, ST_AMER, (uint64_t)n, (uint32_t)(m - 1));
}

// Production normals for the basket: path p, asset a uses normal a of Philox
// path p, stream ST_BASKET.  Z is column-major d x N.
void orc_philox_basket_normals(uint64_t seed, int64_t N, int d, double* Z) {
    for (int64_t p = 0; p < N; ++p)
        for (int a = 0; a < d; ++a) Z[p * d + a] = pnorm(seed, ST_BASKET, (uint64_t)p, (uint32_t)a);
}

// Production normals for an arbitrary stream: out[p*K + k] = normal k of path
// p0+p.  Used to check the device generator bit-for-bit on the uniforms.
void orc_philox_normals(uint64_t seed, uint32_t stream, uint64_t p0, int64_t np, int K, double* out) {
    for (int64_t p = 0; p < np; ++p)
        for (int k = 0; k < K; ++k) out[p * K + k] = pnorm(seed, stream, p0 + (uint64_t)p, (uint32_t)k);
}

// ---------------------------------------------------------------------------
// Closed forms.
static double ncdf(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

// Black-Scholes (the R script depr/eur_analytical.R:1-10 computes the same).
double orc_bs_price(double S0, double E, double r, double sigma, double T, double payoff_fun) {
    double d1 = (log(S0 / E) + (r + 0.5 * sigma * sigma) * T) / (sigma * sqrt(T));
    double d2 = d1 - sigma * sqrt(T);
    if (payoff_fun > 0) return S0 * ncdf(d1) - E * exp(-r * T) * ncdf(d2);
    return E * exp(-r * T) * ncdf(-d2) - S0 * ncdf(-d1);
}

// Discretely monitored geometric-average Asian (monitoring t_i = i T/M,
// i = 1..M): log G ~ N(log S0 + mu_g, v_g) with
//   mu_g = (r - sigma^2/2) T (M+1)/(2M),  v_g = sigma^2 T (M+1)(2M+1)/(6 M^2).
double orc_geo_asian(double S0, double E, double r, double sigma, double T, int M, double payoff_fun) {
    double mg = log(S0) + (r - 0.5 * sigma * sigma) * T * (M + 1) / (2.0 * M);
    double vg = sigma * sigma * T * (M + 1) * (2.0 * M + 1) / (6.0 * M * M);
    double sg = sqrt(vg);
    double fwd = exp(mg + 0.5 * vg);
    double d1 = (mg - log(E) + vg) / sg, d2 = d1 - sg;
    double disc = exp(-r * T);
    if (payoff_fun > 0) return disc * (fwd * ncdf(d1) - E * ncdf(d2));
    return disc * (E * ncdf(-d2) - fwd * ncdf(-d1));
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
