// pcf_dd.hpp -- double-double arithmetic (an unevaluated sum hi + lo of two
// doubles, ~106 significant bits) for the binomial sum's log-space weights.
//
// Every operation is built from error-free transformations: Knuth's TwoSum
// and Dekker's TwoProduct (Veltkamp split, exact without a fused multiply-add),
// with contraction off, so host and device form the same bits from the same
// inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pcf {

struct DD {
    double h, l;
};

__host__ __device__ inline DD dd_two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b, bp = s - a;
    return DD{s, (a - (s - bp)) + (b - bp)};
}

__host__ __device__ inline DD dd_quick(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    return DD{s, b - (s - a)};
}

__host__ __device__ inline DD dd_two_prod(double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    const double ca = 134217729.0 * a, cb = 134217729.0 * b;  // 2^27 + 1
    const double ah = ca - (ca - a), al = a - ah;
    const double bh = cb - (cb - b), bl = b - bh;
    return DD{p, ((ah * bh - p) + ah * bl + al * bh) + al * bl};
}

__host__ __device__ inline DD dd_add(DD a, DD b) {
#pragma clang fp contract(off)
    DD s = dd_two_sum(a.h, b.h);
    s.l += a.l + b.l;
    return dd_quick(s.h, s.l);
}

__host__ __device__ inline DD dd_neg(DD a) { return DD{-a.h, -a.l}; }

__host__ __device__ inline DD dd_mul_d(DD a, double b) {
#pragma clang fp contract(off)
    DD p = dd_two_prod(a.h, b);
    p.l += a.l * b;
    return dd_quick(p.h, p.l);
}

__host__ __device__ inline DD dd_mul(DD a, DD b) {
#pragma clang fp contract(off)
    DD p = dd_two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return dd_quick(p.h, p.l);
}

// a / b: q = a.h / b, then the exact remainder a - q b divided once more.
__host__ __device__ inline DD dd_div_d(DD a, double b) {
#pragma clang fp contract(off)
    const double q = a.h / b;
    const DD qb = dd_two_prod(q, b);
    const double rem = ((a.h - qb.h) - qb.l) + a.l;
    return dd_quick(q, rem / b);
}

__host__ __device__ inline DD dd_div(DD a, DD b) {
#pragma clang fp contract(off)
    const double q1 = a.h / b.h;
    const DD r = dd_add(a, dd_mul_d(dd_neg(b), q1));
    return dd_quick(q1, r.h / b.h);
}

// Host only: 2 atanh(t) = sum_j 2 t^(2j+1) / (2j+1), until a term is below
// 1e-34 of the sum.
inline DD dd_2atanh(DD t) {
    const DD u = dd_mul(t, t);
    DD sum = t, pw = t;
    for (int j = 1; j < 200; ++j) {
        pw = dd_mul(pw, u);
        const DD term = dd_div_d(pw, (double)(2 * j + 1));
        sum = dd_add(sum, term);
        if (std::fabs(term.h) < 1e-34 * std::fabs(sum.h)) break;
    }
    return DD{2.0 * sum.h, 2.0 * sum.l};
}

// log 2 rounded to double-double.
constexpr DD kDdLn2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};

// Host only: log x for a finite x > 0 without libm's log: x = m 2^e with m in
// [0.75, 1.5), log x = e log 2 + 2 atanh((m - 1) / (m + 1)).
inline DD dd_log(double x) {
    int e = 0;
    double m = std::frexp(x, &e);  // [0.5, 1)
    if (m < 0.75) {
        m *= 2.0;
        e -= 1;
    }
    const DD t = dd_div(DD{m - 1.0, 0.0}, dd_two_sum(m, 1.0));
    return dd_add(dd_mul_d(kDdLn2, (double)e), dd_2atanh(t));
}

}  // namespace pcf
