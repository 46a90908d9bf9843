// d3d_atan2.hpp -- atan2 rounded correctly to FP64 (up to cases closer than
// about 2^-100 relative to a rounding midpoint), on the host and the GPU.
//
// The contact-polygon ordering sorts points by np.arctan2, which is the C
// library's atan2: accurate to well under one ulp, so its results are the
// correctly rounded ones in practice.  Near-duplicate polygon points (where
// three or more half-planes meet) are ordered by the last bit of their
// angles, so the device needs the same rounding, not just a close value.
//
// Method: double-double arithmetic (pairs hi + lo, error-free sums and
// products by fma) throughout, one rounding at the end.
//   r = min(|y|,|x|) / max(|y|,|x|) in [0, 1]
//   atan(r) = atan(c) + atan(u),  c = k/16 nearest r,  u = (r - c)/(1 + r c),
//   |u| <= 1/32, atan(c) from a double-double table, atan(u) by its odd
//   series u sum_n (-1)^n u^2n / (2n + 1) to n = 10 (terms below 2^-104)
//   atan2 = atan(r), pi/2 - atan(r) when |y| > |x|, pi - that when x < 0,
//   with the sign of y.  IEEE special cases (zeros, infinities, NaN) as C99.
#pragma once

#include <hip/hip_runtime.h>

namespace d3d {

struct DD {
    double hi, lo;
};

__host__ __device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ DD dd_fast(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return DD{s, b - (s - a)};
}
__host__ __device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast(s.hi, s.lo);
}
__host__ __device__ __forceinline__ DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }
__host__ __device__ __forceinline__ DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return dd_fast(p, e);
}
__host__ __device__ __forceinline__ DD dd_mul_d(DD a, double b) {
    const double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e += a.lo * b;
    return dd_fast(p, e);
}
__host__ __device__ __forceinline__ DD dd_div(DD a, DD b) {
    const double q1 = a.hi / b.hi;
    DD r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    const DD q = dd_fast(q1, q2);
    return dd_add(q, DD{q3, 0.0});
}

// atan(k/16), k = 0..16, as hi + lo
__host__ __device__ __forceinline__ DD atan_table(int k) {
    switch (k) {
    case 0: return DD{0.0, 0.0};
    case 1: return DD{0.06241880999595735, -1.5490756308295046e-18};
    case 2: return DD{0.12435499454676144, -3.1253241424539383e-18};
    case 3: return DD{0.18534794999569476, 4.180692268843079e-18};
    case 4: return DD{0.24497866312686414, 1.0698755618734451e-17};
    case 5: return DD{0.3028848683749714, -1.1010827903001369e-17};
    case 6: return DD{0.35877067027057225, -2.4623815582638635e-17};
    case 7: return DD{0.4124104415973873, -1.587652227770689e-17};
    case 8: return DD{0.4636476090008061, 2.2698777452961687e-17};
    case 9: return DD{0.5123894603107377, -2.5462781472855804e-17};
    case 10: return DD{0.5585993153435624, -5.4556305485916264e-18};
    case 11: return DD{0.6022873461349642, 2.950430737228402e-17};
    case 12: return DD{0.6435011087932844, 1.5834785051444286e-17};
    case 13: return DD{0.6823165548747481, 6.943223671560008e-18};
    case 14: return DD{0.7188299996216245, -2.1478388444456983e-17};
    case 15: return DD{0.7531512809621944, -2.4256934659182068e-17};
    default: return DD{0.7853981633974483, 3.061616997868383e-17};
    }
}

__host__ __device__ inline double atan2_cr(double y, double x) {
    constexpr DD kPi{3.141592653589793, 1.2246467991473532e-16};
    constexpr DD kPi2{1.5707963267948966, 6.123233995736766e-17};
    if (x != x || y != y) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    double res;
    if (ay == 0.0) {
        res = __builtin_signbit(x) ? kPi.hi : 0.0;
    } else if (ax == 0.0) {
        res = kPi2.hi;
    } else if (ay == INFINITY || ax == INFINITY) {
        if (ay == INFINITY && ax == INFINITY) res = x > 0.0 ? 0.7853981633974483 : 2.356194490192345;
        else if (ay == INFINITY) res = kPi2.hi;
        else res = x > 0.0 ? 0.0 : kPi.hi;
    } else {
        const bool swap = ay > ax;
        const double num = swap ? ax : ay, den = swap ? ay : ax;
        const double rh = num / den;
        const DD r = DD{rh, fma(-rh, den, num) / den};
        const int k = (int)(rh * 16.0 + 0.5);
        const double c = (double)k * 0.0625;
        // u = (r - c) / (1 + r c); r.hi - c is exact (Sterbenz, or c = 0)
        const DD nu = dd_fast(r.hi - c, r.lo);
        const DD de = dd_add(DD{1.0, 0.0}, dd_mul_d(r, c));
        const DD u = dd_div(nu, de);
        const DD w = dd_mul(u, u);
        // series tail n = 10..5 in FP64, n = 4..1 in double-double
        double t = 0.047619047619047616;            // 1/21
        t = -0.05263157894736842 + w.hi * t;        // -1/19
        t = 0.058823529411764705 + w.hi * t;        // 1/17
        t = -0.06666666666666667 + w.hi * t;        // -1/15
        t = 0.07692307692307693 + w.hi * t;         // 1/13
        t = -0.09090909090909091 + w.hi * t;        // -1/11
        DD s = dd_add(DD{0.1111111111111111, 6.1679056923619804e-18}, dd_mul_d(w, t));     // 1/9
        s = dd_add(DD{-0.14285714285714285, -7.93016446160826e-18}, dd_mul(w, s));         // -1/7
        s = dd_add(DD{0.2, -1.1102230246251566e-17}, dd_mul(w, s));                        // 1/5
        s = dd_add(DD{-0.3333333333333333, -1.850371707708594e-17}, dd_mul(w, s));         // -1/3
        const DD at_u = dd_add(u, dd_mul(dd_mul(u, w), s));
        DD a = dd_add(atan_table(k), at_u);
        if (swap) a = dd_add(kPi2, dd_neg(a));
        if (x < 0.0) a = dd_add(kPi, dd_neg(a));
        res = a.hi + a.lo;
    }
    return __builtin_signbit(y) ? -res : res;
}

}  // namespace d3d
