// Correctly rounded sin, cos, pow, atan2 and acos in binary64 for the initial
// data (the Fishbone-Moncrief torus and its vector potential,
// src/fluid/solver/initconds.cpp:17-23, 121-225, 431-589, and the Kerr-Schild
// metric components they evaluate, src/fluid/physics/spacetime.cpp:20-34,
// 93-100).  A correctly rounded result is unique, so the device and any other
// correctly rounded implementation agree bit for bit; the C library's
// functions are accurate to about half an ulp and agree except in rare
// near-halfway cases.
//
// Methods (double-double arithmetic of bh_crmath.hpp throughout; each result
// is rounded once from a double-double value with relative error below about
// 2^-96, so it is the correctly rounded one unless the exact value lies that
// close to a rounding boundary):
//  * sin, cos: x = n pi/2 + r, |r| <= pi/4 (Cody-Waite reduction with pi/2
//    in four parts, the first of 33 bits so that n * part is exact for
//    |n| < 2^20), then the Taylor series of sin r and cos r, the terms below
//    2^-58 in binary64 and the rest in double-double.
//  * pow(x, y) = exp(y log x): log x in double-double (one Newton correction of
//    a binary64 estimate, or 2 atanh((x-1)/(x+1)) near 1), the product with y
//    in double-double, exp of a double-double argument.  y = 2, -1, 1 are the
//    exact operations x*x, 1/x, x (whose IEEE results are correctly rounded).
//  * atan2(y, x): one Newton step from the binary64 estimate t0,
//    t = t0 + (y cos t0 - x sin t0) / (x cos t0 + y sin t0), with sin t0 and
//    cos t0 in double-double.
//  * acos(x): one Newton step from the binary64 estimate t0,
//    t = t0 + (cos t0 - x) / sin t0, where for x > 1/2 the numerator is formed
//    as (1 - x) - 2 sin^2(t0/2) (1 - x exact) to keep its relative accuracy
//    near x = 1, and acos(x) = pi - acos(-x) for x < -1/2.
// Special values (zeros, infinities, NaN, |x| > 1 for acos, negative bases
// with non-integer exponents) follow C99 Annex F and are taken from the
// platform's function, which returns them exactly.
#pragma once

#include "bh_crmath.hpp"

namespace bh {

// pi/2 in four parts; the first has 33 significant bits
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a626331p-34;
constexpr double kPio2_3 = 0x1.1701b839a2520p-88;
constexpr double kPio2_4 = 0x1.27044533e63a0p-142;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPiHi = 0x1.921fb54442d18p+1, kPiLo = 0x1.1a62633145c07p-53;

// 1/n! as double-double: kInvFact[n] = {hi, lo}, n = 0..29
static constexpr double kInvFact[30][2] = {
    {1.0, 0.0},
    {1.0, 0.0},
    {0x1.0000000000000p-1, 0.0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107},
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124},
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143},
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149},
    {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153},
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157},
};

BH_HD dd dd_neg(dd a) { return {-a.hi, -a.lo}; }

// sin r and cos r for a double-double |r| <= pi/4 + 2^-40.
// sin r = r * sum_k (-1)^k u^k / (2k+1)!,  cos r = sum_k (-1)^k u^k / (2k)!,
// u = r^2; the terms k >= 9 (below 2^-58 relative) in binary64, k <= 8 in
// double-double.  The first omitted terms are below 2^-110.
BH_HD void sincos_poly(dd r, dd& s, dd& c) {
    const dd u = dd_mul(r, r);
    double qs = 0.0, qc = 0.0;
    for (int k = 13; k >= 9; k--) qs = qs * u.hi + ((k & 1) ? -kInvFact[2 * k + 1][0] : kInvFact[2 * k + 1][0]);
    for (int k = 14; k >= 9; k--) qc = qc * u.hi + ((k & 1) ? -kInvFact[2 * k][0] : kInvFact[2 * k][0]);
    dd ps{qs, 0.0}, pc{qc, 0.0};
    for (int k = 8; k >= 0; k--) {
        const double sgn = (k & 1) ? -1.0 : 1.0;
        ps = dd_add(dd_mul(ps, u), dd{sgn * kInvFact[2 * k + 1][0], sgn * kInvFact[2 * k + 1][1]});
        pc = dd_add(dd_mul(pc, u), dd{sgn * kInvFact[2 * k][0], sgn * kInvFact[2 * k][1]});
    }
    s = dd_mul(r, ps);
    c = pc;
}

// x = n pi/2 + r with |r| <= pi/4 (slightly more at the rounding of n);
// |x| < 2^19.  Returns n mod 4.
BH_HD int reduce_pio2(double x, dd& r) {
    const double nd = rint(x * kTwoOverPi);
    const double r1 = x - nd * kPio2_1;  // exact: n * part 1 exact, Sterbenz
    const dd t = two_prod(nd, kPio2_2);
    r = two_sum(r1, -t.hi);
    r.lo += -t.lo - nd * kPio2_3 - nd * kPio2_4;
    r = fast_two_sum(r.hi, r.lo);
    return ((int)nd) & 3;
}

// sin x and cos x as double-double, x finite, 0 < |x| < 2^19
BH_HD void sincos_dd(double x, dd& s, dd& c) {
    dd r;
    const int q = reduce_pio2(x, r);
    dd sr, cr;
    sincos_poly(r, sr, cr);
    switch (q) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = dd_neg(sr); break;
        case 2: s = dd_neg(sr); c = dd_neg(cr); break;
        default: s = dd_neg(cr); c = sr; break;
    }
}

BH_HD double cr_sin(double x) {
    if (!(fabs(x) < 0x1p19)) return sin(x);  // NaN, inf (and arguments never used here)
    if (fabs(x) < 0x1p-26) return x;          // |sin x - x| < x^3/6 < half an ulp; keeps -0
    dd s, c;
    sincos_dd(x, s, c);
    return s.hi;
}
BH_HD double cr_cos(double x) {
    if (!(fabs(x) < 0x1p19)) return cos(x);
    if (fabs(x) < 0x1p-27) return 1.0;  // 1 - x^2/2 rounds to 1
    dd s, c;
    sincos_dd(x, s, c);
    return c.hi;
}

// log x as double-double, x positive and finite (absolute error about 2^-101)
BH_HD dd log_dd(double x) {
    if (fabs(x - 1.0) < 0x1p-7) {
        // 2 atanh(s), s = (x-1)/(x+1), as in log_near1
        double d = x - 1.0;
        dd den = two_sum(2.0, d);
        double sh = d / den.hi;
        double rr = fma(-sh, den.hi, d);
        rr -= sh * den.lo;
        dd s = fast_two_sum(sh, rr / den.hi);
        dd s2 = dd_mul(s, s);
        const double c[8][2] = {{0x1.0000000000000p+0, 0.0},
                                {0x1.5555555555555p-2, 0x1.5555555555555p-56},
                                {0x1.999999999999ap-3, -0x1.999999999999ap-57},
                                {0x1.2492492492492p-3, 0x1.2492492492492p-57},
                                {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
                                {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
                                {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
                                {0x1.1111111111111p-4, 0x1.1111111111111p-60}};
        dd p = dd{c[7][0], c[7][1]};
        for (int k = 6; k >= 0; k--) p = dd_add(dd_mul(p, s2), dd{c[k][0], c[k][1]});
        dd y = dd_mul(s, p);
        return dd{2.0 * y.hi, 2.0 * y.lo};
    }
    int k = 0;
    if (x < 0x1p-1000) { x *= 0x1p200; k = -200; }
    else if (x > 0x1p1000) { x *= 0x1p-200; k = 200; }
    const double y0 = log(x);
    int m;
    const dd e = exp_dd(y0, m);
    const double xs = ldexp(x, -m);
    const double d = (xs - e.hi) - e.lo;
    const double t = d / e.hi;
    const double corr = t - 0.5 * t * t;
    dd s = two_sum(y0, corr);
    if (k != 0) {
        const double kd = (double)k;
        s = dd_add(s, two_sum(kd * kLn2Hi, kd * kLn2Mid));
        s.lo += kd * kLn2Lo;
        s = fast_two_sum(s.hi, s.lo);
    }
    return s;
}

// exp of a double-double argument, rounded once
BH_HD double exp_of_dd(dd z) {
    if (z.hi != z.hi) return z.hi;
    if (z.hi > 709.79) return HUGE_VAL;
    if (z.hi < -745.14) return 0.0;
    int m;
    const dd e = exp_dd(z.hi, m);
    const double f = z.lo + 0.5 * z.lo * z.lo;  // exp(z.lo) - 1, |z.lo| <= 2^-43
    dd y = dd_add(e, dd_mul_d(e, f));
    if (m < -1021) return scale_subnormal(y, m);
    return ldexp(y.hi, m);
}

BH_HD double cr_pow(double x, double y) {
    if (y == 2.0) return x * x;
    if (y == 1.0) return x;
    if (y == -1.0) return 1.0 / x;
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x == 0.0 || !isfinite(x) || !isfinite(y)) return pow(x, y);
    double sign = 1.0;
    if (x < 0.0) {
        if (rint(y) != y) return pow(x, y);  // NaN
        // odd integer exponent keeps the sign (|y| >= 2^53 is even)
        if (fabs(y) < 0x1p53 && fmod(y, 2.0) != 0.0) sign = -1.0;
        x = -x;
    }
    const dd L = log_dd(x);
    dd z = two_prod(L.hi, y);
    z.lo += L.lo * y;
    z = fast_two_sum(z.hi, z.lo);
    return sign * exp_of_dd(z);
}

BH_HD double cr_atan2(double y, double x) {
    if (x == 0.0 || y == 0.0 || !isfinite(x) || !isfinite(y)) return atan2(y, x);
    // keep the products below in the normal range
    const double big = fmax(fabs(x), fabs(y));
    if (big > 0x1p900) { x *= 0x1p-200; y *= 0x1p-200; }
    else if (big < 0x1p-900) { x *= 0x1p200; y *= 0x1p200; }
    const double t0 = atan2(y, x);
    dd s, c;
    sincos_dd(t0, s, c);
    const dd num = dd_add(dd_mul_d(c, y), dd_neg(dd_mul_d(s, x)));
    const double den = x * c.hi + y * s.hi;
    return t0 + (num.hi + num.lo) / den;
}

// acos x as double-double for -1/2 <= x < 1
BH_HD dd acos_dd(double x) {
    const double t0 = acos(x);
    dd s, c;
    sincos_dd(t0, s, c);
    dd num;
    if (x > 0.5) {
        dd sh, ch;
        sincos_dd(0.5 * t0, sh, ch);
        dd two_s2 = dd_mul(sh, sh);
        two_s2.hi *= 2.0; two_s2.lo *= 2.0;
        num = dd_add(dd{1.0 - x, 0.0}, dd_neg(two_s2));  // (1 - x) - (1 - cos t0)
    } else {
        num = dd_add(c, dd{-x, 0.0});
    }
    const double delta = (num.hi + num.lo) / s.hi;
    return two_sum(t0, delta);
}

BH_HD double cr_acos(double x) {
    if (!(fabs(x) < 1.0)) {
        if (x == 1.0) return 0.0;
        if (x == -1.0) return kPiHi;
        return acos(x);  // NaN
    }
    if (x >= -0.5) {
        const dd a = acos_dd(x);
        return a.hi + a.lo;
    }
    const dd a = acos_dd(-x);
    const dd r = dd_add(dd{kPiHi, kPiLo}, dd_neg(a));
    return r.hi;
}

}  // namespace bh
