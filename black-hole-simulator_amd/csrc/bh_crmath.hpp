// Correctly rounded exp and log in binary64 for the reconstruction of
// log-density and log-energy (the reference reconstructs ln(rho), ln(eps) and
// exponentiates, src/fluid/grid/reconstruction.cpp:49-61).  A correctly
// rounded result is unique, so host and device agree bit for bit, and they
// agree with any other correctly rounded implementation; the C library's
// log/exp are accurate to about half an ulp and agree except in rare
// near-halfway cases.
//
// Method: exp(x) = 2^m * 2^(j/64) * exp(r), x = (64m + j) ln2/64 + r,
// |r| <= ln2/128, with the table 2^(j/64) and the Taylor series of exp(r)
// evaluated in double-double arithmetic (error about 2^-104), then rounded
// once.  log(x) takes a binary64 estimate y0 and applies one Newton
// correction y = y0 + log1p((x - e^y0)/e^y0) with e^y0 in double-double.
// The double-double primitives (two-sum, fast-two-sum, exact product by
// fma) are the standard error-free transformations (Dekker 1971; Knuth).
#pragma once

#include "bh_common.hpp"

namespace bh {

struct dd { double hi, lo; };

BH_HD dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
BH_HD dd fast_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}
BH_HD dd two_prod(double a, double b) {
    double p = a * b;
    double e = fma(a, b, -p);
    return {p, e};
}
BH_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
BH_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
BH_HD dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

// 2^(j/64), j = 0..63, as double-double (exact to ~2^-107)
static constexpr double kExp2Tab[64][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55},
    {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54},
    {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54},
    {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55},
    {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54},
    {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54},
    {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55},
    {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55},
    {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54},
    {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55},
    {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59},
    {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56},
    {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55},
    {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54},
    {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54},
    {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54},
    {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55},
    {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55},
    {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54},
    {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54},
    {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57},
    {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56},
    {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54},
    {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54},
    {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56},
    {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55},
    {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56},
    {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55},
    {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54},
    {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54},
    {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54},
    {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
};

// ln2/64 split in three parts; the first has 32 significant bits so that
// n*C1 is exact for |n| < 2^21.
constexpr double kInvLn2x64 = 0x1.71547652b82fep+6;
constexpr double kC1 = 0x1.62e42ff000000p-7;
constexpr double kC2 = -0x1.718432a1b0e26p-41;
constexpr double kC3 = -0x1.9ff0342542fc3p-96;
constexpr double kLn2Hi = 0x1.62e42ff000000p-1;
constexpr double kLn2Mid = -0x1.718432a1b0e26p-35;
constexpr double kLn2Lo = -0x1.9ff0342542fc3p-90;

// exp(x) = 2^m * y with y in [0.99, 1.99] as double-double; x finite and
// |x| < 746.
BH_HD dd exp_dd(double x, int& m) {
    double nd = rint(x * kInvLn2x64);
    int n = (int)nd;
    int j = n & 63;
    m = (n - j) / 64;
    dd r = two_sum(x, -(nd * kC1));
    r = dd_add(r, two_prod(nd, -kC2));
    r.lo += -(nd * kC3);
    r = fast_two_sum(r.hi, r.lo);
    // degree 6..11 Taylor terms in binary64
    double q = 0x1.ae64567f544e4p-26;      // 1/11!
    q = q * r.hi + 0x1.27e4fb7789f5cp-22;  // 1/10!
    q = q * r.hi + 0x1.71de3a556c734p-19;  // 1/9!
    q = q * r.hi + 0x1.a01a01a01a01ap-16;  // 1/8!
    q = q * r.hi + 0x1.a01a01a01a01ap-13;  // 1/7!
    q = q * r.hi + 0x1.6c16c16c16c17p-10;  // 1/6!
    // degree 0..5 in double-double (Horner)
    dd p = dd_mul_d(r, q);
    p = dd_add(p, dd{0x1.1111111111111p-7, 0x1.1111111111111p-63});  // 1/5!
    p = dd_mul(p, r);
    p = dd_add(p, dd{0x1.5555555555555p-5, 0x1.5555555555555p-59});  // 1/4!
    p = dd_mul(p, r);
    p = dd_add(p, dd{0x1.5555555555555p-3, 0x1.5555555555555p-57});  // 1/3!
    p = dd_mul(p, r);
    p = dd_add(p, dd{0.5, 0.0});
    p = dd_mul(p, r);
    p = dd_add(p, dd{1.0, 0.0});
    p = dd_mul(p, r);
    p = dd_add(p, dd{1.0, 0.0});
    return dd_mul(p, dd{kExp2Tab[j][0], kExp2Tab[j][1]});
}

// Round y*2^m to binary64 when the result is subnormal: scale into the
// integer grid of the subnormal ulp 2^-1074 and round once.
BH_HD double scale_subnormal(dd y, int m) {
    double t = ldexp(y.hi, m + 1074);
    double tl = ldexp(y.lo, m + 1074);
    double q = rint(t);
    double frac = (t - q) + tl;
    if (frac > 0.5) q += 1.0;
    else if (frac < -0.5) q -= 1.0;
    else if (frac == 0.5 || frac == -0.5) {
        double qq = q + (frac > 0 ? 1.0 : -1.0);
        if (fmod(qq, 2.0) == 0.0) q = qq;
    }
    return ldexp(q, -1074);
}

BH_HD double cr_exp(double x) {
    if (x != x) return x + x;
    if (x > 709.79) return HUGE_VAL;
    if (x < -745.14) return 0.0;
    int m;
    dd y = exp_dd(x, m);
    if (m < -1021) return scale_subnormal(y, m);
    return ldexp(y.hi, m);
}

BH_HD double cr_log(double x) {
    if (x != x) return x + x;
    if (x < 0.0) return (x - x) / (x - x);
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    int k = 0;
    if (x < 0x1p-1000) { x *= 0x1p200; k = -200; }
    else if (x > 0x1p1000) { x *= 0x1p-200; k = 200; }
    double y0 = log(x);
    int m;
    dd e = exp_dd(y0, m);
    double xs = ldexp(x, -m);
    double d = (xs - e.hi) - e.lo;
    double t = d / e.hi;
    double corr = t - 0.5 * t * t;
    if (k == 0) return y0 + corr;
    double kd = (double)k;
    dd s = two_sum(kd * kLn2Hi, y0);
    dd b = two_prod(kd, kLn2Mid);
    s = dd_add(s, b);
    s.lo += corr + kd * kLn2Lo;
    return s.hi + s.lo;
}

}  // namespace bh
