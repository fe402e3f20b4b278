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

#include <cstring>

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

// 2^(j/256), j = 0..255, as double-double, for the fast path of cr_exp
static constexpr double kExp2Tab256[256][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.00b1afa5abcbfp+0, -0x1.4f6b2a7609f71p-55},
    {0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54},
    {0x1.02168143b0281p+0, -0x1.2bf310fc54eb6p-55},
    {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.037d42e11bbccp+0, 0x1.56811eeade11ap-57},
    {0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54},
    {0x1.04e5f72f654b1p+0, 0x1.4c3793aa0d08dp-55},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55},
    {0x1.0650a0e3c1f89p+0, -0x1.5cb7b5799c397p-54},
    {0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55},
    {0x1.07bd42b72a836p+0, 0x1.3233454458700p-55},
    {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.092bdf66607e0p+0, -0x1.68063800a3fd1p-54},
    {0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54},
    {0x1.0a9c79b1f3919p+0, 0x1.5d16c873d1d38p-55},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54},
    {0x1.0c0f145e46c85p+0, 0x1.4f98906d21cefp-54},
    {0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54},
    {0x1.0d83b23395decp+0, -0x1.bc14de43f316ap-54},
    {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.0efa55fdfa9c5p+0, -0x1.49db9bc54021bp-54},
    {0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57},
    {0x1.1073028d7233ep+0, 0x1.d46eb1692fdd5p-55},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54},
    {0x1.11edbab5e2ab6p+0, -0x1.ca454f703fb72p-54},
    {0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58},
    {0x1.136a814f204abp+0, -0x1.7108fba48dcf0p-57},
    {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.14e95934f312ep+0, -0x1.b91e839bf44abp-55},
    {0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55},
    {0x1.166a45471c3c2p+0, 0x1.8f23b82ea1a32p-58},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55},
    {0x1.17ed48695bbc0p+0, 0x1.09e3fe2ac5a64p-56},
    {0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54},
    {0x1.1972658375d2fp+0, 0x1.4aadd85f17e08p-54},
    {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1af99f8138a1cp+0, 0x1.7bf85a4b69280p-54},
    {0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54},
    {0x1.1c82f95281c6bp+0, 0x1.009778010f8c9p-54},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54},
    {0x1.1e0e75eb44027p+0, -0x1.6fdd8088cb6dep-54},
    {0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54},
    {0x1.1f9c18438ce4dp+0, -0x1.bf524a097af5cp-54},
    {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.212be3578a819p+0, 0x1.3592d2cfcaac9p-54},
    {0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55},
    {0x1.22bdda27912d1p+0, 0x1.d34fb5577d69fp-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54},
    {0x1.2451ffb82140ap+0, 0x1.acfcc911ca996p-55},
    {0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55},
    {0x1.25e85711ece75p+0, 0x1.3e1a24ac31b2cp-54},
    {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.2780e341ddf29p+0, 0x1.e067c05f9e76cp-54},
    {0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54},
    {0x1.291ba7591bb70p+0, -0x1.2cc7228401cbdp-55},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55},
    {0x1.2ab8a66d10f13p+0, -0x1.95743191690a7p-54},
    {0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59},
    {0x1.2c57e39771b2fp+0, -0x1.50145a6eb5124p-54},
    {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.2df961f641589p+0, 0x1.d16cffbbce198p-54},
    {0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56},
    {0x1.2f9d24abd886bp+0, -0x1.53c55532bda93p-57},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55},
    {0x1.31432edeeb2fdp+0, 0x1.959a3f3f3fcd1p-55},
    {0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55},
    {0x1.32eb83ba8ea32p+0, -0x1.c45e83cb4f318p-54},
    {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.3496266e3fa2dp+0, -0x1.35a75930881a4p-55},
    {0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55},
    {0x1.36431a2de883bp+0, -0x1.c3144a06cb85ep-55},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54},
    {0x1.37f26231e754ap+0, -0x1.9f5ca9eceb23cp-54},
    {0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54},
    {0x1.39a401b7140efp+0, -0x1.9a9a5fc8e2934p-54},
    {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3b57fbfec6cf4p+0, 0x1.54c66e26fff18p-54},
    {0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54},
    {0x1.3d0e544ede173p+0, 0x1.fe8d08c284c71p-56},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55},
    {0x1.3ec70df1c5175p+0, -0x1.af6637b8c9bcap-55},
    {0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54},
    {0x1.40822c367a024p+0, 0x1.bddf8b6f4d048p-55},
    {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.423fb2709468ap+0, -0x1.8462dc0b314ddp-54},
    {0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55},
    {0x1.43ffa3f84b9d4p+0, 0x1.880be9704c003p-55},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59},
    {0x1.45c2042a7d232p+0, -0x1.8641982fb1f8ep-57},
    {0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54},
    {0x1.4786d668b3237p+0, -0x1.c20f0ed445733p-54},
    {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.494e1e192aed2p+0, -0x1.3b2895e499ea0p-55},
    {0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54},
    {0x1.4b17dea6db7d7p+0, -0x1.125b87f2897f0p-55},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56},
    {0x1.4ce41b817c114p+0, 0x1.05e29690abd5dp-54},
    {0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54},
    {0x1.4eb2d81d8abffp+0, -0x1.5257d2e5d7a52p-54},
    {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.508417f4531eep+0, 0x1.a249b49b7465fp-56},
    {0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55},
    {0x1.5257de83f4eefp+0, -0x1.c998d43efef71p-56},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55},
    {0x1.542e2f4f6ad27p+0, 0x1.7926d192d5f7ep-55},
    {0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55},
    {0x1.56070dde910d2p+0, -0x1.0fb6e168eebf0p-54},
    {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.57e27dbe2c4cfp+0, -0x1.0b98c8a57b9c4p-54},
    {0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55},
    {0x1.59c0827ff07ccp+0, -0x1.7e2cee467e60fp-54},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54},
    {0x1.5ba11fba87a03p+0, -0x1.b77a14c233e1ap-54},
    {0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60},
    {0x1.5d84590998b93p+0, -0x1.cd6a7a8b45643p-54},
    {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.5f6a320dceb71p+0, -0x1.9eadde3cdcf92p-55},
    {0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54},
    {0x1.6152ae6cdf6f4p+0, 0x1.e4b3e4ab84c27p-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54},
    {0x1.633dd1d1929fdp+0, 0x1.84710beb964e5p-54},
    {0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55},
    {0x1.652b9febc8fb7p+0, -0x1.ae3d5c9a73e09p-54},
    {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.671c1c70833f6p+0, -0x1.e8732586c6134p-55},
    {0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57},
    {0x1.690f4b19e9538p+0, 0x1.804bd9aeb445dp-55},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54},
    {0x1.6b052fa75173ep+0, 0x1.a38f52c9a9d0ep-56},
    {0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56},
    {0x1.6cfdcddd47645p+0, 0x1.c7aa9b6f17309p-54},
    {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.6ef9298593ae5p+0, -0x1.0b9749e1ac8b2p-54},
    {0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55},
    {0x1.70f7466f42e87p+0, 0x1.9d644d45aa65fp-58},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55},
    {0x1.72f8286ead08ap+0, -0x1.20aa02cd62c72p-54},
    {0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54},
    {0x1.74fbd35d7cbfdp+0, 0x1.047fd618a6e1cp-54},
    {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.77024b1ab6e09p+0, 0x1.b7877169147f8p-54},
    {0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54},
    {0x1.790b938ac1cf6p+0, 0x1.349a862aadd3ep-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55},
    {0x1.7b17b0976cfdbp+0, -0x1.bebb58468dc88p-54},
    {0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54},
    {0x1.7d26a62ff86f0p+0, 0x1.1bddbfb72b8b4p-54},
    {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.7f3878491c491p+0, -0x1.07f11cf9311aep-55},
    {0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54},
    {0x1.814d2add106d9p+0, 0x1.464370d151d4dp-54},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54},
    {0x1.8364c1eb941f7p+0, 0x1.99b9a31df2bd5p-54},
    {0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55},
    {0x1.857f4179f5b21p+0, -0x1.ba748f8b216d0p-58},
    {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.879cad931a436p+0, 0x1.5d2d7d2db47bdp-55},
    {0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54},
    {0x1.89bd0a478580fp+0, 0x1.d53954475202bp-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54},
    {0x1.8be05bad61778p+0, 0x1.ecb5efc43446ep-54},
    {0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56},
    {0x1.8e06a5e0866d9p+0, -0x1.7114a6fc9b2e6p-54},
    {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.902fed0282c8ap+0, 0x1.592ca85fe3fd2p-54},
    {0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54},
    {0x1.925c353aa2fe2p+0, -0x1.3455fa639db7fp-55},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57},
    {0x1.948b82b5f98e5p+0, -0x1.dc3d6797d2d99p-55},
    {0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56},
    {0x1.96bdd9a7670b3p+0, -0x1.ba5967f19c896p-58},
    {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.98f33e47a22a2p+0, 0x1.cabdaa24c78edp-56},
    {0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54},
    {0x1.9b2bb4d53fe0dp+0, -0x1.dd84e4df6d518p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56},
    {0x1.9d674194bb8d5p+0, -0x1.516bea3dd8233p-54},
    {0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56},
    {0x1.9fa5e8d07f29ep+0, -0x1.4a9ceaaf1facep-55},
    {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a1e7aed8eb8bbp+0, 0x1.c6618ee8be70ep-54},
    {0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54},
    {0x1.a42c980460ad8p+0, -0x1.aa780589fb120p-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54},
    {0x1.a674a8af46052p+0, 0x1.50f5630670366p-57},
    {0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54},
    {0x1.a8bfe53c12e59p+0, -0x1.4f867b2ba15a9p-54},
    {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ab0e521356ebap+0, 0x1.89c31dae94545p-55},
    {0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55},
    {0x1.ad5ff3a3c2774p+0, 0x1.7ef3bb6b1b8e5p-54},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54},
    {0x1.afb4ce622f2ffp+0, -0x1.4b2fc0f315ecdp-54},
    {0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54},
    {0x1.b20ce6c9a8952p+0, 0x1.4dd024a0756ccp-54},
    {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b468415b749b1p+0, -0x1.f763de9df7c90p-56},
    {0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54},
    {0x1.b6c6e29f1c52ap+0, 0x1.2a8f352883f6ep-54},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56},
    {0x1.b928cf22749e4p+0, -0x1.b721654cb65c6p-54},
    {0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54},
    {0x1.bb8e0b79a6f1fp+0, -0x1.f52d1c9696205p-60},
    {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.bdf69c3f3a207p+0, -0x1.c262360ea5b52p-60},
    {0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54},
    {0x1.c06286141b33dp+0, -0x1.d8a5aa1fbca34p-55},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55},
    {0x1.c2d1cd9fa652cp+0, -0x1.6e51617c8a5d7p-54},
    {0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54},
    {0x1.c544778fafb22p+0, 0x1.12f072493b5afp-54},
    {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.c7ba88988c933p+0, -0x1.e76bbbe255559p-55},
    {0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56},
    {0x1.ca3405751c4dbp+0, -0x1.7f2bed10d08f5p-55},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56},
    {0x1.ccb0f2e6d1675p+0, -0x1.d220f86009093p-56},
    {0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55},
    {0x1.cf3155b5bab74p+0, -0x1.a08e9b86dff57p-54},
    {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d1b532b08c968p+0, 0x1.55636219a36eep-54},
    {0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54},
    {0x1.d43c8eacaa1d6p+0, 0x1.3db53bf5a1614p-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55},
    {0x1.d6c76e862e6d3p+0, 0x1.fe87a4a8165a0p-58},
    {0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54},
    {0x1.d955d71ff6075p+0, 0x1.a052dbb9af6bep-54},
    {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dbe7cd63a8315p+0, -0x1.b76f1926b8be4p-54},
    {0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54},
    {0x1.de7d5641c0658p+0, -0x1.ca5528e79ba8fp-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54},
    {0x1.e11676b197d17p+0, -0x1.2b529bd5c7f44p-56},
    {0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54},
    {0x1.e3b333b16ee12p+0, -0x1.9f4a431fdc68bp-54},
    {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.e653924676d76p+0, -0x1.63ff87522b735p-55},
    {0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54},
    {0x1.e8f7977cdb740p+0, -0x1.1089480b054b1p-54},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54},
    {0x1.eb9f4867cca6ep+0, 0x1.4832f2293e4f2p-54},
    {0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54},
    {0x1.ee4aaa2188510p+0, 0x1.1c68da487568dp-54},
    {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f0f9c1cb6412ap+0, -0x1.3220065181d45p-54},
    {0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54},
    {0x1.f3ac948dd7274p+0, -0x1.95a5a3ed837dep-56},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54},
    {0x1.f6632798844f8p+0, 0x1.fa37b3539343ep-54},
    {0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54},
    {0x1.f91d802243c89p+0, -0x1.12ea8a779f689p-57},
    {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
    {0x1.fbdba3692d514p+0, -0x1.9677315098eb6p-56},
    {0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57},
    {0x1.fe9d96b2a23d9p+0, 0x1.4a6037442fde3p-56},
};
constexpr double kInvLn2x256 = 0x1.71547652b82fep+8;
constexpr double kD1 = 0x1.62e42ff000000p-9;   // ln2/256 in three parts,
constexpr double kD2 = -0x1.718432a1b0e26p-43;  // the first with 32 bits
constexpr double kD3 = -0x1.9ff0342542fc3p-98;

BH_HD double cr_exp(double x) {
    if (x != x) return x + x;
    if (x > 709.79) return HUGE_VAL;
    if (x < -745.14) return 0.0;
    // fast path: |r| <= ln2/512, exp(r) = 1 + r + r^2 q(r) with q in binary64,
    // total relative error below 2^-70; the result is returned only when it
    // is provably the correctly rounded one (Ziv's test).
    {
        double nd = rint(x * kInvLn2x256);
        int n = (int)nd;
        int j = n & 255;
        int m = (n - j) / 256;
        if (m >= -1021) {
            double rh = x - nd * kD1;  // exact (Sterbenz)
            dd t = two_prod(nd, -kD2);
            dd r = two_sum(rh, t.hi);
            r.lo += t.lo - nd * kD3;
            double q = 0x1.6c16c16c16c17p-10;   // 1/720
            q = q * r.hi + 0x1.1111111111111p-7;  // 1/120
            q = q * r.hi + 0x1.5555555555555p-5;  // 1/24
            q = q * r.hi + 0x1.5555555555555p-3;  // 1/6
            q = q * r.hi + 0.5;
            q = q * (r.hi * r.hi);
            dd p = fast_two_sum(1.0, r.hi);
            p.lo += r.lo + q;
            dd y = dd_mul(dd{kExp2Tab256[j][0], kExp2Tab256[j][1]}, p);
            const double E = 0x1p-66;  // error bound on y (y < 2)
            const double u = (y.hi >= 1.0) ? 0x1p-52 : 0x1p-53;
            const double half_below = (y.hi == 1.0) ? 0x1p-54 : 0.5 * u;
            bool safe = (y.lo >= 0.0) ? (y.lo + E < 0.5 * u) : (E - y.lo < half_below);
            if (safe) return ldexp(y.hi, m);
        }
    }
    int m;
    dd y = exp_dd(x, m);
    if (m < -1021) return scale_subnormal(y, m);
    return ldexp(y.hi, m);
}

// log(x) for |x - 1| < 2^-7: 2 atanh(s), s = (x-1)/(x+1), in double-double
// (relative error about 2^-104 however close x is to 1)
BH_HD double log_near1(double x) {
    double d = x - 1.0;  // exact
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
                            {0x1.1111111111111p-4, 0x1.1111111111111p-60}};  // 1/(2k+1)
    dd p = dd{c[7][0], c[7][1]};
    for (int k = 6; k >= 0; k--) p = dd_add(dd_mul(p, s2), dd{c[k][0], c[k][1]});
    dd y = dd_mul(s, p);
    return 2.0 * y.hi + 2.0 * y.lo;
}

// Fast path of cr_log.  x = 2^e m, m in [1, 2) (j = the top 8 bits of m's
// fraction); for j >= 128 m is halved and e raised, so that m is in
// [0.75, 1.5) and every result away from x = 1 exceeds 2^-9 in magnitude.
// kLogTab[j] = {c_j, -log(c_j) as double-double}: c_j = 1/m_j rounded to 9
// significant bits (m_j the bucket centre; c = 1 for the buckets at 1), so
// that r = m c_j - 1 is exact by one fma and |r| < 2^-8.  Then
// log x = e ln2 - log c_j + log1p(r), log1p(r) = r - r^2/2 + r^3 Q(r) with
// Q to r^7 (truncation below 2^-83 |r|), r^2 exact, the large terms summed
// in double-double: error below 2^-66 |log x|.  The result is returned only
// when the rounding test passes (else the slow path).
static constexpr double kLogTab[256][3] = {
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
    {0x1.fd00000000000p-1, 0x1.8121214586b54p-8, 0x1.c14b9f9377a1dp-65},
    {0x1.fb00000000000p-1, 0x1.41929f96832f0p-7, -0x1.c5517f64bc223p-61},
    {0x1.f900000000000p-1, 0x1.c317384c75f06p-7, 0x1.806208c04c220p-61},
    {0x1.f700000000000p-1, 0x1.228fb1fea2e28p-6, -0x1.cd7b66e01c26dp-61},
    {0x1.f500000000000p-1, 0x1.63d6178690bd6p-6, -0x1.8ed4d357c9c97p-64},
    {0x1.f300000000000p-1, 0x1.a55f548c5c43fp-6, 0x1.ec1a5f86d41f9p-62},
    {0x1.f100000000000p-1, 0x1.e72bf2813ce51p-6, 0x1.75b44595cab18p-60},
    {0x1.f000000000000p-1, 0x1.0415d89e74444p-5, 0x1.c05cf1d753622p-59},
    {0x1.ee00000000000p-1, 0x1.252f32f8d183fp-5, -0x1.947f792615916p-59},
    {0x1.ec00000000000p-1, 0x1.466aed42de3eap-5, -0x1.cdd6f7f4a137ep-59},
    {0x1.ea00000000000p-1, 0x1.67c94f2d4bb58p-5, 0x1.0413e6505e603p-59},
    {0x1.e800000000000p-1, 0x1.894aa149fb343p-5, 0x1.a8be97660a23dp-60},
    {0x1.e600000000000p-1, 0x1.aaef2d0fb10fcp-5, 0x1.a353bb42e0addp-61},
    {0x1.e500000000000p-1, 0x1.bbcebfc68f420p-5, 0x1.e5cf3a0f56f72p-60},
    {0x1.e300000000000p-1, 0x1.dda8adc67ee4ep-5, 0x1.4e6c986f44c55p-59},
    {0x1.e100000000000p-1, 0x1.ffa6911ab9301p-5, -0x1.cd9f1f95c2eedp-59},
    {0x1.df00000000000p-1, 0x1.10e45b3cae831p-4, -0x1.a4a128d192686p-58},
    {0x1.dd00000000000p-1, 0x1.2207b5c78549ep-4, -0x1.cc0fbce104eaap-58},
    {0x1.dc00000000000p-1, 0x1.2aa04a44717a5p-4, -0x1.d15d38d2fa3f7p-58},
    {0x1.da00000000000p-1, 0x1.3bdf5a7d1ee64p-4, 0x1.7a976d3b5b45fp-59},
    {0x1.d800000000000p-1, 0x1.4d3115d207eacp-4, 0x1.769f42c7842ccp-58},
    {0x1.d700000000000p-1, 0x1.55e10050e0384p-4, -0x1.45f9d61c68c1bp-58},
    {0x1.d500000000000p-1, 0x1.674f089365a7ap-4, -0x1.9acd8b33f8fdcp-58},
    {0x1.d300000000000p-1, 0x1.78d02263d82d3p-4, 0x1.abca5b4fdb880p-58},
    {0x1.d200000000000p-1, 0x1.8197e2f40e3f0p-4, 0x1.b9f2dffbeed43p-60},
    {0x1.d000000000000p-1, 0x1.9335e5d594989p-4, -0x1.478a85704ccb7p-58},
    {0x1.ce00000000000p-1, 0x1.a4e7640b1bc38p-4, -0x1.5b5ca203e4259p-58},
    {0x1.cd00000000000p-1, 0x1.adc77ee5aea8cp-4, 0x1.37d8f39bee659p-58},
    {0x1.cb00000000000p-1, 0x1.bf968769fca11p-4, -0x1.cdc9f6f5f38c7p-59},
    {0x1.c900000000000p-1, 0x1.d179788219364p-4, 0x1.9daf7df76ad2ap-59},
    {0x1.c800000000000p-1, 0x1.da727638446a2p-4, 0x1.401fa71733019p-58},
    {0x1.c600000000000p-1, 0x1.ec739830a1120p-4, -0x1.a2bf991780d3fp-59},
    {0x1.c500000000000p-1, 0x1.f57bc7d9005dbp-4, -0x1.9361574fb24e2p-58},
    {0x1.c300000000000p-1, 0x1.03cdc0a51ec0dp-3, 0x1.39e2d3f8b7d10p-57},
    {0x1.c200000000000p-1, 0x1.08598b59e3a07p-3, -0x1.dd7009902bf32p-57},
    {0x1.c000000000000p-1, 0x1.1178e8227e47cp-3, -0x1.0e63a5f01c691p-58},
    {0x1.bf00000000000p-1, 0x1.160c8024b27b1p-3, -0x1.2d56ff61c2bfbp-57},
    {0x1.bd00000000000p-1, 0x1.1f3b925f25d41p-3, 0x1.62c9ef939ac5dp-59},
    {0x1.bc00000000000p-1, 0x1.23d712a49c202p-3, -0x1.6e38161051d69p-57},
    {0x1.ba00000000000p-1, 0x1.2d1610c86813ap-3, -0x1.499a3f25af95fp-58},
    {0x1.b900000000000p-1, 0x1.31b994d3a4f85p-3, -0x1.c4716bdfc0cc9p-58},
    {0x1.b700000000000p-1, 0x1.3b08b6757f2a9p-3, 0x1.70d6cdf05266cp-60},
    {0x1.b600000000000p-1, 0x1.3fb45a59928ccp-3, -0x1.d87e6a354d056p-57},
    {0x1.b400000000000p-1, 0x1.4913d8333b561p-3, -0x1.0d5604930f135p-58},
    {0x1.b300000000000p-1, 0x1.4dc7b897bc1c8p-3, -0x1.927d47803c5f4p-57},
    {0x1.b100000000000p-1, 0x1.5737cc9018cddp-3, 0x1.4f4d710fec38ep-57},
    {0x1.b000000000000p-1, 0x1.5bf406b543db2p-3, -0x1.1f5b44c0df7e7p-61},
    {0x1.ae00000000000p-1, 0x1.6574ebe8c133ap-3, -0x1.d34f0f4621bedp-60},
    {0x1.ad00000000000p-1, 0x1.6a399dabbd383p-3, 0x1.96332bd4b341fp-57},
    {0x1.ac00000000000p-1, 0x1.6f0128b756abcp-3, -0x1.8de59c21e166cp-57},
    {0x1.aa00000000000p-1, 0x1.7898d85444c73p-3, 0x1.ef8f6ebcfb201p-58},
    {0x1.a900000000000p-1, 0x1.7d6903caf5ad0p-3, -0x1.ac5f0c075b847p-59},
    {0x1.a700000000000p-1, 0x1.871213750e994p-3, 0x1.d685f35eea2a0p-57},
    {0x1.a600000000000p-1, 0x1.8beafeb38fe8cp-3, 0x1.55aa8b6997a40p-58},
    {0x1.a500000000000p-1, 0x1.90c6db9fcbcd9p-3, 0x1.054473941ad99p-57},
    {0x1.a300000000000p-1, 0x1.9a8778debaa38p-3, 0x1.f47dfd871f87fp-57},
    {0x1.a200000000000p-1, 0x1.9f6c407089664p-3, 0x1.35a19605e67efp-59},
    {0x1.a100000000000p-1, 0x1.a454082e6ab05p-3, 0x1.df207dc5c34c6p-58},
    {0x1.9f00000000000p-1, 0x1.ae2ca6f672bd4p-3, 0x1.ab5ca9eaa088ap-57},
    {0x1.9e00000000000p-1, 0x1.b31d8575bce3dp-3, -0x1.6353ab386a94dp-57},
    {0x1.9d00000000000p-1, 0x1.b811730b823d2p-3, 0x1.a0ee735d9f0ecp-60},
    {0x1.9c00000000000p-1, 0x1.bd087383bd8adp-3, 0x1.dd355f6a516d7p-60},
    {0x1.9a00000000000p-1, 0x1.c6ffbc6f00f71p-3, -0x1.8e58b2c57a4a5p-57},
    {0x1.9900000000000p-1, 0x1.cc000c9db3c52p-3, 0x1.53d154280394fp-57},
    {0x1.9800000000000p-1, 0x1.d1037f2655e7bp-3, 0x1.60629242471a2p-57},
    {0x1.9600000000000p-1, 0x1.db13db0d48940p-3, 0x1.aa11d49f96cb9p-58},
    {0x1.9500000000000p-1, 0x1.e020cc6235ab5p-3, 0x1.fea48dd7b81d1p-58},
    {0x1.9400000000000p-1, 0x1.e530effe71012p-3, 0x1.2276041f43042p-59},
    {0x1.9300000000000p-1, 0x1.ea4449f04aaf5p-3, -0x1.d33919ab94074p-57},
    {0x1.9100000000000p-1, 0x1.f474b134df229p-3, -0x1.27c77ded76aadp-58},
    {0x1.9000000000000p-1, 0x1.f991c6cb3b379p-3, 0x1.f665066f980a2p-57},
    {0x1.8f00000000000p-1, 0x1.feb2233ea07cdp-3, 0x1.8de00938b4c40p-61},
    {0x1.8e00000000000p-1, 0x1.01eae5626c691p-2, -0x1.18290bd2932e2p-59},
    {0x1.8d00000000000p-1, 0x1.047e60cde83b8p-2, -0x1.0779634061cbcp-56},
    {0x1.8b00000000000p-1, 0x1.09aa572e6c6d4p-2, 0x1.43c2e68684d53p-57},
    {0x1.8a00000000000p-1, 0x1.0c42d676162e3p-2, 0x1.162c79d5d11eep-58},
    {0x1.8900000000000p-1, 0x1.0edd060b78081p-2, -0x1.92b49ef282b09p-57},
    {0x1.8800000000000p-1, 0x1.1178e8227e47cp-2, -0x1.0e63a5f01c691p-57},
    {0x1.8700000000000p-1, 0x1.14167ef367783p-2, 0x1.e0936abd4fa6ep-62},
    {0x1.8600000000000p-1, 0x1.16b5ccbacfb73p-2, 0x1.66fbd28b40935p-56},
    {0x1.8400000000000p-1, 0x1.1bf99635a6b95p-2, -0x1.12aeb84249223p-57},
    {0x1.8300000000000p-1, 0x1.1e9e1678899f4p-2, 0x1.512c3749a1e4ep-56},
    {0x1.8200000000000p-1, 0x1.214456d0eb8d4p-2, 0x1.f7ae91aeba60ap-57},
    {0x1.8100000000000p-1, 0x1.23ec5991eba49p-2, 0x1.bb75d1addf870p-60},
    {0x1.8000000000000p-1, 0x1.269621134db92p-2, 0x1.e0efadd9db02bp-56},
    {0x1.7f00000000000p-1, 0x1.2941afb186b7cp-2, -0x1.856e61c515740p-57},
    {0x1.7e00000000000p-1, 0x1.2bef07cdc9354p-2, -0x1.82dad7fd86088p-56},
    {0x1.7c00000000000p-1, 0x1.314f1e1d35ce4p-2, -0x1.3d69909e5c3dcp-56},
    {0x1.7b00000000000p-1, 0x1.3401e12aecba1p-2, -0x1.cd55b8a4746c0p-58},
    {0x1.7a00000000000p-1, 0x1.36b6776be1117p-2, -0x1.324f0e883858ep-58},
    {0x1.7900000000000p-1, 0x1.396ce359bbf54p-2, -0x1.ce2b31b31e8b0p-58},
    {0x1.7800000000000p-1, 0x1.3c25277333184p-2, -0x1.2ad27e50a8ec6p-56},
    {0x1.7700000000000p-1, 0x1.3edf463c1683ep-2, 0x1.83d680d3c1084p-56},
    {0x1.7600000000000p-1, 0x1.419b423d5e8c7p-2, 0x1.0dbb243827392p-57},
    {0x1.7500000000000p-1, 0x1.44591e0539f49p-2, -0x1.2b125247b0fa5p-56},
    {0x1.7400000000000p-1, 0x1.4718dc271c41bp-2, 0x1.8fb4c14c56eefp-60},
    {0x1.7300000000000p-1, 0x1.49da7f3bcc41fp-2, -0x1.9964a168ccacap-57},
    {0x1.7200000000000p-1, 0x1.4c9e09e172c3cp-2, -0x1.123615b147a5dp-58},
    {0x1.7100000000000p-1, 0x1.4f637ebba9810p-2, -0x1.58cb3124b9245p-56},
    {0x1.7000000000000p-1, 0x1.522ae0738a3d8p-2, -0x1.8f7e9b38a6979p-57},
    {0x1.6f00000000000p-1, 0x1.54f431b7be1a9p-2, -0x1.aacfdbbdab914p-56},
    {0x1.6e00000000000p-1, 0x1.57bf753c8d1fbp-2, -0x1.0908d15f88b63p-57},
    {0x1.6d00000000000p-1, 0x1.5a8cadbbedfa1p-2, -0x1.e6c2bdfb3e037p-58},
    {0x1.6c00000000000p-1, 0x1.5d5bddf595f30p-2, -0x1.6541148cbb8a2p-56},
    {0x1.6b00000000000p-1, 0x1.602d08af091ecp-2, -0x1.6e8920c09b73fp-58},
    {0x1.6a00000000000p-1, 0x1.630030b3aac49p-2, 0x1.dc18ce51fff99p-57},
    {0x1.6900000000000p-1, 0x1.65d558d4ce00bp-2, -0x1.7605a4748480ap-56},
    {0x1.6800000000000p-1, 0x1.68ac83e9c6a14p-2, 0x1.a64eadd740178p-58},
    {0x1.6700000000000p-1, 0x1.6b85b4cffa3fdp-2, -0x1.8af2c8dafcb08p-57},
    {0x1.6600000000000p-1, 0x1.6e60ee6af1972p-2, 0x1.657c222d868cdp-58},
    {0x1.6500000000000p-1, 0x1.713e33a46a17cp-2, -0x1.9367a05ae38d3p-56},
    {0x1.6400000000000p-1, 0x1.741d876c67bb1p-2, 0x1.84a4ee3059583p-56},
    {0x1.6300000000000p-1, 0x1.76feecb947175p-2, -0x1.118d9eb4ea362p-56},
    {0x1.6200000000000p-1, 0x1.79e26687cfb3ep-2, -0x1.c168817443f22p-56},
    {0x1.6100000000000p-1, 0x1.7cc7f7db46a0ep-2, -0x1.8438023cdc3d3p-56},
    {0x1.6000000000000p-1, 0x1.7fafa3bd8151cp-2, -0x1.219024acd3b77p-58},
    {0x1.5f00000000000p-1, 0x1.82996d3ef8bcbp-2, -0x1.2aa30536bb6bep-56},
    {0x1.5e00000000000p-1, 0x1.85855776dcbfbp-2, -0x1.486666443b153p-56},
    {0x1.5d00000000000p-1, 0x1.8873658327ccfp-2, -0x1.c43eff892d8d0p-58},
    {0x1.5c00000000000p-1, 0x1.8b639a88b2df5p-2, -0x1.70f2f38238303p-56},
    {0x1.5b00000000000p-1, 0x1.8e55f9b349b83p-2, 0x1.e2763763baffcp-56},
    {0x1.5a00000000000p-1, 0x1.914a8635bf68ap-2, -0x1.ad4bb98c1f2c5p-56},
    {0x1.5900000000000p-1, 0x1.9441434a03259p-2, 0x1.c0e4afffa1b77p-56},
    {0x1.5800000000000p-1, 0x1.973a3431356aep-2, -0x1.89d2816cf838fp-57},
    {0x1.5800000000000p-1, 0x1.973a3431356aep-2, -0x1.89d2816cf838fp-57},
    {0x1.5700000000000p-1, 0x1.9a355c33bd6bap-2, -0x1.959578e82a9d9p-57},
    {0x1.5600000000000p-1, 0x1.9d32bea15ed3bp-2, 0x1.87bcbcfd3e187p-59},
    {0x1.5500000000000p+0, -0x1.2596010df763ap-2, 0x1.0f76c57075e9ep-58},
    {0x1.5400000000000p+0, -0x1.22941fbcf7966p-2, 0x1.76f5eb09628afp-56},
    {0x1.5300000000000p+0, -0x1.1f8ff9e48a2f3p-2, 0x1.c9fdf9a0c4b07p-56},
    {0x1.5200000000000p+0, -0x1.1c898c16999fbp-2, 0x1.0e5c62aff1c44p-60},
    {0x1.5100000000000p+0, -0x1.1980d2dd4236fp-2, -0x1.9d3d1b0e4d147p-56},
    {0x1.5100000000000p+0, -0x1.1980d2dd4236fp-2, -0x1.9d3d1b0e4d147p-56},
    {0x1.5000000000000p+0, -0x1.1675cababa60ep-2, -0x1.ce63eab883717p-61},
    {0x1.4f00000000000p+0, -0x1.136870293a8b0p-2, -0x1.7b66298edd24ap-56},
    {0x1.4e00000000000p+0, -0x1.1058bf9ae4ad5p-2, -0x1.89fa0ab4cb31dp-58},
    {0x1.4d00000000000p+0, -0x1.0d46b579ab74bp-2, -0x1.03ec81c3cbd92p-57},
    {0x1.4c00000000000p+0, -0x1.0a324e27390e3p-2, -0x1.7dcfde8061c03p-56},
    {0x1.4b00000000000p+0, -0x1.071b85fcd590dp-2, -0x1.d1707f97bde80p-58},
    {0x1.4b00000000000p+0, -0x1.071b85fcd590dp-2, -0x1.d1707f97bde80p-58},
    {0x1.4a00000000000p+0, -0x1.0402594b4d041p-2, 0x1.28ec217a5022dp-57},
    {0x1.4900000000000p+0, -0x1.00e6c45ad501dp-2, 0x1.cb9568ff6feadp-57},
    {0x1.4800000000000p+0, -0x1.fb9186d5e3e2bp-3, 0x1.caaae64f21acbp-57},
    {0x1.4700000000000p+0, -0x1.f550a564b7b37p-3, -0x1.c5f6dfd018c37p-61},
    {0x1.4600000000000p+0, -0x1.ef0adcbdc5936p-3, -0x1.48637950dc20dp-57},
    {0x1.4600000000000p+0, -0x1.ef0adcbdc5936p-3, -0x1.48637950dc20dp-57},
    {0x1.4500000000000p+0, -0x1.e8c0252aa5a60p-3, 0x1.6e03a39bfc89bp-59},
    {0x1.4400000000000p+0, -0x1.e27076e2af2e6p-3, 0x1.61578001e0162p-59},
    {0x1.4300000000000p+0, -0x1.dc1bca0abec7dp-3, -0x1.834c51998b6fcp-57},
    {0x1.4200000000000p+0, -0x1.d5c216b4fbb91p-3, -0x1.6e443597e4d40p-57},
    {0x1.4200000000000p+0, -0x1.d5c216b4fbb91p-3, -0x1.6e443597e4d40p-57},
    {0x1.4100000000000p+0, -0x1.cf6354e09c5dcp-3, -0x1.239a07d55b695p-57},
    {0x1.4000000000000p+0, -0x1.c8ff7c79a9a22p-3, 0x1.4f689f8434012p-57},
    {0x1.3f00000000000p+0, -0x1.c2968558c18c1p-3, 0x1.73dee38a3fb6bp-57},
    {0x1.3f00000000000p+0, -0x1.c2968558c18c1p-3, 0x1.73dee38a3fb6bp-57},
    {0x1.3e00000000000p+0, -0x1.bc286742d8cd6p-3, -0x1.4fce744870f55p-58},
    {0x1.3d00000000000p+0, -0x1.b5b519e8fb5a4p-3, -0x1.ba27fdc19e1a0p-57},
    {0x1.3c00000000000p+0, -0x1.af3c94e80bff3p-3, 0x1.398cff3641985p-58},
    {0x1.3b00000000000p+0, -0x1.a8becfc882f19p-3, 0x1.e8c37918c39ebp-58},
    {0x1.3b00000000000p+0, -0x1.a8becfc882f19p-3, 0x1.e8c37918c39ebp-58},
    {0x1.3a00000000000p+0, -0x1.a23bc1fe2b563p-3, -0x1.93711b07a998cp-59},
    {0x1.3900000000000p+0, -0x1.9bb362e7dfb83p-3, -0x1.575e31f003e0cp-57},
    {0x1.3800000000000p+0, -0x1.9525a9cf456b4p-3, -0x1.d904c1d4e2e26p-57},
    {0x1.3800000000000p+0, -0x1.9525a9cf456b4p-3, -0x1.d904c1d4e2e26p-57},
    {0x1.3700000000000p+0, -0x1.8e928de886d41p-3, 0x1.569d851a56770p-57},
    {0x1.3600000000000p+0, -0x1.87fa06520c911p-3, 0x1.bf7fdbfa08d9ap-57},
    {0x1.3500000000000p+0, -0x1.815c0a14357ebp-3, 0x1.4be48073a0564p-58},
    {0x1.3500000000000p+0, -0x1.815c0a14357ebp-3, 0x1.4be48073a0564p-58},
    {0x1.3400000000000p+0, -0x1.7ab890210d909p-3, -0x1.be36b2d6a0608p-59},
    {0x1.3300000000000p+0, -0x1.740f8f54037a5p-3, 0x1.b264062a84cdbp-58},
    {0x1.3300000000000p+0, -0x1.740f8f54037a5p-3, 0x1.b264062a84cdbp-58},
    {0x1.3200000000000p+0, -0x1.6d60fe719d21dp-3, 0x1.caae268ecd179p-57},
    {0x1.3100000000000p+0, -0x1.66acd4272ad51p-3, 0x1.0900e4e1ea8b2p-58},
    {0x1.3000000000000p+0, -0x1.5ff3070a793d4p-3, 0x1.bc60efafc6f6ep-58},
    {0x1.3000000000000p+0, -0x1.5ff3070a793d4p-3, 0x1.bc60efafc6f6ep-58},
    {0x1.2f00000000000p+0, -0x1.59338d9982086p-3, 0x1.65d22aa8ad7cfp-58},
    {0x1.2e00000000000p+0, -0x1.526e5e3a1b438p-3, 0x1.746ff8a470d3ap-57},
    {0x1.2e00000000000p+0, -0x1.526e5e3a1b438p-3, 0x1.746ff8a470d3ap-57},
    {0x1.2d00000000000p+0, -0x1.4ba36f39a55e5p-3, -0x1.68981bcc36756p-57},
    {0x1.2c00000000000p+0, -0x1.44d2b6ccb7d1ep-3, -0x1.9f4f6543e1f88p-57},
    {0x1.2c00000000000p+0, -0x1.44d2b6ccb7d1ep-3, -0x1.9f4f6543e1f88p-57},
    {0x1.2b00000000000p+0, -0x1.3dfc2b0ecc62ap-3, 0x1.ab3a8e7d81017p-58},
    {0x1.2a00000000000p+0, -0x1.371fc201e8f74p-3, -0x1.de6cb62af18a0p-58},
    {0x1.2a00000000000p+0, -0x1.371fc201e8f74p-3, -0x1.de6cb62af18a0p-58},
    {0x1.2900000000000p+0, -0x1.303d718e47fd3p-3, 0x1.6b9c7d96091fap-63},
    {0x1.2800000000000p+0, -0x1.29552f81ff523p-3, -0x1.301771c407dbfp-57},
    {0x1.2800000000000p+0, -0x1.29552f81ff523p-3, -0x1.301771c407dbfp-57},
    {0x1.2700000000000p+0, -0x1.2266f190a5acbp-3, -0x1.f547bf1809e88p-57},
    {0x1.2600000000000p+0, -0x1.1b72ad52f67a0p-3, -0x1.483023472cd74p-58},
    {0x1.2600000000000p+0, -0x1.1b72ad52f67a0p-3, -0x1.483023472cd74p-58},
    {0x1.2500000000000p+0, -0x1.14785846742acp-3, -0x1.a28813e3a7f07p-57},
    {0x1.2400000000000p+0, -0x1.0d77e7cd08e59p-3, -0x1.9a5dc5e9030acp-57},
    {0x1.2400000000000p+0, -0x1.0d77e7cd08e59p-3, -0x1.9a5dc5e9030acp-57},
    {0x1.2300000000000p+0, -0x1.0671512ca596ep-3, -0x1.50c647eb86499p-58},
    {0x1.2200000000000p+0, -0x1.fec9131dbeabbp-4, 0x1.5746b9981b36cp-58},
    {0x1.2200000000000p+0, -0x1.fec9131dbeabbp-4, 0x1.5746b9981b36cp-58},
    {0x1.2100000000000p+0, -0x1.f0a30c01162a6p-4, -0x1.85f325c5bbacdp-58},
    {0x1.2000000000000p+0, -0x1.e27076e2af2e6p-4, 0x1.61578001e0162p-60},
    {0x1.2000000000000p+0, -0x1.e27076e2af2e6p-4, 0x1.61578001e0162p-60},
    {0x1.1f00000000000p+0, -0x1.d4313d66cb35dp-4, -0x1.790dd951d90fap-58},
    {0x1.1e00000000000p+0, -0x1.c5e548f5bc743p-4, -0x1.5d617ef8161b1p-60},
    {0x1.1e00000000000p+0, -0x1.c5e548f5bc743p-4, -0x1.5d617ef8161b1p-60},
    {0x1.1d00000000000p+0, -0x1.b78c82bb0eda1p-4, -0x1.0878cf0327e21p-61},
    {0x1.1d00000000000p+0, -0x1.b78c82bb0eda1p-4, -0x1.0878cf0327e21p-61},
    {0x1.1c00000000000p+0, -0x1.a926d3a4ad563p-4, -0x1.942f48aa70ea9p-58},
    {0x1.1b00000000000p+0, -0x1.9ab42462033adp-4, 0x1.2099e1c184e8ep-59},
    {0x1.1b00000000000p+0, -0x1.9ab42462033adp-4, 0x1.2099e1c184e8ep-59},
    {0x1.1a00000000000p+0, -0x1.8c345d6319b21p-4, 0x1.4a697ab3424a9p-61},
    {0x1.1a00000000000p+0, -0x1.8c345d6319b21p-4, 0x1.4a697ab3424a9p-61},
    {0x1.1900000000000p+0, -0x1.7da766d7b12cdp-4, 0x1.eeedfcdd94131p-58},
    {0x1.1800000000000p+0, -0x1.6f0d28ae56b4cp-4, 0x1.906d99184b992p-58},
    {0x1.1800000000000p+0, -0x1.6f0d28ae56b4cp-4, 0x1.906d99184b992p-58},
    {0x1.1700000000000p+0, -0x1.60658a93750c4p-4, 0x1.388458ec21b6ap-58},
    {0x1.1700000000000p+0, -0x1.60658a93750c4p-4, 0x1.388458ec21b6ap-58},
    {0x1.1600000000000p+0, -0x1.51b073f06183fp-4, -0x1.a49e39a1a8be4p-58},
    {0x1.1500000000000p+0, -0x1.42edcbea646f0p-4, -0x1.ddd4f935996c9p-59},
    {0x1.1500000000000p+0, -0x1.42edcbea646f0p-4, -0x1.ddd4f935996c9p-59},
    {0x1.1400000000000p+0, -0x1.341d7961bd1d1p-4, 0x1.b599f227becbbp-58},
    {0x1.1400000000000p+0, -0x1.341d7961bd1d1p-4, 0x1.b599f227becbbp-58},
    {0x1.1300000000000p+0, -0x1.253f62f0a1417p-4, 0x1.c125963fc4cfdp-62},
    {0x1.1200000000000p+0, -0x1.16536eea37ae1p-4, 0x1.79da3e8c22cdap-60},
    {0x1.1200000000000p+0, -0x1.16536eea37ae1p-4, 0x1.79da3e8c22cdap-60},
    {0x1.1100000000000p+0, -0x1.075983598e471p-4, -0x1.80da5333c45b8p-59},
    {0x1.1100000000000p+0, -0x1.075983598e471p-4, -0x1.80da5333c45b8p-59},
    {0x1.1000000000000p+0, -0x1.f0a30c01162a6p-5, -0x1.85f325c5bbacdp-59},
    {0x1.1000000000000p+0, -0x1.f0a30c01162a6p-5, -0x1.85f325c5bbacdp-59},
    {0x1.0f00000000000p+0, -0x1.d276b8adb0b52p-5, -0x1.1e3c53257fd47p-61},
    {0x1.0f00000000000p+0, -0x1.d276b8adb0b52p-5, -0x1.1e3c53257fd47p-61},
    {0x1.0e00000000000p+0, -0x1.b42dd711971bfp-5, 0x1.eb9759c130499p-60},
    {0x1.0d00000000000p+0, -0x1.95c830ec8e3ebp-5, -0x1.f5a0e80520bf2p-59},
    {0x1.0d00000000000p+0, -0x1.95c830ec8e3ebp-5, -0x1.f5a0e80520bf2p-59},
    {0x1.0c00000000000p+0, -0x1.77458f632dcfcp-5, -0x1.18d3ca87b9296p-59},
    {0x1.0c00000000000p+0, -0x1.77458f632dcfcp-5, -0x1.18d3ca87b9296p-59},
    {0x1.0b00000000000p+0, -0x1.58a5bafc8e4d5p-5, 0x1.ce55c2b4e2b72p-59},
    {0x1.0b00000000000p+0, -0x1.58a5bafc8e4d5p-5, 0x1.ce55c2b4e2b72p-59},
    {0x1.0a00000000000p+0, -0x1.39e87b9febd60p-5, 0x1.5bfa937f551bbp-59},
    {0x1.0a00000000000p+0, -0x1.39e87b9febd60p-5, 0x1.5bfa937f551bbp-59},
    {0x1.0900000000000p+0, -0x1.1b0d98923d980p-5, 0x1.e9ae889bac481p-60},
    {0x1.0900000000000p+0, -0x1.1b0d98923d980p-5, 0x1.e9ae889bac481p-60},
    {0x1.0800000000000p+0, -0x1.f829b0e783300p-6, -0x1.33e3f04f1ef23p-60},
    {0x1.0700000000000p+0, -0x1.b9fc027af9198p-6, 0x1.0ae69229dc868p-64},
    {0x1.0700000000000p+0, -0x1.b9fc027af9198p-6, 0x1.0ae69229dc868p-64},
    {0x1.0600000000000p+0, -0x1.7b91b07d5b11bp-6, 0x1.5b602ace3a510p-60},
    {0x1.0600000000000p+0, -0x1.7b91b07d5b11bp-6, 0x1.5b602ace3a510p-60},
    {0x1.0500000000000p+0, -0x1.3cea44346a575p-6, 0x1.0cb5a902b3a1cp-62},
    {0x1.0500000000000p+0, -0x1.3cea44346a575p-6, 0x1.0cb5a902b3a1cp-62},
    {0x1.0400000000000p+0, -0x1.fc0a8b0fc03e4p-7, 0x1.83092c59642a1p-62},
    {0x1.0400000000000p+0, -0x1.fc0a8b0fc03e4p-7, 0x1.83092c59642a1p-62},
    {0x1.0300000000000p+0, -0x1.7dc475f810a77p-7, 0x1.16d7687d3df21p-62},
    {0x1.0300000000000p+0, -0x1.7dc475f810a77p-7, 0x1.16d7687d3df21p-62},
    {0x1.0200000000000p+0, -0x1.fe02a6b106789p-8, 0x1.e44b7e3711ebfp-67},
    {0x1.0200000000000p+0, -0x1.fe02a6b106789p-8, 0x1.e44b7e3711ebfp-67},
    {0x1.0100000000000p+0, -0x1.ff00aa2b10bc0p-9, -0x1.2821ad5a6d353p-63},
    {0x1.0100000000000p+0, -0x1.ff00aa2b10bc0p-9, -0x1.2821ad5a6d353p-63},
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
};
BH_HD double cr_log_fast(double x, bool& ok) {
    ok = false;
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    int e = (int)(u >> 52) - 1023;
    const int j = (int)((u >> 44) & 0xff);
    const uint64_t mu = (u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    std::memcpy(&m, &mu, sizeof m);
    if (j >= 128) { m *= 0.5; e += 1; }
    const double c = kLogTab[j][0];
    const double r = fma(m, c, -1.0);
    double q = -0x1.999999999999ap-4;        // -1/10
    q = q * r + 0x1.c71c71c71c71cp-4;        // 1/9
    q = q * r - 0x1.0000000000000p-3;        // -1/8
    q = q * r + 0x1.2492492492492p-3;        // 1/7
    q = q * r - 0x1.5555555555555p-3;        // -1/6
    q = q * r + 0x1.999999999999ap-3;        // 1/5
    q = q * r - 0x1.0000000000000p-2;        // -1/4
    q = q * r + 0x1.5555555555555p-2;        // 1/3
    const dd r2 = two_prod(r, r);
    const double tail = (r2.hi * r) * q;
    const double ed = (double)e;
    dd acc = two_sum(ed * kLn2Hi, kLogTab[j][1]);
    acc = dd_add(acc, two_prod(ed, kLn2Mid));
    acc = dd_add(acc, dd{r, 0.0});
    acc = dd_add(acc, dd{-0.5 * r2.hi, -0.5 * r2.lo});
    acc.lo += tail + (kLogTab[j][2] + ed * kLn2Lo);
    const dd y = fast_two_sum(acc.hi, acc.lo);
    const double E = 0x1p-63 * fabs(y.hi);
    const double up = y.hi + (y.lo + E), dn = y.hi + (y.lo - E);
    if (up != dn) return 0.0;
    ok = true;
    return up;
}

BH_HD double cr_log(double x) {
    if (x != x) return x + x;
    if (x < 0.0) return (x - x) / (x - x);
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    if (x >= 0x1p-1022) {
        bool ok;
        const double y = cr_log_fast(x, ok);
        if (ok) return y;
    }
    if (fabs(x - 1.0) < 0x1p-7) return log_near1(x);
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
