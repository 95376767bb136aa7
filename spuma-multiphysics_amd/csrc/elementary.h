// SPDX-License-Identifier: GPL-3.0-or-later
// Elementary functions with one definition for the device and the host.
//
// Every function is a fixed sequence of IEEE-754 double operations (add,
// multiply, divide, square root, fused multiply-add through an explicit
// fma call, exact exponent scaling), so gfx950 and an x86-64 host compiled
// with -ffp-contract=off produce the same bits for the same argument.  The
// flash kernels use them in place of the device math library, and the
// oracle's twin build uses them in place of glibc: the device flash can then
// be compared with its CPU restatement bit for bit.
//
// log, exp, acos and cos are evaluated to about 2^-68 relative before their
// one final rounding, so they return the correctly rounded value except on
// arguments within ~2^-15 ulp of a rounding boundary.  glibc's versions of
// these four functions are correctly rounded on all but a few in 10^4 of the
// flash's arguments, so the two libraries agree wherever glibc does.
// Methods (table-driven reductions with double-double tails; the tables,
// elementary_tables.h, are 200-bit values rounded to nearest):
//   exp   x = (64 m + j) ln2/64 + r, |r| <= ln2/128 (three-part ln2/64, r kept
//         as a pair); e^x = 2^m 2^(j/64) e^r, e^r - 1 = r + r^2/2 + r^3 Q(r)
//         with r^2 exact (fma) and 2^(j/64) a double-double
//   log   x = 2^e m, m in [3/4, 3/2); c = fl(1/m_i) for the nearest centre
//         m_i = 3/4 + i/128; m c - 1 = r exactly as a pair (fma);
//         log x = e ln2 - log c + log1p(r), log1p(r) = r - r^2/2 + r^3 P(r),
//         the leading terms summed without error
//   cos   x = n pi/2 + r (three-part pi/2, r as a pair), |r| <= pi/4;
//         r = a + t with a = j/64: sin(a + t) and cos(a + t) from sin a,
//         cos a (double-double) and short series of t, |t| <= 1/128;
//         accurate for |x| < 2^20
//   acos  asin by one Newton correction of a series estimate against the
//         accurate sine (t - sin y)/cos y; acos x = pi/2 - asin x for
//         |x| <= 1/2, 2 asin(sqrt((1-|x|)/2)) beyond (the square root carried
//         as a pair), pi minus that for x < -1/2
//   cbrt  exponent third and a quadratic mantissa fit, one Halley step, then
//         a Newton correction from the cube's residual (the square's rounding
//         error recovered by fma) added in one rounding
//   powThird  x^fl(1/3): fl(1/3) = 1/3 - 1/(3 2^54), so
//         x^fl(1/3) = cbrt(x) (1 - ln(x) / (3 2^54)) to within 2^-100
// Measured against 200-bit references (tests/test_elementary.py): log, exp,
// acos and cos correctly rounded on every sampled argument, cbrt and
// powThird within 1 ulp.
#ifndef REACTIVE_AMD_ELEMENTARY_H
#define REACTIVE_AMD_ELEMENTARY_H
#include <cmath>
#include <cstdint>
#include <cstring>

#ifdef __HIPCC__
#define RAMD_ELEM __host__ __device__ inline
#define RAMD_ELEM_TABLE static constexpr
#else
#define RAMD_ELEM inline
#define RAMD_ELEM_TABLE static constexpr
#endif
#include "elementary_tables.h"

namespace ramd {
namespace elem {

RAMD_ELEM uint64_t bits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, sizeof(u));
    return u;
}
RAMD_ELEM double fromBits(uint64_t u) {
    double x;
    std::memcpy(&x, &u, sizeof(x));
    return x;
}
RAMD_ELEM double fmaE(double a, double b, double c) { return ::fma(a, b, c); }
// 2^k for -1022 <= k <= 1023, exactly
RAMD_ELEM double pow2(int k) { return fromBits(uint64_t(int64_t(k) + 1023) << 52); }
// x 2^k for any k that keeps the result normal or zero: two exact steps
RAMD_ELEM double scale(double x, int k) {
    if (k > 1000) return x * pow2(1000) * pow2(k - 1000);
    if (k < -1000) return x * pow2(-1000) * pow2(k + 1000);
    return x * pow2(k);
}

constexpr double kLn2Hi = 0x1.62e42fef00000p-1;   // 32 leading bits of ln 2
constexpr double kLn2Lo = 7.440617110012397e-11;   // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kSqrtHalf = 0.70710678118654757;
constexpr double kPio2Hi = 1.5707963267948966, kPio2Lo = 6.123233995736766e-17;
constexpr double kPiHi = 3.141592653589793, kPiLo = 1.2246467991473532e-16;
constexpr double kPio2A = 1.5707963267341256, kPio2B = 6.077100506303966e-11, kPio2C = 2.0222662487959506e-21;

RAMD_ELEM double logFast(double x) {
    if (!(x > 0)) return x == 0 ? -HUGE_VAL : NAN;  // also NaN in, NaN out
    if (x == HUGE_VAL) return x;
    uint64_t u = bits(x);
    int e = int((u >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: renormalise
        x *= 0x1p54;
        u = bits(x);
        e = int((u >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = fromBits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m > 2 * kSqrtHalf) {
        m *= 0.5;
        ++e;
    }
    const double f = m - 1.0;  // exact
    const double s = f / (2.0 + f);
    const double z = s * s;
    // R(z) = sum_k 2 z^k / (2k + 1), k = 1..13 (z <= 0.0295)
    double r = 0.07407407407407407;
    r = fmaE(r, z, 0.08);
    r = fmaE(r, z, 0.08695652173913043);
    r = fmaE(r, z, 0.09523809523809523);
    r = fmaE(r, z, 0.10526315789473684);
    r = fmaE(r, z, 0.11764705882352941);
    r = fmaE(r, z, 0.13333333333333333);
    r = fmaE(r, z, 0.15384615384615385);
    r = fmaE(r, z, 0.18181818181818182);
    r = fmaE(r, z, 0.2222222222222222);
    r = fmaE(r, z, 0.2857142857142857);
    r = fmaE(r, z, 0.4);
    r = fmaE(r, z, 0.6666666666666666);
    r *= z;
    // log(1 + f) = f - hfsq + s (hfsq + R), hfsq = f^2 / 2
    const double hfsq = 0.5 * f * f;
    const double de = double(e);
    return fmaE(de, kLn2Hi, f - (hfsq - fmaE(s, hfsq + r, de * kLn2Lo)));
}

RAMD_ELEM double expFast(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double kd = ::rint(x * kInvLn2);
    const int k = int(kd);
    // r = x - k ln2 as rh + rl: x - k kLn2Hi is exact, k kLn2Lo is split
    const double r1 = fmaE(-kd, kLn2Hi, x);
    const double t = kd * kLn2Lo, te = fmaE(kd, kLn2Lo, -t);
    const double rh = r1 - t, rl = ((r1 - rh) - t) - te;
    // exp r = 1 + r + r^2/2 + r^3 (1/3! + r/4! + ... + r^10/13!)
    double p = 1.6059043836821613e-10;
    p = fmaE(p, rh, 2.08767569878681e-09);
    p = fmaE(p, rh, 2.505210838544172e-08);
    p = fmaE(p, rh, 2.755731922398589e-07);
    p = fmaE(p, rh, 2.7557319223985893e-06);
    p = fmaE(p, rh, 2.48015873015873e-05);
    p = fmaE(p, rh, 0.0001984126984126984);
    p = fmaE(p, rh, 0.001388888888888889);
    p = fmaE(p, rh, 0.008333333333333333);
    p = fmaE(p, rh, 0.041666666666666664);
    p = fmaE(p, rh, 0.16666666666666666);
    const double zh = rh * rh, zl = fmaE(rh, rh, -zh);  // r^2 exactly
    // leading terms 1 + rh + zh/2 summed with their rounding errors
    const double s1 = 1.0 + rh, e1 = (1.0 - s1) + rh;
    const double half = 0.5 * zh;
    const double s2 = s1 + half, e2 = (s1 - s2) + half;
    const double tail = ((rl + rl * rh) + 0.5 * zl) + (zh * rh) * p;
    return scale(s2 + ((e1 + e2) + tail), k);
}

// cube root of a > 0 as c + corr (c from Halley, corr its Newton correction)
RAMD_ELEM void cbrtParts(double a, double& c, double& corr) {
    int adjust = 0;
    if (a < 0x1p-600) {  // keep a normal and every cube, square and error term in range
        a *= 0x1p900;
        adjust = -300;
    } else if (a > 0x1p600) {
        a *= 0x1p-900;
        adjust = 300;
    }
    const int e = int((bits(a) >> 52) & 0x7ff) - 1023;
    // 2^floor(e/3) 2^(rem/3) times a quadratic fit of the mantissa's cube
    // root on [1, 2) (relative error 1.2e-3)
    const int e3 = e >= 0 ? e / 3 : -((-e + 2) / 3);
    const int rem = e - 3 * e3;  // 0, 1, 2
    const double mant = fromBits((bits(a) & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    const double fit = fmaE(fmaE(-0.05930307106611154, mant, 0.4355650806935765), mant, 0.6249076548324355);
    c = scale((rem == 0 ? 1.0 : rem == 1 ? 1.2599210498948732 : 1.5874010519681994) * fit, e3);
    // one Halley step (cubic convergence: 1.2e-3 -> ~2e-9)
    const double c3 = c * c * c;
    c = c * ((c3 + 2.0 * a) / (2.0 * c3 + a));
    // Newton correction from the residual c^3 - a, with the square's rounding
    // error recovered by fma (quadratic: ~2e-9 -> ~4e-18)
    const double c2 = c * c, c2e = fmaE(c, c, -c2);
    const double residual = fmaE(c2, c, -a) + c2e * c;
    corr = -residual / (3.0 * c2);
    c = scale(c, adjust);
    corr = scale(corr, adjust);
}

// ln x to about 0.007, enough for the size of the 1/3-exponent correction
RAMD_ELEM double logRough(double x) {
    int shift = 0;
    if (x < 0x1p-1022) {  // subnormal
        x *= 0x1p54;
        shift = 54;
    }
    const uint64_t u = bits(x);
    const double t = fromBits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull) - 1.0;
    return double(int((u >> 52) & 0x7ff) - 1023 - shift) * 0.6931471805599453 +
           t * fmaE(-0.25436922244455573, t, 0.940777422303657);
}

RAMD_ELEM double cbrt(double x) {
    if (x == 0 || x != x || x == HUGE_VAL || x == -HUGE_VAL) return x;
    double c, corr;
    cbrtParts(x < 0 ? -x : x, c, corr);
    const double r = c + corr;
    return x < 0 ? -r : r;
}

// x^fl(1/3), the value pow(x, 1.0 / 3.0) denotes, for x >= 0
RAMD_ELEM double powThird(double x) {
    if (x != x || x == HUGE_VAL) return x;
    if (!(x > 0)) return x == 0 ? 0.0 : NAN;
    double c, corr;
    cbrtParts(x, c, corr);
    // x^(1/3 - d) = cbrt(x) (1 - d ln x) with d = 1/(3 2^54): |d ln x| < 2e-14,
    // so ln x needs two digits (logRough; subnormal x only loses accuracy in
    // a term far below the last place)
    return c + (corr - c * (logRough(x) * (1.0 / 54043195528445952.0)));
}

RAMD_ELEM double asinSmall(double t) {  // |t| <= 1/2
    const double z = t * t;
    double p = 0.0017680811205154183;
    p = fmaE(p, z, 0.0018622264064031275);
    p = fmaE(p, z, 0.0019650336162772837);
    p = fmaE(p, z, 0.0020776610325181676);
    p = fmaE(p, z, 0.0022014739737101384);
    p = fmaE(p, z, 0.002338091892111975);
    p = fmaE(p, z, 0.0024894486782468836);
    p = fmaE(p, z, 0.00265787063820729);
    p = fmaE(p, z, 0.002846178401108942);
    p = fmaE(p, z, 0.0030578216492580306);
    p = fmaE(p, z, 0.003297059503473485);
    p = fmaE(p, z, 0.0035692053938259347);
    p = fmaE(p, z, 0.003880964558837669);
    p = fmaE(p, z, 0.004240907093679363);
    p = fmaE(p, z, 0.004660143486915096);
    p = fmaE(p, z, 0.005153309682319905);
    p = fmaE(p, z, 0.005740037670841924);
    p = fmaE(p, z, 0.006447210311889649);
    p = fmaE(p, z, 0.0073125258735988454);
    p = fmaE(p, z, 0.008390335809616815);
    p = fmaE(p, z, 0.009761609529194078);
    p = fmaE(p, z, 0.011551800896139705);
    p = fmaE(p, z, 0.01396484375);
    p = fmaE(p, z, 0.017352764423076924);
    p = fmaE(p, z, 0.022372159090909092);
    p = fmaE(p, z, 0.030381944444444444);
    p = fmaE(p, z, 0.044642857142857144);
    p = fmaE(p, z, 0.075);
    p = fmaE(p, z, 0.16666666666666666);
    return fmaE(t * z, p, t);
}

RAMD_ELEM double acosFast(double x) {
    if (!(x >= -1 && x <= 1)) return NAN;
    if (x >= -0.5 && x <= 0.5) return kPio2Hi - (asinSmall(x) - kPio2Lo);
    if (x > 0) return 2.0 * asinSmall(::sqrt(0.5 * (1.0 - x)));
    return kPiHi - (2.0 * asinSmall(::sqrt(0.5 * (1.0 + x))) - kPiLo);
}

RAMD_ELEM double cosReduced(double r) {  // |r| <= pi/4
    const double z = r * r;
    double p = 4.779477332387385e-14;  // 1/16!
    p = fmaE(p, -z, 1.1470745597729725e-11);
    p = fmaE(p, -z, 2.08767569878681e-09);
    p = fmaE(p, -z, 2.755731922398589e-07);
    p = fmaE(p, -z, 2.48015873015873e-05);
    p = fmaE(p, -z, 0.001388888888888889);
    p = fmaE(p, -z, 0.041666666666666664);
    // cos r = 1 - z/2 + z^2 (1/4! - z/6! + ...)
    const double hz = 0.5 * z, w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + z * z * p);
}
RAMD_ELEM double sinReduced(double r) {  // |r| <= pi/4
    const double z = r * r;
    double p = 2.8114572543455206e-15;  // 1/17!
    p = fmaE(p, -z, 7.647163731819816e-13);
    p = fmaE(p, -z, 1.6059043836821613e-10);
    p = fmaE(p, -z, 2.505210838544172e-08);
    p = fmaE(p, -z, 2.7557319223985893e-06);
    p = fmaE(p, -z, 0.0001984126984126984);
    p = fmaE(p, -z, 0.008333333333333333);
    p = fmaE(p, -z, 0.16666666666666666);
    return fmaE(-r * z, p, r);
}
RAMD_ELEM double cosFast(double x) {
    if (x != x || x == HUGE_VAL || x == -HUGE_VAL) return NAN;
    const double nd = ::rint(x * 0.63661977236758138);  // 2 / pi
    const double r = ((x - nd * kPio2A) - nd * kPio2B) - nd * kPio2C;
    const int n = int(int64_t(nd) & 3);
    if (n == 0) return cosReduced(r);
    if (n == 1) return -sinReduced(r);
    if (n == 2) return -cosReduced(r);
    return sinReduced(r);
}


// ---- accurate log, exp, cos, acos -------------------------------------------
// Error-free transformations: s + e = a + b, p + e = a b exactly.
RAMD_ELEM void twoSum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
RAMD_ELEM void fastTwoSum(double a, double b, double& s, double& e) {  // |a| >= |b| or a == 0
    s = a + b;
    e = b - (s - a);
}
RAMD_ELEM void twoProd(double a, double b, double& p, double& e) {
    p = a * b;
    e = fmaE(a, b, -p);
}

constexpr double kInvLn2x64 = 0x1.71547652b82fep+6;                     // 64 / ln 2
constexpr double kL1 = 0x1.62e42fefa0000p-7, kL2 = 0x1.cf79abc9e3b3ap-46,  // ln2/64: 36 bits,
                 kL3 = -0x1.ff0342542fc33p-100;                           // then two doubles
constexpr double kP1 = 0x1.921fb54400000p+0, kP2 = 0x1.0b4611a626331p-34,  // pi/2: 33 bits,
                 kP3 = 0x1.1701b839a2520p-88;                             // then two doubles

RAMD_ELEM double exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    const double kd = ::rint(x * kInvLn2x64);  // |k| < 2^17, so k kL1 is exact
    const int k = int(kd), j = k & 63, m = (k - j) / 64;
    // r = x - k ln2/64 = rh + rl (x - k kL1 exact by Sterbenz)
    const double r1 = fmaE(-kd, kL1, x);
    double t, te, rh, rhe;
    twoProd(kd, kL2, t, te);
    twoSum(r1, -t, rh, rhe);
    const double rl = rhe - (te + kd * kL3);
    // e^r - 1 = rh + rh^2/2 + [rl (1 + rh) + rh^3 Q(rh)], Q = 1/3! + ... + rh^4/7!
    double sq, sqe;
    twoProd(rh, rh, sq, sqe);
    double Q = 1.9841269841269841e-04;
    Q = fmaE(Q, rh, 1.3888888888888889e-03);
    Q = fmaE(Q, rh, 8.3333333333333332e-03);
    Q = fmaE(Q, rh, 4.1666666666666664e-02);
    Q = fmaE(Q, rh, 1.6666666666666666e-01);
    const double small = rl + (rh * rl + (0.5 * sqe + (rh * sq) * Q));
    double u, ul;
    fastTwoSum(rh, 0.5 * sq, u, ul);  // |rh| >= rh^2/2
    const double Th = kExp2Table[2 * j], Tl = kExp2Table[2 * j + 1];
    double b, bl, s, se;
    twoProd(Th, u, b, bl);
    fastTwoSum(Th, b, s, se);
    const double tail = ul + small;
    return scale(s + (se + (bl + (Th * tail + Tl * ((1.0 + u) + tail)))), m);
}

// e^x - 1 (the adaptive flash's liquid mass, -expm1(ln vapour fraction) q).
// Below ln2/128 in magnitude the reduction is empty (r = x) and the result is
// the series of e^r - 1 itself; otherwise e^x as exp forms it, a double-double
// 2^m (S + E) good to about 2^-68, from which 1 is subtracted exactly.  The
// error stays below 2^-60 of the result (at most 1/128 ulp before rounding).
RAMD_ELEM double expm1(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -37.5) return -1.0;  // e^x < 2^-54: -1 + e^x rounds to -1
    const double kd = ::rint(x * kInvLn2x64);
    const int k = int(kd), j = k & 63, m = (k - j) / 64;
    const double r1 = fmaE(-kd, kL1, x);
    double t, te, rh, rhe;
    twoProd(kd, kL2, t, te);
    twoSum(r1, -t, rh, rhe);
    const double rl = rhe - (te + kd * kL3);
    double sq, sqe;
    twoProd(rh, rh, sq, sqe);
    double Q = 1.9841269841269841e-04;
    Q = fmaE(Q, rh, 1.3888888888888889e-03);
    Q = fmaE(Q, rh, 8.3333333333333332e-03);
    Q = fmaE(Q, rh, 4.1666666666666664e-02);
    Q = fmaE(Q, rh, 1.6666666666666666e-01);
    const double small = rl + (rh * rl + (0.5 * sqe + (rh * sq) * Q));
    double u, ul;
    fastTwoSum(rh, 0.5 * sq, u, ul);
    if (k == 0) return u + (ul + small);  // r = x exactly
    const double Th = kExp2Table[2 * j], Tl = kExp2Table[2 * j + 1];
    double b, bl, s, se;
    twoProd(Th, u, b, bl);
    fastTwoSum(Th, b, s, se);
    const double tail = ul + small;
    const double E = se + (bl + (Th * tail + Tl * ((1.0 + u) + tail)));
    double a, ae;
    twoSum(scale(s, m), -1.0, a, ae);  // 2^m s and 2^m E are exact (m > -56)
    return a + (ae + scale(E, m));
}

RAMD_ELEM double log(double x) {
    if (!(x > 0)) return x == 0 ? -HUGE_VAL : NAN;  // also NaN in, NaN out
    if (x == HUGE_VAL) return x;
    uint64_t u = bits(x);
    int e = int((u >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: renormalise
        x *= 0x1p54;
        u = bits(x);
        e = int((u >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = fromBits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m >= 1.5) {
        m *= 0.5;
        ++e;
    }
    const int i = int((m - 0.75) * 128.0 + 0.5);  // nearest centre 3/4 + i/128, i in [0, 96]
    const double c = kLogTable[3 * i], Ch = kLogTable[3 * i + 1], Cl = kLogTable[3 * i + 2];
    double ph, pl;
    twoProd(m, c, ph, pl);
    const double rh = ph - 1.0, rl = pl;  // m c - 1 exactly; |r| < 0.0053
    // log1p(rh + rl) = rh - rh^2/2 + rh^3 P(rh) + rl (1 - rh), P = 1/3 - rh/4 + ... + rh^8/11
    double sq, sqe;
    twoProd(rh, rh, sq, sqe);
    double P = 9.0909090909090912e-02;
    P = fmaE(P, rh, -0.1);
    P = fmaE(P, rh, 1.1111111111111111e-01);
    P = fmaE(P, rh, -0.125);
    P = fmaE(P, rh, 1.4285714285714285e-01);
    P = fmaE(P, rh, -1.6666666666666666e-01);
    P = fmaE(P, rh, 0.2);
    P = fmaE(P, rh, -0.25);
    P = fmaE(P, rh, 3.3333333333333331e-01);
    const double small = (rl - rl * rh) + ((rh * sq) * P - 0.5 * sqe);
    const double de = double(e);
    double a, ae, b, be, cc, ce;
    twoSum(de * kLn2Hi, Ch, a, ae);  // e kLn2Hi exact (32-bit kLn2Hi)
    twoSum(a, rh, b, be);
    twoSum(b, -0.5 * sq, cc, ce);
    return cc + (((ae + be) + ce) + (Cl + (de * kLn2Lo + small)));
}

// sin (sine) or cos of r = rh + rl, |rh| <= pi/4 + 2^-40, as hi + lo; *other
// (when given) receives the other function to about double precision.
RAMD_ELEM void sinCosPair(double rh, double rl, bool sine, double& hi, double& lo, double* other) {
    const bool negative = rh < 0;
    const double ar = negative ? -rh : rh, al = negative ? -rl : rl;
    const int j = int(ar * 64.0 + 0.5);  // [0, 51]
    const double th = ar - double(j) * 0.015625;  // exact (Sterbenz, or j = 0); |th| <= 1/128
    const double Sh = kSinCosTable[4 * j], Sl = kSinCosTable[4 * j + 1];
    const double Ch = kSinCosTable[4 * j + 2], Cl = kSinCosTable[4 * j + 3];
    double t2, t2e;
    twoProd(th, th, t2, t2e);
    // cos t = 1 - hA - hB (to t^6), sin t = th + sig (to t^7), t = th + al
    const double hA = 0.5 * t2;
    const double hB = (0.5 * t2e + th * al) - (t2 * t2) * fmaE(-t2, 1.3888888888888889e-03, 4.1666666666666664e-02);
    const double sig = al - (th * t2) * fmaE(-t2, fmaE(-t2, 1.9841269841269841e-04, 8.3333333333333332e-03),
                                              1.6666666666666666e-01);
    if (sine) {
        // sin(a + t) = S cos t + C sin t = Sh + Ch th + [Sl + Cl th + C sig - S (hA + hB)]
        double u, ue, w, we, v, ve;
        twoProd(Ch, th, u, ue);
        twoProd(Sh, hA, w, we);
        fastTwoSum(Sh, u, v, ve);
        const double rest = ((Sl + ue) - w) + ((Cl * th - we) + (Ch * sig - (Sh * hB + Sl * hA)));
        hi = v;
        lo = ve + rest;
        if (negative) { hi = -hi; lo = -lo; }
        if (other) *other = Ch - (Sh * th + Ch * hA);
    } else {
        // cos(a + t) = C cos t - S sin t = Ch - Sh th + [Cl - Sl th - S sig - C (hA + hB)]
        double u, ue, w, we, v, ve;
        twoProd(-Sh, th, u, ue);
        twoProd(Ch, hA, w, we);
        fastTwoSum(Ch, u, v, ve);
        const double rest = ((Cl + ue) - w) - ((Sl * th + we) + (Sh * sig + (Ch * hB + Cl * hA)));
        hi = v;
        lo = ve + rest;
        if (other) *other = negative ? -(Sh + Ch * th) : Sh + Ch * th;
    }
}

RAMD_ELEM double cos(double x) {
    if (x != x || x == HUGE_VAL || x == -HUGE_VAL) return NAN;
    const double ax = x < 0 ? -x : x;
    if (!(ax < 0x1p20)) return cosFast(x);
    const double nd = ::rint(ax * 0.63661977236758138);  // 2 / pi
    // r = ax - n pi/2 = rh + rl (ax - n kP1 exact)
    const double r1 = fmaE(-nd, kP1, ax);
    double t, te, r0, r0e, rh, rl;
    twoProd(nd, kP2, t, te);
    twoSum(r1, -t, r0, r0e);
    fastTwoSum(r0, r0e - (te + nd * kP3), rh, rl);
    const int n = int(int64_t(nd) & 3);
    double hi, lo;
    sinCosPair(rh, rl, (n & 1) != 0, hi, lo, nullptr);
    const double v = hi + lo;
    return n == 0 || n == 3 ? v : -v;  // cos r, -sin r, -cos r, sin r
}

// asin t = y + corr for |t| <= 1/2: the series estimate y corrected by one
// Newton step against the accurate sine; *cosine receives cos y
RAMD_ELEM void asinPair(double t, double& y, double& corr, double& cosine) {
    y = asinSmall(t);
    double sh, sl;
    sinCosPair(y, 0.0, true, sh, sl, &cosine);
    corr = ((t - sh) - sl) / cosine;  // t - sh exact (Sterbenz)
}

RAMD_ELEM double acos(double x) {
    if (!(x >= -1 && x <= 1)) return NAN;
    double y, corr, c;
    if (x >= -0.5 && x <= 0.5) {  // pi/2 - asin x
        asinPair(x, y, corr, c);
        double a, ae;
        twoSum(kPio2Hi, -y, a, ae);
        return a + ((ae - corr) + kPio2Lo);
    }
    // 2 asin(s), s = sqrt((1 - |x|)/2) = sh + sl
    const double w = 0.5 * (1.0 - (x > 0 ? x : -x));  // exact
    const double sh = ::sqrt(w);
    const double sl = sh > 0 ? fmaE(-sh, sh, w) / (2.0 * sh) : 0.0;  // |x| = 1: s = 0
    asinPair(sh, y, corr, c);
    const double tail = corr + sl / c;  // asin(sh + sl) = y + corr + sl / cos y
    if (x > 0) return 2.0 * y + 2.0 * tail;
    double a, ae;
    twoSum(kPiHi, -2.0 * y, a, ae);
    return a + ((ae - 2.0 * tail) + kPiLo);
}

}  // namespace elem
}  // namespace ramd
#endif
