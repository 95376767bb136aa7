// SPDX-License-Identifier: GPL-3.0-or-later
// Peng-Robinson / NASA thermodynamics for the gfx950 HEM flash.
//
// Same equations, constants and operation order as the reference evaluator
// (src/reactiveThermo/reactiveDevicePR.h, Cantera 3.2 forms), restructured for
// one-thread-per-cell execution on CDNA4:
//  * tables are compact and templated on the species count (1 or 2 here), so
//    every per-species array lives in registers;
//  * per-species critical temperatures are precomputed once with the
//    reference's own expression (bitwise identical to recomputing them);
//  * Cantera's phaseState density heuristic depends on composition only up to
//    its final (T, rho) comparison, so its composition part is memoised
//    (exact reuse; see PhaseHeuristic).
// All functions are __host__ __device__: the host side of the library uses
// them only to preprocess a model image (table checks, pure-phase constants).
#ifndef REACTIVE_AMD_PR_DEVICE_H
#define REACTIVE_AMD_PR_DEVICE_H
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include "elementary.h"

#define RHD __host__ __device__ inline

namespace ramd {

constexpr double kR = 8314.46261815324;            // reactiveDevicePR.h:31
constexpr double kSmall = 1e-300;
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kOmegaA = 4.5723552892138218e-1;
constexpr double kOmegaB = 7.77960739038885e-2;
constexpr double kOmegaVc = 3.07401308698703833e-1;
constexpr int kMaxRegions = 8;

enum Status : int { kOk = 0, kBadTable, kBadInput, kBadComposition, kNoRoot, kUnavailable, kUnstable, kNonFinite };
enum RootChoice : int { kAutoRoot = 0, kGasRoot = 1, kLiquidRoot = 2 };

// ---- elementary functions, by precision.  FP64: fixed operation sequences
// shared with the oracle's twin build (elementary.h), so results are
// reproducible on the host; square root is IEEE correctly rounded on both.
// FP32 (the mixed-precision search only): the device library.
RHD double m_log(double x) { return elem::log(x); }
RHD double m_exp(double x) { return elem::exp(x); }
RHD double m_third(double x) { return elem::powThird(x); }  // pow(x, 1.0 / 3.0)
RHD double m_acos(double x) { return elem::acos(x); }
RHD double m_cos(double x) { return elem::cos(x); }
RHD double m_sqrt(double x) { return ::sqrt(x); }
RHD double m_expm1(double x) { return elem::expm1(x); }
RHD float m_log(float x) { return ::logf(x); }
RHD float m_exp(float x) { return ::expf(x); }
RHD float m_third(float x) { return ::cbrtf(x); }
RHD float m_acos(float x) { return ::acosf(x); }
RHD float m_cos(float x) { return ::cosf(x); }
RHD float m_sqrt(float x) { return ::sqrtf(x); }

// Precision-dependent constants: the FP64 values are the reference's; the
// FP32 ones scale its absolute thresholds to single precision.
template <class R> struct Real;
template <> struct Real<double> {
    static constexpr double big = 1.7976931348623157e308, eps = DBL_EPSILON, tiny = 1e-300;
    static constexpr double polish = 4e-15;    // cubic-root Newton polish (reactiveDevicePR.h:280-307)
    static constexpr double flatness = 1e-14;  // |f'| below which the polish stops
};
template <> struct Real<float> {
    static constexpr float big = 3.4028234663852886e38f, eps = FLT_EPSILON, tiny = 1e-30f;
    static constexpr float polish = 2e-7f;
    static constexpr float flatness = 1e-30f;
};

template <class R> RHD bool fin(R x) { return x == x && (x < 0 ? -x : x) <= Real<R>::big; }
template <class R> RHD R lo(R a, R b) { return a < b ? a : b; }
template <class R> RHD R hi(R a, R b) { return a > b ? a : b; }
template <class R> RHD R ab(R x) { return x < 0 ? -x : x; }
template <class R> RHD R clamp(R a, R l, R h) { return lo(hi(a, l), h); }

struct Region { double tmin; double c[9]; int poly; };

// Compact PR table of N species (one gas mixture or one pure liquid).
template <int N>
struct Table {
    int ns;
    double refP;
    double W[N], a[N], b[N], kappa[N], Tc[N];
    double aij[N][N];
    double tlo[N], thi[N];  // species fit range: first region min, last region max
    int nreg[N];
    Region region[N][kMaxRegions];
    // Fixed-composition phaseState constants (valid only when N == 1).
    int fixedValid;
    double fixedTc, fixedCritDensity, fixedDensityDelta, fixedTempDelta;
    int fixedStatus;
};

template <class R>
struct PRStateT {  // subset of reactiveDevicePR.h:101-110 used by the flash
    R V, rho, e, h, s, cp, cv, alpha, kappaT, sound, p, dpdT, dpdV;
    int phase;
};
using PRState = PRStateT<double>;

// NASA7 / NASA9 (reactiveDevicePR.h:127-155).
template <class R>
RHD void nasa(const Region& r, R T, R& cp, R& h, R& s) {
    R a[9];
    for (int j = 0; j < 9; ++j) a[j] = R(r.c[j]);
    const R T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    const R invT = R(1) / T, logT = m_log(T);
    if (r.poly == 7) {
        cp = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
        h = a[0] + R(.5) * a[1] * T + a[2] * T2 / R(3) + R(.25) * a[3] * T3 + R(.2) * a[4] * T4 + a[5] * invT;
        s = a[0] * logT + a[1] * T + R(.5) * a[2] * T2 + a[3] * T3 / R(3) + R(.25) * a[4] * T4 + a[6];
    } else {
        const R c0 = a[0] * invT * invT, c1 = a[1] * invT, c2 = a[2], c3 = a[3] * T;
        const R c4 = a[4] * T2, c5 = a[5] * T3, c6 = a[6] * T4;
        cp = c0 + c1 + c2 + c3 + c4 + c5 + c6;
        h = -c0 + logT * c1 + c2 + R(.5) * c3 + c4 / R(3) + R(.25) * c5 + R(.2) * c6 + a[7] * invT;
        s = -R(.5) * c0 - c1 + logT * c2 + c3 + R(.5) * c4 + c5 / R(3) + R(.25) * c6 + a[8];
    }
}

// NASA9 upper region at a shared bound, NASA7 lower (reactiveDevicePR.h:197-212).
template <int N, class R>
RHD const Region& region(const Table<N>& t, int k, R T) {
    int r = 0;
    for (int j = 1; j < t.nreg[k]; ++j) {
        const R bound = R(t.region[k][j].tmin);
        const bool upper = t.region[k][j].poly == 9 ? T >= bound : T > bound;
        if (!upper) break;
        r = j;
    }
    return t.region[k][r];
}

template <class R>
struct MixT { R a, b, aa, da, d2a; };
using Mix = MixT<double>;

// Mixing rule (reactiveDevicePR.h:214-247).
template <int N, class R>
RHD int mixing(const Table<N>& t, R T, const R* x, MixT<R>& m) {
    R q[N], dq[N], d2q[N];
    m.a = 0; m.b = 0; m.aa = 0; m.da = 0; m.d2a = 0;
    for (int i = 0; i < N; ++i) {
        const R Tc = R(t.Tc[i]), kappa = R(t.kappa[i]);
        if (!(Tc > 0) || !fin(Tc)) return kBadTable;
        const R sqrtTr = m_sqrt(T / Tc);
        const R f = R(1) + kappa * (R(1) - sqrtTr);
        if (ab(f) <= R(1e-14)) return kNonFinite;
        const R sign = f > 0 ? R(1) : R(-1);
        q[i] = ab(f);
        dq[i] = -sign * kappa / (R(2) * Tc * sqrtTr);
        d2q[i] = sign * kappa / (R(4) * Tc * Tc * sqrtTr * sqrtTr * sqrtTr);
        m.b += x[i] * R(t.b[i]);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const R w = x[i] * x[j] * R(t.aij[i][j]);
            m.a += w;
            m.aa += w * q[i] * q[j];
            m.da += w * (dq[i] * q[j] + q[i] * dq[j]);
            m.d2a += w * (d2q[i] * q[j] + R(2) * dq[i] * dq[j] + q[i] * d2q[j]);
        }
    return fin(m.a) && fin(m.b) && fin(m.aa) && fin(m.da) && fin(m.d2a) && m.a > 0 && m.b > 0 && m.aa > 0 ? kOk
                                                                                                         : kNonFinite;
}

template <class R> RHD R cubic(R V, R bn, R cn, R dn) { return ((V + bn) * V + cn) * V + dn; }
template <class R> RHD R cbrtSigned(R x) { return x < 0 ? -m_third(-x) : m_third(x); }

// Isolated third root in the near-repeated discriminant band (reactiveDevicePR.h:280-307).
template <class R>
RHD bool isolatedRoot(R center, R delta, R delta2, R q, R bn, R cn, R dn, R b, R RTp, R aap, R& root,
                      bool& gasSide) {
    if (!(delta2 > 0) || !fin(delta) || !fin(q) || q == 0) return false;
    gasSide = q < 0;
    const R pair = center + (gasSide ? R(-1) : R(1)) * delta;
    R x = center + (gasSide ? R(2) : R(-2)) * delta;
    bool converged = false;
    for (int n = 0; n < 24; ++n) {
        const R f = cubic(x, bn, cn, dn), df = (R(3) * x + R(2) * bn) * x + cn;
        if (!fin(f) || !fin(df) || ab(df) <= R(Real<R>::tiny)) return false;
        const R step = f / df;
        x -= step;
        if (ab(step) <= R(Real<R>::polish) * hi(R(1), ab(x))) { converged = true; break; }
    }
    if (!converged || !fin(x) || !(x > b)) return false;
    const R f = cubic(x, bn, cn, dn);
    const R scale = ab(x * x * x) + ab(bn * x * x) + ab(cn * x) + ab(dn);
    if (!(ab(f) <= R(64) * R(Real<R>::eps) * scale)) return false;
    if (!(ab(x - pair) > R(1e-3) * hi(ab(x), ab(pair)))) return false;
    if (gasSide ? !(x > pair) : !(x < pair)) return false;
    const R vb = x - b, den = x * x + R(2) * b * x - b * b;
    if (!(vb > 0) || !(den > 0)) return false;
    if (!(-RTp / (vb * vb) + R(2) * aap * (x + b) / (den * den) < 0)) return false;
    root = x;
    return true;
}

// Single precision: a discriminant within this fraction of q^2 + h^2 (a
// near-repeated root pair) is classified in double precision.  The FP32
// rounding of the discriminant is a few 1e-7 of q^2 + h^2.  On the seeded
// 160^3 box 17 % of the search's cubic solves fell within 1e-4 (a third of
// the wave-level solves had such a lane); 1e-5 makes the search 3 % faster
// with the same results (no flips, same distance to the converged roots).
#ifndef RAMD_CUBIC_FP64_BAND
#define RAMD_CUBIC_FP64_BAND 1e-5
#endif
template <class R>
struct RootsT { R v0, vLast; int count, signedCount, gas, liquid; };
using Roots = RootsT<double>;

// The trigonometric branch's acos and cos: the flash's own (correctly
// rounded in FP64), or, for the FP64 classification inside the FP32 search
// (Exact = false: it only seeds candidates), the shorter ~1-ulp versions.
template <class R, bool Exact>
struct CubicTrig {
    static RHD R acos(R x) { return m_acos(x); }
    static RHD R cos(R x) { return m_cos(x); }
};
template <>
struct CubicTrig<double, false> {
    static RHD double acos(double x) { return elem::acosFast(x); }
    static RHD double cos(double x) { return elem::cosFast(x); }
};

template <class R, bool Exact = true> RHD int roots(R T, R p, R a, R b, R aa, RootsT<R>& out);

// Cubic root set at (T, p) from the mixture coefficients
// (rootsFromMixingCached, reactiveDevicePR.h:701-790).  Only the smallest and
// largest physical roots are kept: the flash selects nothing else.
// In single precision the classification of a near-repeated root pair (the
// discriminant within rounding of zero) is decided in double precision.
template <class R, bool Exact>
RHD int roots(R T, R p, R a, R b, R aa, RootsT<R>& out) {
    if (!fin(T) || !fin(p) || !fin(a) || !fin(b) || !fin(aa) || T <= 0 || p <= 0 || a <= 0 || b <= 0 || aa <= 0)
        return kBadInput;
    const R b2 = b * b;
    const R RTp = R(kR) * T / p, aap = aa / p;
    const R bn = b - RTp;
    const R cn = -(R(2) * RTp * b - aap + R(3) * b2);
    const R dn = b2 * RTp + b2 * b - aap * b;
    const R center = -bn / R(3);
    const R delta2 = (bn * bn - R(3) * cn) / R(9);
    const R q = R(2) * bn * bn * bn / R(27) - bn * cn / R(3) + dn;
    const R delta = delta2 > 0 ? m_sqrt(delta2) : R(0);
    const R h = R(2) * delta * delta2;
    R disc = q * q - h * h;
    if (sizeof(R) < sizeof(double) && ab(disc) <= R(RAMD_CUBIC_FP64_BAND) * (q * q + h * h)) {
        RootsT<double> d;
        const int st = roots<double, false>(double(T), double(p), double(a), double(b), double(aa), d);
        out.v0 = R(d.v0); out.vLast = R(d.vLast);
        out.count = d.count; out.signedCount = d.signedCount; out.gas = d.gas; out.liquid = d.liquid;
        return st;
    }
    if (ab(ab(h) - ab(q)) < R(1e-10)) {
        if (disc > R(1e-10)) return kNonFinite;
        disc = 0;
    }
    R raw[3] = {0, 0, 0};
    int n = 0, isolated = 0;
    if (ab(disc) < R(1e-14)) {
        bool gasSide = false;
        if (!isolatedRoot(center, delta, delta2, q, bn, cn, dn, b, RTp, aap, raw[0], gasSide)) return kNoRoot;
        n = 1;
        isolated = gasSide ? 1 : -1;
    }
    if (isolated) {
    } else if (disc > R(1e-14) || !(delta2 > 0)) {
        const R sd = R(.5) * m_sqrt(hi(R(0), disc));
        raw[0] = center + cbrtSigned(-R(.5) * q + sd) + cbrtSigned(-R(.5) * q - sd);
        n = 1;
    } else if (disc < R(-1e-14)) {
        R arg = -R(.5) * q / (delta * delta * delta);
        arg = hi(R(-1), lo(R(1), arg));
        using Trig = CubicTrig<R, Exact>;
        const R theta = Trig::acos(arg) / R(3);
        raw[0] = center + R(2) * delta * Trig::cos(theta);
        raw[1] = center + R(2) * delta * Trig::cos(theta + R(2) * R(kPi) / R(3));
        raw[2] = center + R(2) * delta * Trig::cos(theta + R(4) * R(kPi) / R(3));
        n = 3;
    } else {
        return kNoRoot;
    }
    if (!isolated)
        for (int i = 0; i < n; ++i)
            for (int it = 0; it < 12; ++it) {
                const R res = cubic(raw[i], bn, cn, dn);
                const R der = (R(3) * raw[i] + R(2) * bn) * raw[i] + cn;
                if (ab(der) <= R(Real<R>::tiny)) break;
                const R step = res / der;
                raw[i] -= step;
                if (ab(step) <= R(Real<R>::polish) * hi(R(1), ab(raw[i]))) break;
            }
    // ascending order, finite roots above the covolume (reactiveDevicePR.h:253-269)
    for (int i = 1; i < n; ++i) {
        const R key = raw[i];
        int j = i - 1;
        while (j >= 0 && raw[j] > key) { raw[j + 1] = raw[j]; --j; }
        raw[j + 1] = key;
    }
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (!fin(raw[i]) || raw[i] <= b) continue;
        raw[m++] = raw[i];
    }
    if (!m) return kNoRoot;
    out.count = m;
    out.v0 = raw[0];
    out.vLast = raw[m - 1];
    const R Tcrit = a * R(kOmegaB) / (b * R(kOmegaA) * R(kR));
    const R Pcrit = R(kOmegaB) * R(kR) * Tcrit / b;
    const R Vcrit = R(kOmegaVc) * R(kR) * Tcrit / Pcrit;
    if (isolated) {
        out.signedCount = isolated;
        out.gas = isolated > 0;
        out.liquid = isolated < 0;
    } else if (m >= 2) {
        out.signedCount = m;
        out.gas = out.liquid = 1;
    } else {
        const bool liquid = T > Tcrit ? raw[0] < Vcrit : raw[0] < center;
        out.signedCount = liquid ? -1 : 1;
        out.liquid = liquid;
        out.gas = !liquid || T > Tcrit;
    }
    return kOk;
}

template <int N, class R>
RHD int massToMole(const Table<N>& t, const R* Y, R* x, R& W) {  // :310-327
    R invW = 0, sumY = 0;
    for (int k = 0; k < N; ++k) {
        if (!fin(Y[k]) || Y[k] < 0) return kBadComposition;
        sumY += Y[k];
        invW += Y[k] / R(t.W[k]);
    }
    if (!(invW > 0) || ab(sumY - R(1)) > (sizeof(R) < sizeof(double) ? R(1e-5) : R(1e-12))) return kBadComposition;
    W = R(1) / invW;
    for (int k = 0; k < N; ++k) x[k] = Y[k] * W / R(t.W[k]);
    return kOk;
}

// Composition part of Cantera's phaseState heuristic (reactiveDevicePR.h:
// 428-471): everything but the final rho < rhoMid(T) comparison.  FP64 for
// every flash that reports states; the FP32 candidate search instantiates it
// in FP32 (its labels only steer which candidates the search finds, and the
// FP64 refinement evaluates every candidate again with the FP64 heuristic).
struct PhaseHeuristic {
    double tc, critDensity, densityDelta, tempDelta;
    int status;  // status of the tmid root loop (reported only when T < tc)
};

template <int N, class R = double>
RHD void phaseHeuristic(const Table<N>& t, const R* x, R W, R mixA, R mixB, PhaseHeuristic& h) {
    const R tc = mixA * R(kOmegaB) / (mixB * R(kOmegaA) * R(kR));
    const R pc = R(kOmegaB) * R(kR) * tc / mixB;
    const R vc = R(kOmegaVc) * R(kR) * tc / pc;
    h.tc = tc;
    R tmid = tc - R(100.0);
    if (tmid < 0) tmid = R(.5) * tc;
    const R ratio = tc / tmid;
    R pp = pc * m_exp(R(-.8734) * ratio * ratio - R(3.4522) * ratio + R(4.2918));
    MixT<R> mm;
    int st = mixing(t, tmid, x, mm);
    if (st != kOk) { h.status = st; return; }
    R liquidVolume = -1;
    bool found = false;
    for (int it = 0; it < 100; ++it) {
        RootsT<R> r;
        st = roots(tmid, pp, mm.a, mm.b, mm.aa, r);
        if (st != kOk) { h.status = st; return; }
        if (r.signedCount == 1 || r.signedCount == 2) {
            found = pp > pc;
            liquidVolume = r.v0;
            pp *= R(1.04);
            if (found) break;
        } else {
            liquidVolume = r.v0;
            found = true;
            break;
        }
    }
    if (!found || !(liquidVolume > 0) || !fin(liquidVolume)) { h.status = kNonFinite; return; }
    const R critDensity = W / vc;
    const R liquidMid = W / liquidVolume;
    const R gasMid = W / (R(kR) * tmid / pp);
    const R midDensity = R(.5) * (liquidMid + gasMid);
    h.critDensity = critDensity;
    h.densityDelta = critDensity - midDensity;
    h.tempDelta = tc - tmid;
    h.status = kOk;
}

// Final comparison: rhoMid = rho_c + (T - Tc) (rho_c - rho_mid) / (Tc - Tmid).
template <class R>
RHD int reportedPhase(const PhaseHeuristic& h, R T, R rho, int& phase) {
    if (T >= R(h.tc)) { phase = -2; return kOk; }
    if (h.status != kOk) return h.status;
    const R rhoMid = R(h.critDensity) + (T - R(h.tc)) * R(h.densityDelta) / R(h.tempDelta);
    phase = rho < rhoMid ? -1 : 0;
    return fin(rhoMid) ? kOk : kNonFinite;
}

// Full PR state at (T, rho) from mixed coefficients
// (evaluateTrhoFromMixingCached, reactiveDevicePR.h:840-930).  mu (J/kmol)
// is computed for every species when mu != nullptr.
// The phaseState comparison and the final finiteness checks follow in
// finishState(), in the reference's order.
template <int N, class R>
RHD int stateTrho(const Table<N>& t, R T, R rho, const R* x, R W, const MixT<R>& mx, PRStateT<R>& s, R* mu) {
    if (!fin(T) || !fin(rho) || T <= 0 || rho <= 0) return kBadInput;
    const R b = mx.b, aa = mx.aa, da = mx.da, d2a = mx.d2a;
    const R Rg = R(kR), sq2 = R(kSqrt2);
    s.rho = rho;
    s.V = W / rho;
    if (!(s.V > b)) return kNoRoot;
    const R den = s.V * s.V + R(2) * s.V * b - b * b;
    const R vmb = s.V - b;
    if (!(den > 0) || !(vmb > 0)) return kNoRoot;
    s.p = Rg * T / vmb - aa / den;
    if (!fin(s.p) || s.p <= 0) return kNoRoot;
    const R dpdV = -Rg * T / (vmb * vmb) + R(2) * aa * (s.V + b) / (den * den);
    const R dpdT = Rg / vmb - da / den;
    s.dpdV = dpdV;
    s.dpdT = dpdT;
    if (!fin(dpdV) || !fin(dpdT) || dpdV >= 0) return kUnstable;
    const R logRatio = m_log((s.V + (R(1) + sq2) * b) / (s.V + (R(1) - sq2) * b));
    const R L = logRatio / (R(2) * sq2 * b);
    R h0 = 0, s0 = 0, cp0 = 0, xlogx = 0;
    R hRT[N], sR[N];
    for (int k = 0; k < N; ++k) {
        R cpR, hk, sk;
        nasa(region(t, k, T), T, cpR, hk, sk);
        hRT[k] = hk; sR[k] = sk;
        cp0 += x[k] * Rg * cpR;
        h0 += x[k] * Rg * T * hk;
        s0 += x[k] * Rg * sk;
        if (x[k] > 0) xlogx += x[k] * m_log(x[k]);
    }
    const R z = s.p * s.V / (Rg * T);
    const R hRes = Rg * T * (z - R(1)) + L * (T * da - aa);
    const R sRes = Rg * m_log(z * (R(1) - b / s.V)) + L * da;
    const R hM = h0 + hRes;
    const R logP = m_log(s.p / R(t.refP));
    const R sM = s0 - Rg * xlogx - Rg * logP + sRes;
    const R dHdT = cp0 + s.V * dpdT - Rg + L * T * d2a;
    const R cpM = dHdT - (s.V + T * dpdT / dpdV) * dpdT;
    const R cvM = cpM + T * dpdT * dpdT / dpdV;
    s.h = hM / W; s.e = (hM - s.p * s.V) / W; s.s = sM / W;
    s.cp = cpM / W; s.cv = cvM / W;
    s.alpha = -dpdT / (s.V * dpdV);
    s.kappaT = -R(1) / (s.V * dpdV);
    const R sound2 = s.V * s.V * (-cpM / cvM * dpdV / W);
    if (!(s.cp > 0) || !(s.cv > 0) || !(s.kappaT > 0) || !(sound2 > 0)) return kUnstable;
    s.sound = m_sqrt(sound2);
    if (mu) {
        const R denomMu = R(2) * sq2 * b * b;
        const R denomMu2 = b * den;
        const R RT = Rg * T;
        R f[N];
        for (int i = 0; i < N; ++i) f[i] = R(1) + R(t.kappa[i]) * (R(1) - m_sqrt(T / R(t.Tc[i])));
        for (int k = 0; k < N; ++k) {
            R pp = 0;
            for (int i = 0; i < N; ++i) pp += x[i] * R(t.aij[k][i]) * ab(f[i] * f[k]);
            const R bk = R(t.b[k]);
            const R num = R(2) * b * pp - aa * bk;
            mu[k] = RT * (hRT[k] - sR[k]) + RT * m_log(hi(R(Real<R>::tiny), x[k])) + RT * logP -
                    RT * m_log(s.p * s.V / RT) + RT * m_log(s.V / vmb) + RT * bk / vmb - num / denomMu * logRatio -
                    aa * bk * s.V / denomMu2;
        }
    }
    return kOk;
}

template <int N, class R>
RHD int finishState(const PhaseHeuristic& ph, R T, PRStateT<R>& s, const R* mu) {
    const int st = reportedPhase(ph, T, s.rho, s.phase);
    if (st != kOk) return st;
    if (!fin(s.rho) || !fin(s.e) || !fin(s.h) || !fin(s.s) || !fin(s.cp) || !fin(s.cv) || !fin(s.alpha) ||
        !fin(s.kappaT) || !fin(s.sound) || !fin(s.p) || !fin(s.dpdT) || !fin(s.dpdV))
        return kNonFinite;
    if (mu)
        for (int k = 0; k < N; ++k)
            if (!fin(mu[k])) return kNonFinite;
    return kOk;
}

// Table and model checks (reactiveDevicePR.h:157-195).
template <int N>
inline int validateTable(const Table<N>& t) {
    if (t.ns != N || !fin(t.refP) || t.refP <= 0) return kBadTable;
    for (int k = 0; k < N; ++k) {
        if (!fin(t.W[k]) || t.W[k] <= 0 || !fin(t.a[k]) || t.a[k] <= 0 || !fin(t.b[k]) || t.b[k] <= 0 ||
            !fin(t.kappa[k]) || t.nreg[k] <= 0 || t.nreg[k] > kMaxRegions)
            return kBadTable;
    }
    return kOk;
}

}  // namespace ramd
#endif
