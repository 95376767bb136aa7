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

// ---- elementary functions: fixed operation sequences shared with the
// oracle's twin build (elementary.h), so results are reproducible on the host;
// square root is IEEE correctly rounded on both.
RHD double m_log(double x) { return elem::log(x); }
RHD double m_exp(double x) { return elem::exp(x); }
RHD double m_third(double x) { return elem::powThird(x); }  // pow(x, 1.0 / 3.0)
RHD double m_acos(double x) { return elem::acos(x); }
RHD double m_cos(double x) { return elem::cos(x); }
RHD double m_sqrt(double x) { return ::sqrt(x); }

RHD bool fin(double x) { return x == x && (x < 0 ? -x : x) <= 1.7976931348623157e308; }
RHD double lo(double a, double b) { return a < b ? a : b; }
RHD double hi(double a, double b) { return a > b ? a : b; }
RHD double ab(double x) { return x < 0 ? -x : x; }
RHD double clamp(double a, double l, double h) { return lo(hi(a, l), h); }

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

struct PRState {  // subset of reactiveDevicePR.h:101-110 used by the flash
    double V, rho, e, h, s, cp, cv, alpha, kappaT, sound, p, dpdT, dpdV;
    int phase;
};

// NASA7 / NASA9 (reactiveDevicePR.h:127-155).
RHD void nasa(const Region& r, double T, double& cp, double& h, double& s) {
    const double* a = r.c;
    const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    const double invT = 1.0 / T, logT = m_log(T);
    if (r.poly == 7) {
        cp = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
        h = a[0] + .5 * a[1] * T + a[2] * T2 / 3.0 + .25 * a[3] * T3 + .2 * a[4] * T4 + a[5] * invT;
        s = a[0] * logT + a[1] * T + .5 * a[2] * T2 + a[3] * T3 / 3.0 + .25 * a[4] * T4 + a[6];
    } else {
        const double c0 = a[0] * invT * invT, c1 = a[1] * invT, c2 = a[2], c3 = a[3] * T;
        const double c4 = a[4] * T2, c5 = a[5] * T3, c6 = a[6] * T4;
        cp = c0 + c1 + c2 + c3 + c4 + c5 + c6;
        h = -c0 + logT * c1 + c2 + .5 * c3 + c4 / 3.0 + .25 * c5 + .2 * c6 + a[7] * invT;
        s = -.5 * c0 - c1 + logT * c2 + c3 + .5 * c4 + c5 / 3.0 + .25 * c6 + a[8];
    }
}

// NASA9 upper region at a shared bound, NASA7 lower (reactiveDevicePR.h:197-212).
template <int N>
RHD const Region& region(const Table<N>& t, int k, double T) {
    int r = 0;
    for (int j = 1; j < t.nreg[k]; ++j) {
        const double bound = t.region[k][j].tmin;
        const bool upper = t.region[k][j].poly == 9 ? T >= bound : T > bound;
        if (!upper) break;
        r = j;
    }
    return t.region[k][r];
}

struct Mix { double a, b, aa, da, d2a; };

// Mixing rule (reactiveDevicePR.h:214-247).
template <int N>
RHD int mixing(const Table<N>& t, double T, const double* x, Mix& m) {
    double q[N], dq[N], d2q[N];
    m.a = 0; m.b = 0; m.aa = 0; m.da = 0; m.d2a = 0;
    for (int i = 0; i < N; ++i) {
        const double Tc = t.Tc[i];
        if (!(Tc > 0) || !fin(Tc)) return kBadTable;
        const double sqrtTr = m_sqrt(T / Tc);
        const double f = 1.0 + t.kappa[i] * (1.0 - sqrtTr);
        if (ab(f) <= 1e-14) return kNonFinite;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[i] = ab(f);
        dq[i] = -sign * t.kappa[i] / (2.0 * Tc * sqrtTr);
        d2q[i] = sign * t.kappa[i] / (4.0 * Tc * Tc * sqrtTr * sqrtTr * sqrtTr);
        m.b += x[i] * t.b[i];
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const double w = x[i] * x[j] * t.aij[i][j];
            m.a += w;
            m.aa += w * q[i] * q[j];
            m.da += w * (dq[i] * q[j] + q[i] * dq[j]);
            m.d2a += w * (d2q[i] * q[j] + 2.0 * dq[i] * dq[j] + q[i] * d2q[j]);
        }
    return fin(m.a) && fin(m.b) && fin(m.aa) && fin(m.da) && fin(m.d2a) && m.a > 0 && m.b > 0 && m.aa > 0 ? kOk
                                                                                                         : kNonFinite;
}

RHD double cubic(double V, double bn, double cn, double dn) { return ((V + bn) * V + cn) * V + dn; }
RHD double cbrtSigned(double x) { return x < 0 ? -m_third(-x) : m_third(x); }

// Isolated third root in the near-repeated discriminant band (reactiveDevicePR.h:280-307).
RHD bool isolatedRoot(double center, double delta, double delta2, double q, double bn, double cn, double dn,
                      double b, double RTp, double aap, double& root, bool& gasSide) {
    if (!(delta2 > 0) || !fin(delta) || !fin(q) || q == 0) return false;
    gasSide = q < 0;
    const double pair = center + (gasSide ? -1.0 : 1.0) * delta;
    double x = center + (gasSide ? 2.0 : -2.0) * delta;
    bool converged = false;
    for (int n = 0; n < 24; ++n) {
        const double f = cubic(x, bn, cn, dn), df = (3.0 * x + 2.0 * bn) * x + cn;
        if (!fin(f) || !fin(df) || ab(df) <= 1e-300) return false;
        const double step = f / df;
        x -= step;
        if (ab(step) <= 4e-15 * hi(1.0, ab(x))) { converged = true; break; }
    }
    if (!converged || !fin(x) || !(x > b)) return false;
    const double f = cubic(x, bn, cn, dn);
    const double scale = ab(x * x * x) + ab(bn * x * x) + ab(cn * x) + ab(dn);
    if (!(ab(f) <= 64.0 * DBL_EPSILON * scale)) return false;
    if (!(ab(x - pair) > 1e-3 * hi(ab(x), ab(pair)))) return false;
    if (gasSide ? !(x > pair) : !(x < pair)) return false;
    const double vb = x - b, den = x * x + 2.0 * b * x - b * b;
    if (!(vb > 0) || !(den > 0)) return false;
    if (!(-RTp / (vb * vb) + 2.0 * aap * (x + b) / (den * den) < 0)) return false;
    root = x;
    return true;
}

struct Roots { double v0, vLast; int count, signedCount, gas, liquid; };

// Cubic root set at (T, p) from the mixture coefficients
// (rootsFromMixingCached, reactiveDevicePR.h:701-790).  Only the smallest and
// largest physical roots are kept: the flash selects nothing else.
RHD int roots(double T, double p, double a, double b, double aa, Roots& out) {
    if (!fin(T) || !fin(p) || !fin(a) || !fin(b) || !fin(aa) || T <= 0 || p <= 0 || a <= 0 || b <= 0 || aa <= 0)
        return kBadInput;
    const double b2 = b * b;
    const double RTp = kR * T / p, aap = aa / p;
    const double bn = b - RTp;
    const double cn = -(2.0 * RTp * b - aap + 3.0 * b2);
    const double dn = b2 * RTp + b2 * b - aap * b;
    const double center = -bn / 3.0;
    const double delta2 = (bn * bn - 3.0 * cn) / 9.0;
    const double q = 2.0 * bn * bn * bn / 27.0 - bn * cn / 3.0 + dn;
    const double delta = delta2 > 0 ? m_sqrt(delta2) : 0;
    const double h = 2.0 * delta * delta2;
    double disc = q * q - h * h;
    if (ab(ab(h) - ab(q)) < 1e-10) {
        if (disc > 1e-10) return kNonFinite;
        disc = 0;
    }
    double raw[3] = {0, 0, 0};
    int n = 0, isolated = 0;
    if (ab(disc) < 1e-14) {
        bool gasSide = false;
        if (!isolatedRoot(center, delta, delta2, q, bn, cn, dn, b, RTp, aap, raw[0], gasSide)) return kNoRoot;
        n = 1;
        isolated = gasSide ? 1 : -1;
    }
    if (isolated) {
    } else if (disc > 1e-14 || !(delta2 > 0)) {
        const double sd = .5 * m_sqrt(hi(0.0, disc));
        raw[0] = center + cbrtSigned(-.5 * q + sd) + cbrtSigned(-.5 * q - sd);
        n = 1;
    } else if (disc < -1e-14) {
        double arg = -.5 * q / (delta * delta * delta);
        arg = hi(-1.0, lo(1.0, arg));
        const double theta = m_acos(arg) / 3.0;
        raw[0] = center + 2.0 * delta * m_cos(theta);
        raw[1] = center + 2.0 * delta * m_cos(theta + 2.0 * kPi / 3.0);
        raw[2] = center + 2.0 * delta * m_cos(theta + 4.0 * kPi / 3.0);
        n = 3;
    } else {
        return kNoRoot;
    }
    if (!isolated)
        for (int i = 0; i < n; ++i)
            for (int it = 0; it < 12; ++it) {
                const double res = cubic(raw[i], bn, cn, dn);
                const double der = (3.0 * raw[i] + 2.0 * bn) * raw[i] + cn;
                if (ab(der) <= 1e-300) break;
                const double step = res / der;
                raw[i] -= step;
                if (ab(step) <= 4e-15 * hi(1.0, ab(raw[i]))) break;
            }
    // ascending order, finite roots above the covolume (reactiveDevicePR.h:253-269)
    for (int i = 1; i < n; ++i) {
        const double key = raw[i];
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
    const double Tcrit = a * kOmegaB / (b * kOmegaA * kR);
    const double Pcrit = kOmegaB * kR * Tcrit / b;
    const double Vcrit = kOmegaVc * kR * Tcrit / Pcrit;
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

template <int N>
RHD int massToMole(const Table<N>& t, const double* Y, double* x, double& W) {  // :310-327
    double invW = 0, sumY = 0;
    for (int k = 0; k < N; ++k) {
        if (!fin(Y[k]) || Y[k] < 0) return kBadComposition;
        sumY += Y[k];
        invW += Y[k] / t.W[k];
    }
    if (!(invW > 0) || ab(sumY - 1.0) > 1e-12) return kBadComposition;
    W = 1.0 / invW;
    for (int k = 0; k < N; ++k) x[k] = Y[k] * W / t.W[k];
    return kOk;
}

// Composition part of Cantera's phaseState heuristic (reactiveDevicePR.h:
// 428-471): everything but the final rho < rhoMid(T) comparison.
struct PhaseHeuristic {
    double tc, critDensity, densityDelta, tempDelta;
    int status;  // status of the tmid root loop (reported only when T < tc)
};

template <int N>
RHD void phaseHeuristic(const Table<N>& t, const double* x, double W, double mixA, double mixB, PhaseHeuristic& h) {
    const double tc = mixA * kOmegaB / (mixB * kOmegaA * kR);
    const double pc = kOmegaB * kR * tc / mixB;
    const double vc = kOmegaVc * kR * tc / pc;
    h.tc = tc;
    double tmid = tc - 100.0;
    if (tmid < 0) tmid = .5 * tc;
    const double ratio = tc / tmid;
    double pp = pc * m_exp(-.8734 * ratio * ratio - 3.4522 * ratio + 4.2918);
    Mix mm;
    int st = mixing(t, tmid, x, mm);
    if (st != kOk) { h.status = st; return; }
    double liquidVolume = -1;
    bool found = false;
    for (int it = 0; it < 100; ++it) {
        Roots r;
        st = roots(tmid, pp, mm.a, mm.b, mm.aa, r);
        if (st != kOk) { h.status = st; return; }
        if (r.signedCount == 1 || r.signedCount == 2) {
            found = pp > pc;
            liquidVolume = r.v0;
            pp *= 1.04;
            if (found) break;
        } else {
            liquidVolume = r.v0;
            found = true;
            break;
        }
    }
    if (!found || !(liquidVolume > 0) || !fin(liquidVolume)) { h.status = kNonFinite; return; }
    const double critDensity = W / vc;
    const double liquidMid = W / liquidVolume;
    const double gasMid = W / (kR * tmid / pp);
    const double midDensity = .5 * (liquidMid + gasMid);
    h.critDensity = critDensity;
    h.densityDelta = critDensity - midDensity;
    h.tempDelta = tc - tmid;
    h.status = kOk;
}

// Final comparison: rhoMid = rho_c + (T - Tc) (rho_c - rho_mid) / (Tc - Tmid).
RHD int reportedPhase(const PhaseHeuristic& h, double T, double rho, int& phase) {
    if (T >= h.tc) { phase = -2; return kOk; }
    if (h.status != kOk) return h.status;
    const double rhoMid = h.critDensity + (T - h.tc) * h.densityDelta / h.tempDelta;
    phase = rho < rhoMid ? -1 : 0;
    return fin(rhoMid) ? kOk : kNonFinite;
}

// Full PR state at (T, rho) from mixed coefficients
// (evaluateTrhoFromMixingCached, reactiveDevicePR.h:840-930).  mu (J/kmol)
// is computed for every species when mu != nullptr.
// The phaseState comparison and the final finiteness checks follow in
// finishState(), in the reference's order.
template <int N>
RHD int stateTrho(const Table<N>& t, double T, double rho, const double* x, double W, const Mix& mx,
                  PRState& s, double* mu) {
    if (!fin(T) || !fin(rho) || T <= 0 || rho <= 0) return kBadInput;
    const double b = mx.b, aa = mx.aa, da = mx.da, d2a = mx.d2a;
    s.rho = rho;
    s.V = W / rho;
    if (!(s.V > b)) return kNoRoot;
    const double den = s.V * s.V + 2.0 * s.V * b - b * b;
    const double vmb = s.V - b;
    if (!(den > 0) || !(vmb > 0)) return kNoRoot;
    s.p = kR * T / vmb - aa / den;
    if (!fin(s.p) || s.p <= 0) return kNoRoot;
    const double dpdV = -kR * T / (vmb * vmb) + 2.0 * aa * (s.V + b) / (den * den);
    const double dpdT = kR / vmb - da / den;
    s.dpdV = dpdV;
    s.dpdT = dpdT;
    if (!fin(dpdV) || !fin(dpdT) || dpdV >= 0) return kUnstable;
    const double logRatio = m_log((s.V + (1.0 + kSqrt2) * b) / (s.V + (1.0 - kSqrt2) * b));
    const double L = logRatio / (2.0 * kSqrt2 * b);
    double h0 = 0, s0 = 0, cp0 = 0, xlogx = 0;
    double hRT[N], sR[N];
    for (int k = 0; k < N; ++k) {
        double cpR, hk, sk;
        nasa(region(t, k, T), T, cpR, hk, sk);
        hRT[k] = hk; sR[k] = sk;
        cp0 += x[k] * kR * cpR;
        h0 += x[k] * kR * T * hk;
        s0 += x[k] * kR * sk;
        if (x[k] > 0) xlogx += x[k] * m_log(x[k]);
    }
    const double z = s.p * s.V / (kR * T);
    const double hRes = kR * T * (z - 1.0) + L * (T * da - aa);
    const double sRes = kR * m_log(z * (1.0 - b / s.V)) + L * da;
    const double hM = h0 + hRes;
    const double logP = m_log(s.p / t.refP);
    const double sM = s0 - kR * xlogx - kR * logP + sRes;
    const double dHdT = cp0 + s.V * dpdT - kR + L * T * d2a;
    const double cpM = dHdT - (s.V + T * dpdT / dpdV) * dpdT;
    const double cvM = cpM + T * dpdT * dpdT / dpdV;
    s.h = hM / W; s.e = (hM - s.p * s.V) / W; s.s = sM / W;
    s.cp = cpM / W; s.cv = cvM / W;
    s.alpha = -dpdT / (s.V * dpdV);
    s.kappaT = -1.0 / (s.V * dpdV);
    const double sound2 = s.V * s.V * (-cpM / cvM * dpdV / W);
    if (!(s.cp > 0) || !(s.cv > 0) || !(s.kappaT > 0) || !(sound2 > 0)) return kUnstable;
    s.sound = m_sqrt(sound2);
    if (mu) {
        const double denomMu = 2.0 * kSqrt2 * b * b;
        const double denomMu2 = b * den;
        const double RT = kR * T;
        double f[N];
        for (int i = 0; i < N; ++i) f[i] = 1.0 + t.kappa[i] * (1.0 - m_sqrt(T / t.Tc[i]));
        for (int k = 0; k < N; ++k) {
            double pp = 0;
            for (int i = 0; i < N; ++i) pp += x[i] * t.aij[k][i] * ab(f[i] * f[k]);
            const double bk = t.b[k];
            const double num = 2.0 * b * pp - aa * bk;
            mu[k] = RT * (hRT[k] - sR[k]) + RT * m_log(hi(kSmall, x[k])) + RT * logP - RT * m_log(s.p * s.V / RT) +
                    RT * m_log(s.V / vmb) + RT * bk / vmb - num / denomMu * logRatio - aa * bk * s.V / denomMu2;
        }
    }
    return kOk;
}

template <int N>
RHD int finishState(const PhaseHeuristic& ph, double T, PRState& s, const double* mu) {
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
