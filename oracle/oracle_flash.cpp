// SPDX-License-Identifier: GPL-3.0-or-later
// CPU ORACLE — test infrastructure only (see oracle.h).  Serial FP64
// restatement of the reference's Peng-Robinson thermodynamics and HEM UV
// flash.  Every function names the reference lines it follows; expression
// grouping follows the reference so that, compiled with -ffp-contract=off
// against glibc, results equal the compiled reference bit for bit.
//
// Scope: pure-liquid condensables (condensedKind 0); a capillary interface
// (color, pressureJump: the gas at p - color J, the liquids at
// p + (1 - color) J) and the bounded vapour-log ("adaptive") boundary search
// of boundaryRecovery = 1 models and capillary equilibrium cells
// (reactiveDeviceFlash.h:327-448,544-590); the finite-difference Jacobian or,
// with analyticJacobian = 1, the analytic one for one active liquid
// (reactiveDeviceFlashJacobian.h).  Other model images are rejected by
// oracle_model_supported().
#include "oracle.h"
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef ORACLE_TWIN
#include "../spuma-multiphysics_amd/csrc/elementary.h"
#endif

namespace oracle {
namespace {

// Elementary functions.  The default build calls glibc, as the reference's
// host build does; the twin build (ORACLE_TWIN) calls the exact-sequence
// functions the device flash uses (csrc/elementary.h), so that the device
// flash can be compared with this restatement bit for bit.
#ifdef ORACLE_TWIN
inline double olog(double x) { return ramd::elem::log(x); }
inline double oexp(double x) { return ramd::elem::exp(x); }
inline double othird(double x) { return ramd::elem::powThird(x); }
inline double oacos(double x) { return ramd::elem::acos(x); }
inline double ocos(double x) { return ramd::elem::cos(x); }
inline double oexpm1(double x) { return ramd::elem::expm1(x); }
#else
inline double olog(double x) { return std::log(x); }
inline double oexp(double x) { return std::exp(x); }
inline double othird(double x) { return std::pow(x, 1.0 / 3.0); }
inline double oacos(double x) { return std::acos(x); }
inline double ocos(double x) { return std::cos(x); }
inline double oexpm1(double x) { return std::expm1(x); }
#endif

constexpr double R = 8314.46261815324;            // reactiveDevicePR.h:31
constexpr double kSmall = 1e-300;                 // :32
constexpr double kSqrt2 = 1.41421356237309504880168872420969808;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kOmegaA = 4.5723552892138218e-1; // :35-37
constexpr double kOmegaB = 7.77960739038885e-2;
constexpr double kOmegaVc = 3.07401308698703833e-1;

enum Status { OK = 0, BadTable, BadInput, BadComposition, NoRoot, Unavailable, Unstable, NonFinite };
enum Root { AutoRoot = 0, GasRoot = 1, LiquidRoot = 2 };

inline bool fin(double x) { return x == x && (x < 0 ? -x : x) <= 1.7976931348623157e308; }
inline double lo(double a, double b) { return a < b ? a : b; }
inline double hi(double a, double b) { return a > b ? a : b; }
inline double ab(double x) { return x < 0 ? -x : x; }
inline double clamp(double a, double l, double h) { return lo(hi(a, l), h); }

struct PRState {  // reactiveDevicePR.h:101-110
    double T, p, W, V, rho, e, h, s, cp, cv, alpha, kappaT, sound, dpdT, dpdV;
    double mu[kMaxSpecies];
    int rootCount, selectedRoot, phase;
};
struct Roots {  // reactiveDevicePR.h:88-99
    double volume[3];
    int count, signedCount, gasAvailable, liquidAvailable;
    double a, b, aa;
};

// NASA7/NASA9 reference-state polynomials (reactiveDevicePR.h:127-155).
void nasa(const NasaRegion& r, double T, double& cp, double& h, double& s) {
    const double* a = r.c;
    const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
    const double invT = 1.0 / T, logT = olog(T);
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

// Region selection: NASA9 takes the upper region at a shared bound (:197-212).
const NasaRegion& region(const Table& t, int k, double T) {
    const Species& sp = t.sp[k];
    int r = 0;
    for (int j = 1; j < sp.nreg; ++j) {
        const double bound = sp.region[j].tmin;
        const bool upper = sp.region[j].poly == 9 ? T >= bound : T > bound;
        if (!upper) break;
        r = j;
    }
    return sp.region[r];
}

// PR mixing rule with alpha = |1 + kappa (1 - sqrt(T/Tc))|^2 (:214-247).
Status mixing(const Table& t, double T, const double* x, double& a, double& b, double& aa, double& da,
              double& d2a) {
    double q[kMaxSpecies], dq[kMaxSpecies], d2q[kMaxSpecies];
    a = 0; b = 0; aa = 0; da = 0; d2a = 0;
    for (int i = 0; i < t.ns; ++i) {
        const Species& sp = t.sp[i];
        const double Tc = sp.a * kOmegaB / (sp.b * kOmegaA * R);
        if (!(Tc > 0) || !fin(Tc)) return BadTable;
        const double sqrtTr = std::sqrt(T / Tc);
        const double f = 1.0 + sp.kappa * (1.0 - sqrtTr);
        if (ab(f) <= 1e-14) return NonFinite;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[i] = ab(f);
        dq[i] = -sign * sp.kappa / (2.0 * Tc * sqrtTr);
        d2q[i] = sign * sp.kappa / (4.0 * Tc * Tc * sqrtTr * sqrtTr * sqrtTr);
        b += x[i] * sp.b;
    }
    for (int i = 0; i < t.ns; ++i)
        for (int j = 0; j < t.ns; ++j) {
            const double w = x[i] * x[j] * t.aij[i][j];
            a += w;
            aa += w * q[i] * q[j];
            da += w * (dq[i] * q[j] + q[i] * dq[j]);
            d2a += w * (d2q[i] * q[j] + 2.0 * dq[i] * dq[j] + q[i] * d2q[j]);
        }
    return fin(a) && fin(b) && fin(aa) && fin(da) && fin(d2a) && a > 0 && b > 0 && aa > 0 ? OK : NonFinite;
}

inline double cubic(double V, double bn, double cn, double dn) { return ((V + bn) * V + cn) * V + dn; }
inline double cbrtSigned(double x) { return x < 0 ? -othird(-x) : othird(x); }

// Separated third root inside the near-repeated discriminant band (:280-307).
bool isolatedRoot(double center, double delta, double delta2, double q, double bn, double cn, double dn, double b,
                  double RTp, double aap, double& root, bool& gasSide) {
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

// Cubic in molar volume at (T, p) from mixed a, b, a*alpha
// (rootsFromMixingCached, :701-790; identical arithmetic to rootsTP :332-422).
Status roots(double T, double p, double a, double b, double aa, Roots& out) {
    if (!fin(T) || !fin(p) || !fin(a) || !fin(b) || !fin(aa) || T <= 0 || p <= 0 || a <= 0 || b <= 0 || aa <= 0)
        return BadInput;
    Roots r{};
    r.a = a; r.b = b; r.aa = aa;
    const double b2 = b * b;
    const double RTp = R * T / p, aap = aa / p;
    const double bn = b - RTp;
    const double cn = -(2.0 * RTp * b - aap + 3.0 * b2);
    const double dn = b2 * RTp + b2 * b - aap * b;
    const double center = -bn / 3.0;
    const double delta2 = (bn * bn - 3.0 * cn) / 9.0;
    const double q = 2.0 * bn * bn * bn / 27.0 - bn * cn / 3.0 + dn;
    const double delta = delta2 > 0 ? std::sqrt(delta2) : 0;
    const double h = 2.0 * delta * delta2;
    double disc = q * q - h * h;
    if (ab(ab(h) - ab(q)) < 1e-10) {
        if (disc > 1e-10) return NonFinite;
        disc = 0;
    }
    double raw[3] = {0, 0, 0};
    int n = 0, isolated = 0;
    if (ab(disc) < 1e-14) {
        bool gasSide = false;
        if (!isolatedRoot(center, delta, delta2, q, bn, cn, dn, b, RTp, aap, raw[0], gasSide)) return NoRoot;
        n = 1;
        isolated = gasSide ? 1 : -1;
    }
    if (isolated) {
    } else if (disc > 1e-14 || !(delta2 > 0)) {
        const double sd = .5 * std::sqrt(hi(0.0, disc));
        raw[0] = center + cbrtSigned(-.5 * q + sd) + cbrtSigned(-.5 * q - sd);
        n = 1;
    } else if (disc < -1e-14) {
        double arg = -.5 * q / (delta * delta * delta);
        arg = hi(-1.0, lo(1.0, arg));
        const double theta = oacos(arg) / 3.0;
        raw[0] = center + 2.0 * delta * ocos(theta);
        raw[1] = center + 2.0 * delta * ocos(theta + 2.0 * kPi / 3.0);
        raw[2] = center + 2.0 * delta * ocos(theta + 4.0 * kPi / 3.0);
        n = 3;
    } else {
        return NoRoot;
    }
    for (int i = 0; !isolated && i < n; ++i)
        for (int it = 0; it < 12; ++it) {
            const double res = cubic(raw[i], bn, cn, dn);
            const double der = (3.0 * raw[i] + 2.0 * bn) * raw[i] + cn;
            if (ab(der) <= 1e-300) break;
            const double step = res / der;
            raw[i] -= step;
            if (ab(step) <= 4e-15 * hi(1.0, ab(raw[i]))) break;
        }
    // ascending insertion sort, keep finite roots above the covolume (:253-269)
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
    r.count = m;
    if (!m) return NoRoot;
    for (int i = 0; i < m; ++i) r.volume[i] = raw[i];
    const double Tcrit = r.a * kOmegaB / (b * kOmegaA * R);
    const double Pcrit = kOmegaB * R * Tcrit / b;
    const double Vcrit = kOmegaVc * R * Tcrit / Pcrit;
    if (isolated) {
        r.signedCount = isolated;
        r.gasAvailable = isolated > 0;
        r.liquidAvailable = isolated < 0;
    } else if (r.count >= 2) {
        r.signedCount = r.count;
        r.gasAvailable = r.liquidAvailable = 1;
    } else {
        const bool liquid = T > Tcrit ? r.volume[0] < Vcrit : r.volume[0] < center;
        r.signedCount = liquid ? -1 : 1;
        r.liquidAvailable = liquid;
        r.gasAvailable = !liquid || T > Tcrit;
    }
    out = r;
    return OK;
}

Status massToMole(const Table& t, const double* Y, double* x, double& W) {  // :310-327
    double invW = 0, sumY = 0;
    for (int k = 0; k < t.ns; ++k) {
        if (!fin(Y[k]) || Y[k] < 0) return BadComposition;
        sumY += Y[k];
        invW += Y[k] / t.sp[k].W;
    }
    if (!(invW > 0) || ab(sumY - 1.0) > 1e-12) return BadComposition;
    W = 1.0 / invW;
    for (int k = 0; k < t.ns; ++k) x[k] = Y[k] * W / t.sp[k].W;
    return OK;
}

// Cantera's phaseState density heuristic (reportedPhase, :428-471).
Status reportedPhase(const Table& t, double T, const double* x, double W, double rho, double mixA, double mixB,
                     int& phase) {
    const double tc = mixA * kOmegaB / (mixB * kOmegaA * R);
    const double pc = kOmegaB * R * tc / mixB;
    const double vc = kOmegaVc * R * tc / pc;
    if (T >= tc) { phase = -2; return OK; }
    double tmid = tc - 100.0;
    if (tmid < 0) tmid = .5 * tc;
    const double ratio = tc / tmid;
    double pp = pc * oexp(-.8734 * ratio * ratio - 3.4522 * ratio + 4.2918);
    double ma, mb, maa, mda, md2a;
    Status st = mixing(t, tmid, x, ma, mb, maa, mda, md2a);
    if (st != OK) return st;
    double liquidVolume = -1;
    bool found = false;
    for (int it = 0; it < 100; ++it) {
        Roots r{};
        st = roots(tmid, pp, ma, mb, maa, r);
        if (st != OK) return st;
        if (r.signedCount == 1 || r.signedCount == 2) {
            found = pp > pc;
            liquidVolume = r.volume[0];
            pp *= 1.04;
            if (found) break;
        } else {
            liquidVolume = r.volume[0];
            found = true;
            break;
        }
    }
    if (!found || !(liquidVolume > 0) || !fin(liquidVolume)) return NonFinite;
    const double critDensity = W / vc;
    const double liquidMid = W / liquidVolume;
    const double gasMid = W / (R * tmid / pp);
    const double midDensity = .5 * (liquidMid + gasMid);
    const double rhoMid = critDensity + (T - tc) * (critDensity - midDensity) / (tc - tmid);
    phase = rho < rhoMid ? -1 : 0;
    return fin(rhoMid) ? OK : NonFinite;
}

// Full PR state at (T, rho) from mixed coefficients
// (evaluateTrhoTrusted :473-562 / evaluateTrhoFromMixingCached :840-930).
Status stateTrho(const Table& t, double T, double rho, const double* x, double W, double a, double b, double aa,
                 double da, double d2a, PRState& out, bool withMu) {
    PRState s{};
    s.T = T; s.W = W; s.rho = rho; s.V = W / rho;
    if (!(s.V > b)) return NoRoot;
    const double den = s.V * s.V + 2.0 * s.V * b - b * b;
    const double vmb = s.V - b;
    if (!(den > 0) || !(vmb > 0)) return NoRoot;
    s.p = R * T / vmb - aa / den;
    if (!fin(s.p) || s.p <= 0) return NoRoot;
    s.dpdV = -R * T / (vmb * vmb) + 2.0 * aa * (s.V + b) / (den * den);
    s.dpdT = R / vmb - da / den;
    if (!fin(s.dpdV) || !fin(s.dpdT) || s.dpdV >= 0) return Unstable;
    const double logRatio = olog((s.V + (1.0 + kSqrt2) * b) / (s.V + (1.0 - kSqrt2) * b));
    const double L = logRatio / (2.0 * kSqrt2 * b);
    double h0 = 0, s0 = 0, cp0 = 0, xlogx = 0;
    double hRT[kMaxSpecies], sR[kMaxSpecies];
    for (int k = 0; k < t.ns; ++k) {
        double cpR, hk, sk;
        nasa(region(t, k, T), T, cpR, hk, sk);
        hRT[k] = hk; sR[k] = sk;
        cp0 += x[k] * R * cpR;
        h0 += x[k] * R * T * hk;
        s0 += x[k] * R * sk;
        if (x[k] > 0) xlogx += x[k] * olog(x[k]);
    }
    const double z = s.p * s.V / (R * T);
    const double hRes = R * T * (z - 1.0) + L * (T * da - aa);
    const double sRes = R * olog(z * (1.0 - b / s.V)) + L * da;
    const double hM = h0 + hRes;
    const double sM = s0 - R * xlogx - R * olog(s.p / t.refP) + sRes;
    const double dHdT = cp0 + s.V * s.dpdT - R + L * T * d2a;
    const double cpM = dHdT - (s.V + T * s.dpdT / s.dpdV) * s.dpdT;
    const double cvM = cpM + T * s.dpdT * s.dpdT / s.dpdV;
    s.h = hM / W; s.e = (hM - s.p * s.V) / W; s.s = sM / W;
    s.cp = cpM / W; s.cv = cvM / W;
    s.alpha = -s.dpdT / (s.V * s.dpdV);
    s.kappaT = -1.0 / (s.V * s.dpdV);
    const double sound2 = s.V * s.V * (-cpM / cvM * s.dpdV / W);
    if (!(s.cp > 0) || !(s.cv > 0) || !(s.kappaT > 0) || !(sound2 > 0)) return Unstable;
    s.sound = std::sqrt(sound2);
    if (withMu) {
        const double denomMu = 2.0 * kSqrt2 * b * b;
        const double denomMu2 = b * den;
        for (int k = 0; k < t.ns; ++k) {
            double pp = 0;
            for (int i = 0; i < t.ns; ++i) {
                const double Tci = t.sp[i].a * kOmegaB / (t.sp[i].b * kOmegaA * R);
                const double Tck = t.sp[k].a * kOmegaB / (t.sp[k].b * kOmegaA * R);
                const double fi = 1.0 + t.sp[i].kappa * (1.0 - std::sqrt(T / Tci));
                const double fk = 1.0 + t.sp[k].kappa * (1.0 - std::sqrt(T / Tck));
                pp += x[i] * t.aij[k][i] * ab(fi * fk);
            }
            const double bk = t.sp[k].b;
            const double num = 2.0 * b * pp - aa * bk;
            const double RT = R * T;
            s.mu[k] = RT * (hRT[k] - sR[k]) + RT * olog(hi(kSmall, x[k])) + RT * olog(s.p / t.refP) -
                      RT * olog(s.p * s.V / RT) + RT * olog(s.V / vmb) + RT * bk / vmb -
                      num / denomMu * logRatio - aa * bk * s.V / denomMu2;
        }
    }
    Status st = reportedPhase(t, T, x, W, rho, a, b, s.phase);
    if (st != OK) return st;
    const double values[] = {s.rho, s.e, s.h, s.s, s.cp, s.cv, s.alpha, s.kappaT, s.sound, s.p, s.dpdT, s.dpdV};
    for (double v : values)
        if (!fin(v)) return NonFinite;
    if (withMu)
        for (int k = 0; k < t.ns; ++k)
            if (!fin(s.mu[k])) return NonFinite;
    out = s;
    return OK;
}

// PR state at (T, rho, Y) (evaluateTrhoTrusted, :473-562).
Status evaluateTrho(const Table& t, double T, double rho, const double* Y, PRState& out, bool withMu) {
    if (!fin(T) || !fin(rho) || T <= 0 || rho <= 0) return BadInput;
    double x[kMaxSpecies] = {}, W = 0;
    Status st = massToMole(t, Y, x, W);
    if (st != OK) return st;
    double a, b, aa, da, d2a;
    st = mixing(t, T, x, a, b, aa, da, d2a);
    if (st != OK) return st;
    return stateTrho(t, T, rho, x, W, a, b, aa, da, d2a, out, withMu);
}

// PR state at (T, p, Y) on the requested cubic root (evaluateTPTrusted, :574-609).
Status evaluateTP(const Table& t, double T, double p, const double* Y, Root choice, PRState& out, bool withMu) {
    if (!fin(T) || !fin(p) || T <= 0 || p <= 0) return BadInput;
    double x[kMaxSpecies] = {}, W = 0;
    Status st = massToMole(t, Y, x, W);
    if (st != OK) return st;
    double a, b, aa, da, d2a;
    st = mixing(t, T, x, a, b, aa, da, d2a);
    if (st != OK) return st;
    Roots r{};
    st = roots(T, p, a, b, aa, r);
    if (st != OK) return st;
    int selected = r.count - 1;
    if (choice == LiquidRoot) {
        if (!r.liquidAvailable) return Unavailable;
        selected = 0;
    } else if (choice == GasRoot) {
        if (!r.gasAvailable) return Unavailable;
        selected = r.count - 1;
    }
    PRState s{};
    st = stateTrho(t, T, W / r.volume[selected], x, W, a, b, aa, da, d2a, s, withMu);
    if (st != OK) return st;
    if (ab(s.p - p) > 1e-7 * hi(1.0, p)) return NonFinite;
    s.p = p;
    s.rootCount = r.count;
    s.selectedRoot = selected;
    if (choice == GasRoot && s.phase >= 0) return Unavailable;
    if (choice == LiquidRoot && s.phase < 0) return Unavailable;
    out = s;
    return OK;
}

Status validateTable(const Table& t) {  // :157-195
    if (t.ns <= 0 || t.ns > kMaxSpecies || !fin(t.refP) || t.refP <= 0) return BadTable;
    for (int k = 0; k < t.ns; ++k) {
        const Species& sp = t.sp[k];
        if (!fin(sp.W) || sp.W <= 0 || !fin(sp.a) || sp.a <= 0 || !fin(sp.b) || sp.b <= 0 || !fin(sp.kappa) ||
            sp.nreg <= 0 || sp.nreg > kMaxRegions)
            return BadTable;
        for (int r = 0; r < sp.nreg; ++r) {
            const NasaRegion& g = sp.region[r];
            if ((g.poly != 7 && g.poly != 9) || !fin(g.tmin) || !fin(g.tmax) || g.tmin <= 0 || g.tmax <= g.tmin ||
                (r && g.tmin < sp.region[r - 1].tmin))
                return BadTable;
            for (int j = 0; j < g.poly; ++j)
                if (!fin(g.c[j])) return BadTable;
        }
        for (int j = 0; j < t.ns; ++j)
            if (!fin(t.aij[k][j]) || t.aij[k][j] != t.aij[j][k] || (k == j && (t.aij[k][j] <= 0 || t.aij[k][j] != sp.a)))
                return BadTable;
    }
    return OK;
}

// ---------------------------------------------------------------------------
// Analytic active-set Jacobian (reactiveDeviceFlashJacobian.h), for one
// active liquid in the flat (no-capillary), non-adaptive flash.

// Partial molar h, u, v of every gas species (phasePartials, :68-132).
bool phasePartials(const Table& t, double T, const double* Y, const PRState& state, double* hbar, double* ubar,
                   double* vbar, double* x, double& W) {
    if (massToMole(t, Y, x, W) != OK) return false;
    double a = 0, b = 0, aa = 0, da = 0, d2a = 0;
    if (mixing(t, T, x, a, b, aa, da, d2a) != OK) return false;
    const double V = state.V, p = state.p, RT = R * T;
    const double vmb = V - b;
    const double den = V * V + 2.0 * V * b - b * b;
    if (!(vmb > 0) || !(den > 0) || !(state.dpdV < 0)) return false;
    const double logRatio = olog((V + (1.0 + kSqrt2) * b) / (V + (1.0 - kSqrt2) * b));
    const double L = logRatio / (2.0 * kSqrt2 * b);
    const double fac = T * da - aa;
    const double fac2 = V + T * state.dpdT / state.dpdV;
    double q[kMaxSpecies]{}, dq[kMaxSpecies]{}, pp[kMaxSpecies]{}, tmp[kMaxSpecies]{};
    for (int k = 0; k < t.ns; ++k) {
        const Species& sp = t.sp[k];
        const double Tc = sp.a * kOmegaB / (sp.b * kOmegaA * R);
        const double sqrtTr = std::sqrt(T / Tc);
        const double f = 1.0 + sp.kappa * (1.0 - sqrtTr);
        if (ab(f) <= 1e-14) return false;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[k] = ab(f);
        dq[k] = -sign * sp.kappa / (2.0 * Tc * sqrtTr);
    }
    for (int k = 0; k < t.ns; ++k)
        for (int i = 0; i < t.ns; ++i) {
            const double pair = t.aij[k][i] * q[k] * q[i];
            pp[k] += x[i] * pair;
            tmp[k] += x[i] * pair * (2.0 * dq[i] / q[i] + 2.0 * dq[k] / q[k]);
        }
    for (int k = 0; k < t.ns; ++k) {
        double cpR = 0, hRT = 0, sR = 0;
        nasa(region(t, k, T), T, cpR, hRT, sR);
        const double bk = t.sp[k].b;
        const double dpdni = RT / vmb + RT * bk / (vmb * vmb) - 2.0 * pp[k] / den + 2.0 * vmb * aa * bk / (den * den);
        const double hEv = V * dpdni - RT - bk / (2.0 * kSqrt2 * b * b) * logRatio * fac + V * bk / (b * den) * fac +
                           L * (T * tmp[k] - 2.0 * pp[k]);
        hbar[k] = RT * hRT + hEv - fac2 * dpdni;
        const double volumeNumerator = RT + RT * b / vmb + RT * bk / vmb + RT * b * bk / (vmb * vmb) -
                                       2.0 * V * pp[k] / den + 2.0 * V * vmb * aa * bk / (den * den);
        const double volumeDenominator = p + RT * b / (vmb * vmb) + aa / den - 2.0 * V * (V + b) * aa / (den * den);
        vbar[k] = volumeNumerator / volumeDenominator;
        ubar[k] = hbar[k] - p * vbar[k];
        if (!fin(hbar[k]) || !fin(vbar[k]) || !fin(ubar[k])) return false;
    }
    return true;
}

// d(mu_k/RT)/dn_j at fixed (T, p) for k = j = the condensable
// (gasCompositionPartials, :136-198, one row and column).
bool gasCompositionPartial(const Table& t, double T, const double* x, double totalMoles, const PRState& state, int k,
                           double& dpsi) {
    if (!(totalMoles > 0)) return false;
    double a = 0, b = 0, aa = 0, da = 0, d2a = 0;
    if (mixing(t, T, x, a, b, aa, da, d2a) != OK) return false;
    const double V = state.V, RT = R * T;
    const double vmb = V - b, den = V * V + 2.0 * V * b - b * b;
    if (!(vmb > 0) || !(den > 0) || !(state.dpdV < 0)) return false;
    const double c1 = 1.0 + kSqrt2, c2 = 1.0 - kSqrt2;
    const double A = V + c1 * b, B = V + c2 * b;
    const double logRatio = olog(A / B);
    const double C = 2.0 * kSqrt2;
    double q[kMaxSpecies]{}, pp[kMaxSpecies]{};
    for (int i = 0; i < t.ns; ++i) {
        const Species& sp = t.sp[i];
        const double Tc = sp.a * kOmegaB / (sp.b * kOmegaA * R);
        q[i] = ab(1.0 + sp.kappa * (1.0 - std::sqrt(T / Tc)));
        if (!(q[i] > 1e-14)) return false;
    }
    for (int i = 0; i < t.ns; ++i)
        for (int j = 0; j < t.ns; ++j) pp[i] += x[j] * t.aij[i][j] * q[i] * q[j];
    const double Pb = RT / (vmb * vmb) + 2.0 * aa * vmb / (den * den);
    const int j = k;
    const double db = t.sp[j].b - b;
    const double daa = 2.0 * (pp[j] - aa);
    const double dV = -(Pb * db - daa / den) / state.dpdV;
    const double dDen = 2.0 * (V + b) * dV + 2.0 * (V - b) * db;
    const double dLogRatio = (dV + c1 * db) / A - (dV + c2 * db) / B;
    if (!(x[k] > 0)) return false;
    const double dxk = 1.0 - x[k];
    const double dpp = t.aij[k][j] * q[k] * q[j] - pp[k];
    const double bk = t.sp[k].b;
    const double num = 2.0 * b * pp[k] - aa * bk;
    const double dnum = 2.0 * db * pp[k] + 2.0 * b * dpp - daa * bk;
    const double K = num / (C * b * b);
    const double dK = (dnum - 2.0 * num * db / b) / (C * b * b);
    const double dResidual = -RT * dV / V + RT * (dV / V - (dV - db) / vmb) - RT * bk * (dV - db) / (vmb * vmb) -
                             dK * logRatio - K * dLogRatio +
                             (-aa * bk * V / (b * den)) * (daa / aa + dV / V - db / b - dDen / den);
    const double dPsiDirection = dxk / x[k] + dResidual / RT;
    dpsi = dPsiDirection / totalMoles;
    return fin(dpsi);
}

// The condensable's partial h, u, v and d(mu/RT)/dn in the gas, from the
// gas root's molar volume at (T, p) (activeGasDerivatives, :203-308, n = 1).
bool activeGasDerivatives(const Table& t, double T, double p, double molarVolume, const double* Y, double gasMass,
                          int k, double& hbar, double& ubar, double& vbar, double& dpsi) {
    if (!(T > 0) || !(p > 0) || !(molarVolume > 0) || !(gasMass > 0)) return false;
    double x[kMaxSpecies]{}, W = 0;
    if (massToMole(t, Y, x, W) != OK) return false;
    const double totalMoles = gasMass / W;
    const double V = molarVolume, RT = R * T;
    double q[kMaxSpecies]{}, dq[kMaxSpecies]{};
    double b = 0, aa = 0, da = 0;
    for (int i = 0; i < t.ns; ++i) {
        const Species& sp = t.sp[i];
        const double Tc = sp.a * kOmegaB / (sp.b * kOmegaA * R);
        const double sqrtTr = std::sqrt(T / Tc);
        const double f = 1.0 + sp.kappa * (1.0 - sqrtTr);
        if (!(Tc > 0) || !(sqrtTr > 0) || ab(f) <= 1e-14) return false;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[i] = ab(f);
        dq[i] = -sign * sp.kappa / (2.0 * Tc * sqrtTr);
        b += x[i] * sp.b;
    }
    for (int i = 0; i < t.ns; ++i)
        for (int j = 0; j < t.ns; ++j) {
            const double w = x[i] * x[j] * t.aij[i][j];
            aa += w * q[i] * q[j];
            da += w * (dq[i] * q[j] + q[i] * dq[j]);
        }
    const double vmb = V - b, den = V * V + 2.0 * V * b - b * b;
    if (!(b > 0) || !(aa > 0) || !(vmb > 0) || !(den > 0)) return false;
    const double calculatedP = RT / vmb - aa / den;
    const double dpdV = -RT / (vmb * vmb) + 2.0 * aa * (V + b) / (den * den);
    const double dpdT = R / vmb - da / den;
    if (!(dpdV < 0) || !fin(calculatedP) || !fin(dpdT) || ab(calculatedP - p) > 1e-7 * hi(1.0, p)) return false;
    const double c1 = 1.0 + kSqrt2, c2 = 1.0 - kSqrt2;
    const double A = V + c1 * b, B = V + c2 * b;
    const double logRatio = olog(A / B);
    const double L = logRatio / (2.0 * kSqrt2 * b);
    const double fac = T * da - aa;
    const double fac2 = V + T * dpdT / dpdV;
    if (!(x[k] > 0)) return false;
    double pp = 0, tmp = 0;
    for (int i = 0; i < t.ns; ++i) {
        const double pair = t.aij[k][i] * q[k] * q[i];
        pp += x[i] * pair;
        tmp += x[i] * pair * (2.0 * dq[i] / q[i] + 2.0 * dq[k] / q[k]);
    }
    double cpR = 0, hRT = 0, sR = 0;
    nasa(region(t, k, T), T, cpR, hRT, sR);
    const double bk = t.sp[k].b;
    const double dpdni = RT / vmb + RT * bk / (vmb * vmb) - 2.0 * pp / den + 2.0 * vmb * aa * bk / (den * den);
    const double hEv = V * dpdni - RT - bk / (2.0 * kSqrt2 * b * b) * logRatio * fac + V * bk / (b * den) * fac +
                       L * (T * tmp - 2.0 * pp);
    hbar = RT * hRT + hEv - fac2 * dpdni;
    const double volumeNumerator = RT + RT * b / vmb + RT * bk / vmb + RT * b * bk / (vmb * vmb) -
                                   2.0 * V * pp / den + 2.0 * V * vmb * aa * bk / (den * den);
    const double volumeDenominator = p + RT * b / (vmb * vmb) + aa / den - 2.0 * V * (V + b) * aa / (den * den);
    vbar = volumeNumerator / volumeDenominator;
    ubar = hbar - p * vbar;
    if (!fin(hbar) || !fin(ubar) || !fin(vbar)) return false;
    const double Pb = RT / (vmb * vmb) + 2.0 * aa * vmb / (den * den);
    const double C = 2.0 * kSqrt2;
    const int j = k;
    const double db = t.sp[j].b - b;
    const double daa = 2.0 * (pp - aa);
    const double dV = -(Pb * db - daa / den) / dpdV;
    const double dDen = 2.0 * (V + b) * dV + 2.0 * (V - b) * db;
    const double dLogRatio = (dV + c1 * db) / A - (dV + c2 * db) / B;
    const double dxk = 1.0 - x[k];
    const double dpp = t.aij[k][j] * q[k] * q[j] - pp;
    const double num = 2.0 * b * pp - aa * bk;
    const double dnum = 2.0 * db * pp + 2.0 * b * dpp - daa * bk;
    const double K = num / (C * b * b);
    const double dK = (dnum - 2.0 * num * db / b) / (C * b * b);
    const double dResidual = -RT * dV / V + RT * (dV / V - (dV - db) / vmb) - RT * bk * (dV - db) / (vmb * vmb) -
                             dK * logRatio - K * dLogRatio +
                             (-aa * bk * V / (b * den)) * (daa / aa + dV / V - db / b - dDen / den);
    dpsi = (dxk / x[k] + dResidual / RT) / totalMoles;
    return fin(dpsi);
}

// ---------------------------------------------------------------------------
// HEM UV flash (reactiveDeviceFlash.h)
struct Evaluation {  // :64-73
    ThermoState state;
    double volume, energy, entropyDensity, Vp, VT, Ep, ET, cpDensity;
    double muGas[2], muLiquid[2], gasCondensable[2];
    bool explicitGas;  // the gas inventory was given (vapour coordinates), not q - liquid
    double liquidHbar[2], liquidVbar[2];
    unsigned liquidMask;
    double gasMolarVolume;
};

// Full-pivot LU with Eigen's rank rule, n <= 4 (:111-139).
bool solve(const double m[4][4], const double* rhs, int n, double* x) {
    double a[4][4]{}, b[4]{}, largest = 0;
    int column[4]{0, 1, 2, 3};
    for (int i = 0; i < n; ++i) {
        b[i] = rhs[i];
        if (!fin(b[i])) return false;
        for (int j = 0; j < n; ++j) {
            a[i][j] = m[i][j];
            if (!fin(a[i][j])) return false;
        }
    }
    for (int k = 0; k < n; ++k) {
        int row = k, col = k;
        double pivot = 0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i)
                if (std::fabs(a[i][j]) > pivot) { pivot = std::fabs(a[i][j]); row = i; col = j; }
        largest = hi(largest, pivot);
        if (!(pivot > largest * n * DBL_EPSILON)) return false;
        if (row != k) {
            for (int j = 0; j < n; ++j) { const double t = a[k][j]; a[k][j] = a[row][j]; a[row][j] = t; }
            const double t = b[k]; b[k] = b[row]; b[row] = t;
        }
        if (col != k) {
            for (int i = 0; i < n; ++i) { const double t = a[i][k]; a[i][k] = a[i][col]; a[i][col] = t; }
            const int t = column[k]; column[k] = column[col]; column[col] = t;
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i][k] / a[k][k];
            a[i][k] = f;
            for (int j = k + 1; j < n; ++j) a[i][j] -= f * a[k][j];
            b[i] -= f * b[k];
        }
    }
    double z[4]{};
    for (int i = n - 1; i >= 0; --i) {
        double v = b[i];
        for (int j = i + 1; j < n; ++j) v -= a[i][j] * z[j];
        z[i] = v / a[i][i];
        if (!fin(z[i])) return false;
    }
    for (int i = 0; i < n; ++i) x[column[i]] = z[i];
    return true;
}

inline double maxNorm(const double* a, int n) {
    double r = 0;
    for (int i = 0; i < n; ++i) r = hi(r, std::fabs(a[i]));
    return r;
}

struct Flash {
    const Model& m;
    const double* q;  // ns species masses
    double energy;
    const ThermoState& guess;
    Counters& count;
    double cacheP = 0, cacheT = 0;  // one unavailable liquid-0 result (:142-145)
    int cacheStatus = -1;
    bool cacheVirtual = false;
    bool adaptive = false;               // vapour-log coordinates of the boundary search (:145)
    double color = 0, pressureJump = 0;  // capillary interface, J fixed during the solve (:146)

    Flash(const Model& model, const double* mass, double e, const ThermoState& g, Counters& c, double geometricColor = 0,
          double jump = 0)
        : m(model), q(mass), energy(e), guess(g), count(c), color(geometricColor), pressureJump(jump) {}
    double gasPressure(double pbar) const { return pbar - color * pressureJump; }           // :152
    double liquidPressure(double pbar) const { return pbar + (1 - color) * pressureJump; }  // :153

    double density() const {
        double r = 0;
        for (int k = 0; k < m.ns; ++k) r += q[k];
        return r;
    }
    bool bounds(double p, double T) const {
        return fin(p) && fin(T) && p >= m.pmin && p <= m.pmax && T >= m.Tmin && T <= m.Tmax;
    }
    static bool unavailable(Status s) { return s == Unavailable || s == NoRoot; }

    // :157-196 (pure-liquid condensables; generic PR path for ns != 4)
    Status phaseStatus(int index, double p, double T, const double* Y, PRState& out, bool withMu = true) {
        ++count.phases;
        if (!bounds(p, T)) return BadInput;
        if (index >= 0 && T < m.liquidTmin[index]) return BadInput;
        if (index >= 0 && T >= m.liquidTc[index] * (1 - 1e-10)) return Unavailable;
        if (index < 0 && m.enforceSpeciesTemperatureBounds)
            for (int k = 0; k < m.ns; ++k)
                if (Y[k] > 0) {
                    const Species& sp = m.phase[0].sp[k];
                    if (T < sp.region[0].tmin || T > sp.region[sp.nreg - 1].tmax) return BadInput;
                }
        double pure[kMaxSpecies]{};
        pure[0] = 1;
        if (index >= 0) return evaluateTP(m.phase[index + 1], T, p, pure, LiquidRoot, out, withMu);
        return evaluateTP(m.phase[0], T, p, Y, GasRoot, out, withMu);
    }
    bool phase(int index, double p, double T, const double* Y, PRState& out, bool withMu = true) {
        return phaseStatus(index, p, T, Y, out, withMu) == OK;
    }

    static void accumulate(Evaluation& v, double mass, const PRState& p, double phaseP) {  // :203-207
        const double volume = mass / p.rho, vp = -volume * p.kappaT, vt = volume * p.alpha;
        v.volume += volume; v.energy += mass * p.e; v.entropyDensity += mass * p.s; v.Vp += vp; v.VT += vt;
        v.Ep += -v.state.T * vt - phaseP * vp;
        v.ET += mass * p.cp - phaseP * vt;
        v.cpDensity += mass * p.cp;
    }

    // Mixture of a gas phase and pure liquids at (p, T) (:208-262); vapor (the
    // adaptive coordinates) gives each condensable's gas inventory directly.
    bool evaluate(const double* mass, double p, double T, Evaluation& v, bool virtualLiquids = false,
                  const double* vapor = nullptr) {
        v = Evaluation{};
        if (!bounds(p, T)) return false;
        if (m.nl > 0 && mass[0] > 0 && cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual)
            return false;
        ThermoState& s = v.state;
        s.p = p; s.T = T; s.rho = density();
        if (!(s.rho > 0)) return false;
        double gas[kMaxSpecies]{};
        for (int k = 0; k < m.ns; ++k) {
            if (!fin(q[k]) || q[k] < 0) return false;
            gas[k] = q[k];
        }
        for (int i = 0; i < m.nl; ++i) {
            const int c = m.condensable[i];
            if (!fin(mass[i]) || mass[i] < 0 || mass[i] > q[c]) return false;
            gas[c] = vapor ? vapor[i] : gas[c] - mass[i];
            if (!fin(gas[c]) || gas[c] < 0 || gas[c] > q[c]) return false;
        }
        for (int i = 0; i < m.nl; ++i) v.gasCondensable[i] = gas[m.condensable[i]];
        v.explicitGas = vapor != nullptr;
        double mg = 0;
        for (int k = 0; k < m.ns; ++k) mg += gas[k];
        s.gasMass = mg;
        const double pg = gasPressure(p), pl = liquidPressure(p);
        if (mg > 0) {
            for (int k = 0; k < m.ns; ++k) gas[k] /= mg;
            PRState prop{};
            if (!phase(-1, pg, T, gas, prop, virtualLiquids)) return false;
            v.gasMolarVolume = prop.V;
            accumulate(v, mg, prop, pg);
            s.alphaGas = mg / prop.rho;
            s.rhoGas = prop.rho;
            if (virtualLiquids)
                for (int i = 0; i < m.nl; ++i) v.muGas[i] = prop.mu[m.condensable[i]] / (8314.46261815324 * T);
        }
        for (int i = 0; i < m.nl; ++i) {
            s.liquidMass[i] = mass[i];
            if (mass[i] > 0 || (virtualLiquids && q[m.condensable[i]] > 0)) {
                PRState prop{};
                Status st;
                if (i == 0 && cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual) {
                    ++count.phases;
                    st = static_cast<Status>(cacheStatus);
                } else {
                    st = phaseStatus(i, pl, T, nullptr, prop, virtualLiquids);
                    if (i == 0 && unavailable(st) && T >= m.liquidTmin[i] && T < m.liquidTc[i] * (1 - 1e-10)) {
                        cacheP = p; cacheT = T; cacheVirtual = virtualLiquids; cacheStatus = static_cast<int>(st);
                    }
                }
                if (st != OK) {
                    if (mass[i] > 0 || !unavailable(st)) return false;
                    v.muLiquid[i] = HUGE_VAL;
                    continue;
                }
                v.liquidHbar[i] = prop.h * m.phase[i + 1].sp[0].W;
                v.liquidVbar[i] = prop.V;
                v.liquidMask |= 1u << i;
                if (virtualLiquids) v.muLiquid[i] = prop.mu[0] / (8314.46261815324 * T);
                if (mass[i] > 0) {
                    accumulate(v, mass[i], prop, pl);
                    s.rhoLiquid[i] = prop.rho;
                    s.alphaLiquid[i] = mass[i] / prop.rho;
                    s.activeLiquids |= 1 << i;
                }
            }
        }
        s.e = v.energy / s.rho; s.entropy = v.entropyDensity / s.rho; s.cp = v.cpDensity / s.rho;
        s.cv = (v.cpDensity + T * v.VT * v.VT / v.Vp) / s.rho;
        if (!fin(s.cv) || s.cv <= 0) return false;
        const double specific = v.volume / s.rho, vp = v.Vp / s.rho, vt = v.VT / s.rho;
        const double adiabatic = vp + T * vt * vt / s.cp;
        if (!(adiabatic < 0)) return false;
        s.soundFrozen = std::sqrt(-specific * specific / adiabatic);
        s.soundEquilibrium = s.soundFrozen;
        return fin(s.soundFrozen) && s.soundFrozen > 0;
    }
    double energyScale(const Evaluation& v) const { return hi(1e5 * v.state.rho, v.cpDensity * v.state.T); }

    // Single-phase gas residual F = rho e(T, rho) - e_in (:264-275).
    bool gasProbe(double T, double& F, double& cv, double& p, double* cp = nullptr) {
        const double rho = density();
        double Y[kMaxSpecies]{};
        for (int k = 0; k < m.ns; ++k) Y[k] = q[k] / rho;
        PRState prop{};
        ++count.phases;
        if (evaluateTrho(m.phase[0], T, rho, Y, prop, false) != OK) return false;
        const double gasP = prop.p;
        p = gasP + color * pressureJump;
        if (!bounds(p, T)) return false;
        PRState gas{};
        if (!phase(-1, gasP, T, Y, gas, false) || std::fabs(gas.rho - rho) > 1e-8 * rho) return false;
        F = rho * prop.e - energy;
        cv = rho * prop.cv;
        if (cp) *cp = prop.cp;
        return fin(F) && fin(cv) && cv > 0;
    }
    // Bracketed scalar Newton in T at fixed rho (:276-299).
    bool scalarGas(Evaluation& result) {
        const double rho = density();
        double T = clamp(guess.T, m.Tmin, m.Tmax), F, cv, p, cp;
        if (!gasProbe(T, F, cv, p, &cp)) return false;
        const double scale = hi(1e5 * rho, rho * cp * T);
        double lower = T, upper = T, fl = F, fh = F;
        if (std::fabs(F) > m.etol * scale) {
            double radius = hi(1., .01 * T);
            for (int j = 0; j < 24 && (fl > 0 || fh < 0); ++j, radius *= 2) {
                if (fl > 0) {
                    double candidate = hi(m.Tmin, T - radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        if (gasProbe(candidate, value, dcv, dp)) { fl = value; lower = candidate; found = true; }
                        else candidate = .5 * (lower + candidate);
                    }
                    if (!found) return false;
                }
                if (fh < 0) {
                    double candidate = lo(m.Tmax, T + radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        if (gasProbe(candidate, value, dcv, dp)) { fh = value; upper = candidate; found = true; }
                        else candidate = .5 * (upper + candidate);
                    }
                    if (!found) return false;
                }
                if (std::fabs(fl) <= m.etol * scale) { T = lower; if (!gasProbe(T, F, cv, p)) return false; break; }
                if (std::fabs(fh) <= m.etol * scale) { T = upper; if (!gasProbe(T, F, cv, p)) return false; break; }
                if (lower == m.Tmin && upper == m.Tmax) break;
            }
            if (!(std::fabs(F) <= m.etol * scale || (fl <= 0 && fh >= 0 && upper > lower))) return false;
        }
        int it = 0;
        for (; it < 70 && std::fabs(F) > m.etol * scale; ++it) {
            if (F > 0) upper = T; else lower = T;
            const double next = T - F / cv;
            T = fin(next) && next > lower && next < upper ? next : .5 * (lower + upper);
            if (!gasProbe(T, F, cv, p)) return false;
        }
        if (std::fabs(F) > m.etol * scale) return false;
        double zero[2]{};
        if (!evaluate(zero, p, T, result)) return false;
        result.state.volumeResidual = std::fabs(result.volume - 1);
        result.state.energyResidual = std::fabs(result.energy - energy) / energyScale(result);
        result.state.iterations = it;
        return result.state.volumeResidual <= m.vtol && result.state.energyResidual <= m.etol;
    }
    // Fixed phase split, 2x2 Newton in (ln p, ln T) with line search (:300-326).
    bool frozen(const double* mass, Evaluation& value, int stallLimit = 0) {
        if (!fin(energy)) return false;
        if (m.scalarRecovery && mass[0] == 0 && mass[1] == 0 && scalarGas(value)) return true;
        double p = clamp(guess.p, m.pmin, m.pmax), T = clamp(guess.T, m.Tmin, m.Tmax);
        if (!evaluate(mass, p, T, value)) return false;
        int stalled = 0;
        for (int it = 0; it < 70; ++it) {
            const double scale = energyScale(value), rv = value.volume - 1, re = (value.energy - energy) / scale;
            value.state.volumeResidual = std::fabs(rv);
            value.state.energyResidual = std::fabs(re);
            value.state.iterations = it;
            if (std::fabs(rv) <= m.vtol && std::fabs(re) <= m.etol) return true;
            double jac[4][4]{};
            jac[0][0] = value.Vp * p; jac[0][1] = value.VT * T;
            jac[1][0] = value.Ep * p / scale; jac[1][1] = value.ET * T / scale;
            double rhs[4]{-rv, -re}, step[4]{};
            if (!solve(jac, rhs, 2, step)) return false;
            const double limit = hi(1., hi(std::fabs(step[0]) / .7, std::fabs(step[1]) / .2));
            step[0] /= limit; step[1] /= limit;
            const double previous = hi(std::fabs(rv), std::fabs(re));
            bool accepted = false;
            for (double fraction = 1; fraction > 1e-9; fraction *= .5) {
                Evaluation trial;
                const double pp = p * oexp(fraction * step[0]), tt = T * oexp(fraction * step[1]);
                if (evaluate(mass, pp, tt, trial)) {
                    const double next = hi(std::fabs(trial.volume - 1), std::fabs(trial.energy - energy) / scale);
                    if (next < previous * (1 - 1e-4 * fraction) || next < lo(m.vtol, m.etol)) {
                        value = trial; p = pp; T = tt; accepted = true;
                        stalled = fraction < 1. / 32 ? stalled + 1 : 0;
                        break;
                    }
                }
            }
            if (!accepted) return false;
            if (stallLimit > 0 && stalled >= stallLimit &&
                !(std::fabs(value.volume - 1) <= m.vtol && std::fabs(value.energy - energy) / energyScale(value) <= m.etol))
                return false;
        }
        return false;
    }
    // Active-set residual [V-1, (E-e)/scale, mu_l/RT - mu_g/RT] (:327-337).
    bool residual(const int* active, int n, const double* x, double scale, double* f, Evaluation* out = nullptr) {
        ++count.residuals;
        double mass[2]{}, vapor[2]{};
        for (int i = 0; i < m.nl; ++i) vapor[i] = q[m.condensable[i]];
        for (int j = 0; j < n; ++j) {
            const int i = active[j];
            const double qq = q[m.condensable[i]];
            if (adaptive) {  // x = ln(vapour fraction) <= 0
                if (!fin(x[j + 2]) || x[j + 2] > 0) return false;
                vapor[i] = oexp(x[j + 2]) * qq;
                if (!(vapor[i] > 0)) return false;
                mass[i] = -oexpm1(x[j + 2]) * qq;
            } else {
                if (!(x[j + 2] >= 0 && x[j + 2] < 1)) return false;
                mass[i] = x[j + 2] * qq;
            }
        }
        Evaluation v;
        if (!evaluate(mass, oexp(x[0]), oexp(x[1]), v, true, adaptive ? vapor : nullptr) || !(v.state.gasMass > 0))
            return false;
        f[0] = v.volume - 1;
        f[1] = (v.energy - energy) / scale;
        for (int j = 0; j < n; ++j) {
            f[2 + j] = v.muLiquid[active[j]] - v.muGas[active[j]];
            if (!fin(f[2 + j])) return false;
        }
        if (out) *out = v;
        return true;
    }
    // Central finite-difference Jacobian with step halving (:338-373).
    bool jacobian(const int* active, int n, const double* x, double scale, const double* f, double jac[4][4]) {
        ++count.fd;
        const int size = n + 2;
        for (int j = 0; j < size; ++j) {
            double h = j < 2 ? 2e-6 : 2e-5;
            bool success = false;
            for (int attempt = 0; attempt < 12 && !success; ++attempt, h *= .5) {
                double xp[4]{}, xm[4]{}, fp[4]{}, fm[4]{};
                for (int i = 0; i < size; ++i) xp[i] = xm[i] = x[i];
                xp[j] += h;
                xm[j] -= h;
                if (adaptive) {  // a step lost to rounding moves by one ulp instead
                    if (xp[j] == x[j]) xp[j] = std::nextafter(x[j], HUGE_VAL);
                    if (xm[j] == x[j]) xm[j] = std::nextafter(x[j], -HUGE_VAL);
                }
                const bool plus = residual(active, n, xp, scale, fp), minus = residual(active, n, xm, scale, fm);
                if (!plus && !minus) continue;
                const double hp = xp[j] - x[j], hm = x[j] - xm[j];
                double maximum = 0;
                success = true;
                for (int i = 0; i < size; ++i) {
                    double v;
                    if (adaptive)  // one-sided differences weighted by the actual steps
                        v = plus && minus ? (hm / (hp + hm)) * ((fp[i] - f[i]) / hp) + (hp / (hp + hm)) * ((f[i] - fm[i]) / hm)
                            : plus ? (fp[i] - f[i]) / hp : (f[i] - fm[i]) / hm;
                    else
                        v = plus && minus ? (fp[i] - fm[i]) / (2 * h) : plus ? (fp[i] - f[i]) / h : (f[i] - fm[i]) / h;
                    jac[i][j] = v;
                    maximum = hi(maximum, std::fabs(v));
                    success = success && fin(v);
                }
                if (adaptive) success = success && maximum > 0;
            }
            if (!success) return false;
        }
        return true;
    }
    // Analytic Jacobian from the residual point's own properties
    // (buildCached, reactiveDeviceFlashJacobian.h:535-640; one active liquid):
    // the gas at its own pressure pg, the liquid's molar energy at pl.
    bool buildCached(const int* active, int n, const double* x, double scale, const Evaluation& v, double jac[4][4]) {
        const double p = v.state.p, T = v.state.T;
        const double gp = gasPressure(p), lp = liquidPressure(p);
        const double pg = gp == 0 ? p : gp, pl = lp == 0 ? p : lp;
        if (!(pg > 0) || !(pl > 0) || !fin(pg) || !fin(pl)) return false;
        if (n != 1 || !(p > 0) || !(T > 0) || !(scale > 0) || !(v.state.gasMass > 0) || !(v.gasMolarVolume > 0) ||
            !fin(p) || !fin(T) || !fin(scale) || !fin(v.Vp) || !fin(v.VT) || !fin(v.Ep) || !fin(v.ET) ||
            !fin(v.state.gasMass) || !fin(v.gasMolarVolume) || v.explicitGas != adaptive)
            return false;
        double Y[kMaxSpecies]{};
        for (int k = 0; k < m.ns; ++k) {
            if (!fin(q[k]) || q[k] < 0) return false;
            Y[k] = q[k];
        }
        for (int i = 0; i < m.nl; ++i) {
            const int k = m.condensable[i];
            if (k < 0 || k >= m.ns || !fin(v.gasCondensable[i]) || v.gasCondensable[i] < 0 ||
                v.gasCondensable[i] > q[k])
                return false;
            Y[k] = v.gasCondensable[i];
        }
        const int i = active[0];
        if (i < 0 || i >= m.nl) return false;
        const int k = m.condensable[i];
        if (!(q[k] > 0)) return false;
        double expectedGas, dLiquid;
        if (adaptive) {
            if (!(x[2] <= 0) || !fin(x[2])) return false;
            expectedGas = oexp(x[2]) * q[k];
            dLiquid = -expectedGas;
        } else {
            if (!(x[2] >= 0 && x[2] < 1)) return false;
            const double product = x[2] * q[k];  // the residual's own two roundings (:35-47)
            expectedGas = q[k] - product;
            dLiquid = q[k];
        }
        if (!(expectedGas > 0) || v.gasCondensable[i] != expectedGas) return false;
        if (!(v.liquidMask & (1u << i)) || !fin(v.liquidHbar[i]) || !fin(v.liquidVbar[i]) || !(v.liquidVbar[i] > 0))
            return false;
        for (int l = 0; l < m.nl; ++l)
            if (l != i && v.gasCondensable[l] != q[m.condensable[l]]) return false;
        double reconstructed = 0;
        for (int s = 0; s < m.ns; ++s) reconstructed += Y[s];
        if (reconstructed != v.state.gasMass) return false;
        for (int s = 0; s < m.ns; ++s) Y[s] /= v.state.gasMass;
        double hbar, ubar, vbar, dpsi;
        if (!activeGasDerivatives(m.phase[0], T, pg, v.gasMolarVolume, Y, v.state.gasMass, k, hbar, ubar, vbar, dpsi))
            return false;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) jac[r][c] = 0;
        jac[0][0] = v.Vp * p;
        jac[0][1] = v.VT * T;
        jac[1][0] = v.Ep * p / scale;
        jac[1][1] = v.ET * T / scale;
        const double Wk = m.phase[0].sp[k].W, Wl = m.phase[i + 1].sp[0].W;
        const double liquidUbar = v.liquidHbar[i] - pl * v.liquidVbar[i];
        jac[0][2] = dLiquid * (v.liquidVbar[i] / Wl - vbar / Wk);
        jac[1][2] = dLiquid * (liquidUbar / Wl - ubar / Wk) / scale;
        jac[2][0] = p * (v.liquidVbar[i] - vbar) / (R * T);
        jac[2][1] = -(v.liquidHbar[i] - hbar) / (R * T);
        jac[2][2] = dpsi * dLiquid / Wk;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                if (!fin(jac[r][c])) return false;
        return true;
    }
    // Analytic Jacobian from fresh gas and liquid states at (p, T)
    // (build, reactiveDeviceFlashJacobian.h:436-530; one active liquid).
    bool buildFresh(const int* active, int n, const double* x, double p, double T, double scale, const Evaluation& v,
                    double jac[4][4]) {
        if (n != 1 || !(p > 0) || !(T > 0) || !(scale > 0)) return false;
        const double gp = gasPressure(p), lp = liquidPressure(p);
        const double pg = gp == 0 ? p : gp, pl = lp == 0 ? p : lp;
        if (!(pg > 0) || !(pl > 0) || !fin(pg) || !fin(pl)) return false;
        double gasMass[kMaxSpecies]{}, Y[kMaxSpecies]{}, total = 0;
        for (int s = 0; s < m.ns; ++s) gasMass[s] = q[s];
        const int i = active[0];
        if (i < 0 || i >= m.nl) return false;
        const int k = m.condensable[i];
        if (k < 0 || k >= m.ns || !(q[k] > 0)) return false;
        double dLiquid;
        if (adaptive) {
            if (!(x[2] <= 0) || !fin(x[2])) return false;
            gasMass[k] = oexp(x[2]) * q[k];
            if (!(gasMass[k] > 0)) return false;
            dLiquid = -gasMass[k];
        } else {
            if (!(x[2] >= 0 && x[2] < 1)) return false;
            gasMass[k] -= x[2] * q[k];
            dLiquid = q[k];
        }
        for (int s = 0; s < m.ns; ++s) total += gasMass[s];
        if (!(total > 0)) return false;
        for (int s = 0; s < m.ns; ++s) Y[s] = gasMass[s] / total;
        PRState gas{};
        ++count.phases;
        if (evaluateTP(m.phase[0], T, pg, Y, GasRoot, gas, false) != OK) return false;
        double hbar[kMaxSpecies]{}, ubar[kMaxSpecies]{}, vbar[kMaxSpecies]{}, xm[kMaxSpecies]{}, W = 0;
        if (!phasePartials(m.phase[0], T, Y, gas, hbar, ubar, vbar, xm, W)) return false;
        const double totalMoles = total / W;
        PRState liquid{};
        double pure[kMaxSpecies]{};
        pure[0] = 1;
        ++count.phases;
        if (evaluateTP(m.phase[i + 1], T, pl, pure, LiquidRoot, liquid, false) != OK) return false;
        double dpsi;
        if (!gasCompositionPartial(m.phase[0], T, xm, totalMoles, gas, k, dpsi)) return false;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) jac[r][c] = 0;
        jac[0][0] = v.Vp * p;
        jac[0][1] = v.VT * T;
        jac[1][0] = v.Ep * p / scale;
        jac[1][1] = v.ET * T / scale;
        const double Wk = m.phase[0].sp[k].W;
        jac[0][2] = dLiquid * (1.0 / liquid.rho - vbar[k] / Wk);
        jac[1][2] = dLiquid * (liquid.e - ubar[k] / Wk) / scale;
        const double Wl = m.phase[i + 1].sp[0].W;
        const double liquidHbar = liquid.h * Wl;
        jac[2][0] = p * (liquid.V - vbar[k]) / (R * T);
        jac[2][1] = -(liquidHbar - hbar[k]) / (R * T);
        jac[2][2] = dpsi * dLiquid / Wk;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                if (!fin(jac[r][c])) return false;
        return true;
    }
    // The Jacobian at a residual point (:338-352): analytic when the model asks
    // for it and it can be built, else central differences.
    bool jacobianAt(const int* active, int n, const double* x, double scale, const double* f, double jac[4][4],
                    const Evaluation& point) {
        if (m.analyticJacobian &&
            (buildCached(active, n, x, scale, point, jac) ||
             buildFresh(active, n, x, point.state.p, point.state.T, scale, point, jac))) {
            ++count.analytic;
            return true;
        }
        return jacobian(active, n, x, scale, f, jac);
    }
    // Damped Newton on the active set, seeded by liquid fraction (:374-448);
    // in adaptive coordinates by ln(vapour fraction), where seed -3 (and a
    // previous state with almost no vapour) starts from a trace of vapour set
    // by the noncondensable inventory.  secondSeed (-2: none) seeds the second
    // active liquid.
    bool activeFlash(const int* active, int n, double seed, Evaluation& value, double secondSeed = -2,
                     const ThermoState* restart = nullptr) {
        ++count.candidates;
        const int size = n + 2;
        const double rho = density();
        const ThermoState& g = restart ? *restart : guess;
        const double scale = hi(1e5 * rho, hi(std::fabs(energy), rho * 2000 * hi(g.T, m.Tmin)));
        double x[4]{olog(clamp(g.p, m.pmin, m.pmax)), olog(clamp(g.T, m.Tmin, m.Tmax))};
        for (int j = 0; j < n; ++j) {
            const int i = active[j];
            const double selected = j == 1 && secondSeed != -2 ? secondSeed : seed;
            const double previous = g.liquidMass[i] / q[m.condensable[i]];
            x[j + 2] = selected < 0 ? clamp(previous, 1e-8, 1 - 1e-8) : selected;
            if (adaptive) {
                double vapor = selected < 0 ? 1 - previous : 1 - selected;
                if (selected == -3 || (selected < 0 && vapor < 1e-6)) {
                    double other = 0;
                    for (int k = 0; k < m.ns; ++k) {
                        bool selectedCondensable = false;
                        for (int a = 0; a < n; ++a) selectedCondensable |= k == m.condensable[active[a]];
                        if (!selectedCondensable) other += q[k];
                    }
                    vapor = other > 0 ? lo(1e-6, .01 * (other / q[m.condensable[i]])) : 1e-8;
                }
                if (!(vapor > 0 && fin(vapor))) return false;
                x[j + 2] = olog(lo(1., vapor));
            }
        }
        double f[4]{};
        if (!residual(active, n, x, scale, f, &value)) return false;
        for (int it = 0; it < 90; ++it) {
            const double chemical = maxNorm(f + 2, n);
            double jac[4][4]{};
            if (std::fabs(f[0]) <= m.vtol && std::fabs(f[1]) <= m.etol && chemical <= m.mutol) {
                value.state.volumeResidual = std::fabs(f[0]);
                value.state.energyResidual = std::fabs(f[1]);
                value.state.chemicalResidual = chemical;
                value.state.iterations = it;
                if (!jacobianAt(active, n, x, scale, f, jac, value)) return false;
                double rhs[4]{-1 / rho, value.state.p / (rho * scale)}, response[4]{};
                if (!solve(jac, rhs, size, response)) return false;
                const double squared = value.state.p * response[0];
                if (!fin(squared) || squared <= 0) return false;
                value.state.soundEquilibrium = std::sqrt(squared);
                return value.state.soundEquilibrium <= value.state.soundFrozen * (1 + 2e-3);
            }
            if (!jacobianAt(active, n, x, scale, f, jac, value)) return false;
            double rhs[4]{}, step[4]{};
            for (int j = 0; j < size; ++j) rhs[j] = -f[j];
            if (!solve(jac, rhs, size, step)) return false;
            double limiter = hi(1., hi(std::fabs(step[0]) / .7, std::fabs(step[1]) / .18));
            for (int j = 2; j < size; ++j) limiter = hi(limiter, std::fabs(step[j]) / (adaptive ? 2. : .3));
            for (int j = 0; j < size; ++j) step[j] /= limiter;
            bool accepted = false;
            const double previous = maxNorm(f, size);
            double fraction = 1;
            for (int index = 0; index < 30; ++index, fraction *= .5) {
                double tx[4]{}, tf[4]{};
                for (int j = 0; j < size; ++j) tx[j] = x[j] + fraction * step[j];
                bool rejected = false;
                for (int j = 0; j < n; ++j)
                    rejected |= adaptive ? (!fin(tx[j + 2]) || tx[j + 2] > 0) : !(tx[j + 2] >= 0 && tx[j + 2] < 1);
                if (rejected) {
                    // rejected fractions form a prefix: bisect to its end (:419-438)
                    int lower = index, upper = 30;
                    while (upper - lower > 1) {
                        const int mid = (lower + upper) / 2;
                        const double fm = std::ldexp(1.0, -mid);
                        bool bad = false;
                        for (int j = 0; j < n; ++j) {
                            const double y = x[j + 2] + fm * step[j + 2];
                            bad |= adaptive ? (!fin(y) || y > 0) : !(y >= 0 && y < 1);
                        }
                        if (bad) lower = mid; else upper = mid;
                    }
                    count.residuals += uint64_t(upper - index);
                    if (upper == 30) break;
                    index = upper;
                    fraction = std::ldexp(1.0, -upper);
                    for (int j = 0; j < size; ++j) tx[j] = x[j] + fraction * step[j];
                }
                Evaluation trial;
                if (residual(active, n, tx, scale, tf, &trial) && maxNorm(tf, size) < previous * (1 - 1e-4 * fraction)) {
                    for (int j = 0; j < size; ++j) { x[j] = tx[j]; f[j] = tf[j]; }
                    value = trial;
                    accepted = true;
                    break;
                }
            }
            if (!accepted) return false;
        }
        return false;
    }
    // Gibbs tangent-plane distance of a trial vapour (:474-485).
    double tangent(double p, double T, const int* present, int n, const double* mu, double fraction) {
        double mole[kMaxSpecies]{}, Y[kMaxSpecies]{}, sum = 0;
        mole[m.condensable[present[0]]] = n == 1 ? 1 : fraction;
        if (n == 2) mole[m.condensable[present[1]]] = 1 - fraction;
        for (int k = 0; k < m.ns; ++k) { Y[k] = mole[k] * m.weights[k]; sum += Y[k]; }
        for (int k = 0; k < m.ns; ++k) Y[k] /= sum;
        PRState prop{};
        const Status st = phaseStatus(-1, gasPressure(p), T, Y, prop, false);
        if (st != OK) return unavailable(st) ? HUGE_VAL : NAN;
        double reference = 0, W = 0;
        for (int i = 0; i < n; ++i) reference += mole[m.condensable[present[i]]] * mu[present[i]];
        for (int k = 0; k < m.ns; ++k) W += mole[k] * m.weights[k];
        return (prop.h - T * prop.s) * W / (8314.46261815324 * T) - reference;
    }
    // Phase stability of a converged candidate (:449-473).
    bool stable(Evaluation& value) {
        const double p = value.state.p, T = value.state.T;
        if (value.state.gasMass > 0) {
            double mass[2]{value.state.liquidMass[0], value.state.liquidMass[1]};
            Evaluation comparison;
            if (!evaluate(mass, p, T, comparison, true, value.explicitGas ? value.gasCondensable : nullptr))
                return false;
            for (int i = 0; i < m.nl; ++i) {
                if (q[m.condensable[i]] == 0) continue;
                const double affinity = comparison.muLiquid[i] - comparison.muGas[i];
                if (mass[i] > 0) {
                    if (!fin(affinity) || std::fabs(affinity) > m.mutol * 2) return false;
                } else if (affinity < -m.mutol) return false;
            }
            return true;
        }
        int present[2]{}, n = 0;
        double mu[2]{};
        for (int i = 0; i < m.nl; ++i)
            if (value.state.liquidMass[i] > 0) {
                present[n++] = i;
                PRState prop{};
                if (!phase(i, liquidPressure(p), T, nullptr, prop)) return false;
                mu[i] = prop.mu[0] / (8314.46261815324 * T);
            }
        if (!n) return false;
        if (n == 1) return tangent(p, T, present, n, mu, 1) >= -m.mutol;
        double best = HUGE_VAL;
        int location = 0;
        for (int j = 0; j <= 32; ++j) {
            const double v = tangent(p, T, present, n, mu, double(j) / 32);
            if (v != v) return false;
            if (v < best) { best = v; location = j; }
        }
        double left = hi(0., double(location - 1) / 32), right = lo(1., double(location + 1) / 32);
        for (int j = 0; j < 55; ++j) {
            const double a = left + (right - left) * .38196601125, b = left + (right - left) * .61803398875;
            const double va = tangent(p, T, present, n, mu, a), vb = tangent(p, T, present, n, mu, b);
            if (va != va || vb != vb) return false;
            best = lo(best, lo(va, vb));
            if (va < vb) right = b; else left = a;
        }
        return best >= -m.mutol;
    }
    void consider(Evaluation& candidate, Evaluation& best, bool& have) {  // :486-489
        if (stable(candidate)) {
            ++count.stable;
            if (!have || candidate.entropyDensity > best.entropyDensity) {
                best = candidate;
                have = true;
            }
        } else {
            ++count.failures;
        }
    }
    bool temperatureSearch(Evaluation& best) {  // :490-508
        bool have = false;
        Evaluation candidate;
        const double seeds[5]{-1, .5, .95, .1, .9999};
        for (int mask = 1; mask < (1 << m.nl); ++mask) {
            int active[2]{}, n = 0;
            bool possible = true;
            double lower = m.Tmin, upper = m.Tmax;
            for (int i = 0; i < m.nl; ++i)
                if (mask & (1 << i)) {
                    if (q[m.condensable[i]] <= 0) possible = false;
                    else {
                        active[n++] = i;
                        lower = hi(lower, m.liquidTmin[i]);
                        upper = lo(upper, m.liquidTc[i] * (1 - 1e-10));
                    }
                }
            if (!possible || !(upper > lower)) continue;
            for (int t = 0; t < 4; ++t) {
                ThermoState restart = guess;
                const double frac = t == 0 ? 0.02 : t == 1 ? 0.25 : t == 2 ? 0.5 : 0.75;  // reactiveFlashSeeds.h:13-16
                restart.T = lower + (upper - lower) * frac;
                for (int j = 0; j < 5; ++j) {
                    if (activeFlash(active, n, seeds[j], candidate, -2, &restart)) consider(candidate, best, have);
                    else ++count.failures;
                }
            }
        }
        return have;
    }
    bool search(Evaluation& best) {  // :509-543
        bool have = false;
        Evaluation candidate;
        double mass[2]{};
        bool gasStable = false;
        const bool prunable =
            m.stableGasPrune && m.nl == 1 && m.condensedKind[0] == 0 && m.condensedKind[1] == 0 && !adaptive;
        if (frozen(mass, candidate, prunable ? 4 : 0)) {
            const uint64_t before = count.stable;
            consider(candidate, best, have);
            gasStable = count.stable > before && candidate.state.gasMass > 0;
        } else {
            ++count.failures;
        }
        const bool prune = prunable && gasStable;
        double other = 0;
        for (int k = 0; k < m.ns; ++k) {
            bool cond = false;
            for (int i = 0; i < m.nl; ++i) cond |= k == m.condensable[i];
            if (!cond) other += q[k];
        }
        if (other == 0 && m.nl) {
            for (int i = 0; i < m.nl; ++i) mass[i] = q[m.condensable[i]];
            if (frozen(mass, candidate)) consider(candidate, best, have);
            else ++count.failures;
        }
        const double seeds[5]{-1, .5, .95, .1, .9999};
        for (int mask = 1; !prune && mask < (1 << m.nl); ++mask) {
            int active[2]{}, n = 0;
            bool possible = true;
            for (int i = 0; i < m.nl; ++i)
                if (mask & (1 << i)) {
                    if (q[m.condensable[i]] <= 0) possible = false;
                    else active[n++] = i;
                }
            if (!possible) continue;
            for (int j = 0; j < 5; ++j) {
                if (activeFlash(active, n, seeds[j], candidate)) consider(candidate, best, have);
                else ++count.failures;
            }
            if (adaptive) {  // the trace-vapour start (:532)
                if (activeFlash(active, n, -3, candidate)) consider(candidate, best, have);
                else ++count.failures;
            }
            if (adaptive && n == 2)  // off-diagonal seed pairs in the order .1, .5, .95, .9999 (:533-536)
                for (int first = 1; first < 5; ++first)
                    for (int second = 1; second < 5; ++second)
                        if (first != second) {
                            const double off[4]{.1, .5, .95, .9999};
                            if (activeFlash(active, n, off[first - 1], candidate, off[second - 1]))
                                consider(candidate, best, have);
                            else ++count.failures;
                        }
        }
        if (!have) have = temperatureSearch(best);
        return have;
    }
    // :544-560.  With boundaryRecovery (or a capillary equilibrium cell), a
    // reference result whose liquid holds nearly all of a condensable with gas
    // present, or no result, runs the search again in vapour coordinates; the
    // higher entropy density wins.
    bool run(Evaluation& result, bool capillaryBoundary = false) {
        if (m.ns <= 0 || m.ns > kMaxSpecies || m.nl < 0 || m.nl > 2 || !fin(energy) || !(density() > 0)) return false;
        for (int k = 0; k < m.ns; ++k)
            if (!fin(q[k]) || q[k] < 0) return false;
        Evaluation reference;
        const bool have = search(reference);
        if (!m.boundaryRecovery && !capillaryBoundary) {
            if (have) result = reference;
            return have;
        }
        if (have) {
            bool boundary = false;
            for (int i = 0; i < m.nl; ++i)
                if (q[m.condensable[i]] > 0 && reference.state.gasMass > 0)
                    boundary |= reference.state.liquidMass[i] / q[m.condensable[i]] > 1 - 1e-4;
            if (!boundary) {
                result = reference;
                return true;
            }
        }
        adaptive = true;
        Evaluation recovered;
        const bool found = search(recovered);
        adaptive = false;
        if (!found && !have) return false;
        result = have && (!found || reference.entropyDensity >= recovered.entropyDensity) ? reference : recovered;
        return true;
    }
};

bool validModel(const Model& m) {  // reactiveDeviceFlash.h:80-108
    if (m.ns < 1 || m.ns > kMaxSpecies || m.nl < 0 || m.nl > 2 || m.phase[0].ns != m.ns ||
        (m.boundaryRecovery != 0 && m.boundaryRecovery != 1) || (m.scalarRecovery != 0 && m.scalarRecovery != 1) ||
        (m.analyticJacobian != 0 && m.analyticJacobian != 1) || (m.stableGasPrune != 0 && m.stableGasPrune != 1) ||
        (m.stableGasPrune && !(m.nl == 1 && m.condensedKind[0] == 0 && m.condensedKind[1] == 0)) ||
        (m.enforceSpeciesTemperatureBounds != 0 && m.enforceSpeciesTemperatureBounds != 1))
        return false;
    if (!fin(m.Tmin) || !fin(m.Tmax) || !(m.Tmin > 0 && m.Tmax > m.Tmin) || !fin(m.pmin) || !fin(m.pmax) ||
        !(m.pmin > 0 && m.pmax > m.pmin) || !fin(m.vtol) || !fin(m.etol) || !fin(m.mutol) ||
        !(m.vtol > 0 && m.etol > 0 && m.mutol > 0))
        return false;
    for (int i = 0; i <= m.nl; ++i)
        if (validateTable(m.phase[i]) != OK) return false;
    for (int k = 0; k < m.ns; ++k)
        if (!fin(m.weights[k]) || m.weights[k] != m.phase[0].sp[k].W) return false;
    for (int i = 0; i < m.nl; ++i) {
        if (m.condensable[i] < 0 || m.condensable[i] >= m.ns || m.phase[i + 1].ns != 1 || !fin(m.liquidTmin[i]) ||
            !fin(m.liquidTc[i]) || !(m.liquidTmin[i] > 0 && m.liquidTc[i] > m.liquidTmin[i]) ||
            m.phase[i + 1].sp[0].W != m.weights[m.condensable[i]])
            return false;
        if (m.condensedKind[i] != 0 && m.condensedKind[i] != 1) return false;
        for (int j = 0; j < i; ++j)
            if (m.condensable[j] == m.condensable[i] && m.condensedKind[j] == m.condensedKind[i]) return false;
    }
    return true;
}

void copyOut(const PRState& s, double* out) {
    const double v[] = {s.T, s.p, s.W, s.V, s.rho, s.e, s.h, s.s, s.cp, s.cv, s.alpha, s.kappaT, s.sound, s.dpdT, s.dpdV};
    for (int i = 0; i < 15; ++i) out[i] = v[i];
    for (int k = 0; k < kMaxSpecies; ++k) out[15 + k] = s.mu[k];
    out[31] = s.phase;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// 1 when the image is valid and inside the oracle's scope (see file header).
int oracle_model_supported(const void* image) {
    const Model& m = *static_cast<const Model*>(image);
    if (!validModel(m)) return 0;
    for (int i = 0; i < m.nl; ++i)
        if (m.condensedKind[i]) return 0;
    return 1;
}

// One UV flash (reactiveDeviceFlash.h:564-567).  out is written on success only.
int oracle_flash_recover(const void* image, const double* q, double energy, const ThermoState* guess,
                         ThermoState* out, uint64_t* counters) {
    const Model& m = *static_cast<const Model*>(image);
    Counters c{};
    Flash flash(m, q, energy, *guess, c);
    Evaluation result;
    const bool ok = flash.run(result);
    if (ok) *out = result.state;
    if (counters) std::memcpy(counters, &c, sizeof(c));
    return ok ? 1 : 0;
}

// Frozen-inventory recovery at the guess's liquid masses (recoverCapillary
// with equilibrium = 0, reactiveDeviceFlash.h:572-590).
int oracle_flash_frozen(const void* image, const double* q, double energy, const ThermoState* guess,
                        ThermoState* out, uint64_t* counters) {
    const Model& m = *static_cast<const Model*>(image);
    Counters c{};
    Flash flash(m, q, energy, *guess, c);
    Evaluation result;
    const double mass[2]{guess->liquidMass[0], guess->liquidMass[1]};
    const bool ok = flash.frozen(mass, result);
    if (ok) *out = result.state;
    if (counters) std::memcpy(counters, &c, sizeof(c));
    return ok ? 1 : 0;
}

// One capillary-interface recovery (recoverCapillary, reactiveDeviceFlash.h:
// 572-590): equilibrium 1 is the UV equilibrium with the boundary search,
// 0 the frozen inventory at the guess's liquid masses; the interface's color
// and pressure jump hold during the solve.  Returns 1 on success, 0 on a
// failed recovery, -2 for an invalid color or jump (the reference's status 2).
int oracle_flash_capillary(const void* image, const double* q, double energy, const ThermoState* guess, double color,
                           double pressureJump, int equilibrium, ThermoState* out, uint64_t* counters) {
    const Model& m = *static_cast<const Model*>(image);
    if (!fin(color) || color < 0 || color > 1 || !fin(pressureJump) ||
        (pressureJump != 0 && (m.nl != 1 || m.condensedKind[0] != 0)))
        return -2;
    Counters c{};
    Flash flash(m, q, energy, *guess, c, color, pressureJump);
    Evaluation result;
    const double mass[2]{guess->liquidMass[0], guess->liquidMass[1]};
    const bool ok = equilibrium ? flash.run(result, true) : flash.frozen(mass, result);
    if (ok) *out = result.state;
    if (counters) std::memcpy(counters, &c, sizeof(c));
    return ok ? 1 : 0;
}
// Batch of capillary recoveries with per-cell color, jump and equilibrium.
int oracle_flash_capillary_batch(const void* image, size_t count, const double* q, const double* energy,
                                 ThermoState* states, const double* color, const double* jump, const int* equilibrium,
                                 int* success, uint64_t* counters, int threads) {
    const Model& m = *static_cast<const Model*>(image);
    long ok = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1) reduction(+ : ok)
    for (long c = 0; c < long(count); ++c) {
        ThermoState result{};
        const int s = oracle_flash_capillary(image, q + size_t(c) * m.ns, energy[c], &states[c], color[c], jump[c],
                                             equilibrium[c], &result, counters ? counters + 9 * size_t(c) : nullptr);
        success[c] = s > 0 ? 1 : 0;
        if (s > 0) states[c] = result;
        ok += s > 0;
    }
    return int(ok);
}

// Batch over cells, threads = OpenMP workers (1 = serial).  q is cell-major
// with ns masses per cell; states hold the seeds on input.
int oracle_flash_batch(const void* image, size_t count, const double* q, const double* energy, ThermoState* states,
                       int* success, uint64_t* counters, int threads) {
    const Model& m = *static_cast<const Model*>(image);
    long ok = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1) reduction(+ : ok)
    for (long c = 0; c < long(count); ++c) {
        ThermoState result{};
        const int s = oracle_flash_recover(image, q + size_t(c) * m.ns, energy[c], &states[c], &result,
                                           counters ? counters + 9 * size_t(c) : nullptr);
        success[c] = s;
        if (s) states[c] = result;
        ok += s;
    }
    return int(ok);
}

// PR properties on phase table `phase` at (T, p) with root `choice`
// (1 gas, 2 liquid) or at (T, rho).  out layout:
// T p W V rho e h s cp cv alpha kappaT sound dpdT dpdV mu[16] phase.
int oracle_pr_tp(const void* image, int phase, double T, double p, const double* Y, int choice, double* out) {
    const Model& m = *static_cast<const Model*>(image);
    PRState s{};
    const Status st = evaluateTP(m.phase[phase], T, p, Y, static_cast<Root>(choice), s, true);
    if (st != OK) return st;
    copyOut(s, out);
    return 0;
}
int oracle_pr_trho(const void* image, int phase, double T, double rho, const double* Y, double* out) {
    const Model& m = *static_cast<const Model*>(image);
    PRState s{};
    const Status st = evaluateTrho(m.phase[phase], T, rho, Y, s, true);
    if (st != OK) return st;
    copyOut(s, out);
    return 0;
}
int oracle_pr_roots(const void* image, int phase, double T, double p, const double* x, double* out) {
    const Model& m = *static_cast<const Model*>(image);
    double a, b, aa, da, d2a;
    Status st = mixing(m.phase[phase], T, x, a, b, aa, da, d2a);
    if (st != OK) return st;
    Roots r{};
    st = roots(T, p, a, b, aa, r);
    if (st != OK) return st;
    out[0] = r.volume[0]; out[1] = r.volume[1]; out[2] = r.volume[2];
    out[3] = r.count; out[4] = r.signedCount; out[5] = r.gasAvailable; out[6] = r.liquidAvailable;
    return 0;
}
// 4x4 full-pivot solve (test hook for the LU against numpy.linalg.solve).
int oracle_solve(const double* matrix, const double* rhs, int n, double* x) {
    double m[4][4]{};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) m[i][j] = matrix[i * n + j];
    return solve(m, rhs, n, x) ? 1 : 0;
}
// NASA9 polynomial at T for species k of phase table `phase`: cp/R, h/RT, s/R.
int oracle_nasa(const void* image, int phase, int k, double T, double* out) {
    const Model& m = *static_cast<const Model*>(image);
    nasa(region(m.phase[phase], k, T), T, out[0], out[1], out[2]);
    return 0;
}

}

// Elementary functions of this build, for tests: which = 0 log, 1 exp,
// 2 x^(1/3) as pow(x, 1.0 / 3.0), 3 acos, 4 cos, 5 expm1.
extern "C" void oracle_elementary(int which, size_t n, const double* x, double* y) {
    for (size_t i = 0; i < n; ++i) {
        const double v = x[i];
        y[i] = which == 0 ? oracle::olog(v) : which == 1 ? oracle::oexp(v) : which == 2 ? oracle::othird(v)
             : which == 3 ? oracle::oacos(v) : which == 4 ? oracle::ocos(v) : oracle::oexpm1(v);
    }
}
