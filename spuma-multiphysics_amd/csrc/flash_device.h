// SPDX-License-Identifier: GPL-3.0-or-later
// Homogeneous-equilibrium UV flash for one cell per work-item on gfx950.
//
// The algorithm is the reference's ReactiveDeviceFlash::Flash
// (src/reactiveThermo/reactiveDeviceFlash.h:140-560): the same candidate
// order (all-gas frozen/scalar recovery, all-liquid, five liquid-fraction
// seeds per active mask, then four restart temperatures), the same Newton
// step limits, line-search fractions, stability tests and maximum-entropy
// selection, and the same finite-difference Jacobian.  Expression grouping
// follows the reference so that FP64 results match the CPU path.
//
// GPU-specific structure: the gas species count NS and the liquid count NL
// (0 or 1) are template parameters, so every per-species and per-unknown
// array is register resident; the evaluation record keeps only the fields
// the flash reads; the model is a compact table image in global memory.
#ifndef REACTIVE_AMD_FLASH_DEVICE_H
#define REACTIVE_AMD_FLASH_DEVICE_H
#include "pr_device.h"

#ifndef RAMD_NOINLINE
#define RAMD_NOINLINE __device__ __noinline__
#endif
#define RD __device__ inline

namespace ramd {

template <int NS>
struct HemModel {
    Table<NS> gas;
    Table<1> liquid;
    PhaseHeuristic gasFixed;     // composition-independent heuristic when NS == 1
    PhaseHeuristic liquidFixed;  // pure liquid table
    int nl, condensable, scalarRecovery, stableGasPrune, enforceBounds;
    double weights[NS];
    double liquidTmin, liquidTc, Tmin, Tmax, pmin, pmax, vtol, etol, mutol;
};

// Thermo output, byte-compatible with ReactiveThermoState (reactiveThermo.h:11-19).
struct ThermoOut {
    double p, T, rho, e, entropy, soundFrozen, soundEquilibrium;
    double volumeResidual, energyResidual, chemicalResidual;
    double alphaGas, alphaLiquid[2], rhoGas, rhoLiquid[2], liquidMass[2];
    double gasMass, cp, cv;
    int activeLiquids, iterations;
};
static_assert(sizeof(ThermoOut) == 176, "ReactiveThermoState layout");

struct Counts {
    uint32_t phases, residuals, candidates, stable, failures, fd;
};

// Evaluation record (reactiveDeviceFlash.h:64-73) for at most one liquid.
struct Eval {
    double p, T, rho, e, entropy, cf, ceq, volRes, enRes, chemRes;
    double alphaGas, alphaLiq, rhoGas, rhoLiq, liqMass, gasMass, cp, cv;
    int active, iterations;
    double volume, energy, entropyDensity, Vp, VT, Ep, ET, cpDensity, muGas, muLiquid;
};

RD void zeroEval(Eval& v) {
    v.p = v.T = v.rho = v.e = v.entropy = v.cf = v.ceq = v.volRes = v.enRes = v.chemRes = 0;
    v.alphaGas = v.alphaLiq = v.rhoGas = v.rhoLiq = v.liqMass = v.gasMass = v.cp = v.cv = 0;
    v.active = v.iterations = 0;
    v.volume = v.energy = v.entropyDensity = v.Vp = v.VT = v.Ep = v.ET = v.cpDensity = v.muGas = v.muLiquid = 0;
}

RD void writeOut(const Eval& v, ThermoOut& o) {
    o.p = v.p; o.T = v.T; o.rho = v.rho; o.e = v.e; o.entropy = v.entropy;
    o.soundFrozen = v.cf; o.soundEquilibrium = v.ceq;
    o.volumeResidual = v.volRes; o.energyResidual = v.enRes; o.chemicalResidual = v.chemRes;
    o.alphaGas = v.alphaGas; o.alphaLiquid[0] = v.alphaLiq; o.alphaLiquid[1] = 0;
    o.rhoGas = v.rhoGas; o.rhoLiquid[0] = v.rhoLiq; o.rhoLiquid[1] = 0;
    o.liquidMass[0] = v.liqMass; o.liquidMass[1] = 0;
    o.gasMass = v.gasMass; o.cp = v.cp; o.cv = v.cv;
    o.activeLiquids = v.active; o.iterations = v.iterations;
}

// Full-pivot LU with Eigen's rank rule (reactiveDeviceFlash.h:111-139), n fixed.
template <int n>
RD bool solveN(const double (&m)[3][3], const double* rhs, double* x) {
    double a[n][n], b[n], largest = 0;
    int column[n];
    for (int i = 0; i < n; ++i) {
        column[i] = i;
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
                if (::fabs(a[i][j]) > pivot) { pivot = ::fabs(a[i][j]); row = i; col = j; }
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
    double z[n];
    for (int i = n - 1; i >= 0; --i) {
        double v = b[i];
        for (int j = i + 1; j < n; ++j) v -= a[i][j] * z[j];
        z[i] = v / a[i][i];
        if (!fin(z[i])) return false;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (column[i] == j) x[j] = z[i];
    return true;
}

template <int NS, int NL>
struct Flash {
    static_assert(NL == 0 || NL == 1, "one pure liquid at most");
    const HemModel<NS>& m;
    double q[NS];
    double energy;
    double guessP, guessT, guessLiquid;
    Counts& n;
    double cacheP, cacheT;  // one unavailable liquid result (reactiveDeviceFlash.h:142-145)
    int cacheStatus;
    bool cacheVirtual;
    double memoKey[NS];     // composition of the memoised gas heuristic (NS > 1)
    PhaseHeuristic memo;
    bool memoValid;

    RD Flash(const HemModel<NS>& model, Counts& counts) : m(model), n(counts) {
        cacheP = cacheT = 0;
        cacheStatus = -1;
        cacheVirtual = false;
        memoValid = false;
    }

    RD double density() const {
        double r = 0;
        for (int k = 0; k < NS; ++k) r += q[k];
        return r;
    }
    RD bool bounds(double p, double T) const {
        return fin(p) && fin(T) && p >= m.pmin && p <= m.pmax && T >= m.Tmin && T <= m.Tmax;
    }
    static RD bool unavailable(int s) { return s == kUnavailable || s == kNoRoot; }

    RD const PhaseHeuristic& gasHeuristic(const double* x, double W, const Mix& mx) {
        if (NS == 1) return m.gasFixed;
        bool hit = memoValid;
        for (int k = 0; k < NS; ++k) hit = hit && memoKey[k] == x[k];
        if (!hit) {
            phaseHeuristic(m.gas, x, W, mx.a, mx.b, memo);
            for (int k = 0; k < NS; ++k) memoKey[k] = x[k];
            memoValid = true;
        }
        return memo;
    }

    // Gas PR state at (T, p) on the gas root (evaluateTPTrusted, reactiveDevicePR.h:574-609).
    RAMD_NOINLINE int gasTP(double T, double p, const double* Y, PRState& s, double* mu) {
        if (!fin(T) || !fin(p) || T <= 0 || p <= 0) return kBadInput;
        double x[NS], W = 0;
        int st = massToMole(m.gas, Y, x, W);
        if (st != kOk) return st;
        Mix mx;
        st = mixing(m.gas, T, x, mx);
        if (st != kOk) return st;
        Roots r;
        st = roots(T, p, mx.a, mx.b, mx.aa, r);
        if (st != kOk) return st;
        if (!r.gas) return kUnavailable;
        st = stateTrho(m.gas, T, W / r.vLast, x, W, mx, s, mu);
        if (st != kOk) return st;
        st = finishState<NS>(gasHeuristic(x, W, mx), T, s, mu);
        if (st != kOk) return st;
        if (ab(s.p - p) > 1e-7 * hi(1.0, p)) return kNonFinite;
        s.p = p;
        if (s.phase >= 0) return kUnavailable;
        return kOk;
    }
    // Pure-liquid PR state at (T, p) on the liquid root.
    RAMD_NOINLINE int liquidTP(double T, double p, PRState& s, double* mu) {
        if (!fin(T) || !fin(p) || T <= 0 || p <= 0) return kBadInput;
        const double Y[1] = {1.0};
        double x[1], W = 0;
        int st = massToMole(m.liquid, Y, x, W);
        if (st != kOk) return st;
        Mix mx;
        st = mixing(m.liquid, T, x, mx);
        if (st != kOk) return st;
        Roots r;
        st = roots(T, p, mx.a, mx.b, mx.aa, r);
        if (st != kOk) return st;
        if (!r.liquid) return kUnavailable;
        st = stateTrho(m.liquid, T, W / r.v0, x, W, mx, s, mu);
        if (st != kOk) return st;
        st = finishState<1>(m.liquidFixed, T, s, mu);
        if (st != kOk) return st;
        if (ab(s.p - p) > 1e-7 * hi(1.0, p)) return kNonFinite;
        s.p = p;
        if (s.phase < 0) return kUnavailable;
        return kOk;
    }

    // reactiveDeviceFlash.h:157-196 for index -1 (gas) and 0 (liquid).
    RD int phaseStatus(int index, double p, double T, const double* Y, PRState& out, double* mu) {
        ++n.phases;
        if (!bounds(p, T)) return kBadInput;
        if (index >= 0 && T < m.liquidTmin) return kBadInput;
        if (index >= 0 && T >= m.liquidTc * (1 - 1e-10)) return kUnavailable;
        if (index < 0 && m.enforceBounds)
            for (int k = 0; k < NS; ++k)
                if (Y[k] > 0 && (T < m.gas.tlo[k] || T > m.gas.thi[k])) return kBadInput;
        if (index >= 0) return liquidTP(T, p, out, mu);
        return gasTP(T, p, Y, out, mu);
    }

    static RD void accumulate(Eval& v, double mass, const PRState& p, double phaseP) {  // :203-207
        const double volume = mass / p.rho, vp = -volume * p.kappaT, vt = volume * p.alpha;
        v.volume += volume; v.energy += mass * p.e; v.entropyDensity += mass * p.s; v.Vp += vp; v.VT += vt;
        v.Ep += -v.T * vt - phaseP * vp;
        v.ET += mass * p.cp - phaseP * vt;
        v.cpDensity += mass * p.cp;
    }

    // Gas + liquid mixture at common (p, T) (reactiveDeviceFlash.h:208-262).
    RAMD_NOINLINE bool evaluate(double mass0, double p, double T, Eval& v, bool virtualLiquids) {
        zeroEval(v);
        if (!bounds(p, T)) return false;
        if (NL > 0 && mass0 > 0 && cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual)
            return false;
        v.p = p; v.T = T; v.rho = density();
        if (!(v.rho > 0)) return false;
        double gas[NS];
        for (int k = 0; k < NS; ++k) {
            if (!fin(q[k]) || q[k] < 0) return false;
            gas[k] = q[k];
        }
        const int c = NL ? m.condensable : 0;
        if (NL) {
            if (!fin(mass0) || mass0 < 0 || mass0 > q[c]) return false;
            gas[c] = gas[c] - mass0;
            if (!fin(gas[c]) || gas[c] < 0 || gas[c] > q[c]) return false;
        }
        double mg = 0;
        for (int k = 0; k < NS; ++k) mg += gas[k];
        v.gasMass = mg;
        if (mg > 0) {
            for (int k = 0; k < NS; ++k) gas[k] /= mg;
            PRState prop;
            double mu[NS];
            if (phaseStatus(-1, p, T, gas, prop, virtualLiquids ? mu : nullptr) != kOk) return false;
            accumulate(v, mg, prop, p);
            v.alphaGas = mg / prop.rho;
            v.rhoGas = prop.rho;
            if (NL && virtualLiquids) v.muGas = mu[c] / (8314.46261815324 * T);
        }
        if (NL) {
            v.liqMass = mass0;
            if (mass0 > 0 || (virtualLiquids && q[c] > 0)) {
                PRState prop;
                double mu[1];
                int st;
                if (cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual) {
                    ++n.phases;
                    st = cacheStatus;
                } else {
                    st = phaseStatus(0, p, T, nullptr, prop, virtualLiquids ? mu : nullptr);
                    if (unavailable(st) && T >= m.liquidTmin && T < m.liquidTc * (1 - 1e-10)) {
                        cacheP = p; cacheT = T; cacheVirtual = virtualLiquids; cacheStatus = st;
                    }
                }
                if (st != kOk) {
                    if (mass0 > 0 || !unavailable(st)) return false;
                    v.muLiquid = __builtin_huge_val();
                } else {
                    if (virtualLiquids) v.muLiquid = mu[0] / (8314.46261815324 * T);
                    if (mass0 > 0) {
                        accumulate(v, mass0, prop, p);
                        v.rhoLiq = prop.rho;
                        v.alphaLiq = mass0 / prop.rho;
                        v.active |= 1;
                    }
                }
            }
        }
        v.e = v.energy / v.rho; v.entropy = v.entropyDensity / v.rho; v.cp = v.cpDensity / v.rho;
        v.cv = (v.cpDensity + T * v.VT * v.VT / v.Vp) / v.rho;
        if (!fin(v.cv) || v.cv <= 0) return false;
        const double specific = v.volume / v.rho, vp = v.Vp / v.rho, vt = v.VT / v.rho;
        const double adiabatic = vp + T * vt * vt / v.cp;
        if (!(adiabatic < 0)) return false;
        v.cf = m_sqrt(-specific * specific / adiabatic);
        v.ceq = v.cf;
        return fin(v.cf) && v.cf > 0;
    }
    RD double energyScale(const Eval& v) const { return hi(1e5 * v.rho, v.cpDensity * v.T); }

    // Single-phase gas energy residual at fixed rho (reactiveDeviceFlash.h:264-275).
    RAMD_NOINLINE bool gasProbe(double T, double& F, double& cv, double& p, double* cp) {
        const double rho = density();
        double Y[NS];
        for (int k = 0; k < NS; ++k) Y[k] = q[k] / rho;
        ++n.phases;
        PRState prop;
        {
            if (!fin(T) || !fin(rho) || T <= 0 || rho <= 0) return false;
            double x[NS], W = 0;
            if (massToMole(m.gas, Y, x, W) != kOk) return false;
            Mix mx;
            if (mixing(m.gas, T, x, mx) != kOk) return false;
            if (stateTrho(m.gas, T, rho, x, W, mx, prop, nullptr) != kOk) return false;
            if (finishState<NS>(gasHeuristic(x, W, mx), T, prop, nullptr) != kOk) return false;
        }
        const double gasP = prop.p;
        p = gasP + 0.0 * 0.0;
        if (!bounds(p, T)) return false;
        PRState gas;
        if (phaseStatus(-1, gasP, T, Y, gas, nullptr) != kOk || ::fabs(gas.rho - rho) > 1e-8 * rho) return false;
        F = rho * prop.e - energy;
        cv = rho * prop.cv;
        if (cp) *cp = prop.cp;
        return fin(F) && fin(cv) && cv > 0;
    }
    // Bracketed scalar Newton in T (reactiveDeviceFlash.h:276-299).
    RAMD_NOINLINE bool scalarGas(Eval& result) {
        const double rho = density();
        double T = clamp(guessT, m.Tmin, m.Tmax), F, cv, p, cp;
        if (!gasProbe(T, F, cv, p, &cp)) return false;
        const double scale = hi(1e5 * rho, rho * cp * T);
        double lower = T, upper = T, fl = F, fh = F;
        if (::fabs(F) > m.etol * scale) {
            double radius = hi(1., .01 * T);
            for (int j = 0; j < 24 && (fl > 0 || fh < 0); ++j, radius *= 2) {
                if (fl > 0) {
                    double candidate = hi(m.Tmin, T - radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        if (gasProbe(candidate, value, dcv, dp, nullptr)) { fl = value; lower = candidate; found = true; }
                        else candidate = .5 * (lower + candidate);
                    }
                    if (!found) return false;
                }
                if (fh < 0) {
                    double candidate = lo(m.Tmax, T + radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        if (gasProbe(candidate, value, dcv, dp, nullptr)) { fh = value; upper = candidate; found = true; }
                        else candidate = .5 * (upper + candidate);
                    }
                    if (!found) return false;
                }
                if (::fabs(fl) <= m.etol * scale) { T = lower; if (!gasProbe(T, F, cv, p, nullptr)) return false; break; }
                if (::fabs(fh) <= m.etol * scale) { T = upper; if (!gasProbe(T, F, cv, p, nullptr)) return false; break; }
                if (lower == m.Tmin && upper == m.Tmax) break;
            }
            if (!(::fabs(F) <= m.etol * scale || (fl <= 0 && fh >= 0 && upper > lower))) return false;
        }
        int it = 0;
        for (; it < 70 && ::fabs(F) > m.etol * scale; ++it) {
            if (F > 0) upper = T; else lower = T;
            const double next = T - F / cv;
            T = fin(next) && next > lower && next < upper ? next : .5 * (lower + upper);
            if (!gasProbe(T, F, cv, p, nullptr)) return false;
        }
        if (::fabs(F) > m.etol * scale) return false;
        if (!evaluate(0.0, p, T, result, false)) return false;
        result.volRes = ::fabs(result.volume - 1);
        result.enRes = ::fabs(result.energy - energy) / energyScale(result);
        result.iterations = it;
        return result.volRes <= m.vtol && result.enRes <= m.etol;
    }
    // Fixed split, 2x2 Newton in (ln p, ln T) with line search (reactiveDeviceFlash.h:300-326).
    RAMD_NOINLINE bool frozen(double mass0, Eval& value, int stallLimit, double mass1 = 0.0) {
        if (!fin(energy)) return false;
        if (m.scalarRecovery && mass0 == 0 && mass1 == 0 && scalarGas(value)) return true;
        double p = clamp(guessP, m.pmin, m.pmax), T = clamp(guessT, m.Tmin, m.Tmax);
        if (!evaluate(mass0, p, T, value, false)) return false;
        int stalled = 0;
        for (int it = 0; it < 70; ++it) {
            const double scale = energyScale(value), rv = value.volume - 1, re = (value.energy - energy) / scale;
            value.volRes = ::fabs(rv);
            value.enRes = ::fabs(re);
            value.iterations = it;
            if (::fabs(rv) <= m.vtol && ::fabs(re) <= m.etol) return true;
            double jac[3][3] = {};
            jac[0][0] = value.Vp * p; jac[0][1] = value.VT * T;
            jac[1][0] = value.Ep * p / scale; jac[1][1] = value.ET * T / scale;
            double rhs[2] = {-rv, -re}, step[2];
            if (!solveN<2>(jac, rhs, step)) return false;
            const double limit = hi(1., hi(::fabs(step[0]) / .7, ::fabs(step[1]) / .2));
            step[0] /= limit; step[1] /= limit;
            const double previous = hi(::fabs(rv), ::fabs(re));
            bool accepted = false;
            for (double fraction = 1; fraction > 1e-9; fraction *= .5) {
                Eval trial;
                const double pp = p * m_exp(fraction * step[0]), tt = T * m_exp(fraction * step[1]);
                if (evaluate(mass0, pp, tt, trial, false)) {
                    const double next = hi(::fabs(trial.volume - 1), ::fabs(trial.energy - energy) / scale);
                    if (next < previous * (1 - 1e-4 * fraction) || next < lo(m.vtol, m.etol)) {
                        value = trial; p = pp; T = tt; accepted = true;
                        stalled = fraction < 1. / 32 ? stalled + 1 : 0;
                        break;
                    }
                }
            }
            if (!accepted) return false;
            if (stallLimit > 0 && stalled >= stallLimit &&
                !(::fabs(value.volume - 1) <= m.vtol && ::fabs(value.energy - energy) / energyScale(value) <= m.etol))
                return false;
        }
        return false;
    }
    // Residual [V-1, (E-e)/scale, mu_l/RT - mu_g/RT] in x = [ln p, ln T, beta]
    // (reactiveDeviceFlash.h:327-337).
    RAMD_NOINLINE bool residual(const double* x, double scale, double* f, Eval* out) {
        ++n.residuals;
        if (!(x[2] >= 0 && x[2] < 1)) return false;
        const double mass0 = x[2] * q[m.condensable];
        Eval v;
        if (!evaluate(mass0, m_exp(x[0]), m_exp(x[1]), v, true) || !(v.gasMass > 0)) return false;
        f[0] = v.volume - 1;
        f[1] = (v.energy - energy) / scale;
        f[2] = v.muLiquid - v.muGas;
        if (!fin(f[2])) return false;
        if (out) *out = v;
        return true;
    }
    // Central finite differences with step halving (reactiveDeviceFlash.h:338-373).
    RAMD_NOINLINE bool jacobian(const double* x, double scale, const double* f, double (&jac)[3][3]) {
        ++n.fd;
        for (int j = 0; j < 3; ++j) {
            double h = j < 2 ? 2e-6 : 2e-5;
            bool success = false;
            for (int attempt = 0; attempt < 12 && !success; ++attempt, h *= .5) {
                double xp[3], xm[3], fp[3] = {}, fm[3] = {};
                for (int i = 0; i < 3; ++i) xp[i] = xm[i] = x[i];
                xp[j] += h;
                xm[j] -= h;
                const bool plus = residual(xp, scale, fp, nullptr), minus = residual(xm, scale, fm, nullptr);
                if (!plus && !minus) continue;
                success = true;
                for (int i = 0; i < 3; ++i) {
                    const double v = plus && minus ? (fp[i] - fm[i]) / (2 * h) : plus ? (fp[i] - f[i]) / h : (f[i] - fm[i]) / h;
                    jac[i][j] = v;
                    success = success && fin(v);
                }
            }
            if (!success) return false;
        }
        return true;
    }
    // Damped Newton on the one-liquid active set (reactiveDeviceFlash.h:374-448).
    RAMD_NOINLINE bool activeFlash(double seed, Eval& value, bool restart, double restartT) {
        ++n.candidates;
        const double rho = density();
        const double gT = restart ? restartT : guessT;
        const double scale = hi(1e5 * rho, hi(::fabs(energy), rho * 2000 * hi(gT, m.Tmin)));
        double x[3];
        x[0] = m_log(clamp(guessP, m.pmin, m.pmax));
        x[1] = m_log(clamp(gT, m.Tmin, m.Tmax));
        const double previousFraction = guessLiquid / q[m.condensable];
        x[2] = seed < 0 ? clamp(previousFraction, 1e-8, 1 - 1e-8) : seed;
        double f[3];
        if (!residual(x, scale, f, &value)) return false;
        for (int it = 0; it < 90; ++it) {
            const double chemical = hi(0.0, ::fabs(f[2]));
            double jac[3][3];
            if (::fabs(f[0]) <= m.vtol && ::fabs(f[1]) <= m.etol && chemical <= m.mutol) {
                value.volRes = ::fabs(f[0]);
                value.enRes = ::fabs(f[1]);
                value.chemRes = chemical;
                value.iterations = it;
                if (!jacobian(x, scale, f, jac)) return false;
                double rhs[3] = {-1 / rho, value.p / (rho * scale), 0.0}, response[3];
                if (!solveN<3>(jac, rhs, response)) return false;
                const double squared = value.p * response[0];
                if (!fin(squared) || squared <= 0) return false;
                value.ceq = m_sqrt(squared);
                return value.ceq <= value.cf * (1 + 2e-3);
            }
            if (!jacobian(x, scale, f, jac)) return false;
            double rhs[3] = {-f[0], -f[1], -f[2]}, step[3];
            if (!solveN<3>(jac, rhs, step)) return false;
            double limiter = hi(1., hi(::fabs(step[0]) / .7, ::fabs(step[1]) / .18));
            limiter = hi(limiter, ::fabs(step[2]) / .3);
            for (int j = 0; j < 3; ++j) step[j] /= limiter;
            bool accepted = false;
            const double previous = hi(hi(hi(0.0, ::fabs(f[0])), ::fabs(f[1])), ::fabs(f[2]));
            double fraction = 1;
            for (int index = 0; index < 30; ++index, fraction *= .5) {
                double tx[3], tf[3];
                for (int j = 0; j < 3; ++j) tx[j] = x[j] + fraction * step[j];
                if (!(tx[2] >= 0 && tx[2] < 1)) {
                    // rejected fractions form a prefix; bisect to its end (:419-438)
                    int lower = index, upper = 30;
                    while (upper - lower > 1) {
                        const int mid = (lower + upper) / 2;
                        const double y = x[2] + ::ldexp(1.0, -mid) * step[2];
                        if (!(y >= 0 && y < 1)) lower = mid; else upper = mid;
                    }
                    n.residuals += uint32_t(upper - index);
                    if (upper == 30) break;
                    index = upper;
                    fraction = ::ldexp(1.0, -upper);
                    for (int j = 0; j < 3; ++j) tx[j] = x[j] + fraction * step[j];
                }
                Eval trial;
                if (residual(tx, scale, tf, &trial) &&
                    hi(hi(hi(0.0, ::fabs(tf[0])), ::fabs(tf[1])), ::fabs(tf[2])) < previous * (1 - 1e-4 * fraction)) {
                    for (int j = 0; j < 3; ++j) { x[j] = tx[j]; f[j] = tf[j]; }
                    value = trial;
                    accepted = true;
                    break;
                }
            }
            if (!accepted) return false;
        }
        return false;
    }
    // Stability (reactiveDeviceFlash.h:449-485) for at most one liquid.
    RAMD_NOINLINE bool stable(Eval& value) {
        const double p = value.p, T = value.T;
        if (value.gasMass > 0) {
            Eval comparison;
            if (!evaluate(value.liqMass, p, T, comparison, true)) return false;
            if (!NL || q[m.condensable] == 0) return true;
            const double affinity = comparison.muLiquid - comparison.muGas;
            if (value.liqMass > 0) {
                if (!fin(affinity) || ::fabs(affinity) > m.mutol * 2) return false;
            } else if (affinity < -m.mutol) return false;
            return true;
        }
        if (!NL || !(value.liqMass > 0)) return false;
        PRState prop;
        double mu[1];
        if (phaseStatus(0, p, T, nullptr, prop, mu) != kOk) return false;
        const double muL = mu[0] / (8314.46261815324 * T);
        // Tangent-plane distance of the pure vapour of the condensable (:474-485, n = 1).
        double mole[NS], Y[NS], sum = 0;
        for (int k = 0; k < NS; ++k) mole[k] = 0;
        mole[m.condensable] = 1;
        for (int k = 0; k < NS; ++k) { Y[k] = mole[k] * m.weights[k]; sum += Y[k]; }
        for (int k = 0; k < NS; ++k) Y[k] /= sum;
        PRState gas;
        const int st = phaseStatus(-1, p, T, Y, gas, nullptr);
        double tangent;
        if (st != kOk) tangent = unavailable(st) ? __builtin_huge_val() : __builtin_nan("");
        else {
            const double reference = 0 + mole[m.condensable] * muL;
            double W = 0;
            for (int k = 0; k < NS; ++k) W += mole[k] * m.weights[k];
            tangent = (gas.h - T * gas.s) * W / (8314.46261815324 * T) - reference;
        }
        return tangent >= -m.mutol;
    }
    RD void consider(Eval& candidate, Eval& best, bool& have) {  // :486-489
        if (stable(candidate)) {
            ++n.stable;
            if (!have || candidate.entropyDensity > best.entropyDensity) { best = candidate; have = true; }
        } else {
            ++n.failures;
        }
    }
    // Candidate search (reactiveDeviceFlash.h:490-543).
    RAMD_NOINLINE bool search(Eval& best) {
        bool have = false;
        Eval candidate;
        bool gasStable = false;
        const bool prunable = m.stableGasPrune != 0;
        if (frozen(0.0, candidate, prunable ? 4 : 0)) {
            const uint32_t before = n.stable;
            consider(candidate, best, have);
            gasStable = n.stable > before && candidate.gasMass > 0;
        } else {
            ++n.failures;
        }
        if (!NL) return have;
        const bool prune = prunable && gasStable;
        const int c = m.condensable;
        double other = 0;
        for (int k = 0; k < NS; ++k)
            if (k != c) other += q[k];
        if (other == 0) {
            if (frozen(q[c], candidate, 0)) consider(candidate, best, have);
            else ++n.failures;
        }
        const double seeds[5] = {-1, .5, .95, .1, .9999};
        if (!prune && q[c] > 0)
            for (int j = 0; j < 5; ++j) {
                if (activeFlash(seeds[j], candidate, false, 0.0)) consider(candidate, best, have);
                else ++n.failures;
            }
        if (!have && q[c] > 0) {
            const double lower = hi(m.Tmin, m.liquidTmin), upper = lo(m.Tmax, m.liquidTc * (1 - 1e-10));
            if (upper > lower)
                for (int t = 0; t < 4; ++t) {
                    const double frac = t == 0 ? 0.02 : t == 1 ? 0.25 : t == 2 ? 0.5 : 0.75;  // reactiveFlashSeeds.h:13-16
                    const double restartT = lower + (upper - lower) * frac;
                    for (int j = 0; j < 5; ++j) {
                        if (activeFlash(seeds[j], candidate, true, restartT)) consider(candidate, best, have);
                        else ++n.failures;
                    }
                }
        }
        return have;
    }
    RD bool run(Eval& result) {  // :544-560, boundaryRecovery = 0
        if (!fin(energy) || !(density() > 0)) return false;
        for (int k = 0; k < NS; ++k)
            if (!fin(q[k]) || q[k] < 0) return false;
        return search(result);
    }
};

}  // namespace ramd
#endif
