// SPDX-License-Identifier: GPL-3.0-or-later
// The reference's analytic active-set Jacobian of the UV flash
// (src/reactiveThermo/reactiveDeviceFlashJacobian.h), for the case this build
// runs: one active liquid, a flat interface (gas and liquid at the same p)
// and normal coordinates [ln p, ln T, liquid fraction].  These functions give
// the condensable's partial molar enthalpy, energy and volume in the gas and
// d(mu_c/RT)/dn_c at fixed (T, p):
//   refActiveGasDerivatives   from the residual point's gas molar volume
//                             (activeGasDerivatives, :203-308, n = 1)
//   refPhasePartials          from a fresh gas state (phasePartials, :68-132)
//   refGasCompositionPartial  (gasCompositionPartials, :136-198, one entry)
// Expressions and their operation order are the reference's (FP64, no
// contraction), so the flash that uses them takes the reference's Newton
// iterates.  The critical temperatures are the table's, computed once with
// the reference's own expression (a OmegaB / (b OmegaA R)).
#ifndef REACTIVE_AMD_FLASH_JACOBIAN_H
#define REACTIVE_AMD_FLASH_JACOBIAN_H
#include "pr_device.h"

namespace ramd {

struct GasPartials {
    double hbar, ubar, vbar, dpsi;  // J/kmol, J/kmol, m^3/kmol, 1/kmol
};

// activeGasDerivatives (:203-308) for the single species k.
template <int NS>
RHD bool refActiveGasDerivatives(const Table<NS>& t, double T, double p, double molarVolume, const double* Y,
                                 double gasMass, int k, GasPartials& o) {
    if (!(T > 0) || !(p > 0) || !(molarVolume > 0) || !(gasMass > 0)) return false;
    double x[NS], W = 0;
    if (massToMole(t, Y, x, W) != kOk) return false;
    const double totalMoles = gasMass / W;
    const double V = molarVolume, RT = kR * T;
    double q[NS], dq[NS];
    double b = 0, aa = 0, da = 0;
    for (int i = 0; i < NS; ++i) {
        const double Tc = t.Tc[i];
        const double sqrtTr = m_sqrt(T / Tc);
        const double f = 1.0 + t.kappa[i] * (1.0 - sqrtTr);
        if (!(Tc > 0) || !(sqrtTr > 0) || ab(f) <= 1e-14) return false;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[i] = ab(f);
        dq[i] = -sign * t.kappa[i] / (2.0 * Tc * sqrtTr);
        b += x[i] * t.b[i];
    }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            const double w = x[i] * x[j] * t.aij[i][j];
            aa += w * q[i] * q[j];
            da += w * (dq[i] * q[j] + q[i] * dq[j]);
        }
    const double vmb = V - b, den = V * V + 2.0 * V * b - b * b;
    if (!(b > 0) || !(aa > 0) || !(vmb > 0) || !(den > 0)) return false;
    const double calculatedP = RT / vmb - aa / den;
    const double dpdV = -RT / (vmb * vmb) + 2.0 * aa * (V + b) / (den * den);
    const double dpdT = kR / vmb - da / den;
    if (!(dpdV < 0) || !fin(calculatedP) || !fin(dpdT) || ab(calculatedP - p) > 1e-7 * hi(1.0, p)) return false;
    const double c1 = 1.0 + kSqrt2, c2 = 1.0 - kSqrt2;
    const double A = V + c1 * b, B = V + c2 * b;
    const double logRatio = m_log(A / B);
    const double L = logRatio / (2.0 * kSqrt2 * b);
    const double fac = T * da - aa;
    const double fac2 = V + T * dpdT / dpdV;
    if (!(x[k] > 0)) return false;
    double pp = 0, tmp = 0;
    for (int i = 0; i < NS; ++i) {
        const double pair = t.aij[k][i] * q[k] * q[i];
        pp += x[i] * pair;
        tmp += x[i] * pair * (2.0 * dq[i] / q[i] + 2.0 * dq[k] / q[k]);
    }
    double cpR, hRT, sR;
    nasa(region(t, k, T), T, cpR, hRT, sR);
    const double bk = t.b[k];
    const double dpdni = RT / vmb + RT * bk / (vmb * vmb) - 2.0 * pp / den + 2.0 * vmb * aa * bk / (den * den);
    const double hEv = V * dpdni - RT - bk / (2.0 * kSqrt2 * b * b) * logRatio * fac + V * bk / (b * den) * fac +
                       L * (T * tmp - 2.0 * pp);
    o.hbar = RT * hRT + hEv - fac2 * dpdni;
    const double volumeNumerator = RT + RT * b / vmb + RT * bk / vmb + RT * b * bk / (vmb * vmb) -
                                   2.0 * V * pp / den + 2.0 * V * vmb * aa * bk / (den * den);
    const double volumeDenominator = p + RT * b / (vmb * vmb) + aa / den - 2.0 * V * (V + b) * aa / (den * den);
    o.vbar = volumeNumerator / volumeDenominator;
    o.ubar = o.hbar - p * o.vbar;
    if (!fin(o.hbar) || !fin(o.ubar) || !fin(o.vbar)) return false;
    const double Pb = RT / (vmb * vmb) + 2.0 * aa * vmb / (den * den);
    const double C = 2.0 * kSqrt2;
    const int j = k;
    const double db = t.b[j] - b;
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
    o.dpsi = (dxk / x[k] + dResidual / RT) / totalMoles;
    return fin(o.dpsi);
}

// phasePartials (:68-132): partial molar h, u, v of every gas species from
// the gas state s at (T, s.p); x and W receive the mole fractions and the
// mean molar mass.
template <int NS>
RHD bool refPhasePartials(const Table<NS>& t, double T, const double* Y, const PRState& s, double* hbar,
                          double* ubar, double* vbar, double* x, double& W) {
    if (massToMole(t, Y, x, W) != kOk) return false;
    MixT<double> mx;
    if (mixing(t, T, x, mx) != kOk) return false;
    const double b = mx.b, aa = mx.aa, da = mx.da;
    const double V = s.V, p = s.p, RT = kR * T;
    const double vmb = V - b;
    const double den = V * V + 2.0 * V * b - b * b;
    if (!(vmb > 0) || !(den > 0) || !(s.dpdV < 0)) return false;
    const double logRatio = m_log((V + (1.0 + kSqrt2) * b) / (V + (1.0 - kSqrt2) * b));
    const double L = logRatio / (2.0 * kSqrt2 * b);
    const double fac = T * da - aa;
    const double fac2 = V + T * s.dpdT / s.dpdV;
    double q[NS], dq[NS], pp[NS], tmp[NS];
    for (int k = 0; k < NS; ++k) {
        const double Tc = t.Tc[k];
        const double sqrtTr = m_sqrt(T / Tc);
        const double f = 1.0 + t.kappa[k] * (1.0 - sqrtTr);
        if (ab(f) <= 1e-14) return false;
        const double sign = f > 0 ? 1.0 : -1.0;
        q[k] = ab(f);
        dq[k] = -sign * t.kappa[k] / (2.0 * Tc * sqrtTr);
        pp[k] = tmp[k] = 0;
    }
    for (int k = 0; k < NS; ++k)
        for (int i = 0; i < NS; ++i) {
            const double pair = t.aij[k][i] * q[k] * q[i];
            pp[k] += x[i] * pair;
            tmp[k] += x[i] * pair * (2.0 * dq[i] / q[i] + 2.0 * dq[k] / q[k]);
        }
    for (int k = 0; k < NS; ++k) {
        double cpR, hRT, sR;
        nasa(region(t, k, T), T, cpR, hRT, sR);
        const double bk = t.b[k];
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

// gasCompositionPartials (:136-198) for target = source = species k.
template <int NS>
RHD bool refGasCompositionPartial(const Table<NS>& t, double T, const double* x, double totalMoles, const PRState& s,
                                  int k, double& dpsi) {
    if (!(totalMoles > 0)) return false;
    MixT<double> mx;
    if (mixing(t, T, x, mx) != kOk) return false;
    const double b = mx.b, aa = mx.aa;
    const double V = s.V, RT = kR * T;
    const double vmb = V - b, den = V * V + 2.0 * V * b - b * b;
    if (!(vmb > 0) || !(den > 0) || !(s.dpdV < 0)) return false;
    const double c1 = 1.0 + kSqrt2, c2 = 1.0 - kSqrt2;
    const double A = V + c1 * b, B = V + c2 * b;
    const double logRatio = m_log(A / B);
    const double C = 2.0 * kSqrt2;
    double q[NS], pp[NS];
    for (int i = 0; i < NS; ++i) {
        q[i] = ab(1.0 + t.kappa[i] * (1.0 - m_sqrt(T / t.Tc[i])));
        if (!(q[i] > 1e-14)) return false;
        pp[i] = 0;
    }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) pp[i] += x[j] * t.aij[i][j] * q[i] * q[j];
    const double Pb = RT / (vmb * vmb) + 2.0 * aa * vmb / (den * den);
    const int j = k;
    const double db = t.b[j] - b;
    const double daa = 2.0 * (pp[j] - aa);
    const double dV = -(Pb * db - daa / den) / s.dpdV;
    const double dDen = 2.0 * (V + b) * dV + 2.0 * (V - b) * db;
    const double dLogRatio = (dV + c1 * db) / A - (dV + c2 * db) / B;
    if (!(x[k] > 0)) return false;
    const double dxk = 1.0 - x[k];
    const double dpp = t.aij[k][j] * q[k] * q[j] - pp[k];
    const double bk = t.b[k];
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

}  // namespace ramd
#endif
