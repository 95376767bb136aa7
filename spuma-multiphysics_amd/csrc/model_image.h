// SPDX-License-Identifier: GPL-3.0-or-later
// Host view of the opaque HEM model image passed through the ABI
// (ReactiveDeviceFlash::Model, src/reactiveThermo/reactiveDeviceFlash.h:39-53,
// tables from reactiveDevicePR.h:57-99, solid block reactiveSolidThermo.h:21-26),
// its validation (validModel, reactiveDeviceFlash.h:80-108) and its
// compaction into the device tables of flash_device.h.
#ifndef REACTIVE_AMD_MODEL_IMAGE_H
#define REACTIVE_AMD_MODEL_IMAGE_H
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include "flash_device.h"
#include "hip_common.h"

namespace ramd {
namespace image {

struct NasaRegion { double tmin, tmax; int poly; double c[9]; };
struct Species { double W, a, b, kappa; int nreg; NasaRegion region[8]; };
struct PRTable { int ns; double refP; Species sp[16]; double aij[16][16]; };
struct Solid {
    int enabled, n;
    double Tmin, Tmax, Tref, pref, pmax, molarVolume, molecularWeight, hRef, sRef;
    double T[64], cp[64];
};
struct Model {
    PRTable phase[3];
    int ns, nl, condensable[2], boundaryRecovery, scalarRecovery, analyticJacobian, stableGasPrune;
    double weights[16], liquidTmin[2], liquidTc[2];
    double Tmin, Tmax, pmin, pmax, vtol, etol, mutol;
    int condensedKind[2], enforceSpeciesTemperatureBounds;
    Solid solid[2];
};
static_assert(sizeof(NasaRegion) == 96 && sizeof(Species) == 808 && sizeof(PRTable) == 14992, "PR table layout");
static_assert(sizeof(Model) == 47448, "HEM model image size (reactiveDeviceFlash.h:52)");
static_assert(offsetof(Model, stableGasPrune) == 45004 && offsetof(Model, weights) == 45008, "HEM model offsets");

// Table checks (reactiveDevicePR.h:157-195).
inline bool validTable(const PRTable& t) {
    if (t.ns <= 0 || t.ns > 16 || !fin(t.refP) || t.refP <= 0) return false;
    for (int k = 0; k < t.ns; ++k) {
        const Species& sp = t.sp[k];
        if (!fin(sp.W) || sp.W <= 0 || !fin(sp.a) || sp.a <= 0 || !fin(sp.b) || sp.b <= 0 || !fin(sp.kappa) ||
            sp.nreg <= 0 || sp.nreg > 8)
            return false;
        for (int r = 0; r < sp.nreg; ++r) {
            const NasaRegion& g = sp.region[r];
            if ((g.poly != 7 && g.poly != 9) || !fin(g.tmin) || !fin(g.tmax) || g.tmin <= 0 || g.tmax <= g.tmin ||
                (r && g.tmin < sp.region[r - 1].tmin))
                return false;
            for (int j = 0; j < g.poly; ++j)
                if (!fin(g.c[j])) return false;
        }
        for (int j = 0; j < t.ns; ++j)
            if (!fin(t.aij[k][j]) || t.aij[k][j] != t.aij[j][k] ||
                (k == j && (t.aij[k][j] <= 0 || t.aij[k][j] != sp.a)))
                return false;
    }
    return true;
}

// validModel (reactiveDeviceFlash.h:80-108) without the solid-model branch,
// which this build rejects separately.
inline bool validModel(const Model& m) {
    if (m.ns < 1 || m.ns > 16 || m.nl < 0 || m.nl > 2 || m.phase[0].ns != m.ns ||
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
        if (!validTable(m.phase[i])) return false;
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

template <int N>
inline void compactTable(const PRTable& src, Table<N>& t) {
    std::memset(static_cast<void*>(&t), 0, sizeof(t));
    t.ns = src.ns;
    t.refP = src.refP;
    for (int k = 0; k < N; ++k) {
        const Species& sp = src.sp[k];
        t.W[k] = sp.W; t.a[k] = sp.a; t.b[k] = sp.b; t.kappa[k] = sp.kappa;
        t.Tc[k] = sp.a * kOmegaB / (sp.b * kOmegaA * kR);  // reactiveDevicePR.h:223
        t.nreg[k] = sp.nreg;
        t.tlo[k] = sp.region[0].tmin;
        t.thi[k] = sp.region[sp.nreg - 1].tmax;
        for (int r = 0; r < sp.nreg; ++r) {
            t.region[k][r].tmin = sp.region[r].tmin;
            t.region[k][r].poly = sp.region[r].poly;
            for (int j = 0; j < 9; ++j) t.region[k][r].c[j] = sp.region[r].c[j];
        }
        for (int j = 0; j < N; ++j) t.aij[k][j] = src.aij[k][j];
    }
}

// Composition-fixed phaseState constants of a one-species table: the
// composition is exactly Y = {1} whenever such a table is evaluated.
template <int N>
inline void fixedHeuristic(const Table<N>& t, PhaseHeuristic& h) {
    const double Y[1] = {1.0};
    double x[1], W = 0;
    h.status = kBadComposition;
    h.tc = __builtin_huge_val();
    if (N != 1 || massToMole(t, Y, x, W) != kOk) return;
    // a and b of the mixing rule do not depend on temperature (mixing(), same sums)
    double a = 0, b = 0;
    b += x[0] * t.b[0];
    a += x[0] * x[0] * t.aij[0][0];
    phaseHeuristic(t, x, W, a, b, h);
}

template <int NS>
inline void buildDeviceModel(const Model& m, HemModel<NS>& d) {
    std::memset(static_cast<void*>(&d), 0, sizeof(d));
    compactTable<NS>(m.phase[0], d.gas);
    if (NS == 1) fixedHeuristic<NS>(d.gas, d.gasFixed);
    if (m.nl) {
        compactTable<1>(m.phase[1], d.liquid);
        fixedHeuristic<1>(d.liquid, d.liquidFixed);
    }
    d.nl = m.nl;
    d.condensable = m.nl ? m.condensable[0] : 0;
    d.scalarRecovery = m.scalarRecovery;
    d.stableGasPrune = m.stableGasPrune;
    d.analyticJacobian = m.analyticJacobian;
    d.boundaryRecovery = m.boundaryRecovery;
    d.enforceBounds = m.enforceSpeciesTemperatureBounds;
    for (int k = 0; k < NS; ++k) d.weights[k] = m.weights[k];
    d.liquidTmin = m.nl ? m.liquidTmin[0] : 0;
    d.liquidTc = m.nl ? m.liquidTc[0] : 0;
    d.Tmin = m.Tmin; d.Tmax = m.Tmax; d.pmin = m.pmin; d.pmax = m.pmax;
    d.vtol = m.vtol; d.etol = m.etol; d.mutol = m.mutol;
}

}  // namespace image
}  // namespace ramd
#endif
