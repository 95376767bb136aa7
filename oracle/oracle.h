// SPDX-License-Identifier: GPL-3.0-or-later
// CPU ORACLE — test infrastructure only.  Serial FP64 restatement of the
// reference's transport operators and HEM UV flash, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// The product library never includes or links anything under oracle/.
//
// Rounding policy: IEEE binary64, -ffp-contract=off, glibc libm, exactly the
// reference's host build (tools/build_reactive_transport.sh:8-11), so the
// restatement reproduces the compiled reference bit for bit.
#ifndef REACTIVE_ORACLE_H
#define REACTIVE_ORACLE_H
#include <cstddef>
#include <cstdint>

namespace oracle {

// ---- ABI layouts (reactiveTransport.h:10-23, reactiveThermo.h:11-19) ----
struct Face { int64_t owner, neighbour, fixed; double normal[3], area, distance, ownerWeight; int kind; };
struct TState { double p, T, rho, cv, sound, gasMass, dilatation; };
struct Config {
    size_t cells, species, variables, faces, fixed;
    double viscosity, conductivity, diffusivity, waveFactor, maxBytes;
    int mechanical;
};
struct ThermoState {
    double p, T, rho, e, entropy, soundFrozen, soundEquilibrium;
    double volumeResidual, energyResidual, chemicalResidual;
    double alphaGas, alphaLiquid[2], rhoGas, rhoLiquid[2], liquidMass[2];
    double gasMass, cp, cv;
    int activeLiquids, iterations;
};
static_assert(sizeof(Face) == 80 && sizeof(TState) == 56 && sizeof(Config) == 88, "transport ABI");
static_assert(sizeof(ThermoState) == 176, "thermo state ABI");

// ---- HEM model image (reactiveDeviceFlash.h:39-53, reactiveDevicePR.h:57-99) ----
constexpr int kMaxSpecies = 16;
constexpr int kMaxRegions = 8;
struct NasaRegion { double tmin, tmax; int poly; double c[9]; };
struct Species { double W, a, b, kappa; int nreg; NasaRegion region[kMaxRegions]; };
struct Table { int ns; double refP; Species sp[kMaxSpecies]; double aij[kMaxSpecies][kMaxSpecies]; };
struct Solid {
    int enabled, n;
    double Tmin, Tmax, Tref, pref, pmax, molarVolume, molecularWeight, hRef, sRef;
    double T[64], cp[64];
};
struct Model {
    Table phase[3];
    int ns, nl, condensable[2], boundaryRecovery, scalarRecovery, analyticJacobian, stableGasPrune;
    double weights[kMaxSpecies], liquidTmin[2], liquidTc[2];
    double Tmin, Tmax, pmin, pmax, vtol, etol, mutol;
    int condensedKind[2], enforceSpeciesTemperatureBounds;
    Solid solid[2];
};
static_assert(sizeof(Model) == 47448, "HEM model image size");
static_assert(offsetof(Model, stableGasPrune) == 45004 && offsetof(Model, weights) == 45008, "HEM model offsets");

struct Counters { uint64_t phases, residuals, candidates, stable, failures, analytic, fd, mixedAttempts, mixedAccepted; };

}  // namespace oracle
#endif
