// SPDX-License-Identifier: GPL-3.0-or-later
// Thin C entry points around the reference's host-compilable device flash
// (src/reactiveThermo/reactiveDeviceFlash.h, reactiveDevicePR.h), compiled
// from the reference sources where they lie.  Test infrastructure only: this
// translation unit is built into oracle/_ref/libref.so by oracle/Makefile
// (target `ref`) and is never linked into the product library.
#include "reactiveDeviceFlash.h"
#include "reactiveBalancedCapillary.h"
#include <cstring>
#include <cstdint>

extern "C" {

size_t ref_model_bytes(void) { return sizeof(ReactiveDeviceFlash::Model); }

int ref_model_valid(const void* model) {
    ReactiveDeviceFlash::Model m;
    std::memcpy(static_cast<void*>(&m), model, sizeof(m));
    return ReactiveDeviceFlash::validModel(m) ? 1 : 0;
}

// One UV flash, exactly ReactiveDeviceFlash::recover (reactiveDeviceFlash.h:564-567).
// useCache=1 builds the pure-liquid reported-phase caches the CUDA plugin uses
// (reactiveGpuHem.cuh:31-36) and passes them to recover().
// counters: phases, residuals, candidates, stable, failures, analytic, fd, mixedAttempts, mixedAccepted
int ref_flash_recover(const void* model, const double* q, double energy, const void* guess,
                      int useCache, void* out, uint64_t* counters) {
    static thread_local ReactiveDeviceFlash::Model m;
    std::memcpy(static_cast<void*>(&m), model, sizeof(m));
    ReactiveDeviceFlash::Input in{};
    for (int k = 0; k < m.ns; ++k) in.q[k] = q[k];
    in.energy = energy;
    std::memcpy(&in.guess, guess, sizeof(ReactiveThermoState));
    ReactiveDevicePRPurePhaseCache::PureReportedPhaseCache cache[2]{};
    if (useCache)
        for (int i = 0; i < 2; ++i)
            if (i < m.nl) ReactiveDevicePRPurePhaseCache::buildPureReportedPhaseCache(m.phase[i + 1], cache[i]);
    const auto o = ReactiveDeviceFlash::recover(m, in, useCache ? cache : nullptr);
    if (o.success) std::memcpy(out, &o.state, sizeof(ReactiveThermoState));
    if (counters) {
        counters[0] = o.counters.phases; counters[1] = o.counters.residuals;
        counters[2] = o.counters.candidates; counters[3] = o.counters.stable;
        counters[4] = o.counters.failures; counters[5] = o.counters.analyticJacobians;
        counters[6] = o.counters.finiteDifferenceJacobians;
        counters[7] = o.counters.mixedLinearAttempts; counters[8] = o.counters.mixedLinearAccepted;
    }
    return o.success;
}

// One capillary-interface recovery, exactly ReactiveDeviceFlash::recoverCapillary
// (reactiveDeviceFlash.h:572-590): equilibrium=0 is the frozen condensed
// inventory solve from guess.liquidMass, equilibrium=1 the UV equilibrium
// search with the boundary recovery.  Same counters layout as ref_flash_recover.
int ref_flash_capillary(const void* model, const double* q, double energy, const void* guess, double color,
                        double pressureJump, int equilibrium, void* out, uint64_t* counters) {
    static thread_local ReactiveDeviceFlash::Model m;
    std::memcpy(static_cast<void*>(&m), model, sizeof(m));
    ReactiveDeviceFlash::Input in{};
    for (int k = 0; k < m.ns; ++k) in.q[k] = q[k];
    in.energy = energy;
    std::memcpy(&in.guess, guess, sizeof(ReactiveThermoState));
    const auto o = ReactiveDeviceFlash::recoverCapillary(m, in, color, pressureJump, equilibrium != 0, nullptr);
    if (o.success) std::memcpy(out, &o.state, sizeof(ReactiveThermoState));
    if (counters) {
        counters[0] = o.counters.phases; counters[1] = o.counters.residuals;
        counters[2] = o.counters.candidates; counters[3] = o.counters.stable;
        counters[4] = o.counters.failures; counters[5] = o.counters.analyticJacobians;
        counters[6] = o.counters.finiteDifferenceJacobians;
        counters[7] = o.counters.mixedLinearAttempts; counters[8] = o.counters.mixedLinearAccepted;
    }
    return o.success;
}

int ref_flash_capillary_batch(const void* model, size_t count, const double* q, const double* energy,
                              const void* guess, int equilibrium, void* out, int* success, uint64_t* counters) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    int ok = 0;
    for (size_t c = 0; c < count; ++c) {
        const char* g = static_cast<const char*>(guess) + c * sizeof(ReactiveThermoState);
        char* o = static_cast<char*>(out) + c * sizeof(ReactiveThermoState);
        std::memcpy(o, g, sizeof(ReactiveThermoState));
        success[c] = ref_flash_capillary(model, q + c * m.ns, energy[c], g, 0.0, 0.0, equilibrium, o,
                                         counters ? counters + 9 * c : nullptr);
        ok += success[c];
    }
    return ok;
}

// Per-cell color, pressure jump and equilibrium flag (recoverCapillary for
// each cell).  A cell whose interface inputs the reference rejects (status 2)
// reports success 0.
int ref_flash_capillary_cells(const void* model, size_t count, const double* q, const double* energy,
                              const void* guess, const double* color, const double* jump, const int* equilibrium,
                              void* out, int* success, uint64_t* counters) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    int ok = 0;
    for (size_t c = 0; c < count; ++c) {
        const char* g = static_cast<const char*>(guess) + c * sizeof(ReactiveThermoState);
        char* o = static_cast<char*>(out) + c * sizeof(ReactiveThermoState);
        std::memcpy(o, g, sizeof(ReactiveThermoState));
        success[c] = ref_flash_capillary(model, q + c * m.ns, energy[c], g, color[c], jump[c], equilibrium[c], o,
                                         counters ? counters + 9 * c : nullptr);
        ok += success[c];
    }
    return ok;
}

// Batch form for fixture generation.
int ref_flash_batch(const void* model, size_t count, const double* q, const double* energy,
                    const void* guess, int useCache, void* out, int* success, uint64_t* counters) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    int ok = 0;
    for (size_t c = 0; c < count; ++c) {
        std::memcpy(static_cast<char*>(out) + c * sizeof(ReactiveThermoState),
                    static_cast<const char*>(guess) + c * sizeof(ReactiveThermoState), sizeof(ReactiveThermoState));
        success[c] = ref_flash_recover(model, q + c * m.ns, energy[c],
                                       static_cast<const char*>(guess) + c * sizeof(ReactiveThermoState), useCache,
                                       static_cast<char*>(out) + c * sizeof(ReactiveThermoState),
                                       counters ? counters + 9 * c : nullptr);
        ok += success[c];
    }
    return ok;
}

// PR phase property at (T, p) on phase table `phase` (0 gas, 1.. liquids).
// choice: 1 gas root, 2 liquid root.  out: T p W V rho e h s cp cv alpha kappaT sound dpdT dpdV mu[16] phase
int ref_pr_tp(const void* model, int phase, double T, double p, const double* Y, int choice, double* out) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    ReactiveDevicePR::State s{};
    const auto st = ReactiveDevicePR::evaluateTPTrusted(m.phase[phase], T, p, Y,
        static_cast<ReactiveDevicePR::RootChoice>(choice), s, true);
    if (st != ReactiveDevicePR::Status::Success) return static_cast<int>(st);
    const double v[] = {s.T, s.p, s.W, s.V, s.rho, s.e, s.h, s.s, s.cp, s.cv, s.alpha, s.kappaT, s.sound, s.dpdT, s.dpdV};
    for (int i = 0; i < 15; ++i) out[i] = v[i];
    for (int k = 0; k < 16; ++k) out[15 + k] = s.mu[k];
    out[31] = s.phase;
    return 0;
}

// PR properties at (T, rho) on phase table `phase`.  Same output layout.
int ref_pr_trho(const void* model, int phase, double T, double rho, const double* Y, double* out) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    ReactiveDevicePR::State s{};
    const auto st = ReactiveDevicePR::evaluateTrhoTrusted(m.phase[phase], T, rho, Y, s, true);
    if (st != ReactiveDevicePR::Status::Success) return static_cast<int>(st);
    const double v[] = {s.T, s.p, s.W, s.V, s.rho, s.e, s.h, s.s, s.cp, s.cv, s.alpha, s.kappaT, s.sound, s.dpdT, s.dpdV};
    for (int i = 0; i < 15; ++i) out[i] = v[i];
    for (int k = 0; k < 16; ++k) out[15 + k] = s.mu[k];
    out[31] = s.phase;
    return 0;
}

// Cubic root set at (T, p, x) (reactiveDevicePR.h:332-422).  out: V0 V1 V2 count signed gas liquid
int ref_pr_roots(const void* model, int phase, double T, double p, const double* x, double* out) {
    const auto& m = *static_cast<const ReactiveDeviceFlash::Model*>(model);
    ReactiveDevicePR::RootSet r{};
    const auto st = ReactiveDevicePR::rootsTP(m.phase[phase], T, p, x, r);
    if (st != ReactiveDevicePR::Status::Success) return static_cast<int>(st);
    out[0] = r.volume[0]; out[1] = r.volume[1]; out[2] = r.volume[2];
    out[3] = r.count; out[4] = r.signedCount; out[5] = r.gasAvailable; out[6] = r.liquidAvailable;
    return 0;
}


// The reference's contact-preserving face flux
// (reactiveBalancedCapillary.h:55-142) with sigma = curvature = colorFace = 0
// and no surface stress.  left/right = {rho, u0, u1, u2, p, a, E}; out =
// {advectLeft, advectRight, pressureMomentum, pressureEnergy}.
int ref_hllc_face(const double* left, const double* right, const double* normal, double* out) {
    ReactiveBalancedCapillary::State l{}, r{};
    l.rho = left[0]; r.rho = right[0];
    for (int d = 0; d < 3; ++d) { l.velocity[d] = left[1 + d]; r.velocity[d] = right[1 + d]; }
    l.pressure = left[4]; r.pressure = right[4];
    l.sound = left[5]; r.sound = right[5];
    l.totalEnergy = left[6]; r.totalEnergy = right[6];
    ReactiveBalancedCapillary::Face f{};
    for (int d = 0; d < 3; ++d) f.normal[d] = normal[d];
    ReactiveBalancedCapillary::Flux flux{};
    if (!ReactiveBalancedCapillary::faceFlux(l, r, f, flux)) return 0;
    out[0] = flux.advectLeft; out[1] = flux.advectRight;
    out[2] = flux.pressureMomentum; out[3] = flux.pressureEnergy;
    return 1;
}

}
