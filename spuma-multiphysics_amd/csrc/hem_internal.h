// SPDX-License-Identifier: GPL-3.0-or-later
// HEM handle shared by the flash plugin and the device-resident step.
#ifndef REACTIVE_AMD_HEM_INTERNAL_H
#define REACTIVE_AMD_HEM_INTERNAL_H
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "flash_device.h"

namespace ramd {

// Flash output record, byte-compatible with ReactiveGpuHemOutputV1 (184 B).
struct HemOutputDev {
    ThermoOut state;
    int success, reserved;
};

// Per-cell capillary input, byte-compatible with ReactiveGpuHemCapillaryInputV2
// (24 B): the interface's color and pressure jump, and equilibrium 1 for the
// UV equilibrium with the boundary search, 0 for the frozen inventory.
struct HemCapillaryDev {
    double color, pressureJump;
    int equilibrium, pad;
};

struct HemHandle {
    int ns = 0, nl = 0, device = 0;
    bool analytic = false;  // the model's analytic-Jacobian mode (FP64 flash)
    bool boundary = false;  // the model's boundary search (boundaryRecovery = 1)
    int arithmetic = 0;     // equilibrium flash of host batches: REACTIVE_AMD_FLASH_FP64 or _MIXED
    size_t capacity = 0, deviceBytes = 0, modelBytes = 0;
    void* model = nullptr;  // HemModel<ns> in device memory
    // phases residuals candidates stable failures fd successes analytic, then
    // (mixed arithmetic) FP32 phases and deferred cells
    unsigned long long* totals = nullptr;
    // mixed arithmetic: candidate sets and the deferred-refinement queue
    // (allocated on the first mixed batch), the last batch's FP32 phase
    // evaluations and deferred cells
    CandidateSet* sets = nullptr;
    uint32_t* queue = nullptr;
    unsigned* queueCount = nullptr;
    uint64_t lastSinglePhases = 0, lastQueued = 0;
    // host-path staging (allocated on the first host batch)
    bool hostReady = false;
    double *dq = nullptr, *de = nullptr;
    ThermoOut* dguess = nullptr;
    HemOutputDev* dout = nullptr;
    uint32_t* dorder = nullptr;
    HemCapillaryDev* dcap = nullptr;      // host-path capillary staging
    std::vector<uint32_t> hostOrder;
    std::vector<uint8_t> hostKey;
    std::vector<HemOutputDev> hostOut;
    hipStream_t stream = nullptr;
    hipEvent_t events[4] = {};
    std::vector<void*> allocations;

    ~HemHandle();
    void* allocate(size_t bytes);
    void ensureHostPath();
    // Flash count cells named by order (or 0..count-1) on stream s; q, e,
    // guess, capillary and out are indexed by cell.  Totals: phases,
    // residuals, candidates, stable, failures, fd Jacobians, successes,
    // analytic Jacobians.
    // With partials (one slot of 8 counters per 64-cell workgroup), the
    // workgroups write their sums there instead of adding them to totals.
    void launch(const double* q, const double* e, const ThermoOut* guess, const uint32_t* order, size_t count,
                const HemCapillaryDev* capillary, HemOutputDev* out, uint32_t* cellPhases, hipStream_t s,
                unsigned long long* partials = nullptr);
    // The equilibrium flash of the same cells in mixed arithmetic: FP32
    // candidate search, FP64 refinement and the deferred queue; totals as
    // launch's plus totals[8] FP32 phases and totals[9] deferred cells.
    void launchMixed(const double* q, const double* e, const ThermoOut* guess, const uint32_t* order, size_t count,
                     HemOutputDev* out, hipStream_t s);
};

HemHandle* createHem(const void* raw, size_t bytes, size_t capacity, bool hostPath);

}  // namespace ramd
#endif
