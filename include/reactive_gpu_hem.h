/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * Drop-in C ABI of the batched homogeneous-equilibrium (HEM) UV flash plugin,
 * exported by the same libreactiveTransport.so as the transport entry points
 * (the reference builds both into one library: reactiveTransport.cpp:448-451).
 *
 * Semantics kept from the reference (src/reactiveThermo/reactiveGpuHem.h):
 *  - the model is an opaque POD image of exactly REACTIVE_GPU_HEM_MODEL_BYTES
 *    bytes (ReactiveDeviceFlash::Model, reactiveDeviceFlash.h:39-53);
 *  - q holds ns contiguous species masses per cell, energy the internal
 *    energy density, state[] the seed on input and the result on success;
 *  - a failed cell leaves its state untouched and success[c]=0; rc stays 0
 *    and profile->deviceFailures counts it; API or device errors give rc 1
 *    with the message in the error buffer.  There is never a host fallback.
 */
#ifndef REACTIVE_AMD_GPU_HEM_H
#define REACTIVE_AMD_GPU_HEM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define REACTIVE_GPU_HEM_MODEL_BYTES 47448u
/* reactiveGpuHem.h:10 */
#define REACTIVE_GPU_HEM_PHASE_CACHE_BYTES_V1 80u
/* reactiveGpuHem.h:31-37: the recognised numerical-policy strings. */
#define REACTIVE_GPU_HEM_FP64_FMA_POLICY_V1 "fp64-fma-rn-guard-v1:prec-div=true:prec-sqrt=true:fast-math=false"
#define REACTIVE_GPU_HEM_INT_LINEAR_POLICY_V1 "int32-q26-newton-step-fp64-accept-v2:experimental"
#define REACTIVE_GPU_HEM_FP32_NASA_POLICY_V1 "fp32-nasa-fp64-eos-v1:experimental"
#define REACTIVE_GPU_HEM_DS_NASA_POLICY_V1 "two-fp32-nasa-fp64-eos-v1:experimental"
#define REACTIVE_GPU_HEM_FP32_SEED_POLICY_V1 "fp32-cubic-seeds-fp64-refine-v2:experimental"
#define REACTIVE_GPU_HEM_FP32_LINEAR_POLICY_V1 "fp32-newton-step-fp64-accept-v2:experimental"
#define REACTIVE_GPU_HEM_FP32_COMBINED_POLICY_V1 "fp32-cubic-newton-fp64-accept-v2:experimental"

/* reactiveThermo.h:11-19 (176 B) */
typedef struct ReactiveThermoState {
    double p, T, rho, e, entropy;
    double soundFrozen, soundEquilibrium;
    double volumeResidual, energyResidual, chemicalResidual;
    double alphaGas, alphaLiquid[2];
    double rhoGas, rhoLiquid[2], liquidMass[2];
    double gasMass, cp, cv;
    int activeLiquids, iterations;
} ReactiveThermoState;

/* reactiveGpuHem.h:14-21 (144 B) */
typedef struct ReactiveGpuHemProfileV1 {
    uint32_t abiVersion, structBytes;
    uint64_t batches, submitted, succeeded, cpuFallbacks, deviceFailures;
    uint64_t phaseEvaluations, residualEvaluations, flashCandidates, stableCandidates;
    uint64_t analyticJacobians, finiteDifferenceJacobians;
    uint64_t transferBytes, hostBytes, deviceBytes;
    double wallSeconds, kernelSeconds, copySeconds;
} ReactiveGpuHemProfileV1;
/* reactiveGpuHem.h:59-62 (24 B) */
typedef struct ReactiveGpuHemCapillaryInputV2 {
    double color, pressureJump;
    int equilibrium;
} ReactiveGpuHemCapillaryInputV2;
/* reactiveGpuHem.h:78-80 (184 B) */
typedef struct ReactiveGpuHemOutputV1 {
    ReactiveThermoState state;
    int success, reserved;
} ReactiveGpuHemOutputV1;

/* reactiveGpuHem.h:40.  Returns "" for plain FP64 (separate rounding, no
 * contraction), which the reference backend accepts as legacy, or, when the
 * process was started with REACTIVE_AMD_HEM_ARITHMETIC=mixed,
 * REACTIVE_GPU_HEM_FP32_COMBINED_POLICY_V1 ("fp32-cubic-newton-fp64-accept-v2:
 * experimental", reactiveGpuHem.h:37): handles created in such a process run
 * the equilibrium flash of reactive_gpu_hem_run_v1 in the mixed arithmetic of
 * REACTIVE_AMD_FLASH_MIXED (include/reactive_amd.h).  The variable is read
 * once per process.  Capillary batches (run_v2, run_resident_v1) and
 * boundaryRecovery models run in FP64 either way. */
const char* reactive_gpu_hem_numerical_policy_v1(void);
/* reactiveGpuHem.h:41.  Model images with one pure liquid at most, no
 * solids; boundaryRecovery = 1 selects the boundary search in every
 * equilibrium flash (reactiveDeviceFlash.h:544-560). */
void* reactive_gpu_hem_create_v1(const void* model, size_t modelBytes, size_t capacity, char* error, size_t errorSize);
/* reactiveGpuHem.h:44 */
void* reactive_gpu_hem_create_resident_v1(const void* model, size_t modelBytes, size_t capacity, char* error, size_t errorSize);
/* reactiveGpuHem.h:46-47 */
size_t reactive_gpu_hem_device_bytes_v1(void* hem);
void reactive_gpu_hem_destroy_v1(void* hem);
/* reactiveGpuHem.h:51-52: host arrays, count <= capacity */
int reactive_gpu_hem_run_v1(void* hem, const double* q, const double* energy, size_t count,
    ReactiveThermoState* state, int* success, ReactiveGpuHemProfileV1* profile, char* error, size_t errorSize);
/* reactiveGpuHem.h:63-65: capillary input per cell, recovered as
 * recoverCapillary (reactiveDeviceFlash.h:572-590): the interface's color and
 * pressure jump held during the solve; equilibrium 0 the frozen condensed
 * inventory at the seed's liquid masses, 1 the UV equilibrium with the
 * boundary search in vapour coordinates.  A batch with a color outside
 * [0, 1], a non-finite jump or an equilibrium flag other than 0 / 1 is
 * rejected with rc 1 before any flash runs (reactiveGpuHem.cuh:237-243); a
 * cell with a jump but no liquid fails on its own (success 0). */
int reactive_gpu_hem_run_v2(void* hem, const double* q, const double* energy,
    const ReactiveGpuHemCapillaryInputV2* capillary, size_t count,
    ReactiveThermoState* state, int* success, ReactiveGpuHemProfileV1* profile, char* error, size_t errorSize);
/* reactiveGpuHem.h:70-72: capillary primitive initialisation; stub (rc 1). */
int reactive_gpu_hem_initialize_tp_v1(void* hem, double* q, double* energy,
    const ReactiveGpuHemCapillaryInputV2* capillary, size_t count,
    ReactiveThermoState* state, int* success, ReactiveGpuHemProfileV1* profile, char* error, size_t errorSize);
/* reactiveGpuHem.h:81-84: every pointer is DEVICE memory; order[0..count)
 * names the cells, output[cell] receives each result.  Capillary input as
 * for run_v2; as in the reference, the inputs are not checked as a batch: a
 * cell whose color or jump recoverCapillary rejects fails on its own. */
int reactive_gpu_hem_run_resident_v1(void* hem, const double* q, const double* energy,
    const ReactiveThermoState* seed, const ReactiveGpuHemCapillaryInputV2* capillary,
    const uint32_t* order, size_t count, ReactiveGpuHemOutputV1* output,
    ReactiveGpuHemProfileV1* profile, char* error, size_t errorSize);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Same arguments,
 * semantics and results as reactive_gpu_hem_run_resident_v1; no device memory
 * is allocated or freed.  The work counters go to a static device scratch of
 * 1 MiB + 64 B per device (held until the process ends), which limits a batch
 * to 1,048,576 cells (rc 1 with a message beyond that).  Calls that overlap
 * on one device share that scratch and race: serialise them. */
int reactive_gpu_hem_run_resident_noalloc_v1(void* hem, const double* q, const double* energy,
    const ReactiveThermoState* seed, const ReactiveGpuHemCapillaryInputV2* capillary,
    const uint32_t* order, size_t count, ReactiveGpuHemOutputV1* output,
    ReactiveGpuHemProfileV1* profile, char* error, size_t errorSize);

#ifdef __cplusplus
}
#endif
#endif
