/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * Drop-in C ABI of the conservative transport library (libreactiveTransport.so)
 * for the MI355X build.  Every type and entry point below keeps the byte layout,
 * name, argument meaning and return convention of the ReactiveFoam transport
 * library it replaces; each declaration cites the reference declaration.
 *
 * Return codes (reference src/reactiveTransport/reactiveTransport.cpp:1471-1491):
 *   0 success, 1 failure (message via reactive_transport_error), 2 concurrent
 *   entry on the same handle rejected, -1 null handle.  create* return NULL and
 *   write the reason into the caller's error buffer.
 *
 * Backends: backend=1 is the HIP device path on gfx950 and has no host
 * fallback.  backend=0 (the reference's serial host operators,
 * reactiveTransport.h:116-117) is rejected with an error: the MI355X library
 * ships no host execution of the operators.
 */
#ifndef REACTIVE_AMD_TRANSPORT_H
#define REACTIVE_AMD_TRANSPORT_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* reactiveGasThermo.h:12-21 */
typedef struct ReactiveGasThermoSpecies {
    uint64_t regionOffset, regionCount;
    double gasConstant;
    int polynomial;
} ReactiveGasThermoSpecies;
typedef struct ReactiveGasThermoRegion {
    double minimumTemperature, maximumTemperature, coefficient[9];
} ReactiveGasThermoRegion;
typedef struct ReactiveGasPartition { double liquidMass[2]; } ReactiveGasPartition;

/* reactiveTransport.h:10-14 (80 B).  kind: 0 internal, 1 slipWall,
 * 2 extrapolate (zero gradient), 3 fixedState. */
typedef struct ReactiveTransportFace {
    int64_t owner, neighbour, fixed;
    double normal[3], area, distance, ownerWeight;
    int kind;
} ReactiveTransportFace;
/* reactiveTransport.h:15-17 (56 B) */
typedef struct ReactiveTransportState {
    double p, T, rho, cv, sound, gasMass, dilatation;
} ReactiveTransportState;
/* reactiveTransport.h:18-23 (88 B) */
typedef struct ReactiveTransportConfig {
    size_t cells, species, variables, faces, fixed;
    double viscosity, conductivity, diffusivity, waveFactor, maxBytes;
    int mechanical;
} ReactiveTransportConfig;
/* reactiveTransport.h:24-29 (88 B) */
typedef struct ReactiveTransportOptionsV2 {
    uint32_t abiVersion, structBytes;
    size_t bridgeCells;
    int recomputeGas;
    char physicalModelHash[65];
} ReactiveTransportOptionsV2;
/* reactiveTransport.h:30-33 (64 B) */
typedef struct ReactiveTransportMemoryV2 {
    uint64_t liveBytes, peakBytes, bridgeBytes, conservedWorkspaceBytes, gasWorkspaceBytes;
    uint64_t rhsWorkspaceBytes, deviceToDeviceBytes, synchronizations;
} ReactiveTransportMemoryV2;
/* reactiveTransport.h:34 (24 B) */
typedef struct ReactiveTransportToken { uint64_t attemptId, stageId, contentVersion; } ReactiveTransportToken;
/* reactiveTransport.h:38-47 */
typedef struct ReactiveCapillaryOptionsV1 {
    uint32_t abiVersion, structBytes;
    double sigma, capillaryCfl, geometryEpsilon;
    int64_t condensableSpecies;
} ReactiveCapillaryOptionsV1;
typedef struct ReactiveCapillaryProfileV1 {
    uint32_t abiVersion, structBytes;
    uint64_t geometryBuilds, geometryCells, geometryKernels, faceFluxBuilds;
    uint64_t colorUploadBytes, geometryDownloadBytes, capillaryWorkspaceBytes;
} ReactiveCapillaryProfileV1;
/* reactiveTransport.h:73-77 */
typedef struct ReactiveTransportStats {
    uint64_t allocatedBytes, uploadedBytes, downloadedBytes, kernelLaunches;
    uint64_t stages, stepQueries;
} ReactiveTransportStats;
typedef struct ReactiveTransportPrimitive { double rho, u[3]; } ReactiveTransportPrimitive;
/* reactiveTransport.h:79-89 */
typedef struct ReactiveTransportProfile {
    uint64_t conservedUploads, conservedUploadBytes, conservedDownloadBytes;
    uint64_t stateUploadBytes, primitiveUploadBytes, gasUploadBytes;
    uint64_t boundaryPartitions, residentStages;
} ReactiveTransportProfile;
typedef struct ReactiveTransportDeviceProfile {
    uint64_t gasPropertyBuilds, gasPropertyCells, partitionUploadBytes, thermoTableBytes;
    uint64_t gasStatusChecks, cflFaceLaunches, transportFaceLaunches, faceWorkspaceBytes;
} ReactiveTransportDeviceProfile;
/* reactiveTransportV21.h:5-17 */
typedef struct ReactiveTransportOptionsV21 {
    uint32_t abiVersion, structBytes;
    size_t slotBytes, pinnedBudgetBytes, bridgeCellsOverride;
    uint32_t slots, blockThreads;
    int recomputeGas, detailedGasCounters;
    char physicalModelHash[65];
} ReactiveTransportOptionsV21;
typedef struct ReactiveTransportProfileV21 {
    uint64_t layoutTiles, layoutLaunches, memcpyCalls, payloadBytes;
    uint64_t nasaValidationEvaluations, nasaFaceEvaluations, temperatureBasisBuilds;
    uint64_t pinnedBytes, deviceStagingBytes, temperatureBasisBytes, eventWaits;
    uint64_t slotCells, slots, blockThreads, failedCalls;
} ReactiveTransportProfileV21;

/* ---- Core entry points, built on the device ---------------------------- */
/* reactiveTransport.h:118-121.  Physics: HLL (or the HLLC option) face flux,
 * constant viscosity (stress and its work) and conductivity (heat flux) with
 * the CFL diffusion number, and the WALE closure below.  A configuration with
 * diffusivity > 0 or the mechanical closure is rejected.  Viscous or
 * conductive transports run the generic face-graph kernels even on a
 * structured box. */
void* reactive_transport_create(int backend, const ReactiveTransportConfig* config,
    const double* volume, const ReactiveTransportFace* faces, const double* fixedQ,
    const ReactiveTransportState* fixedStates, const double* fixedGasY,
    const double* fixedGasH, char* error, size_t errorSize);
/* reactiveTransport.h:60-63 */
void* reactive_transport_create_v2(int backend, const ReactiveTransportConfig* config,
    const ReactiveTransportOptionsV2* options, const double* volume, const ReactiveTransportFace* faces,
    const double* fixedQ, const ReactiveTransportState* fixedStates, const double* fixedY, const double* fixedH,
    char* error, size_t errorSize);
/* reactiveTransportV21.h:25-28 */
void* reactive_transport_create_v21(int backend, const ReactiveTransportConfig* config,
    const ReactiveTransportOptionsV21* options, const double* volume, const ReactiveTransportFace* faces,
    const double* fixedQ, const ReactiveTransportState* fixedStates, const double* fixedY, const double* fixedH,
    char* error, size_t errorSize);
/* reactiveTransport.h:122-125 */
void reactive_transport_destroy(void* transport);
const char* reactive_transport_error(void* transport);
int reactive_transport_is_cuda(void* transport);
int reactive_transport_stats(void* transport, ReactiveTransportStats* stats);
/* reactiveTransport.h:126-127: CFL dt = min(maximum, cfl V / sum_f A_f s_f) */
int reactive_transport_stable_step(void* transport, const double* q,
    const ReactiveTransportState* state, double cfl, double maximum, double* dt);
/* reactiveTransport.h:108-110 */
int reactive_transport_stable_step_primitives(void* transport,
    const ReactiveTransportPrimitive* primitive, const ReactiveTransportState* state,
    double cfl, double maximum, double* dt);
/* reactiveTransport.h:135-137: rhs = L(q), cell-major; boundaryRate[nv] */
int reactive_transport_rhs(void* transport, const double* q,
    const ReactiveTransportState* state, const double* gasY, const double* gasH,
    double* rhs, double* boundaryRate);
/* reactiveTransport.h:131-133: SSPRK2 stage 0 / 1 on host q (in place) */
int reactive_transport_stage(void* transport, double* q,
    const ReactiveTransportState* state, const double* gasY, const double* gasH,
    double dt, int stage, double* boundaryRate);
/* reactiveTransport.h:106 */
int reactive_transport_upload_conserved(void* transport, const double* q, uint64_t version);
/* reactiveTransport.h:113-115 */
int reactive_transport_stage_resident(void* transport, const ReactiveTransportState* state,
    const double* gasY, const double* gasH, double dt, int stage,
    uint64_t inputVersion, uint64_t outputVersion, double* qOutput, double* boundaryRate);
/* reactiveTransport.h:67-72 */
int reactive_transport_begin_attempt(void* transport, const char* modelHash, uint64_t attemptId);
int reactive_transport_end_attempt(void* transport, uint64_t attemptId, int commit);
int reactive_transport_advance_resident_v2(void* transport, ReactiveTransportToken input,
    ReactiveTransportToken output, const ReactiveTransportState* state, const ReactiveGasPartition* partition,
    const double* gasY, const double* gasH, double dt, double* boundaryRate);
int reactive_transport_download_conserved(void* transport, ReactiveTransportToken token, double* output);
/* reactiveTransport.h:64, 84, 90; reactiveTransportV21.h:29 */
int reactive_transport_memory_v2(void* transport, ReactiveTransportMemoryV2* result);
int reactive_transport_profile(void* transport, ReactiveTransportProfile* profile);
int reactive_transport_device_profile(void* transport, ReactiveTransportDeviceProfile* profile);
int reactive_transport_profile_v21(void* transport, ReactiveTransportProfileV21* result);

/* ---- Exported stubs: physics that is off on this path ------------------
 * Each returns 1 and sets "<entry>: not provided by the MI355X transport
 * build" as the handle error (or -1 on a null handle).  Pointer parameters are
 * declared as opaque pointers: the reference passes them by address only. */
/* reactiveTransport.h:93-103 (device NASA gas properties, diffusive flux) */
int reactive_transport_set_gas_thermo(void* transport, const ReactiveGasThermoSpecies* species,
    size_t count, const ReactiveGasThermoRegion* regions, size_t regionCount,
    const int64_t* liquidSpecies, size_t liquids);
int reactive_transport_stage_resident_gas(void* transport, const ReactiveTransportState* state,
    const ReactiveGasPartition* partition, double dt, int stage,
    uint64_t inputVersion, uint64_t outputVersion, double* qOutput, double* boundaryRate);
int reactive_transport_rhs_gas(void* transport, const double* q, const ReactiveTransportState* state,
    const ReactiveGasPartition* partition, double* rhs, double* boundaryRate);
int reactive_transport_gas_properties_resident(void* transport, const ReactiveTransportState* state,
    const ReactiveGasPartition* partition, double* gasY, double* gasH);
/* reactiveTransport.h:48-59 (capillary) */
int reactive_transport_set_capillary_v1(void* transport, const ReactiveCapillaryOptionsV1* options);
int reactive_transport_capillary_geometry_v1(void* transport, const double* cellColor,
    const double* fixedColor, double* surfaceEnergy, double* curvature, double* normalXYZ);
int reactive_transport_capillary_profile_v1(void* transport, ReactiveCapillaryProfileV1* profile);
int reactive_transport_replace_stage_state_v1(void* transport, ReactiveTransportToken before,
    uint64_t replacementVersion, const double* conserved);
/* reactiveCapillaryResident.h:26-41 */
int reactive_transport_capillary_resident_begin_v1(void* transport, size_t species,
    const double* q, const double* liquid, const double* bulk, const void* states,
    const double* fixedColor, double sigma, int equilibrium, int* fallback);
int reactive_transport_capillary_resident_select_v1(void* transport, int iteration,
    void* view, int* fallback);
int reactive_transport_capillary_resident_update_v1(void* transport, size_t count,
    double* residual, int* fallback);
int reactive_transport_capillary_resident_relax_v1(void* transport);
int reactive_transport_capillary_resident_commit_v1(void* transport, void* states,
    double* color, double* surfaceEnergy, double* curvature);
/* reactiveCartesianTransport.h:34-49 */
int reactive_transport_install_cartesian_v1(void* transport, const void* mesh);
int reactive_transport_cartesian_info_v1(void* transport, void* info);
int reactive_transport_set_implicit_v1(void* transport, const void* options);
int reactive_transport_implicit_coefficients_v1(void* transport, double* output,
    const double* restore, size_t count, size_t* required);
/* reactiveGeometricTransport.h:42-47 */
int reactive_transport_geometric_diagnostic_v1(void* transport, const void* options,
    const double* q, const ReactiveTransportState* state, const double* cellColor,
    const void* cellGeometry, const void* faceGeometry, double* rhs, double* boundaryRate,
    void* residual);
/* reactiveTurbulence.h:12-26 (WALE LES stress closure, Delta = cbrt(cell
 * volume)).  set_wale installs the closure once, before any upload or step;
 * the eddy viscosity is rebuilt from each call's state (Gradients +
 * WaleViscosities, reactiveTransportKernels.h:290-313) and adds rho nu_t to
 * the face viscosity of the stress, its work and the CFL diffusion number.
 * wale_primitives returns nu_t (m^2/s) of the given primitives and states.
 * Generic face graphs only: the resident step and the stage adjoint reject a
 * transport with the closure installed. */
typedef struct ReactiveWaleOptionsV1 {
    uint32_t abiVersion, structBytes;
    double Cw;
} ReactiveWaleOptionsV1;
typedef struct ReactiveWaleProfileV1 {
    uint32_t abiVersion, structBytes;
    uint64_t gradientBuilds, viscosityBuilds, cellsEvaluated, workspaceBytes;
} ReactiveWaleProfileV1;
int reactive_transport_set_wale_v1(void* transport, const ReactiveWaleOptionsV1* options);
int reactive_transport_wale_profile_v1(void* transport, ReactiveWaleProfileV1* profile);
int reactive_transport_wale_primitives_v1(void* transport, const ReactiveTransportPrimitive* primitive,
    const ReactiveTransportState* state, double* nut);
/* reactiveTurbulence.h:27-81: the scalar closure (turbulent Prandtl and
 * Schmidt fluxes) and the PR property epochs are not provided (stubs that
 * return 1 with an error). */
int reactive_transport_set_wale_scalars_v1(void* transport, const void* options);
int reactive_transport_wale_scalar_profile_v1(void* transport, void* profile);
int reactive_transport_wale_scalar_fields_v1(void* transport, const double* cellCp,
    const double* cellSpeciesH, const double* fixedCp, const double* fixedSpeciesH);
int reactive_transport_set_wale_pr_model_v1(void* transport, const void* options);
int reactive_transport_wale_pr_properties_v1(void* transport, const void* epoch,
    const double* q, size_t qStride, const void* state);
int reactive_transport_wale_pr_properties_v2(void* transport, const void* epoch,
    const double* q, size_t qStride, const void* state, int statesUnchanged);
int reactive_transport_wale_pr_profile_v1(void* transport, void* profile);
/* reactiveClosureAcceleration.h:39-42 (scalar closure candidate).  create
 * returns NULL with an error message; run returns 1. */
void* reactive_closure_scalar_create_v1(size_t capacity, char* error, size_t errorSize);
void reactive_closure_scalar_destroy_v1(void* handle);
int reactive_closure_scalar_run_v1(void* handle, const void* input, size_t count,
    void* output, void* profile, char* error, size_t errorSize);

#ifdef __cplusplus
}
#endif
#endif
