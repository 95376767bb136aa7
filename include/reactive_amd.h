/* SPDX-License-Identifier: GPL-3.0-or-later
 *
 * Additive entry points of the MI355X build (no counterpart in the reference
 * ABI).  They keep one full SSPRK2 step resident on the device: the
 * reference driver's Flow::step (src/reactiveFoam/ReactiveFoam.C:961-1019,
 * without chemistry) runs there as a host loop with a conserved download
 * after every stage and a host-packed flash batch per stage
 * (ReactiveFoam.C:651-661, 689-724).  Here the whole step is one stream of
 * kernels over device-resident q and states: CFL, stage 0, flash, CFL
 * recheck, stage 1, flash, conservation totals; dt never leaves the device.
 *
 * The step uses an existing transport handle (structured box, see
 * reactive_transport_create) and HEM handle (reactive_gpu_hem_create_v1) on
 * the same device; both must describe the same species.  Return codes
 * follow the transport ABI: 0 ok, 1 failure (reactive_amd_step_error), -1
 * null handle.
 */
#ifndef REACTIVE_AMD_H
#define REACTIVE_AMD_H
#include <stddef.h>
#include <stdint.h>
#include "reactive_gpu_hem.h"
#include "reactive_transport.h"
#ifdef __cplusplus
extern "C" {
#endif

#define REACTIVE_AMD_MAX_VARIABLES 32

typedef struct ReactiveAmdStepReport {
    uint32_t abiVersion, structBytes;   /* 1, sizeof */
    uint64_t steps;                     /* steps run by the call */
    uint64_t flashFailures[2];          /* cells whose flash #1 / #2 failed, summed over the call */
    uint64_t cflViolations;             /* steps whose post-stage-0 CFL recheck failed */
    uint64_t phaseEvaluations;          /* PR phase evaluations, summed over the call */
    double dt, recheckDt;               /* last step: dt and the post-stage-0 stable step */
    double totals[REACTIVE_AMD_MAX_VARIABLES];           /* sum_c q_ck V_c after the last step */
    double boundaryIntegral[REACTIVE_AMD_MAX_VARIABLES]; /* .5 dt (bA + bB) of the last step */
    double seconds;                     /* device time of the call (HIP events) */
    /* per-launch device times (HIP events on the step stream), summed over the call */
    double stageSeconds[2];             /* fused transport stage 0 / stage 1 kernels */
    double flashSeconds[2];             /* flash #1 / flash #2 kernels */
    double cflSeconds;                  /* both CFL sweeps (+ their final minimum) */
    uint64_t fluxFailures;              /* face evaluations the HLLC flux rejected (0 under HLL) */
    uint64_t singlePhaseEvaluations;    /* FP32 phase evaluations of the mixed flash (phaseEvaluations: FP64) */
    double searchSeconds;               /* mixed flash: the FP32 search kernels' share of flashSeconds */
    uint64_t queuedCells;               /* mixed flash: cells whose FP64 refinement needed a full FP64 search
                                           (or the FP64 restart search), run from a compacted queue */
    uint32_t tileDepth;                 /* launch shapes in use: cells per column of a stage workgroup, */
    uint32_t flashBlock;                /* lanes per flash workgroup, */
    uint32_t stateBins;                 /* state-order bins per axis (see REACTIVE_AMD_STEP_AUTOTUNE) */
    uint32_t reserved;                  /* state-order cost levels per bin */
    uint64_t retries;                   /* REACTIVE_AMD_STEP_RETRY: attempts rolled back and repeated */
    uint64_t reusedCells;               /* REACTIVE_AMD_STEP_REUSE: cells given another cell's flash result
                                           (the flash counters count the flashed cells only) */
} ReactiveAmdStepReport;

/* Flash arithmetic of a step.  FP64: the reference's arithmetic.  MIXED: the
 * candidate search runs in FP32 (Newton iterations, cubic solves, line
 * searches; the cubic's near-repeated-root classification in FP64) and
 * records each converged candidate; FP64 Newton corrections then finish them
 * (to 1e-3 of the tolerances), and stability, maximum entropy and the
 * equilibrium sound speed are decided in FP64.  Cells whose FP32 decisions
 * FP64 contradicts are searched again in FP64. */
#define REACTIVE_AMD_FLASH_FP64 0
#define REACTIVE_AMD_FLASH_MIXED 1

/* Flash arithmetic of a HEM handle's host batches (reactive_gpu_hem_run_v1,
 * equilibrium flash), REACTIVE_AMD_FLASH_FP64 or _MIXED; a new handle takes
 * the process default (REACTIVE_AMD_HEM_ARITHMETIC, see
 * reactive_gpu_hem_numerical_policy_v1).  No reference counterpart: the
 * reference selects its FP32 policies by building another library.  Returns
 * 0, 1 for an unknown mode, -1 for a null handle. */
int reactive_amd_hem_set_arithmetic(void* hem, int mode);
/* The last host batch's FP32 phase evaluations and deferred cells (zero after
 * an FP64 batch).  The profile's phaseEvaluations counts both precisions. */
int reactive_amd_hem_mixed_report(void* hem, uint64_t* singlePhaseEvaluations, uint64_t* queuedCells);
/* OR-ed into flashMode: replay each step from a captured HIP graph (after a
 * first step on the stream that records the flash scheduling costs).  The
 * report then holds totals and the call's device time only, no per-kernel
 * times. */
#define REACTIVE_AMD_STEP_GRAPH 0x100
/* OR-ed into flashMode: the first run with this flag (per flash arithmetic)
 * times the candidate launch shapes on the current state, before its steps
 * when a previous flash's costs are known and otherwise after its first step,
 * and keeps the fastest for the handle's later steps: stage chunk depth 2 / 4
 * / 8 cells, flash workgroups of 64 / 128 / 256 lanes, state-order bins of
 * 32, 64 or 128 per axis, or 64 per axis split into 4 levels of the previous
 * flash's cost (the choices are reported in ReactiveAmdStepReport, the cost
 * levels in its `reserved` field; nothing the trials compute is kept). */
#define REACTIVE_AMD_STEP_AUTOTUNE 0x200
/* OR-ed into flashMode: the driver's attempt loop around each step
 * (ReactiveFoam.C:1443-1494).  The step-start q, p, c_frozen and states are
 * kept; an attempt in which any cell's flash fails or the post-stage-0 CFL
 * recheck fails is rolled back and repeated with half its dt, at most 12
 * times and while dt / 2 > 1e-15 (then the call fails with the step-start
 * state restored).  In a batch each case retries with its own halved dt (the
 * cases that passed repeat their step unchanged, bit for bit).  Steps run on
 * the stream (not with REACTIVE_AMD_STEP_GRAPH); the boundary-corrected
 * conservation check of the driver stays with the caller, from the report's
 * totals and boundary integral. */
#define REACTIVE_AMD_STEP_RETRY 0x400
/* OR-ed into flashMode: exact reuse within each flash (the reference pool's
 * exactBatchReuse, reactiveClosurePool.inc:3-54).  Cells whose flash inputs
 * are bitwise equal (species masses, internal energy, seed state) take the
 * result of the lowest such cell, which alone is flashed; results are those
 * of flashing every cell, bit for bit.  Uniform regions (an ambient at rest,
 * a reservoir's plume far from the front) collapse to one flash each. */
#define REACTIVE_AMD_STEP_REUSE 0x800

/* Face flux of a transport handle (applies to reactive_transport_rhs/stage,
 * the resident stages and the resident step): REACTIVE_AMD_FLUX_HLL (the
 * reference's default, reactiveTransportKernels.h:343-356) or
 * REACTIVE_AMD_FLUX_HLLC, the reference's contact-preserving flux
 * (src/reactiveInterface/reactiveBalancedCapillary.h:55-142) with zero surface
 * tension, which only its capillary model reaches
 * (reactiveTransportKernels.h:358-417).  A face the HLLC flux rejects fails
 * the stage ("HLLC face flux failed"); the resident step counts it in
 * ReactiveAmdStepReport.fluxFailures.  Returns 0, 1 (unknown scheme), 2
 * (concurrent entry) or -1 (null handle). */
#define REACTIVE_AMD_FLUX_HLL 0
#define REACTIVE_AMD_FLUX_HLLC 1
int reactive_amd_transport_set_flux(void* transport, int scheme);

void* reactive_amd_step_create(void* transport, void* hem, char* error, size_t errorSize);
/* Host cell-major q (cells x variables) and the step-start thermo states. */
int reactive_amd_step_load(void* step, const double* q, const ReactiveThermoState* states);
/* Runs `steps` SSPRK2 steps with CFL number cfl (0 < cfl <= 0.5, the
 * reactive_transport_stable_step contract) and dt <= maximum.  A flash
 * failure keeps that cell's step-start state and is counted. */
int reactive_amd_step_run(void* step, double cfl, double maximum, int steps, int flashMode,
                          ReactiveAmdStepReport* report);
/* Copies the current q (cell-major) and states to host memory. */
int reactive_amd_step_read(void* step, double* q, ReactiveThermoState* states);

/* Batch of independent cases: `cases` structured-box transports (each created
 * by reactive_transport_create* on its own mesh, with its own cell volume,
 * face area and wave factor, the same species and flux scheme, and no
 * fixed-state slots) stepped together with one HEM handle.  The returned step
 * takes the cases' cells concatenated in case order in reactive_amd_step_load
 * / reactive_amd_step_read; reactive_amd_step_run advances every case by its
 * own CFL dt: one flux launch per SSPRK2 stage and one CFL sweep for all
 * cases, one flash over every cell (the boundary-flux integrals are launched
 * per case), and each case's results equal those of the case stepped alone,
 * bit for bit.  The run's
 * report then holds the smallest dt and recheck step and the totals and
 * boundary integrals summed over the cases; the run fails if any case's dt
 * is invalid.  Serves the reference's uneven multi-case workloads (no
 * counterpart in the reference ABI, whose driver steps one case per
 * handle). */
void* reactive_amd_step_create_batch(void* const* transports, int cases, void* hem, char* error, size_t errorSize);
/* Case `box` of the last run: dt, recheck step, CFL-recheck violations,
 * totals and boundary integral (other fields zero). */
int reactive_amd_step_report_box(void* step, int box, ReactiveAmdStepReport* report);
const char* reactive_amd_step_error(void* step);
void reactive_amd_step_destroy(void* step);

/* ---- Adjoint of a step's pieces (no counterpart in the reference ABI; for
 * gradients of one step's output with respect to the PR parameters and the
 * reservoir states, behind reactive_amd/adjoint.py's torch.autograd.Function).
 *
 * Flash backward by the implicit-function theorem at the converged states of
 * reactive_gpu_hem_run_v1 (one gas species): given y_bar = (p_bar, c_bar) per
 * cell (adjoints of p and c_frozen), returns q_bar (count x 5, cell-major, the
 * adjoint of the conserved inputs rho, rho u, rho E) and theta_bar[2] = the
 * adjoint of (ln a-scale, ln b-scale), summed over the cells.  `scaled` holds
 * four HEM handles created from the same model with the PR a (then b) of every
 * species multiplied by 1 + relStep, 1 - relStep (a then b).  Cells with
 * success[c] == 0 (may be null) contribute nothing; cells whose local system is
 * singular are skipped and counted in *skipped.  Host arrays.  Returns 0 or 1
 * (message in `error`). */
int reactive_amd_flash_adjoint(void* hem, void* const* scaled, double relStep, const double* q, int nv,
                               const ReactiveThermoState* states, const int* success, size_t count,
                               const double* pBar, const double* cBar, double* qBar, double* thetaBar,
                               uint64_t* skipped, char* error, size_t errorSize);
/* Transposed HLL stage rates of a structured-box transport (one species):
 * given rate_bar (cells x 5, cell-major: the adjoint of the rates L the stage
 * adds dt times, reactive_transport_rhs), the cells' transport states and the
 * handle's own reservoir states, returns the adjoints of each cell's q
 * (cells x 5), p and c_frozen (cells), and of each reservoir's (q[5], p,
 * c_frozen) in fixedBar (fixed x 7).  Host arrays.  Returns 0, 1
 * (reactive_transport_error), 2 (concurrent entry) or -1 (null handle). */
int reactive_amd_transport_stage_adjoint(void* transport, const double* q, const ReactiveTransportState* states,
                                         const double* rateBar, double* qBar, double* pBar, double* cBar,
                                         double* fixedBar);

#ifdef __cplusplus
}
#endif
#endif
