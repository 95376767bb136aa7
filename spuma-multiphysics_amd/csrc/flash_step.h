// SPDX-License-Identifier: GPL-3.0-or-later
// Per-cell helpers shared by the resident step's flash kernels (step.hip)
// and the single-precision search kernel (search.hip).
#ifndef REACTIVE_AMD_FLASH_STEP_H
#define REACTIVE_AMD_FLASH_STEP_H
#include <hip/hip_runtime.h>
#include "hip_common.h"
#include "flash_device.h"

// Tuning knobs (compile-time): the model image in LDS, and the minimum waves
// per SIMD the register allocation targets for the FP64 flash kernels (3:
// 168 VGPRs, at 160^3 17 % faster than 2) and for the mixed flash's FP64
// refinement kernels (2: at 3 the refinement ran 8 % slower).
#ifndef RAMD_MODEL_LDS
#define RAMD_MODEL_LDS 1
#endif
#ifndef RAMD_FLASH_MIN_WAVES
#define RAMD_FLASH_MIN_WAVES 3
#endif
#ifndef RAMD_REFINE_MIN_WAVES
#define RAMD_REFINE_MIN_WAVES 2
#endif
// Persistent grid of the queue kernels (the step's and the HEM ABI's mixed
// flash): 4096 single-wave workgroups (sixteen
// per CU, more than stay resident, so a wave that finishes its cell is
// replaced at once) take up to 4096 deferred cells per round one cell per
// wave, or 262,144 one cell per lane; workgroups beyond the queue leave at
// once.  Against 1024: 160^3 pure N2O step 135.8 -> 132.5 ms (2,537 deferred
// cells per flash), binary N2O/N2 96^3 2242 -> 1788 ms, bit for bit the same.
#ifndef RAMD_QUEUE_BLOCKS
#define RAMD_QUEUE_BLOCKS 4096
#endif
// Deferred refinements up to this many per flash run one cell per wave with
// the Newton line searches spread over its lanes (the flash's WAVE mode);
// more run one cell per lane.  Measured on the step's 4096-wave queue grid:
// one cell per wave takes a binary N2O/N2 160^3 step (75k deferred cells per
// flash) from 6.51 to 4.85 s; on a 1024-wave grid it had been slower there
// (6.74 -> 7.18 s) and faster at 96^3 (16k cells, 4.55 -> 2.23 s).
#ifndef RAMD_WAVE_QUEUE_CELLS
#define RAMD_WAVE_QUEUE_CELLS 131072
#endif

namespace ramd {

// Flashes of the kernels read the workgroup's LDS model copy directly.
constexpr bool kModelLds = RAMD_MODEL_LDS != 0;

// The workgroup's copy of the model image: tables are read from LDS rather
// than global memory.  Every thread of the workgroup must call this (it
// synchronises the workgroup).
template <int NS>
__device__ inline const HemModel<NS>& workgroupModel(const HemModel<NS>* __restrict__ global) {
#if RAMD_MODEL_LDS
    static_assert(sizeof(HemModel<NS>) % 4 == 0, "model image in 4-byte words");
    HemModel<NS>& local = sharedModel<NS>();
    const uint32_t* src = reinterpret_cast<const uint32_t*>(global);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&local);
    for (uint32_t w = threadIdx.x; w < sizeof(HemModel<NS>) / 4; w += blockDim.x) dst[w] = src[w];
    __syncthreads();
    return local;
#else
    return *global;
#endif
}

// Loads cell c into a flash: species masses, e = rhoE - |rho u|^2 / (2 rho)
// (ReactiveFoam.C:727-728, in FP64) and the step-start seed.
template <class F>
__device__ inline void loadCell(F& f, const double* __restrict__ q, size_t stride, const ThermoOut& g, uint32_t c,
                                int ns) {
    using R = decltype(f.energy);
    double rho = 0, mass[4];
    for (int k = 0; k < ns; ++k) {
        mass[k] = q[size_t(k) * stride + c];
        f.q[k] = R(mass[k]);
        rho += mass[k];
    }
    const double m0 = q[size_t(ns) * stride + c], m1 = q[size_t(ns + 1) * stride + c];
    const double m2 = q[size_t(ns + 2) * stride + c], E = q[size_t(ns + 3) * stride + c];
    f.energy = R(E - .5 * (m0 * m0 + m1 * m1 + m2 * m2) / rho);
    f.guessP = R(g.p);
    f.guessT = R(g.T);
    f.guessLiquid = R(g.liquidMass[0]);
}

// Loads cell c of a HEM batch into a flash: species masses q[c * NS + k], the
// given internal energy and the seed.
template <int NS, class F>
__device__ inline void loadHemCell(F& f, const double* __restrict__ q, const double* __restrict__ energy,
                                   const ThermoOut& g, uint32_t c) {
    using R = decltype(f.energy);
    for (int k = 0; k < NS; ++k) f.q[k] = R(q[size_t(c) * NS + k]);
    f.energy = R(energy[c]);
    f.guessP = R(g.p);
    f.guessT = R(g.T);
    f.guessLiquid = R(g.liquidMass[0]);
}

// Commits a flash result, or keeps the step-start state of a failed cell
// (the reference commits nothing for it); p, c_frozen and rho go to the SoA
// arrays the next transport stage reads.
template <bool A, bool RJ>
__device__ inline void storeCell(bool ok, const EvalT<double, A, RJ>& result, const ThermoOut& g, uint32_t c,
                                 ThermoOut* __restrict__ out, double* __restrict__ pOut, double* __restrict__ cOut,
                                 double* __restrict__ rOut) {
    if (ok) {
        writeOut(result, out[c]);
        pOut[c] = result.p;
        cOut[c] = result.cf;
        rOut[c] = result.rho;
    } else {
        out[c] = g;
        pOut[c] = g.p;
        cOut[c] = g.soundFrozen;
        rOut[c] = g.rho;
    }
}

__device__ inline void countFlash(uint32_t failed, uint32_t phases, unsigned long long* __restrict__ counters) {
    const uint32_t sumFailed = waveSum(failed), sumPhases = waveSum(phases);
    if ((threadIdx.x & 63) == 0) {
        if (sumFailed) atomicAdd(&counters[0], (unsigned long long)sumFailed);
        atomicAdd(&counters[1], (unsigned long long)sumPhases);
    }
}

// Single-precision candidate search of the mixed flash (search.hip): one
// launch over the cells in `order` in workgroups of `block` lanes (64, 128 or
// 256), candidate sets and FP32 costs by cell; with rep (may be null), only
// the cells that represent themselves (rep[c] == c) are searched.
void launchSearch(int ns, int nl, const void* model, const double* q, size_t stride, const ThermoOut* seeds,
                  const uint32_t* order, uint32_t count, const double* dtDev, CandidateSet* sets, uint16_t* cost,
                  unsigned long long* phases, hipStream_t stream, unsigned block, const uint32_t* rep);
// The same search over a HEM batch (loadHemCell's layout), 64-lane workgroups.
void launchHemSearch(int ns, int nl, const void* model, const double* q, const double* energy,
                     const ThermoOut* guess, const uint32_t* order, uint32_t count, CandidateSet* sets,
                     unsigned long long* phases, hipStream_t stream);

}  // namespace ramd
#endif
