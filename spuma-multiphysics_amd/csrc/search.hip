// SPDX-License-Identifier: GPL-3.0-or-later
// Single-precision candidate search of the mixed flash.  This file alone is
// compiled with -freciprocal-math -fapprox-func (build.sh): the FP32 search
// only chooses and seeds candidates, every accepted result is recomputed by
// the FP64 refinement, so it may use the hardware reciprocal and
// transcendental approximations.  Non-finite values keep their IEEE meaning
// (no finite-math assumption), so the search's NaN and range checks hold.
#include <hip/hip_runtime.h>
#include "flash_step.h"

// Minimum waves per SIMD the register allocation of the search kernel targets
// (5: 96 VGPRs; at 160^3 3-4 % faster than 4, 6 and 8 slower: more stack).
#ifndef RAMD_SEARCH_MIN_WAVES
#define RAMD_SEARCH_MIN_WAVES 5
#endif

namespace ramd {

// Mixed precision, first half: the candidate search in FP32, recording the
// converged candidates of every cell.
template <int NS, int NL>
__global__ __launch_bounds__(256, RAMD_SEARCH_MIN_WAVES) void stepSearchKernel(const HemModel<NS>* __restrict__ globalModel,
                                                       const double* __restrict__ q, size_t stride,
                                                       const ThermoOut* __restrict__ seeds,
                                                       const uint32_t* __restrict__ order, uint32_t count,
                                                       const double* __restrict__ dtDev,
                                                       CandidateSet* __restrict__ sets, uint16_t* __restrict__ cost,
                                                       unsigned long long* __restrict__ phases,
                                                       const uint32_t* __restrict__ rep) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0};
    if (i < count && *dtDev > 0 && !(rep && rep[order[i]] != order[i])) {
        const uint32_t c = order[i];
        Flash<NS, NL, float, true, false, false, kModelLds> f(model, n);
        loadCell(f, q, stride, seeds[c], c, NS);
        CandidateSet set;
        f.searchRecord(set);
        sets[c] = set;
        if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
    }
    const uint32_t sum = waveSum(n.phases);
    if ((threadIdx.x & 63) == 0) atomicAdd(phases, (unsigned long long)sum);
}

// The same search for a HEM batch (reactive_gpu_hem_run_v1 with the mixed
// arithmetic): species masses q[c * NS + k], internal energy energy[c] and the
// seed guess[c] by cell, cells in `order`.
template <int NS, int NL>
__global__ __launch_bounds__(256, RAMD_SEARCH_MIN_WAVES) void hemSearchKernel(const HemModel<NS>* __restrict__ globalModel,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ energy,
                                                      const ThermoOut* __restrict__ guess,
                                                      const uint32_t* __restrict__ order, uint32_t count,
                                                      CandidateSet* __restrict__ sets,
                                                      unsigned long long* __restrict__ phases) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0, 0};
    if (i < count) {
        const uint32_t c = order[i];
        Flash<NS, NL, float, true, false, false, kModelLds> f(model, n);
        loadHemCell<NS>(f, q, energy, guess[c], c);
        CandidateSet set;
        f.searchRecord(set);
        sets[c] = set;
    }
    const uint32_t sum = waveSum(n.phases);
    if ((threadIdx.x & 63) == 0) atomicAdd(phases, (unsigned long long)sum);
}

void launchHemSearch(int ns, int nl, const void* model, const double* q, const double* energy,
                     const ThermoOut* guess, const uint32_t* order, uint32_t count, CandidateSet* sets,
                     unsigned long long* phases, hipStream_t stream) {
    const unsigned blocks = unsigned((size_t(count) + 63) / 64);
#define RAMD_HEM_SEARCH(NS_, NL_)                                                                                 \
    hipLaunchKernelGGL((hemSearchKernel<NS_, NL_>), dim3(blocks), dim3(64), 0, stream,                             \
                       static_cast<const HemModel<NS_>*>(model), q, energy, guess, order, count, sets, phases)
    if (ns == 1 && nl == 0) { RAMD_HEM_SEARCH(1, 0); }
    else if (ns == 1) { RAMD_HEM_SEARCH(1, 1); }
    else if (nl == 0) { RAMD_HEM_SEARCH(2, 0); }
    else { RAMD_HEM_SEARCH(2, 1); }
#undef RAMD_HEM_SEARCH
    RAMD_HIP(hipGetLastError());
}

void launchSearch(int ns, int nl, const void* model, const double* q, size_t stride, const ThermoOut* seeds,
                  const uint32_t* order, uint32_t count, const double* dtDev, CandidateSet* sets, uint16_t* cost,
                  unsigned long long* phases, hipStream_t stream, unsigned block, const uint32_t* rep) {
    const unsigned blocks = unsigned((size_t(count) + block - 1) / block);
#define RAMD_SEARCH(NS_, NL_)                                                                                     \
    hipLaunchKernelGGL((stepSearchKernel<NS_, NL_>), dim3(blocks), dim3(block), 0, stream,                         \
                       static_cast<const HemModel<NS_>*>(model), q, stride, seeds, order, count, dtDev, sets, cost, \
                       phases, rep)
    if (ns == 1 && nl == 0) { RAMD_SEARCH(1, 0); }
    else if (ns == 1) { RAMD_SEARCH(1, 1); }
    else if (nl == 0) { RAMD_SEARCH(2, 0); }
    else { RAMD_SEARCH(2, 1); }
#undef RAMD_SEARCH
    RAMD_HIP(hipGetLastError());
}

}  // namespace ramd
