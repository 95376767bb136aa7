// SPDX-License-Identifier: GPL-3.0-or-later
// Batched HEM UV flash plugin for gfx950: the reactive_gpu_hem_* ABI
// (src/reactiveThermo/reactiveGpuHem.h:40-84) over the device flash of
// flash_device.h.
//
// Execution: one work-item per cell in 64-lane workgroups (one wavefront),
// cells visited in heavy-first bucket order (the reference's scheduling key,
// src/reactiveTransport/reactiveHemBucket.h:13-18): liquid-active cells run
// the long seeded searches, so they go first and cheap gas cells fill the
// tail.  Work counters are reduced per wavefront and accumulated with one
// 64-bit atomic per counter per wavefront.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "hip_common.h"
#include "flash_device.h"
#include "model_image.h"
#include "hem_internal.h"
#include "flash_step.h"
#include "reactive_gpu_hem.h"
#include "reactive_amd.h"

namespace ramd {

static_assert(sizeof(ThermoOut) == sizeof(ReactiveThermoState), "state layout");
static_assert(sizeof(HemCapillaryDev) == sizeof(ReactiveGpuHemCapillaryInputV2) &&
                  offsetof(HemCapillaryDev, equilibrium) == offsetof(ReactiveGpuHemCapillaryInputV2, equilibrium),
              "capillary input layout");
static_assert(sizeof(HemOutputDev) == sizeof(ReactiveGpuHemOutputV1) &&
                  offsetof(HemOutputDev, success) == offsetof(ReactiveGpuHemOutputV1, success),
              "output layout");

// Adds one wavefront's work counters (phases, residuals, candidates, stable,
// failures, fd Jacobians, successes, analytic Jacobians) to totals: one
// 64-bit atomic per counter per wavefront.
__device__ inline void addTotals(const Counts& n, uint32_t succeeded, unsigned long long* __restrict__ totals) {
    const uint32_t s0 = waveSum(n.phases), s1 = waveSum(n.residuals), s2 = waveSum(n.candidates);
    const uint32_t s3 = waveSum(n.stable), s4 = waveSum(n.failures), s5 = waveSum(n.fd);
    const uint32_t s6 = waveSum(succeeded), s7 = waveSum(n.analytic);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&totals[0], (unsigned long long)s0);
        atomicAdd(&totals[1], (unsigned long long)s1);
        atomicAdd(&totals[2], (unsigned long long)s2);
        atomicAdd(&totals[3], (unsigned long long)s3);
        atomicAdd(&totals[4], (unsigned long long)s4);
        atomicAdd(&totals[5], (unsigned long long)s5);
        atomicAdd(&totals[6], (unsigned long long)s6);
        atomicAdd(&totals[7], (unsigned long long)s7);
    }
}

// A flash result into the batch's output record (zero state where it failed).
template <class E>
__device__ inline void storeHemCell(bool ok, const E& result, HemOutputDev& o) {
    if (ok) writeOut(result, o.state);
    else std::memset(&o.state, 0, sizeof(o.state));
    o.success = ok ? 1 : 0;
    o.reserved = 0;
}

// Without capillary input: equilibrium UV flash (recover,
// reactiveDeviceFlash.h:564-567).  With it (recoverCapillary, :572-590): the
// cell's interface (color, pressure jump) held during the solve, then the
// UV equilibrium with the boundary search (equilibrium = 1) or the frozen
// inventory at the seed's liquid masses (0); a cell whose color or jump the
// reference rejects (its status 2) fails.  RJ: the model's analytic-Jacobian
// mode (analyticJacobian = 1); CAP: the capillary / boundary-search flash,
// for capillary input and boundaryRecovery models.
template <int NS, int NL, bool RJ, bool CAP>
__global__ __launch_bounds__(64, RAMD_FLASH_MIN_WAVES) void hemFlashKernel(const HemModel<NS>* __restrict__ globalModel,
                                                     const double* __restrict__ q, const double* __restrict__ energy,
                                                     const ThermoOut* __restrict__ guess,
                                                     const uint32_t* __restrict__ order, uint32_t count,
                                                     const HemCapillaryDev* __restrict__ capillary,
                                                     HemOutputDev* __restrict__ out,
                                                     unsigned long long* __restrict__ totals,
                                                     uint32_t* __restrict__ cellPhases,
                                                     unsigned long long* __restrict__ partials) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0, 0};
    uint32_t succeeded = 0;
    if (i < count) {
        const uint32_t c = order ? order[i] : i;
        using F = Flash<NS, NL, double, false, RJ, CAP, kModelLds>;
        F f(model, n);
        for (int k = 0; k < NS; ++k) f.q[k] = q[size_t(c) * NS + k];
        f.energy = energy[c];
        const ThermoOut& g = guess[c];
        f.guessP = g.p;
        f.guessT = g.T;
        f.guessLiquid = g.liquidMass[0];
        typename F::E result;
        zeroEval(result);
        bool ok;
        if (capillary) {
            const HemCapillaryDev& x = capillary[c];
            if constexpr (CAP) {
                if (!fin(x.color) || x.color < 0 || x.color > 1 || !fin(x.pressureJump) ||
                    (x.pressureJump != 0 && NL != 1)) {
                    ok = false;
                } else {
                    f.color = x.color;
                    f.jump = x.pressureJump;
                    ok = x.equilibrium != 0 ? f.run(result, true)
                                            : f.frozen(NL ? g.liquidMass[0] : 0.0, result, 0, g.liquidMass[1]);
                }
            } else {
                ok = f.frozen(NL ? g.liquidMass[0] : 0.0, result, 0, g.liquidMass[1]);
            }
        } else {
            ok = f.run(result);
        }
        storeHemCell(ok, result, out[c]);
        succeeded = ok ? 1u : 0u;
        if (cellPhases) cellPhases[c] = n.phases;
    }
    if (partials) {  // one 64-lane workgroup: its eight sums go to its own slot
        const uint32_t s0 = waveSum(n.phases), s1 = waveSum(n.residuals), s2 = waveSum(n.candidates);
        const uint32_t s3 = waveSum(n.stable), s4 = waveSum(n.failures), s5 = waveSum(n.fd);
        const uint32_t s6 = waveSum(succeeded), s7 = waveSum(n.analytic);
        if (threadIdx.x == 0) {
            unsigned long long* slot = partials + size_t(blockIdx.x) * 8;
            slot[0] = s0; slot[1] = s1; slot[2] = s2; slot[3] = s3;
            slot[4] = s4; slot[5] = s5; slot[6] = s6; slot[7] = s7;
        }
    } else if (totals) {
        addTotals(n, succeeded, totals);
    }
}

// Mixed arithmetic, second half (after hemSearchKernel): FP64 refinement,
// stability and selection of each cell's FP32 candidates; cells the
// refinement defers (a full FP64 search or FP64 restarts) go to the queue.
template <int NS, int NL>
__global__ __launch_bounds__(64, RAMD_REFINE_MIN_WAVES) void hemRefineKernel(const HemModel<NS>* __restrict__ globalModel,
                                                     const double* __restrict__ q, const double* __restrict__ energy,
                                                     const ThermoOut* __restrict__ guess,
                                                     const uint32_t* __restrict__ order, uint32_t count,
                                                     const CandidateSet* __restrict__ sets,
                                                     HemOutputDev* __restrict__ out,
                                                     unsigned long long* __restrict__ totals,
                                                     uint32_t* __restrict__ queue, unsigned* __restrict__ queueCount) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0, 0};
    uint32_t succeeded = 0;
    if (i < count) {
        const uint32_t c = order[i];
        Flash<NS, NL, double, true, false, false, kModelLds> f(model, n);
        loadHemCell<NS>(f, q, energy, guess[c], c);
        EvalT<double, true> result;
        zeroEval(result);
        bool deferred = false;
        const bool ok = f.refine(sets[c], result, &deferred);
        if (deferred) {
            queue[atomicAdd(queueCount, 1u)] = c;
        } else {
            storeHemCell(ok, result, out[c]);
            succeeded = ok ? 1u : 0u;
        }
    }
    addTotals(n, succeeded, totals);
}

// The deferred cells, refined without deferral by a persistent grid of
// single-wave workgroups: one queued cell per wave at a time with its line
// searches spread over the lanes (the flash's WAVE mode) while the queue is
// short, one cell per lane for a long queue (RAMD_WAVE_QUEUE_CELLS; the count
// is read on the device).
template <int NS, int NL, bool WAVE>
__device__ inline void hemRefineQueued(const HemModel<NS>& model, const double* __restrict__ q,
                                       const double* __restrict__ energy, const ThermoOut* __restrict__ guess,
                                       uint32_t c, const CandidateSet& set, bool lead,
                                       HemOutputDev* __restrict__ out, unsigned long long* __restrict__ totals) {
    Counts n = {0, 0, 0, 0, 0, 0, 0};
    Flash<NS, NL, double, true, false, false, kModelLds, WAVE> f(model, n);
    loadHemCell<NS>(f, q, energy, guess[c], c);
    EvalT<double, true> result;
    zeroEval(result);
    const bool ok = f.refine(set, result);
    if (lead) storeHemCell(ok, result, out[c]);
    const Counts none = {0, 0, 0, 0, 0, 0, 0};
    addTotals(lead ? n : none, lead && ok ? 1u : 0u, totals);
}
template <int NS, int NL>
__global__ __launch_bounds__(64, RAMD_REFINE_MIN_WAVES) void hemRefineQueueKernel(
    const HemModel<NS>* __restrict__ globalModel, const double* __restrict__ q, const double* __restrict__ energy,
    const ThermoOut* __restrict__ guess, const uint32_t* __restrict__ queue, const unsigned* __restrict__ queueCount,
    const CandidateSet* __restrict__ sets, HemOutputDev* __restrict__ out, unsigned long long* __restrict__ totals) {
    const unsigned count = *queueCount;
    const bool perWave = count <= RAMD_WAVE_QUEUE_CELLS;  // uniform
    if (blockIdx.x * (perWave ? 1u : 64u) >= count) return;  // uniform per workgroup
    const HemModel<NS>& model = workgroupModel(globalModel);
    if (perWave) {
        for (unsigned i = blockIdx.x; i < count; i += gridDim.x)  // uniform per workgroup
            hemRefineQueued<NS, NL, true>(model, q, energy, guess, queue[i], sets[queue[i]], threadIdx.x == 0, out,
                                          totals);
    } else {
        for (unsigned base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u) {  // uniform per workgroup
            const unsigned i = base + threadIdx.x;
            if (i < count) {
                hemRefineQueued<NS, NL, false>(model, q, energy, guess, queue[i], sets[queue[i]], true, out, totals);
            } else {
                const Counts none = {0, 0, 0, 0, 0, 0, 0};
                addTotals(none, 0u, totals);
            }
        }
    }
}

__global__ void clearCountKernel(unsigned* __restrict__ count) {
    if (threadIdx.x == 0) *count = 0;
}

// totals[k] = sum over b < blocks of partials[b * width + k], k < width: one
// 64-lane workgroup, a fixed summation order (no atomics, nothing cleared).
__global__ __launch_bounds__(64) void reducePartialsKernel(const unsigned long long* __restrict__ partials,
                                                          unsigned blocks, int width,
                                                          unsigned long long* __restrict__ totals) {
    for (int k = 0; k < width; ++k) {
        unsigned long long sum = 0;
        for (unsigned b = threadIdx.x; b < blocks; b += 64) sum += partials[size_t(b) * width + k];
        for (int offset = 32; offset > 0; offset >>= 1) sum += __shfl_xor(sum, offset, 64);
        if (threadIdx.x == 0) totals[k] = sum;
    }
}

__global__ void zeroTotalsKernel(unsigned long long* totals, int n) {
    if (threadIdx.x < n) totals[threadIdx.x] = 0;
}
__global__ void queueCountKernel(const unsigned* __restrict__ count, unsigned long long* __restrict__ total) {
    if (threadIdx.x == 0) *total = *count;
}

// Workgroups of the persistent deferred-refinement grid (64 lanes each).
constexpr unsigned kQueueBlocks = RAMD_QUEUE_BLOCKS;

// Scheduling key (reactiveHemBucket.h:13-18).
inline unsigned bucketKey(const double* q, int ns, int activeLiquids) {
    unsigned present = 0, dominant = 0;
    for (int k = 0; k < ns; ++k) {
        if (q[k] > 0) ++present;
        if (q[k] > q[dominant]) dominant = unsigned(k);
    }
    const unsigned liquids = activeLiquids < 0 ? 0u : (activeLiquids > 3 ? 3u : unsigned(activeLiquids));
    return (liquids << 6) | ((present > 7 ? 7u : present) << 3) | (dominant > 7 ? 7u : dominant);
}

HemHandle::~HemHandle() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto e : events)
        if (e) (void)hipEventDestroy(e);
    for (void* p : allocations) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

void* HemHandle::allocate(size_t bytes) {
    void* p = nullptr;
    RAMD_HIP(hipMalloc(&p, bytes));
    allocations.push_back(p);
    deviceBytes += bytes;
    return p;
}

void HemHandle::ensureHostPath() {
    if (hostReady) return;
    dq = static_cast<double*>(allocate(capacity * ns * sizeof(double)));
    de = static_cast<double*>(allocate(capacity * sizeof(double)));
    dguess = static_cast<ThermoOut*>(allocate(capacity * sizeof(ThermoOut)));
    dout = static_cast<HemOutputDev*>(allocate(capacity * sizeof(HemOutputDev)));
    dorder = static_cast<uint32_t*>(allocate(capacity * sizeof(uint32_t)));
    hostOrder.resize(capacity);
    hostKey.resize(capacity);
    hostOut.resize(capacity);
    hostReady = true;
}

void HemHandle::launch(const double* q, const double* e, const ThermoOut* g, const uint32_t* order, size_t count,
                       const HemCapillaryDev* capillary, HemOutputDev* out, uint32_t* cellPhases, hipStream_t s,
                       unsigned long long* partials) {
    if (!count) return;
    const unsigned blocks = unsigned((count + 63) / 64);
    if (!partials) hipLaunchKernelGGL(zeroTotalsKernel, dim3(1), dim3(64), 0, s, totals, 8);
#define RAMD_HEM_LAUNCH(NS_, NL_, RJ_, CAP_)                                                                       \
    hipLaunchKernelGGL((hemFlashKernel<NS_, NL_, RJ_, CAP_>), dim3(blocks), dim3(64), 0, s,                        \
                       static_cast<const HemModel<NS_>*>(model), q, e, g, order, uint32_t(count), capillary, out, totals, \
                       cellPhases, partials)
#define RAMD_HEM_SPECIES(CAP_)                                                                                    \
    if (ns == 1 && nl == 0) RAMD_HEM_LAUNCH(1, 0, false, CAP_);                                                   \
    else if (ns == 1 && !analytic) RAMD_HEM_LAUNCH(1, 1, false, CAP_);                                            \
    else if (ns == 1) RAMD_HEM_LAUNCH(1, 1, true, CAP_);                                                          \
    else if (nl == 0) RAMD_HEM_LAUNCH(2, 0, false, CAP_);                                                         \
    else if (!analytic) RAMD_HEM_LAUNCH(2, 1, false, CAP_);                                                       \
    else RAMD_HEM_LAUNCH(2, 1, true, CAP_);
    // the capillary / boundary-search flash for interface input and boundaryRecovery models
    if (capillary || boundary) { RAMD_HEM_SPECIES(true) }
    else { RAMD_HEM_SPECIES(false) }
#undef RAMD_HEM_SPECIES
#undef RAMD_HEM_LAUNCH
    RAMD_HIP(hipGetLastError());
}

void HemHandle::launchMixed(const double* q, const double* e, const ThermoOut* g, const uint32_t* order,
                            size_t count, HemOutputDev* out, hipStream_t s) {
    if (!count) return;
    if (!sets) {
        sets = static_cast<CandidateSet*>(allocate(capacity * sizeof(CandidateSet)));
        queue = static_cast<uint32_t*>(allocate(capacity * sizeof(uint32_t)));
        queueCount = static_cast<unsigned*>(allocate(sizeof(unsigned)));
    }
    const unsigned blocks = unsigned((count + 63) / 64);
    hipLaunchKernelGGL(zeroTotalsKernel, dim3(1), dim3(64), 0, s, totals, 10);
    launchHemSearch(ns, nl, model, q, e, g, order, uint32_t(count), sets, totals + 8, s);
    hipLaunchKernelGGL(clearCountKernel, dim3(1), dim3(64), 0, s, queueCount);
#define RAMD_HEM_MIXED(NS_, NL_)                                                                                  \
    hipLaunchKernelGGL((hemRefineKernel<NS_, NL_>), dim3(blocks), dim3(64), 0, s,                                   \
                       static_cast<const HemModel<NS_>*>(model), q, e, g, order, uint32_t(count), sets, out, totals, \
                       queue, queueCount);                                                                          \
    hipLaunchKernelGGL((hemRefineQueueKernel<NS_, NL_>), dim3(kQueueBlocks), dim3(64), 0, s,                        \
                       static_cast<const HemModel<NS_>*>(model), q, e, g, queue, queueCount, sets, out, totals)
    if (ns == 1 && nl == 0) { RAMD_HEM_MIXED(1, 0); }
    else if (ns == 1) { RAMD_HEM_MIXED(1, 1); }
    else if (nl == 0) { RAMD_HEM_MIXED(2, 0); }
    else { RAMD_HEM_MIXED(2, 1); }
#undef RAMD_HEM_MIXED
    hipLaunchKernelGGL(queueCountKernel, dim3(1), dim3(64), 0, s, queueCount, totals + 9);
    RAMD_HIP(hipGetLastError());
}

// The process's default flash arithmetic: REACTIVE_AMD_HEM_ARITHMETIC=mixed
// selects the mixed flash, anything else (or nothing) plain FP64.  Read once,
// so the numerical policy a driver hashes and the handles it creates agree.
static int defaultArithmetic() {
    static const int mode = [] {
        const char* v = std::getenv("REACTIVE_AMD_HEM_ARITHMETIC");
        return v && std::strcmp(v, "mixed") == 0 ? REACTIVE_AMD_FLASH_MIXED : REACTIVE_AMD_FLASH_FP64;
    }();
    return mode;
}

HemHandle* createHem(const void* raw, size_t bytes, size_t capacity, bool hostPath) {
    using namespace image;
    require(raw && bytes == sizeof(Model) && capacity && capacity <= UINT32_MAX, "Invalid HIP HEM model/capacity");
    Model m;
    std::memcpy(static_cast<void*>(&m), raw, sizeof(m));
    require(validModel(m), "Invalid HIP HEM model tables or domain");
    require(m.ns == 1 || m.ns == 2, "This HEM build supports 1 or 2 gas species");
    require(m.nl <= 1, "This HEM build supports at most one pure liquid");
    require(!(m.nl && m.condensedKind[0]), "Solid condensed phases are not provided by this HEM build");
    int devices = 0;
    RAMD_HIP(hipGetDeviceCount(&devices));
    require(devices > 0, "No HIP device");
    std::unique_ptr<HemHandle> h(new HemHandle);
    h->ns = m.ns;
    h->nl = m.nl;
    h->analytic = m.analyticJacobian != 0;
    h->boundary = m.boundaryRecovery != 0;
    h->arithmetic = defaultArithmetic();
    h->capacity = capacity;
    h->device = currentDevice();
    RAMD_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (auto& e : h->events) RAMD_HIP(hipEventCreate(&e));
    if (m.ns == 1) {
        HemModel<1> d;
        buildDeviceModel<1>(m, d);
        h->model = h->allocate(sizeof(d));
        RAMD_HIP(hipMemcpy(h->model, &d, sizeof(d), hipMemcpyHostToDevice));
    } else {
        HemModel<2> d;
        buildDeviceModel<2>(m, d);
        h->model = h->allocate(sizeof(d));
        RAMD_HIP(hipMemcpy(h->model, &d, sizeof(d), hipMemcpyHostToDevice));
    }
    h->totals = static_cast<unsigned long long*>(h->allocate(10 * sizeof(unsigned long long)));
    h->modelBytes = h->deviceBytes;
    if (hostPath) h->ensureHostPath();
    return h.release();
}

// Host-array batch (reactiveGpuHem.cuh:227-349 semantics).
int runHost(HemHandle* h, const double* q, const double* energy, const ReactiveGpuHemCapillaryInputV2* capillary,
            size_t count, ReactiveThermoState* states, int* success, ReactiveGpuHemProfileV1* profile) {
    require(h && q && energy && states && success && profile && profile->abiVersion == 1 &&
                profile->structBytes == sizeof(*profile),
            "Invalid HIP HEM batch arguments");
    require(count && count <= h->capacity, "HIP HEM batch exceeds capacity");
    if (capillary) {
        for (size_t c = 0; c < count; ++c) {
            const auto& x = capillary[c];
            require(fin(x.color) && x.color >= 0 && x.color <= 1 && fin(x.pressureJump) &&
                        (x.equilibrium == 0 || x.equilibrium == 1),
                    "Invalid HIP HEM capillary cell input");  // reactiveGpuHem.cuh:237-243
        }
    }
    RAMD_HIP(hipSetDevice(h->device));
    h->ensureHostPath();
    if (capillary && !h->dcap)
        h->dcap = static_cast<HemCapillaryDev*>(h->allocate(h->capacity * sizeof(HemCapillaryDev)));
    const auto begin = std::chrono::steady_clock::now();
    // stable counting sort on the scheduling key, heavy buckets first
    std::array<size_t, 256> counts{};
    for (size_t c = 0; c < count; ++c) {
        const unsigned key = bucketKey(q + c * h->ns, h->ns, states[c].activeLiquids);
        h->hostKey[c] = uint8_t(key);
        ++counts[key];
    }
    std::array<size_t, 256> offset{};
    size_t next = 0;
    for (int b = 255; b >= 0; --b) { offset[b] = next; next += counts[b]; }
    for (size_t c = 0; c < count; ++c) h->hostOrder[offset[h->hostKey[c]]++] = uint32_t(c);
    hipStream_t s = h->stream;
    RAMD_HIP(hipEventRecord(h->events[0], s));
    RAMD_HIP(hipMemcpyAsync(h->dq, q, count * h->ns * sizeof(double), hipMemcpyHostToDevice, s));
    RAMD_HIP(hipMemcpyAsync(h->de, energy, count * sizeof(double), hipMemcpyHostToDevice, s));
    RAMD_HIP(hipMemcpyAsync(h->dguess, states, count * sizeof(ThermoOut), hipMemcpyHostToDevice, s));
    RAMD_HIP(hipMemcpyAsync(h->dorder, h->hostOrder.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (capillary)
        RAMD_HIP(hipMemcpyAsync(h->dcap, capillary, count * sizeof(HemCapillaryDev), hipMemcpyHostToDevice, s));
    RAMD_HIP(hipEventRecord(h->events[1], s));
    // the mixed arithmetic runs the equilibrium flash without the boundary search
    const bool mixed = !capillary && !h->boundary && h->arithmetic == REACTIVE_AMD_FLASH_MIXED;
    if (mixed) h->launchMixed(h->dq, h->de, h->dguess, h->dorder, count, h->dout, s);
    else h->launch(h->dq, h->de, h->dguess, h->dorder, count, capillary ? h->dcap : nullptr, h->dout, nullptr, s);
    RAMD_HIP(hipEventRecord(h->events[2], s));
    RAMD_HIP(hipMemcpyAsync(h->hostOut.data(), h->dout, count * sizeof(HemOutputDev), hipMemcpyDeviceToHost, s));
    unsigned long long totals[10] = {};
    RAMD_HIP(hipMemcpyAsync(totals, h->totals, sizeof(totals), hipMemcpyDeviceToHost, s));
    RAMD_HIP(hipEventRecord(h->events[3], s));
    RAMD_HIP(hipEventSynchronize(h->events[3]));
    float upload = 0, kernel = 0, download = 0;
    RAMD_HIP(hipEventElapsedTime(&upload, h->events[0], h->events[1]));
    RAMD_HIP(hipEventElapsedTime(&kernel, h->events[1], h->events[2]));
    RAMD_HIP(hipEventElapsedTime(&download, h->events[2], h->events[3]));
    ReactiveGpuHemProfileV1 p{};
    p.abiVersion = 1;
    p.structBytes = sizeof(p);
    p.batches = 1;
    p.submitted = count;
    for (size_t c = 0; c < count; ++c) {
        const HemOutputDev& o = h->hostOut[c];
        success[c] = o.success;
        if (o.success) {
            std::memcpy(&states[c], &o.state, sizeof(ReactiveThermoState));
            ++p.succeeded;
        }
    }
    p.deviceFailures = count - p.succeeded;
    // the mixed arithmetic counts its FP32 search's phase evaluations too
    p.phaseEvaluations = totals[0] + (mixed ? totals[8] : 0);
    h->lastSinglePhases = mixed ? totals[8] : 0;
    h->lastQueued = mixed ? totals[9] : 0;
    p.residualEvaluations = totals[1];
    p.flashCandidates = totals[2];
    p.stableCandidates = totals[3];
    p.finiteDifferenceJacobians = totals[5];
    p.analyticJacobians = totals[7];
    p.transferBytes = count * (h->ns * sizeof(double) + sizeof(double) + sizeof(ThermoOut) + sizeof(uint32_t) +
                               sizeof(HemOutputDev) + (capillary ? sizeof(int) : 0)) +
                      sizeof(totals);
    p.deviceBytes = h->deviceBytes;
    p.hostBytes = h->capacity * (sizeof(HemOutputDev) + sizeof(uint32_t) + sizeof(uint8_t));
    p.wallSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    p.kernelSeconds = kernel / 1000.;
    p.copySeconds = (upload + download) / 1000.;
    *profile = p;
    return 0;
}

// ---- device scratch of the allocation-free resident batch
// (reactive_gpu_hem_run_resident_noalloc_v1).  One slot of 8 counters per
// 64-cell workgroup of a launch, then the 8 batch totals: 1 MiB + 64 B on
// every device the library runs on, held until the process ends.  Every slot
// a call reads, the same call has written first, so nothing is cleared
// between calls; calls that overlap on one device share it and race.
constexpr unsigned kNoAllocBlocks = 16384;
constexpr size_t kNoAllocMaxCells = size_t(kNoAllocBlocks) * 64;  // 1,048,576 cells per call
static __device__ __align__(256) unsigned long long scratch[size_t(kNoAllocBlocks) * 8 + 8];

// scratch's address on `device` (the host-side name of the array is only a
// stand-in): looked up once per device with that device current.
unsigned long long* scratchAddress(int device) {
    constexpr int kDevices = 64;
    static std::mutex lock;
    static unsigned long long* address[kDevices] = {};
    require(device >= 0 && device < kDevices, "HIP device index beyond the allocation-free batch's table");
    std::lock_guard<std::mutex> guard(lock);
    if (!address[device]) {
        void* ptr = nullptr;
        RAMD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        address[device] = static_cast<unsigned long long*>(ptr);
    }
    return address[device];
}

}  // namespace ramd

using namespace ramd;

extern "C" {

// "" (plain FP64, separate rounding) or, with the mixed arithmetic as the
// process default, the FP32 combined policy (reactiveGpuHem.h:37); both are
// strings reactiveThermo.cpp:294-301 accepts.
const char* reactive_gpu_hem_numerical_policy_v1(void) {
    return defaultArithmetic() == REACTIVE_AMD_FLASH_MIXED ? "fp32-cubic-newton-fp64-accept-v2:experimental" : "";
}

int reactive_amd_hem_set_arithmetic(void* raw, int mode) {
    if (!raw) return -1;
    if (mode != REACTIVE_AMD_FLASH_FP64 && mode != REACTIVE_AMD_FLASH_MIXED) return 1;
    static_cast<HemHandle*>(raw)->arithmetic = mode;
    return 0;
}

int reactive_amd_hem_mixed_report(void* raw, uint64_t* singlePhaseEvaluations, uint64_t* queuedCells) {
    if (!raw) return -1;
    if (!singlePhaseEvaluations || !queuedCells) return 1;
    const HemHandle* h = static_cast<const HemHandle*>(raw);
    *singlePhaseEvaluations = h->lastSinglePhases;
    *queuedCells = h->lastQueued;
    return 0;
}

void* reactive_gpu_hem_create_v1(const void* model, size_t bytes, size_t capacity, char* error, size_t size) {
    try {
        return createHem(model, bytes, capacity, true);
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return nullptr;
    }
}

void* reactive_gpu_hem_create_resident_v1(const void* model, size_t bytes, size_t capacity, char* error, size_t size) {
    try {
        return createHem(model, bytes, capacity, false);
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return nullptr;
    }
}

size_t reactive_gpu_hem_device_bytes_v1(void* raw) { return raw ? static_cast<HemHandle*>(raw)->deviceBytes : 0; }

void reactive_gpu_hem_destroy_v1(void* raw) { delete static_cast<HemHandle*>(raw); }

int reactive_gpu_hem_run_v1(void* raw, const double* q, const double* energy, size_t count, ReactiveThermoState* states,
                            int* success, ReactiveGpuHemProfileV1* profile, char* error, size_t size) {
    try {
        return runHost(static_cast<HemHandle*>(raw), q, energy, nullptr, count, states, success, profile);
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return 1;
    }
}

int reactive_gpu_hem_run_v2(void* raw, const double* q, const double* energy,
                            const ReactiveGpuHemCapillaryInputV2* capillary, size_t count, ReactiveThermoState* states,
                            int* success, ReactiveGpuHemProfileV1* profile, char* error, size_t size) {
    if (!capillary) {
        writeError(error, size, "Null HIP HEM capillary input");
        return 1;
    }
    try {
        return runHost(static_cast<HemHandle*>(raw), q, energy, capillary, count, states, success, profile);
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return 1;
    }
}

int reactive_gpu_hem_initialize_tp_v1(void*, double*, double*, const ReactiveGpuHemCapillaryInputV2*, size_t,
                                      ReactiveThermoState*, int*, ReactiveGpuHemProfileV1*, char* error, size_t size) {
    writeError(error, size, "reactive_gpu_hem_initialize_tp_v1: capillary primitive initialisation is not provided "
                            "by the MI355X HEM build");
    return 1;
}

int reactive_gpu_hem_run_resident_v1(void* raw, const double* q, const double* energy, const ReactiveThermoState* seed,
                                     const ReactiveGpuHemCapillaryInputV2* capillary, const uint32_t* order,
                                     size_t count, ReactiveGpuHemOutputV1* output, ReactiveGpuHemProfileV1* profile,
                                     char* error, size_t size) {
    try {
        require(raw && q && energy && seed && capillary && order && output && profile && profile->abiVersion == 1 &&
                    profile->structBytes == sizeof(*profile),
                "Invalid resident HIP HEM batch arguments");
        HemHandle* h = static_cast<HemHandle*>(raw);
        require(count && count <= h->capacity, "Resident HIP HEM batch exceeds capacity");
        const auto begin = std::chrono::steady_clock::now();
        RAMD_HIP(hipSetDevice(h->device));
        hipStream_t s = h->stream;
        // every input is device memory indexed by global cell; as in the
        // reference (reactiveGpuHem.cuh:360-380), a cell whose interface input
        // recoverCapillary rejects fails on its own
        const auto* cap = reinterpret_cast<const HemCapillaryDev*>(capillary);
        RAMD_HIP(hipEventRecord(h->events[0], s));
        h->launch(q, energy, reinterpret_cast<const ThermoOut*>(seed), order, count, cap,
                  reinterpret_cast<HemOutputDev*>(output), nullptr, s);
        RAMD_HIP(hipEventRecord(h->events[1], s));
        unsigned long long totals[8] = {};
        RAMD_HIP(hipMemcpyAsync(totals, h->totals, sizeof(totals), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipEventRecord(h->events[2], s));
        RAMD_HIP(hipEventSynchronize(h->events[2]));
        float kernel = 0, download = 0;
        RAMD_HIP(hipEventElapsedTime(&kernel, h->events[0], h->events[1]));
        RAMD_HIP(hipEventElapsedTime(&download, h->events[1], h->events[2]));
        ReactiveGpuHemProfileV1 p{};
        p.abiVersion = 1;
        p.structBytes = sizeof(p);
        p.batches = 1;
        p.submitted = count;
        p.succeeded = totals[6];
        p.deviceFailures = count - totals[6];
        p.phaseEvaluations = totals[0];
        p.residualEvaluations = totals[1];
        p.flashCandidates = totals[2];
        p.stableCandidates = totals[3];
        p.finiteDifferenceJacobians = totals[5];
        p.analyticJacobians = totals[7];
        p.transferBytes = sizeof(totals);
        p.deviceBytes = h->deviceBytes;
        p.wallSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
        p.kernelSeconds = kernel / 1000.;
        p.copySeconds = download / 1000.;
        *profile = p;
        return 0;
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return 1;
    }
}

// Same contract and results as reactive_gpu_hem_run_resident_v1, without any
// device allocation: the work counters go to the library's static scratch.
int reactive_gpu_hem_run_resident_noalloc_v1(void* raw, const double* q, const double* energy,
                                             const ReactiveThermoState* seed,
                                             const ReactiveGpuHemCapillaryInputV2* capillary, const uint32_t* order,
                                             size_t count, ReactiveGpuHemOutputV1* output,
                                             ReactiveGpuHemProfileV1* profile, char* error, size_t size) {
    try {
        require(raw && q && energy && seed && capillary && order && output && profile && profile->abiVersion == 1 &&
                    profile->structBytes == sizeof(*profile),
                "Invalid resident HIP HEM batch arguments");
        HemHandle* h = static_cast<HemHandle*>(raw);
        require(count && count <= h->capacity, "Resident HIP HEM batch exceeds capacity");
        require(count <= kNoAllocMaxCells, "reactive_gpu_hem_run_resident_noalloc_v1: a batch holds at most "
                                           "1048576 cells (the size of its static device scratch)");
        const auto begin = std::chrono::steady_clock::now();
        RAMD_HIP(hipSetDevice(h->device));
        unsigned long long* partials = scratchAddress(h->device);
        unsigned long long* totalsDev = partials + size_t(kNoAllocBlocks) * 8;
        hipStream_t s = h->stream;
        const unsigned blocks = unsigned((count + 63) / 64);
        const auto* cap = reinterpret_cast<const HemCapillaryDev*>(capillary);
        RAMD_HIP(hipEventRecord(h->events[0], s));
        h->launch(q, energy, reinterpret_cast<const ThermoOut*>(seed), order, count, cap,
                  reinterpret_cast<HemOutputDev*>(output), nullptr, s, partials);
        hipLaunchKernelGGL(reducePartialsKernel, dim3(1), dim3(64), 0, s, partials, blocks, 8, totalsDev);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipEventRecord(h->events[1], s));
        unsigned long long totals[8] = {};
        RAMD_HIP(hipMemcpyAsync(totals, totalsDev, sizeof(totals), hipMemcpyDeviceToHost, s));
        RAMD_HIP(hipEventRecord(h->events[2], s));
        RAMD_HIP(hipEventSynchronize(h->events[2]));
        float kernel = 0, download = 0;
        RAMD_HIP(hipEventElapsedTime(&kernel, h->events[0], h->events[1]));
        RAMD_HIP(hipEventElapsedTime(&download, h->events[1], h->events[2]));
        ReactiveGpuHemProfileV1 p{};
        p.abiVersion = 1;
        p.structBytes = sizeof(p);
        p.batches = 1;
        p.submitted = count;
        p.succeeded = totals[6];
        p.deviceFailures = count - totals[6];
        p.phaseEvaluations = totals[0];
        p.residualEvaluations = totals[1];
        p.flashCandidates = totals[2];
        p.stableCandidates = totals[3];
        p.finiteDifferenceJacobians = totals[5];
        p.analyticJacobians = totals[7];
        p.transferBytes = sizeof(totals);
        p.deviceBytes = h->deviceBytes;
        p.wallSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
        p.kernelSeconds = kernel / 1000.;
        p.copySeconds = download / 1000.;
        *profile = p;
        return 0;
    } catch (const std::exception& e) {
        writeError(error, size, e.what());
        return 1;
    }
}

}
