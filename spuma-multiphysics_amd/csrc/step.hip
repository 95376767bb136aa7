// SPDX-License-Identifier: GPL-3.0-or-later
// Device-resident SSPRK2 step (reactive_amd_step_*): the reference driver's
// Flow::step (src/reactiveFoam/ReactiveFoam.C:961-1019, no chemistry) as one
// stream of kernels with q, states and dt resident in HBM:
//
//   dt  = CFL(qA, stateA)                        stable_step_primitives
//   qB  = qA + dt L(qA)           [bA]           stage 0
//   sB  = flash(qB | seeds sA)                   recover (ReactiveFoam.C:992)
//   dt <= CFL(qB, sB, max dt) (1 + 1e-10) ?      recheck (ReactiveFoam.C:993)
//   qA  = .5 qA + .5 (qB + dt L(qB))  [bB]       stage 1
//   sA  = flash(qA | seeds sA)                   recover from the step-start
//                                                seeds (ReactiveFoam.C:1011-1013)
//   totals = sum q V, boundary = .5 dt (bA + bB) (ReactiveFoam.C:953-958, 1016)
//
// Flash inputs are read straight from the SoA conserved arrays: species
// masses and e = rhoE - |rho u|^2 / (2 rho) (ReactiveFoam.C:727-728).  The
// flash epilogue writes p and c_frozen as SoA for the next transport stage.
// Cells are flashed in heavy-first bucket order built on the device.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "hip_common.h"
#include "flash_device.h"
#include "hem_internal.h"
#include "transport_internal.h"
#include "cartesian_kernels.h"
#include "flash_step.h"
#include "reactive_amd.h"

namespace ramd {

// ---------------------------------------------------------------- kernels
// Plain FP64 flash of the cells in `order`.
template <int NS, int NL>
__global__ __launch_bounds__(64, RAMD_FLASH_MIN_WAVES) void stepFlashKernel(const HemModel<NS>* __restrict__ globalModel,
                                                      const double* __restrict__ q, size_t stride,
                                                      const ThermoOut* __restrict__ seeds,
                                                      const uint32_t* __restrict__ order, uint32_t count,
                                                      const double* __restrict__ dtDev, ThermoOut* __restrict__ out,
                                                      double* __restrict__ pOut, double* __restrict__ cOut,
                                                      double* __restrict__ rOut, uint16_t* __restrict__ cost,
                                                      unsigned long long* __restrict__ counters) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0};
    uint32_t failed = 0;
    if (i < count && *dtDev > 0) {
        const uint32_t c = order[i];
        const ThermoOut g = seeds[c];
        Flash<NS, NL> f(model, n);
        loadCell(f, q, stride, g, c, NS);
        Eval result;
        zeroEval(result);
        const bool ok = f.run(result);
        storeCell(ok, result, g, c, out, pOut, cOut, rOut);
        failed = ok ? 0u : 1u;
        if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
    }
    countFlash(failed, n.phases, counters);
}

// Mixed precision, second half: FP64 refinement, stability and selection.
template <int NS, int NL>
__global__ __launch_bounds__(64, RAMD_FLASH_MIN_WAVES) void stepRefineKernel(const HemModel<NS>* __restrict__ globalModel,
                                                       const double* __restrict__ q, size_t stride,
                                                       const ThermoOut* __restrict__ seeds,
                                                       const uint32_t* __restrict__ order, uint32_t count,
                                                       const double* __restrict__ dtDev,
                                                       const CandidateSet* __restrict__ sets,
                                                       ThermoOut* __restrict__ out, double* __restrict__ pOut,
                                                       double* __restrict__ cOut, double* __restrict__ rOut,
                                                       uint16_t* __restrict__ cost,
                                                       unsigned long long* __restrict__ counters) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0};
    uint32_t failed = 0;
    if (i < count && *dtDev > 0) {
        const uint32_t c = order[i];
        const ThermoOut g = seeds[c];
        Flash<NS, NL, double, true> f(model, n);
        loadCell(f, q, stride, g, c, NS);
        EvalT<double, true> result;
        zeroEval(result);
        const bool ok = f.refine(sets[c], result);
        storeCell(ok, result, g, c, out, pOut, cOut, rOut);
        failed = ok ? 0u : 1u;
        if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
    }
    countFlash(failed, n.phases, counters);
}

// Heavy-first scheduling and a counting sort: histogram, exclusive scan from
// the heaviest bucket down, scatter.  Without a cost history the key is the
// reference's bucket (reactiveHemBucket.h:13-18): liquids, present species,
// dominant species.  With one it is the cell's phase-evaluation count in its
// latest flash, in bins of 8, so the 64 cells of a wave carry near-equal work
// (the order decides nothing numerical: every cell's flash is independent).
__global__ __launch_bounds__(256) void bucketHistogramKernel(const double* __restrict__ q, size_t stride, int ns,
                                                             const ThermoOut* __restrict__ seeds,
                                                             const uint16_t* __restrict__ cost, size_t cells,
                                                             uint8_t* __restrict__ key,
                                                             unsigned* __restrict__ count) {
    __shared__ unsigned hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < cells) {
        unsigned k;
        if (cost) {
            const unsigned bin = unsigned(cost[c]) >> 3;
            k = bin > 255u ? 255u : bin;
        } else {
            unsigned present = 0, dominant = 0;
            double best = q[c];
            for (int s = 0; s < ns; ++s) {
                const double v = q[size_t(s) * stride + c];
                if (v > 0) ++present;
                if (v > best) { best = v; dominant = unsigned(s); }
            }
            const int al = seeds[c].activeLiquids;
            const unsigned liquids = al < 0 ? 0u : (al > 3 ? 3u : unsigned(al));
            k = (liquids << 6) | ((present > 7 ? 7u : present) << 3) | (dominant > 7 ? 7u : dominant);
        }
        key[c] = uint8_t(k);
        atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    if (hist[threadIdx.x]) atomicAdd(&count[threadIdx.x], hist[threadIdx.x]);
}

__global__ void bucketScanKernel(unsigned* __restrict__ count, unsigned* __restrict__ offset) {
    if (threadIdx.x != 0) return;
    unsigned next = 0;
    for (int b = 255; b >= 0; --b) {
        offset[b] = next;
        next += count[b];
        count[b] = 0;  // leaves the histogram cleared for the next ordering
    }
}

// Each block ranks its cells per bucket in LDS and reserves one run per
// bucket with a single global atomic, so no bucket counter is hit by every
// cell of the grid.
__global__ __launch_bounds__(256) void bucketScatterKernel(const uint8_t* __restrict__ key, size_t cells,
                                                           unsigned* __restrict__ offset,
                                                           uint32_t* __restrict__ order) {
    __shared__ unsigned local[256], base[256];
    local[threadIdx.x] = 0;
    __syncthreads();
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned k = 0, rank = 0;
    if (c < cells) {
        k = key[c];
        rank = atomicAdd(&local[k], 1u);
    }
    __syncthreads();
    if (local[threadIdx.x]) base[threadIdx.x] = atomicAdd(&offset[threadIdx.x], local[threadIdx.x]);
    __syncthreads();
    if (c < cells) order[base[k] + rank] = uint32_t(c);
}

// Neumaier-compensated sum of q_k V over the cells: per-block partials in a
// fixed order, then one block merges them in block order (deterministic).
__device__ inline void twoSumAdd(double& s, double& comp, double x) {
    const double t = s + x;
    comp += ::fabs(s) >= ::fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}
constexpr int kTotalBlocks = 512;
__global__ __launch_bounds__(256) void totalsPartialKernel(const double* __restrict__ q, size_t stride, size_t cells,
                                                           int nv, double volume, double2* __restrict__ partial) {
    __shared__ double ls[256], lc[256];
    for (int k = 0; k < nv; ++k) {
        double s = 0, comp = 0;
        for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells; c += size_t(gridDim.x) * blockDim.x)
            twoSumAdd(s, comp, q[size_t(k) * stride + c] * volume);
        ls[threadIdx.x] = s;
        lc[threadIdx.x] = comp;
        __syncthreads();
        for (unsigned st = 128; st; st >>= 1) {
            if (threadIdx.x < st) {
                double a = ls[threadIdx.x], ac = lc[threadIdx.x] + lc[threadIdx.x + st];
                twoSumAdd(a, ac, ls[threadIdx.x + st]);
                ls[threadIdx.x] = a;
                lc[threadIdx.x] = ac;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[size_t(k) * gridDim.x + blockIdx.x] = make_double2(ls[0], lc[0]);
        __syncthreads();
    }
}
__global__ void totalsFinishKernel(const double2* __restrict__ partial, int blocks, int nv, double* __restrict__ out) {
    const int k = threadIdx.x;
    if (k >= nv) return;
    double s = 0, comp = 0;
    for (int b = 0; b < blocks; ++b) {
        comp += partial[size_t(k) * blocks + b].y;
        twoSumAdd(s, comp, partial[size_t(k) * blocks + b].x);
    }
    out[k] = s + comp;
}

// scalars: [0] dt, [1] recheck dt, [2] cfl-violation counter (as double), [3] max dt of this call
__global__ void finishCflKernel(const double* __restrict__ blockMin, size_t n, double* __restrict__ out) {
    double v = __builtin_huge_val();
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) v = blockMin[i] < v ? blockMin[i] : v;
    __shared__ double lane[4];
    v = waveMin(v);
    if ((threadIdx.x & 63) == 0) lane[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lane[0];
        for (int w = 1; w < int(blockDim.x >> 6); ++w) m = lane[w] < m ? lane[w] : m;
        *out = isfinite(m) && m > 0 ? m : -1;
    }
}
__global__ void recheckKernel(double* __restrict__ scalars, unsigned long long* __restrict__ counters) {
    if (threadIdx.x) return;
    const double dt = scalars[0], stable = scalars[1];
    if (!(dt > 0) || !(dt <= stable * (1 + 1e-10))) atomicAdd(&counters[4], 1ull);
}
__global__ void boundaryIntegralKernel(const double* __restrict__ bA, const double* __restrict__ bB,
                                      const double* __restrict__ scalars, int nv, double* __restrict__ out) {
    const int k = threadIdx.x;
    if (k < nv) out[k] = .5 * scalars[0] * (bA[k] + bB[k]);
}
__global__ void setScalarKernel(double* __restrict__ dst, double value) {
    if (threadIdx.x == 0) *dst = value;
}
__global__ void clearCountersKernel(unsigned long long* __restrict__ c, int n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}
__global__ void toThermoKernel(const ReactiveThermoState* __restrict__ in, size_t n, ThermoOut* __restrict__ out,
                               double* __restrict__ p, double* __restrict__ c, double* __restrict__ r) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    std::memcpy(&out[i], &in[i], sizeof(ThermoOut));
    p[i] = in[i].p;
    c[i] = in[i].soundFrozen;
    r[i] = in[i].rho;
}

inline unsigned blocksFor(size_t n, unsigned b) { return unsigned((n + b - 1) / b); }
constexpr int kMarks = 11;  // events per step

// ---------------------------------------------------------------- handle
struct Step {
    Transport* t = nullptr;
    HemHandle* h = nullptr;
    int ns = 0, nv = 0, device = 0;
    size_t cells = 0, all = 0;
    std::vector<void*> allocations;
    double *qA = nullptr, *qB = nullptr, *pA = nullptr, *cA = nullptr, *pB = nullptr, *cB = nullptr;
    double *rA = nullptr, *rB = nullptr;  // recovered densities (read by the HLLC flux)
    ThermoOut *sA = nullptr, *sB = nullptr;
    uint8_t* key = nullptr;
    uint32_t* order = nullptr;
    uint16_t* cost = nullptr;  // FP64 phase evaluations of each cell in its latest flash
    bool haveCost = false;
    // mixed precision: FP32 search order and cost, candidate sets
    uint32_t* orderSearch = nullptr;
    uint16_t* costSearch = nullptr;
    bool haveCostSearch = false;
    CandidateSet* sets = nullptr;
    int mode = REACTIVE_AMD_FLASH_FP64;
    unsigned *bucketCount = nullptr, *bucketOffset = nullptr;
    double* blockMin = nullptr;
    double* scalars = nullptr;  // dt, recheck
    double *bA = nullptr, *bB = nullptr, *boundary = nullptr, *totals = nullptr;
    double2* totalPartial = nullptr;
    // fail1, phases1, fail2, phases2, violations, HLLC rejections, FP32 phases (mixed)
    unsigned long long* counters = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t events[2] = {};
    // 11 per step: 9 in sequence bracketing cfl, stage 0, flash #1, recheck, stage 1,
    // flash #2, then the end of each flash's FP32 search (mixed mode)
    std::vector<hipEvent_t> marks;
    // captured steps: graph[parity] for the current (sA, sB) assignment, with
    // the settings it was captured for
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graphMode[2] = {-1, -1}, graphScheme[2] = {-1, -1};
    double graphCfl[2] = {0, 0}, graphMax[2] = {0, 0};
    int parity = 0;  // flips with every step (the seed and result buffers swap)
    bool loaded = false;
    char error[1024] = {};

    ~Step() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto e : events)
            if (e) (void)hipEventDestroy(e);
        for (auto e : marks) (void)hipEventDestroy(e);
        for (void* p : allocations) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <class T> T* alloc(size_t n) {
        void* p = nullptr;
        RAMD_HIP(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
        allocations.push_back(p);
        return static_cast<T*>(p);
    }

    // Flash of every cell of q (step-start seeds sA) into out and the SoA
    // transport inputs; costOut (may be null) receives the FP64 cost.
    void flash(const double* q, double* pOut, double* cOut, double* rOut, ThermoOut* out, uint16_t* costOut,
               unsigned long long* counter, hipEvent_t split) {
        const unsigned blocks = blocksFor(cells, 64);
        if (mode == REACTIVE_AMD_FLASH_MIXED) {
#define RAMD_STEP_MIXED(NS_, NL_)                                                                                  \
    launchSearch(NS_, NL_, h->model, q, all, sA, orderSearch, uint32_t(cells), scalars, sets,                     \
                 costOut ? costSearch : nullptr, counters + 6, stream);                                            \
    if (split) RAMD_HIP(hipEventRecord(split, stream));                                                            \
    hipLaunchKernelGGL((stepRefineKernel<NS_, NL_>), dim3(blocks), dim3(64), 0, stream,                             \
                       static_cast<const HemModel<NS_>*>(h->model), q, all, sA, order, uint32_t(cells), scalars,   \
                       sets, out, pOut, cOut, rOut, costOut, counter)
            if (h->ns == 1 && h->nl == 0) { RAMD_STEP_MIXED(1, 0); }
            else if (h->ns == 1) { RAMD_STEP_MIXED(1, 1); }
            else if (h->nl == 0) { RAMD_STEP_MIXED(2, 0); }
            else { RAMD_STEP_MIXED(2, 1); }
#undef RAMD_STEP_MIXED
        } else {
#define RAMD_STEP_FLASH(NS_, NL_)                                                                                 \
    hipLaunchKernelGGL((stepFlashKernel<NS_, NL_>), dim3(blocks), dim3(64), 0, stream,                             \
                       static_cast<const HemModel<NS_>*>(h->model), q, all, sA, order, uint32_t(cells), scalars, out, \
                       pOut, cOut, rOut, costOut, counter)
            if (h->ns == 1 && h->nl == 0) { RAMD_STEP_FLASH(1, 0); }
            else if (h->ns == 1) { RAMD_STEP_FLASH(1, 1); }
            else if (h->nl == 0) { RAMD_STEP_FLASH(2, 0); }
            else { RAMD_STEP_FLASH(2, 1); }
#undef RAMD_STEP_FLASH
        }
        RAMD_HIP(hipGetLastError());
    }

    // Heavy-first order of all cells: by cost when one is known, else by the
    // reference bucket of the step-start state.
    void sortCells(const uint16_t* byCost, uint32_t* out) {
        hipLaunchKernelGGL(bucketHistogramKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, stream, qA, all, ns, sA,
                           byCost, cells, key, bucketCount);
        hipLaunchKernelGGL(bucketScanKernel, dim3(1), dim3(64), 0, stream, bucketCount, bucketOffset);
        hipLaunchKernelGGL(bucketScatterKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, stream, key, cells,
                           bucketOffset, out);
    }

    // One SSPRK2 step; every launch is on `stream` and reads dt from device memory.
    void mark(hipEvent_t* m, int i) {
        if (m) RAMD_HIP(hipEventRecord(m[i], stream));
    }
    void one(double cfl, double maximum, hipEvent_t* m) {
        const CartesianArgs a = makeCartesianArgs(t->cartesian, t->cfg.waveFactor);
        const unsigned nb = unsigned((cells + 255) / 256);
        // CFL on the step-start state
        mark(m, 0);
        launchCartesianCfl(a, ns, qA, all, pA, cA, cfl, maximum, nullptr, blockMin, stream);
        hipLaunchKernelGGL(finishCflKernel, dim3(1), dim3(256), 0, stream, blockMin, size_t(nb), scalars);
        mark(m, 1);
        // heavy-first orders from the latest costs (or the step-start seeds)
        sortCells(haveCost ? cost : nullptr, order);
        if (mode == REACTIVE_AMD_FLASH_MIXED) sortCells(haveCostSearch ? costSearch : nullptr, orderSearch);
        // stage 0 and its boundary rate
        t->boundaryRatesFrom(qA, pA, cA, rA, bA, counters + 5, stream);
        mark(m, 2);
        launchCartesianStage(a, ns, nv, t->scheme, qA, all, pA, cA, rA, 0.0, scalars, 0, qB, counters + 5, stream);
        mark(m, 3);
        // flash #1 seeded from the step-start states
        flash(qB, pB, cB, rB, sB, nullptr, counters, m ? m[9] : nullptr);
        mark(m, 4);
        // CFL recheck after stage 0 (maximum = dt)
        launchCartesianCfl(a, ns, qB, all, pB, cB, cfl, 0.0, scalars, blockMin, stream);
        hipLaunchKernelGGL(finishCflKernel, dim3(1), dim3(256), 0, stream, blockMin, size_t(nb), scalars + 1);
        mark(m, 5);
        hipLaunchKernelGGL(recheckKernel, dim3(1), dim3(64), 0, stream, scalars, counters);
        // stage 1 (in place into qA, which holds q0) and its boundary rate
        t->boundaryRatesFrom(qB, pB, cB, rB, bB, counters + 5, stream);
        mark(m, 6);
        launchCartesianStage(a, ns, nv, t->scheme, qB, all, pB, cB, rB, 0.0, scalars, 1, qA, counters + 5, stream);
        mark(m, 7);
        // flash #2 from the step-start seeds, into the state buffer of stage 1
        flash(qA, pA, cA, rA, sB, cost, counters + 2, m ? m[10] : nullptr);
        haveCost = true;
        if (mode == REACTIVE_AMD_FLASH_MIXED) haveCostSearch = true;
        mark(m, 8);
        std::swap(sA, sB);
        parity ^= 1;
        // conservation totals and boundary integral
        hipLaunchKernelGGL(totalsPartialKernel, dim3(kTotalBlocks), dim3(256), 0, stream, qA, all, cells, nv,
                           t->cartesian.volume, totalPartial);
        hipLaunchKernelGGL(totalsFinishKernel, dim3(1), dim3(64), 0, stream, totalPartial, kTotalBlocks, nv, totals);
        hipLaunchKernelGGL(boundaryIntegralKernel, dim3(1), dim3(64), 0, stream, bA, bB, scalars, nv, boundary);
        RAMD_HIP(hipGetLastError());
    }
};

// One step from a captured graph.  The first step after a load runs on the
// stream (it records the costs that order the flashes); later steps replay the
// graph captured for the current buffer parity, recaptured when the flash
// arithmetic, flux scheme, CFL number or maximum step changes.  The step
// launches only kernels (no memory-set or copy nodes).
void graphStep(Step& s, double cfl, double maximum) {
    if (!s.haveCost || (s.mode == REACTIVE_AMD_FLASH_MIXED && !s.haveCostSearch)) {
        s.one(cfl, maximum, nullptr);
        return;
    }
    const int p = s.parity;
    if (!s.graph[p] || s.graphMode[p] != s.mode || s.graphScheme[p] != s.t->scheme || s.graphCfl[p] != cfl ||
        s.graphMax[p] != maximum) {
        if (s.graph[p]) {
            RAMD_HIP(hipGraphExecDestroy(s.graph[p]));
            s.graph[p] = nullptr;
        }
        hipGraph_t g = nullptr;
        RAMD_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        s.one(cfl, maximum, nullptr);  // recorded, not run: undo its host-side buffer swap
        std::swap(s.sA, s.sB);
        s.parity ^= 1;
        RAMD_HIP(hipStreamEndCapture(s.stream, &g));
        const hipError_t st = hipGraphInstantiate(&s.graph[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RAMD_HIP(st);
        s.graphMode[p] = s.mode;
        s.graphScheme[p] = s.t->scheme;
        s.graphCfl[p] = cfl;
        s.graphMax[p] = maximum;
    }
    RAMD_HIP(hipGraphLaunch(s.graph[p], s.stream));
    std::swap(s.sA, s.sB);
    s.parity ^= 1;
}

template <class F>
int guard(void* handle, F f) {
    if (!handle) return -1;
    Step& s = *static_cast<Step*>(handle);
    try {
        RAMD_HIP(hipSetDevice(s.device));
        f(s);
        s.error[0] = 0;
        return 0;
    } catch (const std::exception& e) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        std::snprintf(s.error, sizeof(s.error), "%s", e.what());
        return 1;
    }
}

}  // namespace ramd

using namespace ramd;

extern "C" {

void* reactive_amd_step_create(void* transport, void* hem, char* error, size_t errorSize) {
    try {
        require(transport && hem, "Null transport or HEM handle");
        std::unique_ptr<Step> s(new Step);
        s->t = static_cast<Transport*>(transport);
        s->h = static_cast<HemHandle*>(hem);
        require(s->t->cartesian.valid, "The resident step needs a structured-box transport mesh");
        require(s->t->device == s->h->device, "Transport and HEM handles are on different devices");
        require(size_t(s->h->ns) == s->t->cfg.species && s->t->cfg.variables == s->t->cfg.species + 4,
                "Transport and HEM species differ, or the transport carries liquid inventories");
        require(s->t->cfg.variables <= REACTIVE_AMD_MAX_VARIABLES, "Too many conserved variables");
        s->device = s->t->device;
        RAMD_HIP(hipSetDevice(s->device));
        s->ns = int(s->t->cfg.species);
        s->nv = int(s->t->cfg.variables);
        s->cells = s->t->cfg.cells;
        s->all = s->t->all;
        require(s->cells <= UINT32_MAX, "Too many cells");
        RAMD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (auto& e : s->events) RAMD_HIP(hipEventCreate(&e));
        s->qA = s->alloc<double>(s->all * s->nv);
        s->qB = s->alloc<double>(s->all * s->nv);
        s->pA = s->alloc<double>(s->all);
        s->cA = s->alloc<double>(s->all);
        s->pB = s->alloc<double>(s->all);
        s->cB = s->alloc<double>(s->all);
        s->rA = s->alloc<double>(s->all);
        s->rB = s->alloc<double>(s->all);
        s->sA = s->alloc<ThermoOut>(s->cells);
        s->sB = s->alloc<ThermoOut>(s->cells);
        s->key = s->alloc<uint8_t>(s->cells);
        s->order = s->alloc<uint32_t>(s->cells);
        s->cost = s->alloc<uint16_t>(s->cells);
        s->orderSearch = s->alloc<uint32_t>(s->cells);
        s->costSearch = s->alloc<uint16_t>(s->cells);
        s->bucketCount = s->alloc<unsigned>(256);
        s->bucketOffset = s->alloc<unsigned>(256);
        s->blockMin = s->alloc<double>((s->cells + 255) / 256 + 1);
        s->scalars = s->alloc<double>(4);
        s->bA = s->alloc<double>(s->nv);
        s->bB = s->alloc<double>(s->nv);
        s->boundary = s->alloc<double>(s->nv);
        s->totals = s->alloc<double>(s->nv);
        s->totalPartial = s->alloc<double2>(size_t(s->nv) * kTotalBlocks);
        s->counters = s->alloc<unsigned long long>(8);
        RAMD_HIP(hipMemsetAsync(s->bucketCount, 0, 256 * sizeof(unsigned), s->stream));
        RAMD_HIP(hipStreamSynchronize(s->stream));
        return s.release();
    } catch (const std::exception& e) {
        writeError(error, errorSize, e.what());
        return nullptr;
    }
}

int reactive_amd_step_load(void* step, const double* q, const ReactiveThermoState* states) {
    return guard(step, [&](Step& s) {
        require(q && states, "Null step input");
        std::vector<double> soa(s.all * s.nv, 0.0);
        for (size_t c = 0; c < s.cells; ++c)
            for (int k = 0; k < s.nv; ++k) soa[size_t(k) * s.all + c] = q[c * s.nv + k];
        RAMD_HIP(hipMemcpyAsync(s.qA, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, s.stream));
        ReactiveThermoState* staging = nullptr;
        RAMD_HIP(hipMalloc(&staging, s.cells * sizeof(ReactiveThermoState)));
        RAMD_HIP(hipMemcpyAsync(staging, states, s.cells * sizeof(ReactiveThermoState), hipMemcpyHostToDevice,
                                s.stream));
        hipLaunchKernelGGL(toThermoKernel, dim3(blocksFor(s.cells, 256)), dim3(256), 0, s.stream, staging, s.cells,
                           s.sA, s.pA, s.cA, s.rA);
        RAMD_HIP(hipGetLastError());
        RAMD_HIP(hipStreamSynchronize(s.stream));
        RAMD_HIP(hipFree(staging));
        s.loaded = true;
        s.haveCost = false;  // new states: schedule by the reference bucket until a flash has run
        s.haveCostSearch = false;
    });
}

int reactive_amd_step_run(void* step, double cfl, double maximum, int steps, int flashMode,
                          ReactiveAmdStepReport* report) {
    return guard(step, [&](Step& s) {
        require(s.loaded, "reactive_amd_step_load must precede reactive_amd_step_run");
        require(report && report->abiVersion == 1 && report->structBytes == sizeof(*report), "Invalid step report ABI");
        require(std::isfinite(cfl) && cfl > 0 && cfl <= .5 && std::isfinite(maximum) && maximum > 0,
                "Invalid transport CFL controls");
        require(steps >= 1, "At least one step");
        const bool useGraph = (flashMode & REACTIVE_AMD_STEP_GRAPH) != 0;
        flashMode &= ~REACTIVE_AMD_STEP_GRAPH;
        require(flashMode == REACTIVE_AMD_FLASH_FP64 || flashMode == REACTIVE_AMD_FLASH_MIXED,
                "Unknown flash arithmetic mode");
        if (flashMode == REACTIVE_AMD_FLASH_MIXED && !s.sets) s.sets = s.alloc<CandidateSet>(s.cells);
        s.mode = flashMode;
        hipLaunchKernelGGL(clearCountersKernel, dim3(1), dim3(64), 0, s.stream, s.counters, 8);
        while (s.marks.size() < size_t(steps) * kMarks) {
            hipEvent_t e;
            RAMD_HIP(hipEventCreate(&e));
            s.marks.push_back(e);
        }
        RAMD_HIP(hipEventRecord(s.events[0], s.stream));
        for (int i = 0; i < steps; ++i) {
            if (useGraph) graphStep(s, cfl, maximum);
            else s.one(cfl, maximum, s.marks.data() + size_t(i) * kMarks);
        }
        RAMD_HIP(hipEventRecord(s.events[1], s.stream));
        unsigned long long counters[8];
        double scalars[4], totals[REACTIVE_AMD_MAX_VARIABLES], boundary[REACTIVE_AMD_MAX_VARIABLES];
        RAMD_HIP(hipMemcpyAsync(counters, s.counters, sizeof(counters), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(scalars, s.scalars, sizeof(scalars), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(totals, s.totals, s.nv * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(boundary, s.boundary, s.nv * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        float ms = 0;
        RAMD_HIP(hipEventElapsedTime(&ms, s.events[0], s.events[1]));
        ReactiveAmdStepReport r{};
        r.abiVersion = 1;
        r.structBytes = sizeof(r);
        r.steps = uint64_t(steps);
        r.flashFailures[0] = counters[0];
        r.flashFailures[1] = counters[2];
        r.phaseEvaluations = counters[1] + counters[3];
        r.cflViolations = counters[4];
        r.fluxFailures = counters[5];
        r.singlePhaseEvaluations = counters[6];
        r.dt = scalars[0];
        r.recheckDt = scalars[1];
        for (int k = 0; k < s.nv; ++k) {
            r.totals[k] = totals[k];
            r.boundaryIntegral[k] = boundary[k];
        }
        r.seconds = ms / 1000.0;
        for (int i = 0; i < steps && !useGraph; ++i) {
            const hipEvent_t* m = s.marks.data() + size_t(i) * kMarks;
            float d[8];
            for (int j = 0; j < 8; ++j) RAMD_HIP(hipEventElapsedTime(&d[j], m[j], m[j + 1]));
            if (s.mode == REACTIVE_AMD_FLASH_MIXED) {
                float a = 0, b = 0;
                RAMD_HIP(hipEventElapsedTime(&a, m[3], m[9]));
                RAMD_HIP(hipEventElapsedTime(&b, m[7], m[10]));
                r.searchSeconds += (a + b) / 1000.0;
            }
            r.cflSeconds += (d[0] + d[4]) / 1000.0;
            r.stageSeconds[0] += d[2] / 1000.0;
            r.flashSeconds[0] += d[3] / 1000.0;
            r.stageSeconds[1] += d[6] / 1000.0;
            r.flashSeconds[1] += d[7] / 1000.0;
        }
        *report = r;
        require(r.dt > 0, "Invalid wave/diffusion timestep");
    });
}

int reactive_amd_step_read(void* step, double* q, ReactiveThermoState* states) {
    return guard(step, [&](Step& s) {
        require(s.loaded, "Nothing loaded");
        std::vector<double> soa(s.all * s.nv);
        RAMD_HIP(hipMemcpyAsync(soa.data(), s.qA, soa.size() * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        if (states)
            RAMD_HIP(hipMemcpyAsync(states, s.sA, s.cells * sizeof(ThermoOut), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        if (q)
            for (size_t c = 0; c < s.cells; ++c)
                for (int k = 0; k < s.nv; ++k) q[c * s.nv + k] = soa[size_t(k) * s.all + c];
    });
}

const char* reactive_amd_step_error(void* step) { return step ? static_cast<Step*>(step)->error : "Null step"; }

void reactive_amd_step_destroy(void* step) { delete static_cast<Step*>(step); }

}
