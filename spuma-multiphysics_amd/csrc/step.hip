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
#include <cstdio>
#include <cstring>
#include <functional>
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
// Plain FP64 flash of the cells in `order`; RJ: the model's
// analytic-Jacobian mode.
template <int NS, int NL, bool RJ>
__global__ __launch_bounds__(256, RAMD_FLASH_MIN_WAVES) void stepFlashKernel(const HemModel<NS>* __restrict__ globalModel,
                                                      const double* __restrict__ q, size_t stride,
                                                      const ThermoOut* __restrict__ seeds,
                                                      const uint32_t* __restrict__ order, uint32_t count,
                                                      const double* __restrict__ dtDev, ThermoOut* __restrict__ out,
                                                      double* __restrict__ pOut, double* __restrict__ cOut,
                                                      double* __restrict__ rOut, uint16_t* __restrict__ cost,
                                                      unsigned long long* __restrict__ counters,
                                                      uint8_t* __restrict__ failMark,
                                                      const uint32_t* __restrict__ rep) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0};
    uint32_t failed = 0;
    if (i < count && *dtDev > 0 && !(rep && rep[order[i]] != order[i])) {
        const uint32_t c = order[i];
        const ThermoOut g = seeds[c];
        using F = Flash<NS, NL, double, false, RJ, false, kModelLds>;
        F f(model, n);
        loadCell(f, q, stride, g, c, NS);
        typename F::E result;
        zeroEval(result);
        const bool ok = f.run(result);
        storeCell(ok, result, g, c, out, pOut, cOut, rOut);
        if (!ok && failMark) failMark[c] = 1;
        failed = ok ? 0u : 1u;
        if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
    }
    countFlash(failed, n.phases, counters);
}

// Mixed precision, second half: FP64 refinement, stability and selection.
template <int NS, int NL>
__global__ __launch_bounds__(256, RAMD_REFINE_MIN_WAVES) void stepRefineKernel(const HemModel<NS>* __restrict__ globalModel,
                                                       const double* __restrict__ q, size_t stride,
                                                       const ThermoOut* __restrict__ seeds,
                                                       const uint32_t* __restrict__ order, uint32_t count,
                                                       const double* __restrict__ dtDev,
                                                       const CandidateSet* __restrict__ sets,
                                                       ThermoOut* __restrict__ out, double* __restrict__ pOut,
                                                       double* __restrict__ cOut, double* __restrict__ rOut,
                                                       uint16_t* __restrict__ cost,
                                                       unsigned long long* __restrict__ counters,
                                                       uint32_t* __restrict__ queue, unsigned* __restrict__ queueCount,
                                                       uint8_t* __restrict__ failMark,
                                                       const uint32_t* __restrict__ rep) {
    const HemModel<NS>& model = workgroupModel(globalModel);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts n = {0, 0, 0, 0, 0, 0};
    uint32_t failed = 0;
    if (i < count && *dtDev > 0 && !(rep && rep[order[i]] != order[i])) {
        const uint32_t c = order[i];
        const ThermoOut g = seeds[c];
        Flash<NS, NL, double, true, false, false, kModelLds> f(model, n);
        loadCell(f, q, stride, g, c, NS);
        EvalT<double, true> result;
        zeroEval(result);
        bool deferred = false;
        const bool ok = f.refine(sets[c], result, queue ? &deferred : nullptr);
        if (deferred) {
            queue[atomicAdd(queueCount, 1u)] = c;
        } else {
            storeCell(ok, result, g, c, out, pOut, cOut, rOut);
            if (!ok && failMark) failMark[c] = 1;
            failed = ok ? 0u : 1u;
            if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
        }
    }
    countFlash(failed, n.phases, counters);
}

// One deferred cell refined without deferral; lead: this lane stores it.
template <int NS, int NL, bool WAVE>
__device__ inline void refineQueued(const HemModel<NS>& model, const double* __restrict__ q, size_t stride,
                                    const ThermoOut* __restrict__ seeds, uint32_t c, const CandidateSet& set,
                                    bool lead, ThermoOut* __restrict__ out, double* __restrict__ pOut,
                                    double* __restrict__ cOut, double* __restrict__ rOut, uint16_t* __restrict__ cost,
                                    uint8_t* __restrict__ failMark, Counts& n, uint32_t& failed) {
    const ThermoOut g = seeds[c];
    Flash<NS, NL, double, true, false, false, kModelLds, WAVE> f(model, n);
    loadCell(f, q, stride, g, c, NS);
    EvalT<double, true> result;
    zeroEval(result);
    const bool ok = f.refine(set, result);
    if (lead) {
        storeCell(ok, result, g, c, out, pOut, cOut, rOut);
        if (!ok && failMark) failMark[c] = 1;
        if (cost) cost[c] = uint16_t(n.phases < 65535u ? n.phases : 65535u);
        failed = ok ? 0u : 1u;
    }
}

// The cells the refinement deferred (full FP64 search, FP64 restarts, or a
// frozen candidate's solve from the cell's guess), refined again without
// deferral by a persistent grid of single-wave workgroups: one queued cell
// per wave at a time (every lane on the cell, its line searches spread over
// the lanes) while the queue is short, so that the few long refinements
// neither share a wave nor halve their steps one evaluation at a time; one
// cell per lane for a long queue.  queueCount is read on the device (the
// launch does not depend on it).
template <int NS, int NL>
__global__ __launch_bounds__(64, RAMD_REFINE_MIN_WAVES) void stepRefineQueueKernel(
    const HemModel<NS>* __restrict__ globalModel, const double* __restrict__ q, size_t stride,
    const ThermoOut* __restrict__ seeds, const uint32_t* __restrict__ queue, const unsigned* __restrict__ queueCount,
    const double* __restrict__ dtDev, const CandidateSet* __restrict__ sets, ThermoOut* __restrict__ out,
    double* __restrict__ pOut, double* __restrict__ cOut, double* __restrict__ rOut, uint16_t* __restrict__ cost,
    unsigned long long* __restrict__ counters, uint8_t* __restrict__ failMark) {
    const unsigned count = *queueCount;
    const bool perWave = count <= RAMD_WAVE_QUEUE_CELLS;  // uniform
    if (blockIdx.x * (perWave ? 1u : 64u) >= count) return;  // uniform per workgroup: usually every workgroup leaves here
    const HemModel<NS>& model = workgroupModel(globalModel);
    if (perWave) {
        const bool lead = threadIdx.x == 0;
        for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {  // uniform per workgroup
            Counts n = {0, 0, 0, 0, 0, 0, 0};
            uint32_t failed = 0;
            if (*dtDev > 0)
                refineQueued<NS, NL, true>(model, q, stride, seeds, queue[i], sets[queue[i]], lead, out, pOut, cOut,
                                           rOut, cost, failMark, n, failed);
            countFlash(failed, lead ? n.phases : 0u, counters);
        }
    } else {
        for (unsigned base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u) {  // uniform per workgroup
            const unsigned i = base + threadIdx.x;
            Counts n = {0, 0, 0, 0, 0, 0, 0};
            uint32_t failed = 0;
            if (i < count && *dtDev > 0)
                refineQueued<NS, NL, false>(model, q, stride, seeds, queue[i], sets[queue[i]], true, out, pOut, cOut,
                                            rOut, cost, failMark, n, failed);
            countFlash(failed, n.phases, counters);
        }
    }
}
__global__ void clearQueueKernel(unsigned* __restrict__ count) {
    if (threadIdx.x == 0) *count = 0;
}
__global__ void queueTotalKernel(const unsigned* __restrict__ count, unsigned long long* __restrict__ total) {
    if (threadIdx.x == 0) atomicAdd(total, (unsigned long long)*count);
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

// ---- thermodynamic-state order ---------------------------------------------
//
// Cells whose (ln rho, e) fall in the same cell of a side x side grid
// spanning the flash input's range are flashed by the same waves: their searches take
// the same branches with similar iteration counts, so the lanes of a wave
// stay converged (cost order alone groups cells by total work, whose paths
// still differ).  The order decides nothing numerical.
#ifndef RAMD_STATE_ORDER
#define RAMD_STATE_ORDER 1
#endif
// side: bins per axis, 32, 64 (default) or 128, chosen by the autotuner
constexpr int kStateMaxBins = 128 * 128, kStateBlocks = 1024;

__device__ inline unsigned long long orderedBits(double x) {
    const unsigned long long u = __double_as_longlong(x);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ inline double fromOrdered(unsigned long long u) {
    return __longlong_as_double((u >> 63) ? (u & ~(1ull << 63)) : ~u);
}
// ln rho and specific internal energy of cell c (non-finite values map to the range ends)
__device__ inline void cellState(const double* __restrict__ q, size_t stride, int ns, size_t c, double& lr,
                                 double& e) {
    double rho = 0;
    for (int k = 0; k < ns; ++k) rho += q[size_t(k) * stride + c];
    const double m0 = q[size_t(ns) * stride + c], m1 = q[size_t(ns + 1) * stride + c];
    const double m2 = q[size_t(ns + 2) * stride + c], E = q[size_t(ns + 3) * stride + c];
    lr = rho > 0 ? ::log(rho) : -1e300;
    e = (E - .5 * (m0 * m0 + m1 * m1 + m2 * m2) / rho) / rho;
    if (!isfinite(lr)) lr = -1e300;
    if (!isfinite(e)) e = -1e300;
}
// range[0..3] = ordered min ln rho, max ln rho, min e, max e
__global__ void stateRangeInitKernel(unsigned long long* __restrict__ range) {
    if (threadIdx.x < 4) range[threadIdx.x] = (threadIdx.x & 1) ? 0ull : ~0ull;
}
__device__ inline unsigned long long waveMinU(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__global__ __launch_bounds__(256) void stateRangeKernel(const double* __restrict__ q, size_t stride, int ns,
                                                        size_t cells, unsigned long long* __restrict__ range) {
    // min of (lr, -lr, e, -e) in ordered bits: entries 1 and 3 hold ~max
    unsigned long long v[4] = {~0ull, ~0ull, ~0ull, ~0ull};
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells; c += size_t(gridDim.x) * blockDim.x) {
        double lr, e;
        cellState(q, stride, ns, c, lr, e);
        if (lr > -1e300 && e > -1e300) {
            const unsigned long long a = orderedBits(lr), b = orderedBits(e);
            v[0] = a < v[0] ? a : v[0];
            v[1] = ~a < v[1] ? ~a : v[1];
            v[2] = b < v[2] ? b : v[2];
            v[3] = ~b < v[3] ? ~b : v[3];
        }
    }
    __shared__ unsigned long long lane[4][4];
    for (int j = 0; j < 4; ++j) {
        v[j] = waveMinU(v[j]);
        if ((threadIdx.x & 63) == 0) lane[j][threadIdx.x >> 6] = v[j];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int j = threadIdx.x;
        unsigned long long m = lane[j][0];
        for (int w = 1; w < 4; ++w) m = lane[j][w] < m ? lane[j][w] : m;
        if (j & 1) atomicMax(&range[j], ~m);
        else atomicMin(&range[j], m);
    }
}
__device__ inline unsigned stateBin(double x, double lo, double hi, int side) {
    if (!(hi > lo)) return 0;
    const double t = (x - lo) / (hi - lo) * side;
    return t <= 0 ? 0u : (t >= side - 1 ? unsigned(side - 1) : unsigned(t));
}
// bin of every cell (key) and the bin counts; grid-stride over kStateBlocks blocks
// With a cost array (levels > 1), each state bin is split by the cell's
// phase-evaluation count in its previous flash: level = min(levels - 1,
// floor(log2(cost / 32))), so cells of one bin with very different work do
// not share waves.
__global__ __launch_bounds__(256) void stateKeyKernel(const double* __restrict__ q, size_t stride, int ns,
                                                      size_t cells, const unsigned long long* __restrict__ range,
                                                      int side, const uint16_t* __restrict__ cost, int levels,
                                                      uint16_t* __restrict__ key, unsigned* __restrict__ count) {
    __shared__ unsigned hist[kStateMaxBins];
    const int bins = side * side * levels;
    for (int b = threadIdx.x; b < bins; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    const double lo0 = fromOrdered(range[0]), hi0 = fromOrdered(range[1]);
    const double lo1 = fromOrdered(range[2]), hi1 = fromOrdered(range[3]);
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells; c += size_t(gridDim.x) * blockDim.x) {
        double lr, e;
        cellState(q, stride, ns, c, lr, e);
        unsigned k = stateBin(lr, lo0, hi0, side) * side + stateBin(e, lo1, hi1, side);
        if (levels > 1) {
            unsigned level = 0;
            for (unsigned w = cost[c] >> 5; w > 1 && level + 1 < unsigned(levels); w >>= 1) ++level;
            k = k * unsigned(levels) + level;
        }
        key[c] = uint16_t(k);
        atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += blockDim.x)
        if (hist[b]) atomicAdd(&count[b], hist[b]);
}
// exclusive scan of the bin counts into offset; clears count (one block of 256)
__global__ __launch_bounds__(256) void stateScanKernel(unsigned* __restrict__ count, unsigned* __restrict__ offset,
                                                       int bins) {
    const int per = (bins + 255) / 256;
    __shared__ unsigned part[256];
    unsigned sum = 0;
    for (int j = 0; j < per; ++j)
        if (threadIdx.x * per + j < bins) sum += count[threadIdx.x * per + j];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int t = 0; t < 256; ++t) {
            const unsigned v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    unsigned run = part[threadIdx.x];
    for (int j = 0; j < per; ++j) {
        const int b = threadIdx.x * per + j;
        if (b >= bins) break;
        offset[b] = run;
        run += count[b];
        count[b] = 0;
    }
}
// each block reserves one run per bin for the cells of its grid-stride set
__global__ __launch_bounds__(256) void stateScatterKernel(const uint16_t* __restrict__ key, size_t cells, int bins,
                                                          unsigned* __restrict__ offset,
                                                          uint32_t* __restrict__ order) {
    __shared__ unsigned local[kStateMaxBins];
    for (int b = threadIdx.x; b < bins; b += blockDim.x) local[b] = 0;
    __syncthreads();
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells; c += size_t(gridDim.x) * blockDim.x)
        atomicAdd(&local[key[c]], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += blockDim.x)
        if (local[b]) local[b] = atomicAdd(&offset[b], local[b]);
    __syncthreads();
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < cells; c += size_t(gridDim.x) * blockDim.x)
        order[atomicAdd(&local[key[c]], 1u)] = uint32_t(c);
}

// Neumaier-compensated sum of q_k V over the cells: per-block partials in a
// fixed order, then one block merges them in block order (deterministic).
__device__ inline void twoSumAdd(double& s, double& comp, double x) {
    const double t = s + x;
    comp += ::fabs(s) >= ::fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}
constexpr int kTotalBlocks = 512;
// Partial sums of block `block` of kTotalBlocks (each box of a batch is summed
// by its own kTotalBlocks blocks, as when it runs alone).
__device__ inline void totalsPartial(const double* __restrict__ q, size_t stride, size_t cells, int nv, double volume,
                                     unsigned block, double2* __restrict__ partial) {
    __shared__ double ls[256], lc[256];
    for (int k = 0; k < nv; ++k) {
        double s = 0, comp = 0;
        for (size_t c = size_t(block) * blockDim.x + threadIdx.x; c < cells; c += size_t(kTotalBlocks) * blockDim.x)
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
        if (threadIdx.x == 0) partial[size_t(k) * kTotalBlocks + block] = make_double2(ls[0], lc[0]);
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void totalsPartialKernel(const double* __restrict__ q, size_t stride, size_t cells,
                                                           int nv, double volume, double2* __restrict__ partial) {
    totalsPartial(q, stride, cells, nv, volume, blockIdx.x, partial);
}
// grid (kTotalBlocks, boxes): box blockIdx.y into partial + nv kTotalBlocks blockIdx.y
__global__ __launch_bounds__(256) void batchTotalsPartialKernel(const BoxDev* __restrict__ boxes,
                                                                const double* __restrict__ q, size_t stride, int nv,
                                                                double2* __restrict__ partial) {
    const BoxDev& box = boxes[blockIdx.y];
    totalsPartial(q + box.offset, stride, box.cells, nv, box.volume, blockIdx.x,
                  partial + size_t(nv) * kTotalBlocks * blockIdx.y);
}
// one block per box (thread k: variable k)
__global__ void totalsFinishKernel(const double2* __restrict__ partial, int blocks, int nv, double* __restrict__ out) {
    const int k = threadIdx.x;
    if (k >= nv) return;
    partial += size_t(nv) * blocks * blockIdx.x;
    double s = 0, comp = 0;
    for (int b = 0; b < blocks; ++b) {
        comp += partial[size_t(k) * blocks + b].y;
        twoSumAdd(s, comp, partial[size_t(k) * blocks + b].x);
    }
    out[size_t(nv) * blockIdx.x + k] = s + comp;
}

// The CFL minimum of box blockIdx.x over its blocks [start[i], start[i + 1]):
// scalars[4 i + slot] (-1 if invalid), at most limit[i] when limit is given
// (a repeated attempt's halved dt).
__global__ void batchFinishCflKernel(const double* __restrict__ blockMin, const long long* __restrict__ start,
                                     double* __restrict__ scalars, int slot, const double* __restrict__ limit) {
    const int i = blockIdx.x;
    double v = __builtin_huge_val();
    for (long long b = start[i] + threadIdx.x; b < start[i + 1]; b += blockDim.x) v = blockMin[b] < v ? blockMin[b] : v;
    __shared__ double lane[4];
    v = waveMin(v);
    if ((threadIdx.x & 63) == 0) lane[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lane[0];
        for (int w = 1; w < int(blockDim.x >> 6); ++w) m = lane[w] < m ? lane[w] : m;
        if (limit && m > 0 && limit[i] < m) m = limit[i];
        scalars[4 * i + slot] = isfinite(m) && m > 0 ? m : -1;
    }
}
// The flash gate of a batch: the smallest dt of its boxes (-1 if any is invalid).
__global__ void batchGateKernel(const double* __restrict__ scalars, int count, double* __restrict__ gate) {
    if (threadIdx.x) return;
    double g = __builtin_huge_val();
    for (int i = 0; i < count; ++i) {
        const double dt = scalars[4 * i];
        g = dt > 0 && dt < g ? dt : (dt > 0 ? g : -1.0);
        if (g < 0) break;
    }
    *gate = g;
}
// thread i: box i (violations may be null)
__global__ void recheckKernel(double* __restrict__ scalars, int count, unsigned long long* __restrict__ counters,
                              unsigned long long* __restrict__ violations) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= count) return;
    const double dt = scalars[4 * i], stable = scalars[4 * i + 1];
    if (!(dt > 0) || !(dt <= stable * (1 + 1e-10))) {
        atomicAdd(&counters[4], 1ull);
        if (violations) atomicAdd(&violations[i], 1ull);
    }
}
// Cells of box blockIdx.y whose flash failed in the attempt (failMark) into boxFail[box].
__global__ __launch_bounds__(256) void boxFailKernel(const BoxDev* __restrict__ boxes,
                                                     const uint8_t* __restrict__ failMark,
                                                     unsigned long long* __restrict__ boxFail) {
    const BoxDev& b = boxes[blockIdx.y];
    uint32_t n = 0;
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < b.cells; c += size_t(gridDim.x) * blockDim.x)
        n += failMark[b.offset + c];
    n = waveSum(n);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&boxFail[blockIdx.y], (unsigned long long)n);
}
// thread (i, k) = (blockIdx.x, threadIdx.x)
__global__ void boundaryIntegralKernel(const double* __restrict__ bA, const double* __restrict__ bB,
                                      const double* __restrict__ scalars, int nv, double* __restrict__ out) {
    const int k = threadIdx.x;
    const size_t j = size_t(nv) * blockIdx.x + k;
    if (k < nv) out[j] = .5 * scalars[4 * blockIdx.x] * (bA[j] + bB[j]);
}
__global__ void setScalarKernel(double* __restrict__ dst, double value) {
    if (threadIdx.x == 0) *dst = value;
}
// ---- exact reuse (REACTIVE_AMD_STEP_REUSE): cells whose flash inputs are
// bitwise equal (species masses, e = rhoE - |rho u|^2 / (2 rho), the 176-byte
// seed state) share one flash, the one of the lowest such cell index, as the
// reference's pool does within a batch (reactiveClosurePool.inc:3-54).  An
// open-addressing table of cell indices (twice the cells, a power of two):
// a slot goes from empty to a cell and then only to a lower cell with the
// same key, so every key ends at its lowest cell whatever the insert order.
constexpr uint32_t kEmptySlot = 0xffffffffu;
__global__ void reuseClearKernel(uint32_t* __restrict__ table, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        table[i] = kEmptySlot;
}
// word w of cell c's flash input key (ns masses, e, the seed's 22 words)
__device__ inline uint64_t reuseWord(const double* __restrict__ q, size_t stride, int ns,
                                     const ThermoOut* __restrict__ seeds, size_t c, int w) {
    if (w < ns) return uint64_t(__double_as_longlong(q[size_t(w) * stride + c]));
    if (w == ns) {
        double rho = 0;
        for (int k = 0; k < ns; ++k) rho += q[size_t(k) * stride + c];
        const double m0 = q[size_t(ns) * stride + c], m1 = q[size_t(ns + 1) * stride + c];
        const double m2 = q[size_t(ns + 2) * stride + c], E = q[size_t(ns + 3) * stride + c];
        return uint64_t(__double_as_longlong(E - .5 * (m0 * m0 + m1 * m1 + m2 * m2) / rho));
    }
    return reinterpret_cast<const uint64_t*>(seeds + c)[w - ns - 1];
}
constexpr int kSeedWords = int(sizeof(ThermoOut) / 8);
__global__ __launch_bounds__(256) void reuseInsertKernel(const double* __restrict__ q, size_t stride, int ns,
                                                         const ThermoOut* __restrict__ seeds, size_t cells,
                                                         uint32_t* __restrict__ table, uint32_t mask,
                                                         uint32_t* __restrict__ slotOf) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int words = ns + 1 + kSeedWords;
    uint64_t h = 0xcbf29ce484222325ull;  // FNV-1a offset basis over 64-bit words, then a 64-bit finaliser
    for (int w = 0; w < words; ++w) {
        h ^= reuseWord(q, stride, ns, seeds, c, w);
        h *= 0x100000001b3ull;
    }
    h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
    uint32_t slot = uint32_t(h) & mask;
    for (;;) {  // the table holds at least twice the cells: an empty slot is always reached
        uint32_t other = table[slot];
        if (other == kEmptySlot) {
            other = atomicCAS(&table[slot], kEmptySlot, uint32_t(c));
            if (other == kEmptySlot) break;
        }
        bool same = true;
        for (int w = 0; w < words && same; ++w)
            same = reuseWord(q, stride, ns, seeds, c, w) == reuseWord(q, stride, ns, seeds, other, w);
        if (same) {
            atomicMin(&table[slot], uint32_t(c));
            break;
        }
        slot = (slot + 1) & mask;
    }
    slotOf[c] = slot;
}
__global__ void reuseResolveKernel(const uint32_t* __restrict__ table, const uint32_t* __restrict__ slotOf,
                                   size_t cells, uint32_t* __restrict__ rep) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < cells) rep[c] = table[slotOf[c]];
}
// after the flash: every other cell takes its representative's results
__global__ __launch_bounds__(256) void reuseCopyKernel(const uint32_t* __restrict__ rep, size_t cells,
                                                       const double* __restrict__ dtDev, ThermoOut* __restrict__ out,
                                                       double* __restrict__ pOut, double* __restrict__ cOut,
                                                       double* __restrict__ rOut, uint16_t* __restrict__ cost,
                                                       uint8_t* __restrict__ failMark,
                                                       unsigned long long* __restrict__ reused) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t copied = 0;
    if (c < cells && *dtDev > 0) {
        const uint32_t r = rep[c];
        if (r != uint32_t(c)) {
            out[c] = out[r];
            pOut[c] = pOut[r];
            cOut[c] = cOut[r];
            rOut[c] = rOut[r];
            if (cost) cost[c] = cost[r];
            if (failMark) failMark[c] = failMark[r];
            copied = 1;
        }
    }
    const uint32_t sum = waveSum(copied);
    if ((threadIdx.x & 63) == 0 && sum) atomicAdd(reused, (unsigned long long)sum);
}

__global__ void clearCountersKernel(unsigned long long* __restrict__ c, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) c[i] = 0;
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

// Fixed-state slots [cells, all) of the transport (reservoir q and its p,
// c_frozen, rho) into both step buffers; no stage or flash writes them.
__global__ void copyFixedKernel(const double* __restrict__ tq, const ReactiveTransportState* __restrict__ ts,
                                size_t cells, size_t all, int nv, double* __restrict__ qA, double* __restrict__ qB,
                                double* __restrict__ pA, double* __restrict__ cA, double* __restrict__ rA,
                                double* __restrict__ pB, double* __restrict__ cB, double* __restrict__ rB) {
    const size_t c = cells + size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= all) return;
    for (int k = 0; k < nv; ++k) qA[size_t(k) * all + c] = qB[size_t(k) * all + c] = tq[size_t(k) * all + c];
    pA[c] = pB[c] = ts[c].p;
    cA[c] = cB[c] = ts[c].sound;
    rA[c] = rB[c] = ts[c].rho;
}

inline unsigned blocksFor(size_t n, unsigned b) { return unsigned((n + b - 1) / b); }
// Deferred refinements go through a compacted queue (knob, compile time).
#ifndef RAMD_REFINE_QUEUE
#define RAMD_REFINE_QUEUE 1
#endif
constexpr unsigned kQueueBlocks = RAMD_QUEUE_BLOCKS;
constexpr int kMarks = 11;  // events per step

// ---------------------------------------------------------------- handle
struct Step {
    Transport* t = nullptr;  // the first box's transport (flux scheme, wave factor)
    HemHandle* h = nullptr;
    int ns = 0, nv = 0, device = 0;
    size_t cells = 0, all = 0;
    // boxes: one per independent case (a single-box step has one); box i
    // owns cells [offset, offset + cells) of every array
    int count = 1;
    std::vector<Transport*> ts;
    std::vector<BoxDev> boxHost;
    BoxDev* boxes = nullptr;
    long long *stageStart = nullptr, *cflStart = nullptr;
    long stageBlocks = 0, cflBlocks = 0;
    // launch shapes (the autotuner's choices): stage chunk depth, flash
    // workgroup size, state-order bins per axis; one stage table per depth
    int tileZ = kChunkZ, stateSide = 64;
    unsigned flashBlock = 64;
    int stateLevels = 1;                  // cost levels per state bin (1: state only)
    int stateLevelsOf[2] = {1, 1};
    int stateSideOf[2] = {64, 64};        // per flash arithmetic (FP64, mixed)
    unsigned flashBlockOf[2] = {64, 64};
    bool tunedOf[2] = {false, false};
    long long* stageStarts[3] = {nullptr, nullptr, nullptr};
    long stageBlocksZ[3] = {0, 0, 0};
    hipEvent_t tuneEvents[2] = {};
    unsigned long long* violations = nullptr;  // per box: failed CFL rechecks
    // attempts (REACTIVE_AMD_STEP_RETRY): the step-start q, p, c_frozen, r and
    // states, per-cell flash failure marks, per-box failed cells and dt limits
    double *qKeep = nullptr, *pKeep = nullptr, *cKeep = nullptr, *rKeep = nullptr;
    ThermoOut* sKeep = nullptr;
    uint8_t* failMark = nullptr;  // non-null while an attempt records its failed cells
    uint8_t* failMarks = nullptr;
    unsigned long long* boxFail = nullptr;
    double* limit = nullptr;      // per box, applied to the step's dt while non-null
    double* limits = nullptr;
    // exact reuse (REACTIVE_AMD_STEP_REUSE): hash table of cell indices, each
    // cell's slot and representative (rep non-null while reuse is on)
    uint32_t *reuseTable = nullptr, *reuseSlot = nullptr, *reuseRep = nullptr, *rep = nullptr;
    uint32_t reuseMask = 0;
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
    uint32_t* queue = nullptr;  // cells the refinement deferred, and their count
    unsigned* queueCount = nullptr;
    int mode = REACTIVE_AMD_FLASH_FP64;
    unsigned *bucketCount = nullptr, *bucketOffset = nullptr;
    // state order: bin keys, bin counts (kept zero between uses) and offsets, range
    uint16_t* stateKey = nullptr;
    unsigned *stateCount = nullptr, *stateOffset = nullptr;
    unsigned long long* stateRange = nullptr;
    double* blockMin = nullptr;
    double* scalars = nullptr;  // per box [4 i]: dt, recheck; [4 count]: flash gate of a batch
    double *bA = nullptr, *bB = nullptr, *boundary = nullptr, *totals = nullptr;  // nv per box
    double2* totalPartial = nullptr;
    // fail1, phases1, fail2, phases2, violations, HLLC rejections, FP32 phases, queued cells (mixed);
    // [8, 16): the autotuner's trial flashes; [16]: cells given a reused flash result ([24]: tuner's)
    unsigned long long* counters = nullptr;
    unsigned long long* flashCounters = nullptr;  // where flash() counts: counters, or counters + 8 while tuning
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
    const uint32_t* graphRep[2] = {nullptr, nullptr};
    int parity = 0;  // flips with every step (the seed and result buffers swap)
    uint64_t lastSteps = 0;
    bool loaded = false;
    char error[1024] = {};

    ~Step() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto e : events)
            if (e) (void)hipEventDestroy(e);
        for (auto e : tuneEvents)
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
        const unsigned blocks = blocksFor(cells, flashBlock);
        unsigned long long* counters = flashCounters;
#if RAMD_STATE_ORDER
        stateOrder(q, order);
        uint32_t* orderSearch = order;
#endif
        if (rep) {
            hipLaunchKernelGGL(reuseClearKernel, dim3(1024), dim3(256), 0, stream, reuseTable, size_t(reuseMask) + 1);
            hipLaunchKernelGGL(reuseInsertKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, stream, q, all, ns, sA,
                               cells, reuseTable, reuseMask, reuseSlot);
            hipLaunchKernelGGL(reuseResolveKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, stream, reuseTable,
                               reuseSlot, cells, rep);
        }
        if (mode == REACTIVE_AMD_FLASH_MIXED) {
#define RAMD_STEP_MIXED(NS_, NL_)                                                                                  \
    launchSearch(NS_, NL_, h->model, q, all, sA, orderSearch, uint32_t(cells), gate(), sets,                      \
                 costOut ? costSearch : nullptr, counters + 6, stream, flashBlock, rep);                           \
    if (split) RAMD_HIP(hipEventRecord(split, stream));                                                            \
    if (RAMD_REFINE_QUEUE) hipLaunchKernelGGL(clearQueueKernel, dim3(1), dim3(64), 0, stream, queueCount);          \
    hipLaunchKernelGGL((stepRefineKernel<NS_, NL_>), dim3(blocks), dim3(flashBlock), 0, stream,                     \
                       static_cast<const HemModel<NS_>*>(h->model), q, all, sA, order, uint32_t(cells), gate(),    \
                       sets, out, pOut, cOut, rOut, costOut, counter, RAMD_REFINE_QUEUE ? queue : nullptr,         \
                       queueCount, failMark, rep);                                                                 \
    if (RAMD_REFINE_QUEUE) {                                                                                       \
        hipLaunchKernelGGL((stepRefineQueueKernel<NS_, NL_>), dim3(kQueueBlocks), dim3(64), 0, stream,              \
                           static_cast<const HemModel<NS_>*>(h->model), q, all, sA, queue, queueCount, gate(),      \
                           sets, out, pOut, cOut, rOut, costOut, counter, failMark);                                \
        hipLaunchKernelGGL(queueTotalKernel, dim3(1), dim3(64), 0, stream, queueCount, counters + 7);             \
    }
            if (h->ns == 1 && h->nl == 0) { RAMD_STEP_MIXED(1, 0); }
            else if (h->ns == 1) { RAMD_STEP_MIXED(1, 1); }
            else if (h->nl == 0) { RAMD_STEP_MIXED(2, 0); }
            else { RAMD_STEP_MIXED(2, 1); }
#undef RAMD_STEP_MIXED
        } else {
#define RAMD_STEP_FLASH(NS_, NL_, RJ_)                                                                            \
    hipLaunchKernelGGL((stepFlashKernel<NS_, NL_, RJ_>), dim3(blocks), dim3(flashBlock), 0, stream,                \
                       static_cast<const HemModel<NS_>*>(h->model), q, all, sA, order, uint32_t(cells), gate(), out, \
                       pOut, cOut, rOut, costOut, counter, failMark, rep)
            if (h->ns == 1 && h->nl == 0) { RAMD_STEP_FLASH(1, 0, false); }
            else if (h->ns == 1 && !h->analytic) { RAMD_STEP_FLASH(1, 1, false); }
            else if (h->ns == 1) { RAMD_STEP_FLASH(1, 1, true); }
            else if (h->nl == 0) { RAMD_STEP_FLASH(2, 0, false); }
            else if (!h->analytic) { RAMD_STEP_FLASH(2, 1, false); }
            else { RAMD_STEP_FLASH(2, 1, true); }
#undef RAMD_STEP_FLASH
        }
        if (rep)
            hipLaunchKernelGGL(reuseCopyKernel, dim3(blocksFor(cells, 256)), dim3(256), 0, stream, rep, cells, gate(),
                               out, pOut, cOut, rOut, costOut, failMark, counters + 16);
        RAMD_HIP(hipGetLastError());
    }

    // The flash runs when every box has a valid dt.
    const double* gate() const { return count > 1 ? scalars + 4 * count : scalars; }

    // CFL sweep of q into scalars[4 i + slot]; maxDev: the maximum step is
    // each box's current dt (the recheck).
    void cflSweep(const double* q, const double* p, const double* c, double cfl, double maxHost, bool maxDev,
                  int slot) {
        (void)p;
        launchBatchCfl(ns, boxes, cflStart, count, cflBlocks, q, all, c, cfl, maxHost, maxDev ? scalars : nullptr,
                       blockMin, stream);
        hipLaunchKernelGGL(batchFinishCflKernel, dim3(unsigned(count)), dim3(256), 0, stream, blockMin, cflStart,
                           scalars, slot, slot == 0 ? limit : nullptr);
        if (count > 1 && slot == 0)
            hipLaunchKernelGGL(batchGateKernel, dim3(1), dim3(64), 0, stream, scalars, count, scalars + 4 * count);
    }
    void stage(const double* q, const double* p, const double* c, const double* r, int which, double* out) {
        const int z = tileZ == 2 ? 0 : (tileZ == 8 ? 2 : 1);
        launchBatchStage(ns, nv, t->scheme, boxes, stageStarts[z], count, stageBlocksZ[z], q, all, p, c, r, 0.0,
                         scalars, which, out, counters + 5, stream, tileZ);
    }
    void boundaryRates(const double* q, const double* p, const double* c, const double* r, double* out) {
        for (int i = 0; i < count; ++i) {
            const size_t o = boxHost[i].offset;
            ts[i]->boundaryRatesFrom(q + o, p + o, c + o, r + o, out + size_t(nv) * i, counters + 5, stream, all);
        }
    }

    // Order of the cells of q by thermodynamic state (into out).
    void stateOrder(const double* q, uint32_t* out) {
        // cost levels: the previous flash's search (mixed) or FP64 phase counts
        const uint16_t* levelCost = mode == REACTIVE_AMD_FLASH_MIXED ? (haveCostSearch ? costSearch : nullptr)
                                                                       : (haveCost ? cost : nullptr);
        const int levels = levelCost && stateLevels > 1 ? stateLevels : 1;
        hipLaunchKernelGGL(stateRangeInitKernel, dim3(1), dim3(64), 0, stream, stateRange);
        hipLaunchKernelGGL(stateRangeKernel, dim3(kStateBlocks), dim3(256), 0, stream, q, all, ns, cells, stateRange);
        const int bins = stateSide * stateSide * levels;
        hipLaunchKernelGGL(stateKeyKernel, dim3(kStateBlocks), dim3(256), 0, stream, q, all, ns, cells, stateRange,
                           stateSide, levelCost, levels, stateKey, stateCount);
        hipLaunchKernelGGL(stateScanKernel, dim3(1), dim3(256), 0, stream, stateCount, stateOffset, bins);
        hipLaunchKernelGGL(stateScatterKernel, dim3(kStateBlocks), dim3(256), 0, stream, stateKey, cells, bins,
                           stateOffset, out);
    }

    // Autotuner: times each candidate launch shape on the loaded state (the
    // step-start q and seeds) and keeps the fastest.  Stage: chunk depth 2,
    // 4, 8 (stage 0 from q0, three launches each after a warm-up).  Flash (the
    // current arithmetic, into the scratch state buffers): state-order bins
    // per axis 32, 64, 128 with 64-lane workgroups, then workgroups of 64,
    // 128, 256 lanes with the chosen bins.  Nothing it computes is kept.
    float timeIt(const std::function<void()>& work, int reps) {
        RAMD_HIP(hipEventRecord(tuneEvents[0], stream));
        for (int r = 0; r < reps; ++r) work();
        RAMD_HIP(hipEventRecord(tuneEvents[1], stream));
        RAMD_HIP(hipEventSynchronize(tuneEvents[1]));
        float ms = 0;
        RAMD_HIP(hipEventElapsedTime(&ms, tuneEvents[0], tuneEvents[1]));
        return ms / reps;
    }
    void tune(double cfl, double maximum) {
        cflSweep(qA, pA, cA, cfl, maximum, false, 0);  // a valid dt for the trial stages
        float best = 0;
        int bestZ = tileZ;
        for (int z : kChunkDepths) {
            tileZ = z;
            stage(qA, pA, cA, rA, 0, qB);
            const float ms = timeIt([&] { stage(qA, pA, cA, rA, 0, qB); }, 3);
            if (z == kChunkDepths[0] || ms < best) { best = ms; bestZ = z; }
        }
        tileZ = bestZ;
        flashCounters = counters + 8;
#if !RAMD_STATE_ORDER
        sortCells(haveCost ? cost : nullptr, order);
        if (mode == REACTIVE_AMD_FLASH_MIXED) sortCells(haveCostSearch ? costSearch : nullptr, orderSearch);
#endif
        auto trial = [&] { flash(qA, pB, cB, rB, sB, nullptr, flashCounters, nullptr); };
        int bestSide = stateSide, bestLevels = 1;
        flashBlock = 64;
        // (bins per axis, cost levels): cost levels need a previous flash's costs
        const int candidates[4][2] = {{32, 1}, {64, 1}, {128, 1}, {64, 4}};
        const bool costKnown = mode == REACTIVE_AMD_FLASH_MIXED ? haveCostSearch : haveCost;
        for (int k = 0; k < 4; ++k) {
            if (candidates[k][1] > 1 && !costKnown) continue;
            stateSide = candidates[k][0];
            stateLevels = candidates[k][1];
            const float ms = timeIt(trial, 1);
            if (k == 0 || ms < best) { best = ms; bestSide = stateSide; bestLevels = stateLevels; }
        }
        stateSide = bestSide;
        stateLevels = bestLevels;
        unsigned bestBlock = 64;
        for (unsigned b : {64u, 128u, 256u}) {
            flashBlock = b;
            const float ms = b == 64u ? best : timeIt(trial, 1);
            if (ms < best || b == 64u) { best = ms; bestBlock = b; }
        }
        flashBlock = bestBlock;
        flashCounters = counters;
        for (auto& g : graph)  // captured with the old shapes
            if (g) { RAMD_HIP(hipGraphExecDestroy(g)); g = nullptr; }
        stateSideOf[mode] = stateSide;
        stateLevelsOf[mode] = stateLevels;
        flashBlockOf[mode] = flashBlock;
        tunedOf[mode] = true;
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
        // CFL on the step-start state
        mark(m, 0);
        cflSweep(qA, pA, cA, cfl, maximum, false, 0);
        mark(m, 1);
#if !RAMD_STATE_ORDER
        // heavy-first orders from the latest costs (or the step-start seeds)
        sortCells(haveCost ? cost : nullptr, order);
        if (mode == REACTIVE_AMD_FLASH_MIXED) sortCells(haveCostSearch ? costSearch : nullptr, orderSearch);
#endif
        // stage 0 and its boundary rate
        boundaryRates(qA, pA, cA, rA, bA);
        mark(m, 2);
        stage(qA, pA, cA, rA, 0, qB);
        mark(m, 3);
        // flash #1 seeded from the step-start states
        flash(qB, pB, cB, rB, sB, nullptr, counters, m ? m[9] : nullptr);
        mark(m, 4);
        // CFL recheck after stage 0 (maximum = dt)
        cflSweep(qB, pB, cB, cfl, 0.0, true, 1);
        mark(m, 5);
        hipLaunchKernelGGL(recheckKernel, dim3(blocksFor(size_t(count), 64)), dim3(64), 0, stream, scalars, count,
                           counters, violations);
        // stage 1 (in place into qA, which holds q0) and its boundary rate
        boundaryRates(qB, pB, cB, rB, bB);
        mark(m, 6);
        stage(qB, pB, cB, rB, 1, qA);
        mark(m, 7);
        // flash #2 from the step-start seeds, into the state buffer of stage 1
        flash(qA, pA, cA, rA, sB, cost, counters + 2, m ? m[10] : nullptr);
        haveCost = true;
        if (mode == REACTIVE_AMD_FLASH_MIXED) haveCostSearch = true;
        mark(m, 8);
        std::swap(sA, sB);
        parity ^= 1;
        // conservation totals and boundary integral of every box
        if (count == 1)
            hipLaunchKernelGGL(totalsPartialKernel, dim3(kTotalBlocks), dim3(256), 0, stream, qA, all, cells, nv,
                               t->cartesian.volume, totalPartial);
        else
            hipLaunchKernelGGL(batchTotalsPartialKernel, dim3(kTotalBlocks, unsigned(count)), dim3(256), 0, stream,
                               boxes, qA, all, nv, totalPartial);
        hipLaunchKernelGGL(totalsFinishKernel, dim3(unsigned(count)), dim3(64), 0, stream, totalPartial, kTotalBlocks,
                           nv, totals);
        hipLaunchKernelGGL(boundaryIntegralKernel, dim3(unsigned(count)), dim3(64), 0, stream, bA, bB, scalars, nv,
                           boundary);
        RAMD_HIP(hipGetLastError());
    }
};

// One step under the driver's attempt loop (ReactiveFoam.C:1443-1494): the
// step-start q, p, c_frozen, r and states are kept; an attempt in which a
// case has a failed flash or a failed CFL recheck is rolled back and every
// case repeats it, the failed cases with half their dt (the others recompute
// the same dt and repeat their step bit for bit), at most 12 times and while
// dt / 2 > 1e-15.  Returns the attempts rolled back; throws with the
// step-start state restored when the step cannot be completed.
uint64_t attemptStep(Step& s, double cfl, double maximum, hipEvent_t* marks) {
    const int n = s.count;
    if (!s.qKeep) {
        s.qKeep = s.alloc<double>(s.all * s.nv);
        s.pKeep = s.alloc<double>(s.all);
        s.cKeep = s.alloc<double>(s.all);
        s.rKeep = s.alloc<double>(s.all);
        s.sKeep = s.alloc<ThermoOut>(s.cells);
        s.failMarks = s.alloc<uint8_t>(s.cells);
        s.boxFail = s.alloc<unsigned long long>(size_t(n));
        s.limits = s.alloc<double>(size_t(n));
    }
    auto copy = [&](void* to, const void* from, size_t bytes) {
        RAMD_HIP(hipMemcpyAsync(to, from, bytes, hipMemcpyDeviceToDevice, s.stream));
    };
    const size_t qBytes = s.all * s.nv * sizeof(double), aBytes = s.all * sizeof(double);
    copy(s.qKeep, s.qA, qBytes);
    copy(s.pKeep, s.pA, aBytes);
    copy(s.cKeep, s.cA, aBytes);
    copy(s.rKeep, s.rA, aBytes);
    copy(s.sKeep, s.sA, s.cells * sizeof(ThermoOut));
    const size_t boxes = size_t(n);
    std::vector<double> limit(boxes, __builtin_huge_val()), scalars(4 * boxes + 1);
    std::vector<unsigned long long> before(boxes), after(boxes), failed(boxes);
    for (uint64_t rolledBack = 0;; ++rolledBack) {
        RAMD_HIP(hipMemcpyAsync(s.limits, limit.data(), n * sizeof(double), hipMemcpyHostToDevice, s.stream));
        RAMD_HIP(hipMemsetAsync(s.failMarks, 0, s.cells, s.stream));
        RAMD_HIP(hipMemsetAsync(s.boxFail, 0, n * sizeof(unsigned long long), s.stream));
        RAMD_HIP(hipMemcpyAsync(before.data(), s.violations, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                s.stream));
        s.limit = s.limits;
        s.failMark = s.failMarks;
        s.one(cfl, maximum, marks);
        s.failMark = nullptr;
        s.limit = nullptr;
        hipLaunchKernelGGL(boxFailKernel, dim3(256, unsigned(n)), dim3(256), 0, s.stream, s.boxes, s.failMarks,
                           s.boxFail);
        RAMD_HIP(hipMemcpyAsync(failed.data(), s.boxFail, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipMemcpyAsync(after.data(), s.violations, n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipMemcpyAsync(scalars.data(), s.scalars, scalars.size() * sizeof(double), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        std::string why;
        bool any = false, invalid = false, exhausted = false;
        for (int i = 0; i < n; ++i) {
            const double dt = scalars[4 * size_t(i)];
            if (!(dt > 0)) invalid = true;
            if (!failed[i] && after[i] == before[i]) continue;
            any = true;
            if (why.empty()) {
                char text[160];
                std::snprintf(text, sizeof text, "case %d: %llu failed flash(es)%s at dt %.6e", i, failed[i],
                              after[i] > before[i] ? ", CFL recheck failed" : "", dt);
                why = text;
            }
            limit[i] = .5 * dt;
            if (!(limit[i] > 1e-15)) exhausted = true;
        }
        if (!any && !invalid) return rolledBack;
        // roll back (the attempt's states are in sA after its buffer swap)
        copy(s.qA, s.qKeep, qBytes);
        copy(s.pA, s.pKeep, aBytes);
        copy(s.cA, s.cKeep, aBytes);
        copy(s.rA, s.rKeep, aBytes);
        copy(s.sA, s.sKeep, s.cells * sizeof(ThermoOut));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        require(!invalid, "Invalid CFL time step (non-finite or non-positive wave speeds)");
        if (rolledBack + 1 > 12 || exhausted)
            throw std::runtime_error("Step failed after " + std::to_string(rolledBack + 1) +
                                     " attempts; the step-start state is restored (" + why + ")");
    }
}

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
        s.graphMax[p] != maximum || s.graphRep[p] != s.rep) {
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
        s.graphRep[p] = s.rep;
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

}  // extern "C"

namespace ramd {

// A step over the boxes of `list` (one or more independent cases), all on the
// HEM handle's device with its species.
static Step* createStep(const std::vector<Transport*>& list, HemHandle* hem) {
    require(!list.empty() && hem, "Null transport or HEM handle");
    require(!hem->boundary, "The resident step's flash has no boundary search (boundaryRecovery = 1): run such "
                            "models through reactive_gpu_hem_run_v1");
    std::unique_ptr<Step> s(new Step);
    s->t = list[0];
    s->h = hem;
    s->ts = list;
    s->count = int(list.size());
    size_t total = 0;
    for (Transport* t : list) {
        require(t, "Null transport handle");
        require(!t->viscous() && !t->capillary,
                "The resident step carries the inviscid flux (viscosity, conductivity, WALE and capillary model off)");
        require(t->cartesian.valid, "The resident step needs a structured-box transport mesh");
        require(t->device == hem->device, "Transport and HEM handles are on different devices");
        require(size_t(hem->ns) == t->cfg.species && t->cfg.variables == t->cfg.species + 4,
                "Transport and HEM species differ, or the transport carries liquid inventories");
        require(t->cfg.variables == list[0]->cfg.variables, "The boxes of a batch carry different variables");
        require(list.size() == 1 || t->all == t->cfg.cells, "A batched box carries fixed-state slots");
        total += t->cfg.cells;
    }
    require(list[0]->cfg.variables <= REACTIVE_AMD_MAX_VARIABLES, "Too many conserved variables");
    s->device = hem->device;
    RAMD_HIP(hipSetDevice(s->device));
    s->ns = int(list[0]->cfg.species);
    s->nv = int(list[0]->cfg.variables);
    s->cells = total;
    s->all = list.size() == 1 ? list[0]->all : total;
    require(s->cells <= UINT32_MAX, "Too many cells");
    RAMD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (auto& e : s->events) RAMD_HIP(hipEventCreate(&e));
    // box table and the block ranges of the batched launches
    std::vector<long long> stageStart(1, 0), cflStart(1, 0), startsZ[3];
    for (auto& v : startsZ) v.assign(1, 0);
    size_t offset = 0;
    for (Transport* t : list) {
        BoxDev b;
        b.a = makeCartesianArgs(t->cartesian, t->cfg.waveFactor);
        b.offset = offset;
        b.cells = t->cfg.cells;
        b.volume = t->cartesian.volume;
        s->boxHost.push_back(b);
        stageStart.push_back(stageStart.back() + tileBlocks(b.a));
        for (int z = 0; z < 3; ++z) startsZ[z].push_back(startsZ[z].back() + tileBlocks(b.a, kChunkDepths[z]));
        cflStart.push_back(cflStart.back() + long((b.cells + 255) / 256));
        offset += t->cfg.cells;
    }
    s->stageBlocks = long(stageStart.back());
    s->cflBlocks = long(cflStart.back());
    const int n = s->count;
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
    s->stateKey = s->alloc<uint16_t>(s->cells);
    s->stateCount = s->alloc<unsigned>(kStateMaxBins);
    s->stateOffset = s->alloc<unsigned>(kStateMaxBins);
    s->stateRange = s->alloc<unsigned long long>(4);
    s->blockMin = s->alloc<double>(size_t(s->cflBlocks) + 1);
    s->scalars = s->alloc<double>(4 * size_t(n) + 1);
    s->bA = s->alloc<double>(size_t(s->nv) * n);
    s->bB = s->alloc<double>(size_t(s->nv) * n);
    s->boundary = s->alloc<double>(size_t(s->nv) * n);
    s->totals = s->alloc<double>(size_t(s->nv) * n);
    s->totalPartial = s->alloc<double2>(size_t(s->nv) * kTotalBlocks * n);
    s->counters = s->alloc<unsigned long long>(32);
    s->flashCounters = s->counters;
    for (int z = 0; z < 3; ++z) {
        s->stageStarts[z] = s->alloc<long long>(size_t(n) + 1);
        s->stageBlocksZ[z] = long(startsZ[z].back());
        RAMD_HIP(hipMemcpyAsync(s->stageStarts[z], startsZ[z].data(), (n + 1) * sizeof(long long),
                                hipMemcpyHostToDevice, s->stream));
    }
    for (auto& e : s->tuneEvents) RAMD_HIP(hipEventCreate(&e));
    s->violations = s->alloc<unsigned long long>(size_t(n));
    s->boxes = s->alloc<BoxDev>(size_t(n));
    s->stageStart = s->alloc<long long>(size_t(n) + 1);
    s->cflStart = s->alloc<long long>(size_t(n) + 1);
    RAMD_HIP(hipMemcpyAsync(s->boxes, s->boxHost.data(), n * sizeof(BoxDev), hipMemcpyHostToDevice, s->stream));
    RAMD_HIP(hipMemcpyAsync(s->stageStart, stageStart.data(), (n + 1) * sizeof(long long), hipMemcpyHostToDevice,
                            s->stream));
    RAMD_HIP(hipMemcpyAsync(s->cflStart, cflStart.data(), (n + 1) * sizeof(long long), hipMemcpyHostToDevice,
                            s->stream));
    RAMD_HIP(hipMemsetAsync(s->bucketCount, 0, 256 * sizeof(unsigned), s->stream));
    RAMD_HIP(hipMemsetAsync(s->stateCount, 0, kStateMaxBins * sizeof(unsigned), s->stream));
    RAMD_HIP(hipMemsetAsync(s->violations, 0, n * sizeof(unsigned long long), s->stream));
    RAMD_HIP(hipStreamSynchronize(s->stream));
    for (Transport* t : list) t->stepAttached = true;  // the WALE closure can no longer be installed
    return s.release();
}

}  // namespace ramd

extern "C" {

void* reactive_amd_step_create(void* transport, void* hem, char* error, size_t errorSize) {
    try {
        return createStep({static_cast<Transport*>(transport)}, static_cast<HemHandle*>(hem));
    } catch (const std::exception& e) {
        writeError(error, errorSize, e.what());
        return nullptr;
    }
}

void* reactive_amd_step_create_batch(void* const* transports, int cases, void* hem, char* error, size_t errorSize) {
    try {
        require(transports && cases >= 1, "A batch needs at least one transport handle");
        std::vector<Transport*> list;
        for (int i = 0; i < cases; ++i) list.push_back(static_cast<Transport*>(transports[i]));
        for (Transport* t : list)
            require(t && t->scheme == list[0]->scheme, "The boxes of a batch use different face fluxes");
        return createStep(list, static_cast<HemHandle*>(hem));
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
        if (s.all > s.cells) {
            hipLaunchKernelGGL(copyFixedKernel, dim3(blocksFor(s.all - s.cells, 64)), dim3(64), 0, s.stream, s.t->q,
                               s.t->state, s.cells, s.all, s.nv, s.qA, s.qB, s.pA, s.cA, s.rA, s.pB, s.cB, s.rB);
            RAMD_HIP(hipGetLastError());
        }
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
        const bool autotune = (flashMode & REACTIVE_AMD_STEP_AUTOTUNE) != 0;
        const bool retry = (flashMode & REACTIVE_AMD_STEP_RETRY) != 0;
        const bool reuse = (flashMode & REACTIVE_AMD_STEP_REUSE) != 0;
        flashMode &= ~(REACTIVE_AMD_STEP_GRAPH | REACTIVE_AMD_STEP_AUTOTUNE | REACTIVE_AMD_STEP_RETRY |
                       REACTIVE_AMD_STEP_REUSE);
        require(!(retry && useGraph), "Retried steps run on the stream, not from a captured graph");
        require(flashMode == REACTIVE_AMD_FLASH_FP64 || flashMode == REACTIVE_AMD_FLASH_MIXED,
                "Unknown flash arithmetic mode");
        if (flashMode == REACTIVE_AMD_FLASH_MIXED && !s.sets) {
            s.sets = s.alloc<CandidateSet>(s.cells);
            s.queue = s.alloc<uint32_t>(s.cells);
            s.queueCount = s.alloc<unsigned>(1);
        }
        s.mode = flashMode;
        if (reuse && !s.reuseTable) {
            uint32_t size = 1;
            while (size < 2 * s.cells) size <<= 1;
            s.reuseTable = s.alloc<uint32_t>(size);
            s.reuseMask = size - 1;
            s.reuseSlot = s.alloc<uint32_t>(s.cells);
            s.reuseRep = s.alloc<uint32_t>(s.cells);
        }
        s.rep = reuse ? s.reuseRep : nullptr;
        s.stateSide = s.stateSideOf[s.mode];
        s.stateLevels = s.stateLevelsOf[s.mode];
        s.flashBlock = s.flashBlockOf[s.mode];
        // tuning wants the costs of a previous flash (for the cost-level
        // candidates): without them it runs after this call's first step
        bool pendingTune = autotune && !s.tunedOf[s.mode];
        const bool costKnown = s.mode == REACTIVE_AMD_FLASH_MIXED ? s.haveCostSearch : s.haveCost;
        if (pendingTune && costKnown) {
            s.tune(cfl, maximum);
            pendingTune = false;
        }
        hipLaunchKernelGGL(clearCountersKernel, dim3(1), dim3(64), 0, s.stream, s.counters, 32);
        hipLaunchKernelGGL(clearCountersKernel, dim3(blocksFor(size_t(s.count), 64)), dim3(64), 0, s.stream,
                           s.violations, s.count);
        while (s.marks.size() < size_t(steps) * kMarks) {
            hipEvent_t e;
            RAMD_HIP(hipEventCreate(&e));
            s.marks.push_back(e);
        }
        RAMD_HIP(hipEventRecord(s.events[0], s.stream));
        uint64_t retries = 0;
        for (int i = 0; i < steps; ++i) {
            if (useGraph) graphStep(s, cfl, maximum);
            else if (retry) retries += attemptStep(s, cfl, maximum, s.marks.data() + size_t(i) * kMarks);
            else s.one(cfl, maximum, s.marks.data() + size_t(i) * kMarks);
            if (pendingTune) {
                s.tune(cfl, maximum);
                pendingTune = false;
            }
        }
        RAMD_HIP(hipEventRecord(s.events[1], s.stream));
        unsigned long long counters[17];
        const size_t n = size_t(s.count);
        std::vector<double> scalars(4 * n + 1), totals(s.nv * n), boundary(s.nv * n);
        RAMD_HIP(hipMemcpyAsync(counters, s.counters, sizeof(counters), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(scalars.data(), s.scalars, scalars.size() * sizeof(double), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipMemcpyAsync(totals.data(), s.totals, totals.size() * sizeof(double), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipMemcpyAsync(boundary.data(), s.boundary, boundary.size() * sizeof(double), hipMemcpyDeviceToHost,
                                s.stream));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        s.lastSteps = uint64_t(steps);
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
        r.queuedCells = counters[7];
        r.tileDepth = uint32_t(s.tileZ);
        r.flashBlock = s.flashBlock;
        r.stateBins = uint32_t(s.stateSide);
        r.reserved = uint32_t(s.stateLevels);
        r.retries = retries;
        r.reusedCells = counters[16];
        // a batch reports its smallest dt and recheck step and the sums over its
        // boxes (box order); reactive_amd_step_report_box gives each box
        r.dt = scalars[0];
        r.recheckDt = scalars[1];
        for (size_t i = 1; i < n; ++i) {
            r.dt = scalars[4 * i] < r.dt || !(scalars[4 * i] > 0) ? scalars[4 * i] : r.dt;
            r.recheckDt = scalars[4 * i + 1] < r.recheckDt ? scalars[4 * i + 1] : r.recheckDt;
        }
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < s.nv; ++k) {
                r.totals[k] += totals[s.nv * i + k];
                r.boundaryIntegral[k] += boundary[s.nv * i + k];
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
        for (size_t i = 0; i < n; ++i) require(scalars[4 * i] > 0, "Invalid wave/diffusion timestep");
    });
}

int reactive_amd_step_report_box(void* step, int box, ReactiveAmdStepReport* report) {
    return guard(step, [&](Step& s) {
        require(report && report->abiVersion == 1 && report->structBytes == sizeof(*report), "Invalid step report ABI");
        require(box >= 0 && box < s.count, "Box index out of range");
        double scalars[2];
        unsigned long long violations = 0;
        std::vector<double> totals(s.nv), boundary(s.nv);
        RAMD_HIP(hipMemcpyAsync(scalars, s.scalars + 4 * box, sizeof(scalars), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(&violations, s.violations + box, sizeof(violations), hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(totals.data(), s.totals + size_t(s.nv) * box, s.nv * sizeof(double),
                                hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipMemcpyAsync(boundary.data(), s.boundary + size_t(s.nv) * box, s.nv * sizeof(double),
                                hipMemcpyDeviceToHost, s.stream));
        RAMD_HIP(hipStreamSynchronize(s.stream));
        ReactiveAmdStepReport r{};
        r.abiVersion = 1;
        r.structBytes = sizeof(r);
        r.steps = s.lastSteps;
        r.cflViolations = violations;
        r.dt = scalars[0];
        r.recheckDt = scalars[1];
        for (int k = 0; k < s.nv; ++k) {
            r.totals[k] = totals[k];
            r.boundaryIntegral[k] = boundary[k];
        }
        *report = r;
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
