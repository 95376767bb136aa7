// SPDX-License-Identifier: GPL-3.0-or-later
// Homogeneous-equilibrium UV flash for one cell per work-item on gfx950.
//
// The algorithm is the reference's ReactiveDeviceFlash::Flash
// (src/reactiveThermo/reactiveDeviceFlash.h:140-560): the same candidate
// order (all-gas frozen/scalar recovery, all-liquid, five liquid-fraction
// seeds per active mask, then four restart temperatures), the same Newton
// step limits, line-search fractions, stability tests and maximum-entropy
// selection, and the same finite-difference Jacobian.  Expression grouping
// follows the reference so that FP64 results match the CPU path.
//
// GPU-specific structure: the gas species count NS and the liquid count NL
// (0 or 1) are template parameters, so every per-species and per-unknown
// array is register resident; the evaluation record keeps only the fields
// the flash reads; the model is a compact table image in global memory.
//
// The scalar type R is double for the reference arithmetic.  R = float is
// the search half of the mixed-precision flash: the same candidate search
// with single-precision tolerances and finite-difference steps, which records
// each converged candidate's (ln p, ln T, liquid fraction) in a CandidateSet;
// Flash<NS, NL, double>::refine then finishes the recorded candidates with
// FP64 Newton corrections and decides stability and maximum entropy in FP64.
#ifndef REACTIVE_AMD_FLASH_DEVICE_H
#define REACTIVE_AMD_FLASH_DEVICE_H
#include <type_traits>
#include "pr_device.h"
#include "flash_jacobian.h"

// Inlining of the phase evaluation (knobs, per arithmetic: the FP32 search
// and the FP64 flash kernels): 1 = the phase-state functions (gasTP,
// liquidTP) into phaseStatus, 2 = also evaluate() into its solvers.  Fewer
// calls mean fewer call frames in scratch: at 96^3 the search runs 26 % faster
// with 2 than with 1; the FP64 step runs 13 % faster with 1 than with none and
// 28 % faster again with 2.
#ifndef RAMD_SEARCH_INLINE
#define RAMD_SEARCH_INLINE 2
#endif
#ifndef RAMD_FLASH_INLINE
#define RAMD_FLASH_INLINE 2
#endif
// Mole-fraction distance within which the FP32 search reuses a mixture's
// gas phase heuristic (FP64 rebuilds it for every new composition, as the
// reference does).  At 96^3 binary N2O/N2 1e-3 takes the search from 3.5 s
// to 2.0 s per flash with the same candidates (no new flips, the same
// distance to the converged roots).
#ifndef RAMD_HEURISTIC_REUSE
#define RAMD_HEURISTIC_REUSE 1e-3f
#endif
#ifdef RAMD_INLINE_EVERYTHING  // tuning experiment: every solver function inlinable
#define RAMD_NOINLINE __device__ inline
#endif
#ifndef RAMD_NOINLINE
#define RAMD_NOINLINE __device__ __noinline__
#endif
#define RD __device__ inline

namespace ramd {

template <int NS>
struct HemModel {
    Table<NS> gas;
    Table<1> liquid;
    PhaseHeuristic gasFixed;     // composition-independent heuristic when NS == 1
    PhaseHeuristic liquidFixed;  // pure liquid table
    int nl, condensable, scalarRecovery, stableGasPrune, enforceBounds, analyticJacobian, boundaryRecovery;
    double weights[NS];
    double liquidTmin, liquidTc, Tmin, Tmax, pmin, pmax, vtol, etol, mutol;
};

// Thermo output, byte-compatible with ReactiveThermoState (reactiveThermo.h:11-19).
struct ThermoOut {
    double p, T, rho, e, entropy, soundFrozen, soundEquilibrium;
    double volumeResidual, energyResidual, chemicalResidual;
    double alphaGas, alphaLiquid[2], rhoGas, rhoLiquid[2], liquidMass[2];
    double gasMass, cp, cv;
    int activeLiquids, iterations;
};
static_assert(sizeof(ThermoOut) == 176, "ReactiveThermoState layout");

struct Counts {
    uint32_t phases, residuals, candidates, stable, failures, fd, analytic;
};

// Per-phase properties the analytic Jacobian reads: the condensable's
// partial specific volume and energy, partial molar enthalpy / RT and
// p v / RT in each phase (for a pure phase the phase's own specific and
// molar values), and muN, the change of the gas's mu_c / RT per unit mass of
// the condensable added to the gas at fixed (p, T) (zero for a pure gas).
// Empty when the flash uses only the reference's finite-difference Jacobian.
template <class R, bool A>
struct PhaseParts {
    RD void clear() {}
};
template <class R>
struct PhaseParts<R, true> {
    R vG, vL, eG, eL, hG, hL, zG, zL, muN;
    RD void clear() { vG = vL = eG = eL = hG = hL = zG = zL = muN = 0; }
};

// The residual point's own properties the reference's analytic Jacobian
// reads (Evaluation::gasCondensable, liquidHbar, liquidVbar,
// liquidThermoMask, gasMolarVolume; reactiveDeviceFlash.h:64-73): kept only
// by the flash that runs the reference's analytic-Jacobian mode.
template <bool RJ>
struct RefParts {
    RD void clear() {}
};
template <>
struct RefParts<true> {
    double gasMolarVolume, gasCondensable, liquidHbar, liquidVbar;
    int liquidMask;
    RD void clear() {
        gasMolarVolume = gasCondensable = liquidHbar = liquidVbar = 0;
        liquidMask = 0;
    }
};

// The condensable's gas inventory of the residual point and whether it was
// given (vapour coordinates, Evaluation::gasCondensable / explicitGas,
// reactiveDeviceFlash.h:67): kept by the capillary / boundary-search flash.
// A base (empty for the other flashes, so the record does not grow).
template <bool C>
struct CapParts {
    RD void clearCap() {}
};
template <>
struct CapParts<true> {
    double capGasCondensable;
    bool capExplicitGas;
    RD void clearCap() {
        capGasCondensable = 0;
        capExplicitGas = false;
    }
};

// The fields of the evaluation record that only the committed state reads
// (ReactiveThermoState's outputs): kept by the FP64 flashes, left out of the
// FP32 search's record (a base, empty there), which never commits a state,
// so its per-lane record is smaller.
template <class R, bool Keep>
struct Outputs {
    R e, entropy, cf, ceq, volRes, enRes, chemRes;
    R alphaGas, alphaLiq, rhoGas, rhoLiq, cp, cv;
    int active, iterations;
    RD void clearOutputs() {
        e = entropy = cf = ceq = volRes = enRes = chemRes = 0;
        alphaGas = alphaLiq = rhoGas = rhoLiq = cp = cv = 0;
        active = iterations = 0;
    }
};
template <class R>
struct Outputs<R, false> {
    RD void clearOutputs() {}
};

// Evaluation record (reactiveDeviceFlash.h:64-73) for at most one liquid;
// A adds the per-phase parts of this build's analytic Jacobian (mixed
// flash), RJ the inputs of the reference's analytic Jacobian, C the
// vapour-coordinate inventory of the capillary / boundary-search flash.
template <class R, bool A = false, bool RJ = false, bool C = false>
struct EvalT : CapParts<C>, Outputs<R, !std::is_same<R, float>::value> {
    static constexpr bool kOutputs = !std::is_same<R, float>::value;
    R p, T, rho, liqMass, gasMass;
    R volume, energy, entropyDensity, Vp, VT, Ep, ET, cpDensity, muGas, muLiquid;
    PhaseParts<R, A> parts;
    RefParts<RJ> ref;
};
using Eval = EvalT<double>;

template <class R, bool A, bool RJ, bool C>
RD void zeroEval(EvalT<R, A, RJ, C>& v) {
    v.p = v.T = v.rho = v.liqMass = v.gasMass = 0;
    v.clearOutputs();
    v.volume = v.energy = v.entropyDensity = v.Vp = v.VT = v.Ep = v.ET = v.cpDensity = v.muGas = v.muLiquid = 0;
    v.parts.clear();
    v.ref.clear();
    v.clearCap();
}

template <bool A, bool RJ, bool C>
RD void writeOut(const EvalT<double, A, RJ, C>& v, ThermoOut& o) {
    o.p = v.p; o.T = v.T; o.rho = v.rho; o.e = v.e; o.entropy = v.entropy;
    o.soundFrozen = v.cf; o.soundEquilibrium = v.ceq;
    o.volumeResidual = v.volRes; o.energyResidual = v.enRes; o.chemicalResidual = v.chemRes;
    o.alphaGas = v.alphaGas; o.alphaLiquid[0] = v.alphaLiq; o.alphaLiquid[1] = 0;
    o.rhoGas = v.rhoGas; o.rhoLiquid[0] = v.rhoLiq; o.rhoLiquid[1] = 0;
    o.liquidMass[0] = v.liqMass; o.liquidMass[1] = 0;
    o.gasMass = v.gasMass; o.cp = v.cp; o.cv = v.cv;
    o.activeLiquids = v.active; o.iterations = v.iterations;
}

// Candidates of the single-precision search, in search order.
enum CandidateKind : int16_t { kGasCandidate = 0, kLiquidCandidate = 1, kSeedCandidate = 2, kRestartCandidate = 3 };
struct Candidate {
    float x[3];     // ln p, ln T, liquid fraction (0 for the gas and all-liquid candidates)
    int16_t kind;   // CandidateKind
    int16_t index;  // seed j, or 5 t + j for restart temperature t
};
constexpr int kMaxCandidates = 6;
enum CandidateFlags : uint32_t {
    kRestartsSearched = 1u,  // the search ran the restart temperatures
    kOverflow = 2u,          // more distinct candidates than the set holds
    kSeedsPruned = 4u,       // the stable-gas pruning skipped the seeds
};
struct CandidateSet {
    uint32_t count, flags;
    Candidate c[kMaxCandidates];
};
static_assert(sizeof(CandidateSet) == 104, "candidate set layout");

// Full-pivot LU with Eigen's rank rule (reactiveDeviceFlash.h:111-139), n fixed.
template <int n, class R>
RD bool solveN(const R (&m)[3][3], const R* rhs, R* x) {
    R a[n][n], b[n], largest = 0;
    int column[n];
    for (int i = 0; i < n; ++i) {
        column[i] = i;
        b[i] = rhs[i];
        if (!fin(b[i])) return false;
        for (int j = 0; j < n; ++j) {
            a[i][j] = m[i][j];
            if (!fin(a[i][j])) return false;
        }
    }
    for (int k = 0; k < n; ++k) {
        int row = k, col = k;
        R pivot = 0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i)
                if (ab(a[i][j]) > pivot) { pivot = ab(a[i][j]); row = i; col = j; }
        largest = hi(largest, pivot);
        if (!(pivot > largest * R(n) * R(Real<R>::eps))) return false;
        if (row != k) {
            for (int j = 0; j < n; ++j) { const R t = a[k][j]; a[k][j] = a[row][j]; a[row][j] = t; }
            const R t = b[k]; b[k] = b[row]; b[row] = t;
        }
        if (col != k) {
            for (int i = 0; i < n; ++i) { const R t = a[i][k]; a[i][k] = a[i][col]; a[i][col] = t; }
            const int t = column[k]; column[k] = column[col]; column[col] = t;
        }
        for (int i = k + 1; i < n; ++i) {
            const R f = a[i][k] / a[k][k];
            a[i][k] = f;
            for (int j = k + 1; j < n; ++j) a[i][j] -= f * a[k][j];
            b[i] -= f * b[k];
        }
    }
    R z[n];
    for (int i = n - 1; i >= 0; --i) {
        R v = b[i];
        for (int j = i + 1; j < n; ++j) v -= a[i][j] * z[j];
        z[i] = v / a[i][i];
        if (!fin(z[i])) return false;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (column[i] == j) x[j] = z[i];
    return true;
}

// Lane w's copy of a trivially copyable value, in every lane of the wave.
template <class T>
__device__ inline T waveBroadcast(const T& v, int w) {
    static_assert(sizeof(T) % 4 == 0, "broadcast in 4-byte words");
    uint32_t words[sizeof(T) / 4];
    __builtin_memcpy(words, &v, sizeof(T));
    for (size_t k = 0; k < sizeof(T) / 4; ++k) words[k] = uint32_t(__shfl(int(words[k]), w, 64));
    T out;
    __builtin_memcpy(&out, words, sizeof(T));
    return out;
}

// Precision-dependent flash settings.  FP64: the reference's values.  FP32:
// tolerances and finite-difference steps at single-precision rounding, and a
// hand-over threshold: an FP32 Newton that stalls (rejected line search)
// below it still records its point for the FP64 refinement.
template <class R> struct FlashPrecision;
template <> struct FlashPrecision<double> {
    static constexpr bool single = false;
    static constexpr double hLog = 2e-6, hFraction = 2e-5;  // reactiveDeviceFlash.h:338-373
    static constexpr double fractionEdge = 1e-8;             // seed clamp (:385)
    static constexpr double rootCheck = 1e-7;                // |p_EOS - p| (reactiveDevicePR.h:574-609)
    static constexpr double densityCheck = 1e-8;             // gas probe root consistency (:264-275)
    static constexpr double vtolFloor = 0, etolFloor = 0, mutolFloor = 0, handOver = 0;
    static constexpr double heuristicReuse = 0;  // the gas phase heuristic for exactly its composition
};
template <> struct FlashPrecision<float> {
    static constexpr bool single = true;
    static constexpr float hLog = 1e-3f, hFraction = 1e-3f;
    static constexpr float fractionEdge = 1e-6f;
    static constexpr float rootCheck = 5e-3f;
    static constexpr float densityCheck = 1e-4f;
    static constexpr float vtolFloor = 3e-5f, etolFloor = 3e-5f, mutolFloor = 3e-4f, handOver = 1e-3f;
    // a mixture's phase heuristic (up to 100 FP64 cubic solves) is reused
    // while no mole fraction moves more than this from the one it was built
    // for: its labels move only for densities within a hair of the mid density
    static constexpr float heuristicReuse = RAMD_HEURISTIC_REUSE;
};

// A (analytic): the active-set Newton steps and the refinement's equilibrium
// sound speed use the analytic Jacobian (a gas mixture through the
// condensable's partial quantities, falling back to the reference's
// finite differences where those are not finite).  On for the FP32 search
// and the FP64 refinement of the mixed flash, off for the reference-arithmetic
// flash.
// RJ: the reference's analytic-Jacobian mode (Model::analyticJacobian = 1;
// FP64 flash only).
// CAP: the capillary / boundary-search flash (FP64 flash only): a fixed
// interface (color, jump: the gas at p - color J, the liquid at
// p + (1 - color) J) and the search in vapour coordinates of boundaryRecovery
// models and capillary equilibrium cells (reactiveDeviceFlash.h:327-448,
// 509-590).  Without it those members stay zero and no code reads them.
// The capillary flash's interface and search mode, a base that is empty for
// the other flashes (their per-lane record does not grow).
template <bool CAP, class R>
struct FlashCap {};
template <class R>
struct FlashCap<true, R> {
    R color = 0, jump = 0;  // capillary interface, J fixed during the solve
    bool adaptive = false;  // vapour-log coordinates of the boundary search
};

// The workgroup's copy of the model image in LDS (filled by the kernels'
// workgroupModel): a flash built with LDS reads its tables through this
// object, whose address space the compiler knows in every solver function,
// so the reads are LDS loads rather than generic (flat) ones.
template <int NS>
__device__ inline HemModel<NS>& sharedModel() {
    __shared__ HemModel<NS> local;
    return local;
}

// LDS: the model is the workgroup's LDS copy (sharedModel); otherwise the
// reference given to the constructor.
// WAVE: every lane of a wave runs this flash on the same cell (so its
// control flow is the wave's), and the Newton line searches evaluate their
// fractions on the wave's lanes at once, lane j the j-th fraction of the
// loop; the first accepted fraction in loop order is taken, which is the
// loop's choice, so the results are those of the one-lane flash bit for bit
// (the rejected evaluations it adds change nothing: the flash's only memo,
// an unavailable liquid at (p, T), repeats what recomputing gives).  For the
// few long refinements: a frozen solve that crawls halves each of its seventy
// Newton steps many times.  Every lane makes the active set's residual call
// (the lanes past the loop's end repeat its last fraction): with the call
// made by only some lanes, the results depended on what earlier work had
// left on the device.
template <int NS, int NL, class R = double, bool A = FlashPrecision<R>::single, bool RJ = false, bool CAP = false,
          bool LDS = false, bool WAVE = false>
struct Flash : FlashCap<CAP, R> {
    static_assert(NL == 0 || NL == 1, "one pure liquid at most");
    static_assert(!RJ || (!A && !FlashPrecision<R>::single), "the reference's analytic mode is an FP64 flash");
    static_assert(!CAP || (!A && !FlashPrecision<R>::single), "the capillary flash is an FP64 flash");
    static_assert(!WAVE || !FlashPrecision<R>::single, "wave-shared line searches are an FP64 refinement");
    using P = FlashPrecision<R>;
    using E = EvalT<R, A, RJ, CAP>;
    using S = PRStateT<R>;
    using M = MixT<R>;
    const HemModel<NS>& m;
    RD const HemModel<NS>& mod() const {
        if constexpr (LDS) return sharedModel<NS>();
        else return m;
    }
    R q[NS];
    R energy;
    R guessP, guessT, guessLiquid;
    Counts& n;
    R cacheP, cacheT;  // one unavailable liquid result (reactiveDeviceFlash.h:142-145)
    int cacheStatus;
    bool cacheVirtual;
    R memoKey[NS];     // composition of the memoised gas heuristic (NS > 1)
    PhaseHeuristic memo;
    bool memoValid;
    R Tmin, Tmax, pmin, pmax, vtol, etol, mutol, liquidTmin, liquidTc;
    // single-precision search: where converged candidates are recorded
    CandidateSet* record;
    bool stalled;  // the last FP32 Newton stopped near convergence (hand-over)

    RD Flash(const HemModel<NS>& model, Counts& counts) : m(model), n(counts) {
        cacheP = cacheT = 0;
        cacheStatus = -1;
        cacheVirtual = false;
        memoValid = false;
        Tmin = R(mod().Tmin); Tmax = R(mod().Tmax); pmin = R(mod().pmin); pmax = R(mod().pmax);
        vtol = hi(R(mod().vtol), R(P::vtolFloor));
        etol = hi(R(mod().etol), R(P::etolFloor));
        mutol = hi(R(mod().mutol), R(P::mutolFloor));
        liquidTmin = R(mod().liquidTmin); liquidTc = R(mod().liquidTc);
        record = nullptr;
        stalled = false;
    }
    // WAVE: after a line search shared by the wave, the counts the one-lane
    // loop makes (the trials up to lane `last`, the winner or all), the same
    // in every lane.
    RD void waveCounts(const Counts& before, int last) {
        const int lane = int(threadIdx.x & 63u);
        uint32_t phases = lane <= last ? n.phases - before.phases : 0u;
        uint32_t residuals = lane <= last ? n.residuals - before.residuals : 0u;
        for (int o = 32; o > 0; o >>= 1) {
            phases += uint32_t(__shfl_xor(int(phases), o, 64));
            residuals += uint32_t(__shfl_xor(int(residuals), o, 64));
        }
        n.phases = before.phases + phases;
        n.residuals = before.residuals + residuals;
    }
    // Phase pressures of a curved interface (reactiveDeviceFlash.h:152-153).
    RD R gasPressure(R p) const {
        if constexpr (CAP) return p - this->color * this->jump;
        else return p;
    }
    RD R liquidPressure(R p) const {
        if constexpr (CAP) return p + (R(1) - this->color) * this->jump;
        else return p;
    }

    RD R density() const {
        R r = 0;
        for (int k = 0; k < NS; ++k) r += q[k];
        return r;
    }
    RD bool bounds(R p, R T) const {
        return fin(p) && fin(T) && p >= pmin && p <= pmax && T >= Tmin && T <= Tmax;
    }
    static RD bool unavailable(int s) { return s == kUnavailable || s == kNoRoot; }

    RD const PhaseHeuristic& gasHeuristic(const R* x, R W, const M& mx) {
        if (NS == 1) return mod().gasFixed;
        bool hit = memoValid;
        for (int k = 0; k < NS; ++k)
            hit = hit && (P::single ? ab(memoKey[k] - x[k]) <= R(P::heuristicReuse) : memoKey[k] == x[k]);
        if (!hit) {
            phaseHeuristic<NS, R>(mod().gas, x, W, mx.a, mx.b, memo);
            for (int k = 0; k < NS; ++k) memoKey[k] = x[k];
            memoValid = true;
        }
        return memo;
    }

    // Gas PR state at (T, p) on the gas root (evaluateTPTrusted, reactiveDevicePR.h:574-609).
    RAMD_NOINLINE int gasTP(R T, R p, const R* Y, S& s, R* mu) { return gasTPBody(T, p, Y, s, mu); }
    __device__ __attribute__((always_inline)) inline int gasTPBody(R T, R p, const R* Y, S& s, R* mu) {
        if (!fin(T) || !fin(p) || T <= 0 || p <= 0) return kBadInput;
        R x[NS], W = 0;
        int st = massToMole(mod().gas, Y, x, W);
        if (st != kOk) return st;
        M mx;
        st = mixing(mod().gas, T, x, mx);
        if (st != kOk) return st;
        RootsT<R> r;
        st = roots(T, p, mx.a, mx.b, mx.aa, r);
        if (st != kOk) return st;
        if (!r.gas) return kUnavailable;
        st = stateTrho(mod().gas, T, W / r.vLast, x, W, mx, s, mu);
        if (st != kOk) return st;
        st = finishState<NS>(gasHeuristic(x, W, mx), T, s, mu);
        if (st != kOk) return st;
        if (ab(s.p - p) > R(P::rootCheck) * hi(R(1), p)) return kNonFinite;
        s.p = p;
        if (s.phase >= 0) return kUnavailable;
        return kOk;
    }
    // Pure-liquid PR state at (T, p) on the liquid root.
    RAMD_NOINLINE int liquidTP(R T, R p, S& s, R* mu) { return liquidTPBody(T, p, s, mu); }
    __device__ __attribute__((always_inline)) inline int liquidTPBody(R T, R p, S& s, R* mu) {
        if (!fin(T) || !fin(p) || T <= 0 || p <= 0) return kBadInput;
        const R Y[1] = {R(1)};
        R x[1], W = 0;
        int st = massToMole(mod().liquid, Y, x, W);
        if (st != kOk) return st;
        M mx;
        st = mixing(mod().liquid, T, x, mx);
        if (st != kOk) return st;
        RootsT<R> r;
        st = roots(T, p, mx.a, mx.b, mx.aa, r);
        if (st != kOk) return st;
        if (!r.liquid) return kUnavailable;
        st = stateTrho(mod().liquid, T, W / r.v0, x, W, mx, s, mu);
        if (st != kOk) return st;
        st = finishState<1>(mod().liquidFixed, T, s, mu);
        if (st != kOk) return st;
        if (ab(s.p - p) > R(P::rootCheck) * hi(R(1), p)) return kNonFinite;
        s.p = p;
        if (s.phase < 0) return kUnavailable;
        return kOk;
    }

    // reactiveDeviceFlash.h:157-196 for index -1 (gas) and 0 (liquid).
    RD int phaseStatus(int index, R p, R T, const R* Y, S& out, R* mu) {
        ++n.phases;
        if (!bounds(p, T)) return kBadInput;
        if (index >= 0 && T < liquidTmin) return kBadInput;
        if (index >= 0 && T >= liquidTc * (R(1) - R(1e-10))) return kUnavailable;
        if (index < 0 && mod().enforceBounds)
            for (int k = 0; k < NS; ++k)
                if (Y[k] > 0 && (T < R(mod().gas.tlo[k]) || T > R(mod().gas.thi[k]))) return kBadInput;
        if constexpr (P::single ? RAMD_SEARCH_INLINE != 0 : RAMD_FLASH_INLINE != 0) {
            if (index >= 0) return liquidTPBody(T, p, out, mu);
            return gasTPBody(T, p, Y, out, mu);
        }
        if (index >= 0) return liquidTP(T, p, out, mu);
        return gasTP(T, p, Y, out, mu);
    }

    static RD void accumulate(E& v, R mass, const S& p, R phaseP) {  // :203-207
        const R volume = mass / p.rho, vp = -volume * p.kappaT, vt = volume * p.alpha;
        v.volume += volume; v.energy += mass * p.e; v.entropyDensity += mass * p.s; v.Vp += vp; v.VT += vt;
        v.Ep += -v.T * vt - phaseP * vp;
        v.ET += mass * p.cp - phaseP * vt;
        v.cpDensity += mass * p.cp;
    }

    // Specific volume and energy, molar enthalpy / RT and compressibility of
    // one phase state (analytic-Jacobian flashes only).
    static RD void recordPhase(const S& s, R p, R T, R& v, R& e, R& h, R& z) {
        const R RT = R(8314.46261815324) * T;
        v = R(1) / s.rho;
        e = s.e;
        h = s.h * s.V * s.rho / RT;  // molar weight = V rho
        z = p * s.V / RT;
    }

    // Condensable's partial quantities in a PR gas mixture at (p, T) (gas
    // mass mg, mass fractions Y, state s from gasTP), by differentiating the
    // residual Helmholtz energy A_r = -n R T ln(1 - B/V) - D ln((V + (1+√2)B) /
    // (V + (1-√2)B)) / (2√2 B), with B = Σ n_i b_i and D = Σ n_i n_j a_ij(T),
    // in the moles n_c at fixed (T, V) and moving to fixed (T, p) with
    // v_c = -(∂p/∂n_c) / (∂p/∂V):
    //   partial volume v_c, energy e_c = e_c^ig + ∂U_r/∂n_c + (T ∂p/∂T - p) v_c,
    //   enthalpy h_c = e_c + p v_c, and
    //   ∂mu_c/∂n_c|_{T,p} = R T / n_c + ∂²A_r/∂n_c² + (∂p/∂n_c)² / (∂p/∂V).
    // mu_c is the reference's gas chemical potential (reactiveDevicePR.h:900-916),
    // whose (T, p, n) derivatives these are.
    RD void gasPartials(R p, R T, R mg, const R* Y, const S& s, PhaseParts<R, true>& out) const {
        const int c = mod().condensable;
        const auto& t = mod().gas;
        R x[NS], W = 0;
        if (massToMole(t, Y, x, W) != kOk) { out.muN = R(Real<R>::big) * R(2); return; }
        R qv[NS], dq[NS];
        R b = 0, aa = 0, da = 0;
        for (int i = 0; i < NS; ++i) {
            const R Tc = R(t.Tc[i]), kappa = R(t.kappa[i]), sqrtTr = m_sqrt(T / Tc);
            const R f = R(1) + kappa * (R(1) - sqrtTr);
            qv[i] = ab(f);
            dq[i] = -(f > 0 ? R(1) : R(-1)) * kappa / (R(2) * Tc * sqrtTr);
            b += x[i] * R(t.b[i]);
        }
        R pp = 0, dpp = 0;
        for (int i = 0; i < NS; ++i) {
            for (int j = 0; j < NS; ++j) {
                const R w = x[i] * x[j] * R(t.aij[i][j]);
                aa += w * qv[i] * qv[j];
                da += w * (dq[i] * qv[j] + qv[i] * dq[j]);
            }
            const R aci = R(t.aij[c][i]);
            pp += x[i] * aci * qv[i] * qv[c];
            dpp += x[i] * aci * (dq[c] * qv[i] + qv[c] * dq[i]);
        }
        const R Rg = R(kR), sq2 = R(kSqrt2), RT = Rg * T, V = s.V, bc = R(t.b[c]);
        const R vmb = V - b, den = V * V + R(2) * V * b - b * b;
        const R L = m_log((V + (R(1) + sq2) * b) / (V + (R(1) - sq2) * b)) / (R(2) * sq2 * b);
        const R pn = RT / vmb + RT * bc / (vmb * vmb) - R(2) * pp / den + R(2) * aa * vmb * bc / (den * den);
        const R vbar = -pn / s.dpdV;
        R cpR, hRT, sR;
        nasa(region(t, c, T), T, cpR, hRT, sR);
        const R dU = R(2) * (T * dpp - pp) * L + (T * da - aa) * (bc / b) * (V / den - L);
        const R ebar = RT * hRT - RT + dU + (T * s.dpdT - p) * vbar;
        const R acc = R(t.aij[c][c]) * qv[c] * qv[c];
        const R d2A = R(2) * RT * bc / vmb + RT * bc * bc / (vmb * vmb) - R(2) * acc * L -
                      R(4) * pp * bc * (V / (b * den) - L / b) -
                      aa * bc * bc * (-R(2) * V / (b * b * den) - R(2) * V * vmb / (b * den * den) + R(2) * L / (b * b));
        const R nmu = RT / x[c] + d2A + pn * pn / s.dpdV;  // n_total * ∂mu_c/∂n_c|_{T,p}
        const R Wc = R(t.W[c]);
        out.vG = vbar / Wc;
        out.eG = ebar / Wc;
        out.hG = (ebar + p * vbar) / RT;
        out.zG = p * vbar / RT;
        out.muN = nmu * W / (RT * mg * Wc);  // n_total = mg / W, dn_c = dm / W_c
    }

    // Analytic Jacobian of the one-liquid residual (mixed flash): the volume
    // and energy rows are the (p, T) derivatives the evaluation accumulates
    // plus the change of one unit of the condensable's mass from gas to
    // liquid (the gas loses its partial volume and energy), and the
    // chemical-potential row follows from d(mu/RT)/d ln p = p v / RT and
    // d(mu/RT)/d ln T = -h/RT (Gibbs-Helmholtz) with partial molar v and h,
    // and, in the liquid amount, from the gas composition's change (muN; a
    // pure gas's mu does not depend on the phase amounts).
    RD void analyticJacobian(const E& v, R scale, R (&jac)[3][3]) const {
        const R qc = q[mod().condensable];
        jac[0][0] = v.Vp * v.p;
        jac[0][1] = v.VT * v.T;
        jac[0][2] = qc * (v.parts.vL - v.parts.vG);
        jac[1][0] = v.Ep * v.p / scale;
        jac[1][1] = v.ET * v.T / scale;
        jac[1][2] = qc * (v.parts.eL - v.parts.eG) / scale;
        jac[2][0] = v.parts.zL - v.parts.zG;
        jac[2][1] = -(v.parts.hL - v.parts.hG);
        jac[2][2] = NS == 1 ? R(0) : qc * v.parts.muN;
    }
    RD bool analyticJacobianFinite(const E& v, R scale, R (&jac)[3][3]) const {
        analyticJacobian(v, scale, jac);
        bool ok = true;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) ok = ok && fin(jac[i][j]);
        return ok;
    }

    // Gas + liquid mixture at common (p, T) (reactiveDeviceFlash.h:208-262).
    // Inlined into the solvers at inlining level 2 (see RAMD_SEARCH_INLINE).
    // vapour >= 0 (CAP, vapour coordinates): the condensable's gas inventory
    // itself rather than q - mass0.
    RD bool evaluate(R mass0, R p, R T, E& v, bool virtualLiquids, R vapour = R(-1)) {
        if constexpr (P::single ? RAMD_SEARCH_INLINE >= 2 : RAMD_FLASH_INLINE >= 2)
            return evaluateBody(mass0, p, T, v, virtualLiquids, vapour);
        else return evaluateCall(mass0, p, T, v, virtualLiquids, vapour);
    }
    RAMD_NOINLINE bool evaluateCall(R mass0, R p, R T, E& v, bool virtualLiquids, R vapour) {
        return evaluateBody(mass0, p, T, v, virtualLiquids, vapour);
    }
    __device__ __attribute__((always_inline)) inline bool evaluateBody(R mass0, R p, R T, E& v, bool virtualLiquids,
                                                                      R vapour) {
        zeroEval(v);
        if (!bounds(p, T)) return false;
        if (NL > 0 && mass0 > 0 && cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual)
            return false;
        v.p = p; v.T = T; v.rho = density();
        if (!(v.rho > 0)) return false;
        R gas[NS];
        for (int k = 0; k < NS; ++k) {
            if (!fin(q[k]) || q[k] < 0) return false;
            gas[k] = q[k];
        }
        const int c = NL ? mod().condensable : 0;
        if (NL) {
            if (!fin(mass0) || mass0 < 0 || mass0 > q[c]) return false;
            if constexpr (CAP) gas[c] = vapour >= 0 ? vapour : gas[c] - mass0;
            else gas[c] = gas[c] - mass0;
            if (!fin(gas[c]) || gas[c] < 0 || gas[c] > q[c]) return false;
            if constexpr (RJ) v.ref.gasCondensable = gas[c];
            if constexpr (CAP) {
                v.capGasCondensable = gas[c];
                v.capExplicitGas = vapour >= 0;
            }
        }
        R mg = 0;
        for (int k = 0; k < NS; ++k) mg += gas[k];
        v.gasMass = mg;
        const R pg = gasPressure(p), pl = liquidPressure(p);
        if (mg > 0) {
            for (int k = 0; k < NS; ++k) gas[k] /= mg;
            S prop;
            R mu[NS];
            if (phaseStatus(-1, pg, T, gas, prop, virtualLiquids ? mu : nullptr) != kOk) return false;
            if constexpr (RJ) v.ref.gasMolarVolume = prop.V;
            accumulate(v, mg, prop, pg);
            if constexpr (E::kOutputs) {
                v.alphaGas = mg / prop.rho;
                v.rhoGas = prop.rho;
            }
            if constexpr (A) {
                if constexpr (NS == 1) recordPhase(prop, p, T, v.parts.vG, v.parts.eG, v.parts.hG, v.parts.zG);
                else gasPartials(p, T, mg, gas, prop, v.parts);
            }
            if (NL && virtualLiquids) v.muGas = mu[c] / (R(8314.46261815324) * T);
        }
        if (NL) {
            v.liqMass = mass0;
            if (mass0 > 0 || (virtualLiquids && q[c] > 0)) {
                S prop;
                R mu[1];
                int st;
                if (cacheStatus >= 0 && p == cacheP && T == cacheT && virtualLiquids == cacheVirtual) {
                    ++n.phases;
                    st = cacheStatus;
                } else {
                    st = phaseStatus(0, pl, T, nullptr, prop, virtualLiquids ? mu : nullptr);
                    if (unavailable(st) && T >= liquidTmin && T < liquidTc * (R(1) - R(1e-10))) {
                        cacheP = p; cacheT = T; cacheVirtual = virtualLiquids; cacheStatus = st;
                    }
                }
                if (st != kOk) {
                    if (mass0 > 0 || !unavailable(st)) return false;
                    v.muLiquid = R(Real<R>::big) * R(2);  // +infinity
                } else {
                    if (virtualLiquids) v.muLiquid = mu[0] / (R(8314.46261815324) * T);
                    if constexpr (A) recordPhase(prop, p, T, v.parts.vL, v.parts.eL, v.parts.hL, v.parts.zL);
                    if constexpr (RJ) {
                        v.ref.liquidHbar = prop.h * mod().liquid.W[0];
                        v.ref.liquidVbar = prop.V;
                        v.ref.liquidMask |= 1;
                    }
                    if (mass0 > 0) {
                        accumulate(v, mass0, prop, pl);
                        if constexpr (E::kOutputs) {
                            v.rhoLiq = prop.rho;
                            v.alphaLiq = mass0 / prop.rho;
                            v.active |= 1;
                        }
                    }
                }
            }
        }
        const R cp = v.cpDensity / v.rho;
        const R cv = (v.cpDensity + T * v.VT * v.VT / v.Vp) / v.rho;
        if constexpr (E::kOutputs) {
            v.e = v.energy / v.rho; v.entropy = v.entropyDensity / v.rho; v.cp = cp;
            v.cv = cv;
        }
        if (!fin(cv) || cv <= 0) return false;
        const R specific = v.volume / v.rho, vp = v.Vp / v.rho, vt = v.VT / v.rho;
        const R adiabatic = vp + T * vt * vt / cp;
        if (!(adiabatic < 0)) return false;
        const R cf = m_sqrt(-specific * specific / adiabatic);
        if constexpr (E::kOutputs) {
            v.cf = cf;
            v.ceq = cf;
        }
        return fin(cf) && cf > 0;
    }
    RD R energyScale(const E& v) const { return hi(R(1e5) * v.rho, v.cpDensity * v.T); }

    // Single-phase gas energy residual at fixed rho (reactiveDeviceFlash.h:264-275).
    // check = false skips the gas-root consistency test (the refinement's
    // polish, whose final evaluation at (p, T) repeats it through the volume
    // residual)
    RAMD_NOINLINE bool gasProbe(R T, R& F, R& cv, R& p, R* cp, bool check = true) {
        const R rho = density();
        R Y[NS];
        for (int k = 0; k < NS; ++k) Y[k] = q[k] / rho;
        ++n.phases;
        S prop;
        {
            if (!fin(T) || !fin(rho) || T <= 0 || rho <= 0) return false;
            R x[NS], W = 0;
            if (massToMole(mod().gas, Y, x, W) != kOk) return false;
            M mx;
            if (mixing(mod().gas, T, x, mx) != kOk) return false;
            if (stateTrho(mod().gas, T, rho, x, W, mx, prop, (R*)nullptr) != kOk) return false;
            if (finishState<NS>(gasHeuristic(x, W, mx), T, prop, (const R*)nullptr) != kOk) return false;
        }
        const R gasP = prop.p;
        if constexpr (CAP) p = gasP + this->color * this->jump;
        else p = gasP + R(0) * R(0);
        if (!bounds(p, T)) return false;
        if (check) {
            S gas;
            if (phaseStatus(-1, gasP, T, Y, gas, nullptr) != kOk || ab(gas.rho - rho) > R(P::densityCheck) * rho)
                return false;
        }
        F = rho * prop.e - energy;
        cv = rho * prop.cv;
        if (cp) *cp = prop.cp;
        return fin(F) && fin(cv) && cv > 0;
    }
    // Bracketed scalar Newton in T (reactiveDeviceFlash.h:276-299).
    RAMD_NOINLINE bool scalarGas(E& result, int polish = 0) {
        const R rho = density();
        R T = clamp(guessT, Tmin, Tmax), F, cv, p, cp;
        if (!gasProbe(T, F, cv, p, &cp)) return false;
        const R scale = hi(R(1e5) * rho, rho * cp * T);
        R lower = T, upper = T, fl = F, fh = F;
        if (ab(F) > etol * scale) {
            R radius = hi(R(1), R(.01) * T);
            for (int j = 0; j < 24 && (fl > 0 || fh < 0); ++j, radius *= R(2)) {
                if (fl > 0) {
                    R candidate = hi(Tmin, T - radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        // FP32 search: the bracket end itself (clamped at Tmin, or reached
                        // by the halving) was probed already, with this result
                        if (P::single && candidate == lower) { found = true; break; }
                        if (gasProbe(candidate, value, dcv, dp, nullptr)) { fl = value; lower = candidate; found = true; }
                        else candidate = R(.5) * (lower + candidate);
                    }
                    if (!found) return false;
                }
                if (fh < 0) {
                    R candidate = lo(Tmax, T + radius), value, dcv, dp;
                    bool found = false;
                    for (int retry = 0; retry < 8 && !found; ++retry) {
                        if (P::single && candidate == upper) { found = true; break; }
                        if (gasProbe(candidate, value, dcv, dp, nullptr)) { fh = value; upper = candidate; found = true; }
                        else candidate = R(.5) * (upper + candidate);
                    }
                    if (!found) return false;
                }
                if (ab(fl) <= etol * scale) { T = lower; if (!gasProbe(T, F, cv, p, nullptr)) return false; break; }
                if (ab(fh) <= etol * scale) { T = upper; if (!gasProbe(T, F, cv, p, nullptr)) return false; break; }
                if (lower == Tmin && upper == Tmax) break;
            }
            if (!(ab(F) <= etol * scale || (fl <= 0 && fh >= 0 && upper > lower))) return false;
        }
        int it = 0;
        for (; it < 70 && ab(F) > etol * scale; ++it) {
            if (F > 0) upper = T; else lower = T;
            const R next = T - F / cv;
            T = fin(next) && next > lower && next < upper ? next : R(.5) * (lower + upper);
            if (!gasProbe(T, F, cv, p, nullptr)) return false;
        }
        if (ab(F) > etol * scale) return false;
        // refinement only: further Newton steps toward 1e-3 of the tolerance;
        // a failed probe keeps the converged point
        for (int extra = 0; extra < polish && ab(F) > R(1e-3) * etol * scale; ++extra) {
            if (F > 0) upper = T; else lower = T;
            const R next = T - F / cv;
            const R trialT = fin(next) && next > lower && next < upper ? next : R(.5) * (lower + upper);
            R trialF, trialCv, trialP;
            if (!gasProbe(trialT, trialF, trialCv, trialP, nullptr)) break;
            T = trialT; F = trialF; cv = trialCv; p = trialP;
        }
        if (!evaluate(R(0), p, T, result, false)) return false;
        const R volRes = ab(result.volume - R(1)), enRes = ab(result.energy - energy) / energyScale(result);
        if constexpr (E::kOutputs) {
            result.volRes = volRes;
            result.enRes = enRes;
            result.iterations = it;
        }
        return volRes <= vtol && enRes <= etol;
    }
    // Fixed split, 2x2 Newton in (ln p, ln T) with line search (reactiveDeviceFlash.h:300-326).
    // polish > 0 (FP64 refinement only): up to that many further Newton steps
    // once the tolerances hold, while the residual is above 1e-3 of them; a
    // rejected polishing step keeps the converged point.
    RAMD_NOINLINE bool frozen(R mass0, E& value, int stallLimit, R mass1 = R(0), int polish = 0) {
        stalled = false;
        if (!fin(energy)) return false;
        if (mod().scalarRecovery && mass0 == 0 && mass1 == 0 && scalarGas(value, polish)) return true;
        R p = clamp(guessP, pmin, pmax), T = clamp(guessT, Tmin, Tmax);
        if (!evaluate(mass0, p, T, value, false)) return false;
        int stalls = 0;
        bool polishing = false;
        for (int it = 0; it < 70; ++it) {
            const R scale = energyScale(value), rv = value.volume - R(1), re = (value.energy - energy) / scale;
            if constexpr (E::kOutputs) {
                value.volRes = ab(rv);
                value.enRes = ab(re);
                value.iterations = it;
            }
            if (ab(rv) <= vtol && ab(re) <= etol) {
                if (polish > 0 && !(ab(rv) <= R(1e-3) * vtol && ab(re) <= R(1e-3) * etol)) {
                    --polish;
                    polishing = true;
                } else {
                    return true;
                }
            }
            R jac[3][3] = {};
            jac[0][0] = value.Vp * p; jac[0][1] = value.VT * T;
            jac[1][0] = value.Ep * p / scale; jac[1][1] = value.ET * T / scale;
            R rhs[2] = {-rv, -re}, step[2];
            if (!solveN<2>(jac, rhs, step)) return polishing;
            const R limit = hi(R(1), hi(ab(step[0]) / R(.7), ab(step[1]) / R(.2)));
            step[0] /= limit; step[1] /= limit;
            const R previous = hi(ab(rv), ab(re));
            bool accepted = false;
            if constexpr (WAVE) {
                // lane j tries fraction 2^-j, the loop's j-th (fractions above 1e-9: j < 30)
                const int lane = int(threadIdx.x & 63u);
                const R fraction = R(::ldexp(1.0, -lane));
                const Counts before = n;
                E trial;
                zeroEval(trial);
                R pp = 0, tt = 0;
                bool ok = false;
                if (fraction > R(1e-9)) {
                    pp = p * m_exp(fraction * step[0]);
                    tt = T * m_exp(fraction * step[1]);
                    if (evaluate(mass0, pp, tt, trial, false)) {
                        const R next = hi(ab(trial.volume - R(1)), ab(trial.energy - energy) / scale);
                        ok = next < previous * (R(1) - R(1e-4) * fraction) || next < lo(vtol, etol);
                    }
                }
                const unsigned long long won = __ballot(ok);
                waveCounts(before, won ? __ffsll(won) - 1 : 63);
                if (won) {
                    const int w = __ffsll(won) - 1;
                    value = waveBroadcast(trial, w);
                    p = waveBroadcast(pp, w);
                    T = waveBroadcast(tt, w);
                    accepted = true;
                    stalls = R(::ldexp(1.0, -w)) < R(1) / R(32) ? stalls + 1 : 0;
                }
            } else
            for (R fraction = 1; fraction > R(1e-9); fraction *= R(.5)) {
                E trial;
                const R pp = p * m_exp(fraction * step[0]), tt = T * m_exp(fraction * step[1]);
                // FP32 search: a fraction too small to move the point (nor
                // any smaller one) would repeat the rejected evaluation, so
                // the halving stops there (13 % of the search at 160^3); the
                // FP64 flashes keep every halving, as the reference counts them
                if constexpr (P::single)
                    if (pp == p && tt == T) break;
                if (evaluate(mass0, pp, tt, trial, false)) {
                    const R next = hi(ab(trial.volume - R(1)), ab(trial.energy - energy) / scale);
                    if (next < previous * (R(1) - R(1e-4) * fraction) || next < lo(vtol, etol)) {
                        value = trial; p = pp; T = tt; accepted = true;
                        stalls = fraction < R(1) / R(32) ? stalls + 1 : 0;
                        break;
                    }
                }
            }
            if (!accepted) {
                stalled = P::single && previous <= R(P::handOver);
                return polishing;
            }
            if (stallLimit > 0 && stalls >= stallLimit &&
                !(ab(value.volume - R(1)) <= vtol && ab(value.energy - energy) / energyScale(value) <= etol))
                return false;
        }
        return polishing;
    }
    // Residual [V-1, (E-e)/scale, mu_l/RT - mu_g/RT] in x = [ln p, ln T, beta]
    // (reactiveDeviceFlash.h:327-337).
    // In vapour coordinates (CAP, this->adaptive) x[2] = ln(vapour fraction) <= 0.
    RAMD_NOINLINE bool residual(const R* x, R scale, R* f, E* out) {
        ++n.residuals;
        R mass0, vapour = R(-1);
        if constexpr (CAP) {
            if (this->adaptive) {
                if (!fin(x[2]) || x[2] > 0) return false;
                vapour = m_exp(x[2]) * q[mod().condensable];
                if (!(vapour > 0)) return false;
                mass0 = -m_expm1(x[2]) * q[mod().condensable];
            } else {
                if (!(x[2] >= 0 && x[2] < 1)) return false;
                mass0 = x[2] * q[mod().condensable];
            }
        } else {
            if (!(x[2] >= 0 && x[2] < 1)) return false;
            mass0 = x[2] * q[mod().condensable];
        }
        E v;
        if (!evaluate(mass0, m_exp(x[0]), m_exp(x[1]), v, true, vapour) || !(v.gasMass > 0)) return false;
        f[0] = v.volume - R(1);
        f[1] = (v.energy - energy) / scale;
        f[2] = v.muLiquid - v.muGas;
        if (!fin(f[2])) return false;
        if (out) *out = v;
        return true;
    }
    // Central finite differences with step halving (reactiveDeviceFlash.h:338-373).
    RAMD_NOINLINE bool jacobian(const R* x, R scale, const R* f, R (&jac)[3][3]) {
        ++n.fd;
        for (int j = 0; j < 3; ++j) {
            R h = j < 2 ? R(P::hLog) : R(P::hFraction);
            bool success = false;
            for (int attempt = 0; attempt < 12 && !success; ++attempt, h *= R(.5)) {
                R xp[3], xm[3], fp[3] = {}, fm[3] = {};
                for (int i = 0; i < 3; ++i) xp[i] = xm[i] = x[i];
                xp[j] += h;
                xm[j] -= h;
                if constexpr (CAP) {
                    if (this->adaptive) {  // a step lost to rounding moves by one ulp instead
                        if (xp[j] == x[j]) xp[j] = ::nextafter(x[j], double(HUGE_VAL));
                        if (xm[j] == x[j]) xm[j] = ::nextafter(x[j], -double(HUGE_VAL));
                    }
                }
                const bool plus = residual(xp, scale, fp, nullptr), minus = residual(xm, scale, fm, nullptr);
                if (!plus && !minus) continue;
                success = true;
                bool weighted = false;
                if constexpr (CAP) weighted = this->adaptive;
                if (weighted) {  // one-sided differences weighted by the actual steps (CAP: FP64)
                    const R hp = xp[j] - x[j], hm = x[j] - xm[j];
                    R maximum = 0;
                    for (int i = 0; i < 3; ++i) {
                        const R v = plus && minus ? (hm / (hp + hm)) * ((fp[i] - f[i]) / hp) +
                                                        (hp / (hp + hm)) * ((f[i] - fm[i]) / hm)
                                    : plus ? (fp[i] - f[i]) / hp : (f[i] - fm[i]) / hm;
                        jac[i][j] = v;
                        maximum = hi(maximum, ab(v));
                        success = success && fin(v);
                    }
                    success = success && maximum > 0;
                    continue;
                }
                for (int i = 0; i < 3; ++i) {
                    const R v = plus && minus ? (fp[i] - fm[i]) / (R(2) * h) : plus ? (fp[i] - f[i]) / h : (f[i] - fm[i]) / h;
                    jac[i][j] = v;
                    success = success && fin(v);
                }
            }
            if (!success) return false;
        }
        return true;
    }
    // The reference's analytic Jacobian at a residual point
    // (reactiveDeviceFlash.h:338-352): from the point's own properties
    // (buildCached, reactiveDeviceFlashJacobian.h:535-640), else from fresh
    // gas and liquid states at its (p, T) (build, :436-530); false (the
    // caller then differences) when neither can be built or the mode is off.
    RD bool referenceJacobian(const R* x, R scale, const E& v, R (&jac)[3][3]) {
        if constexpr (RJ) {
            if (jacobianCached(x, scale, v, jac) || jacobianFresh(x, v.p, v.T, scale, v, jac)) {
                ++n.analytic;
                return true;
            }
        }
        return false;
    }
    RAMD_NOINLINE bool jacobianCached(const R* x, R scale, const E& v, R (&jac)[3][3]) {
        if constexpr (RJ) {
            const double p = v.p, T = v.T, gm = v.gasMass, V = v.ref.gasMolarVolume;
            if (!(p > 0) || !(T > 0) || !(scale > 0) || !(gm > 0) || !(V > 0) || !fin(p) || !fin(T) || !fin(scale) ||
                !fin(v.Vp) || !fin(v.VT) || !fin(v.Ep) || !fin(v.ET) || !fin(gm) || !fin(V))
                return false;
            const int c = mod().condensable;
            double Y[NS];
            for (int k = 0; k < NS; ++k) {
                if (!fin(q[k]) || q[k] < 0) return false;
                Y[k] = q[k];
            }
            const double gc = v.ref.gasCondensable;
            if (!fin(gc) || gc < 0 || gc > q[c]) return false;
            Y[c] = gc;
            // the phase pressures of a curved interface (zero: flat)
            double pg = p, pl = p;
            bool vapourCoordinates = false;
            if constexpr (CAP) {
                const double gp = gasPressure(p), lp = liquidPressure(p);
                pg = gp == 0 ? p : gp;
                pl = lp == 0 ? p : lp;
                if (!(pg > 0) || !(pl > 0) || !fin(pg) || !fin(pl) || v.capExplicitGas != this->adaptive) return false;
                vapourCoordinates = this->adaptive;
            }
            if (!(q[c] > 0)) return false;
            double expectedGas, dLiquid;
            if (vapourCoordinates) {
                if (!(x[2] <= 0) || !fin(x[2])) return false;
                expectedGas = m_exp(x[2]) * q[c];
                dLiquid = -expectedGas;
            } else {
                if (!(x[2] >= 0 && x[2] < 1)) return false;
                const double product = x[2] * q[c];  // the residual's own two roundings (:35-47)
                expectedGas = q[c] - product;
                dLiquid = q[c];
            }
            if (!(expectedGas > 0) || gc != expectedGas) return false;
            const double Hl = v.ref.liquidHbar, Vl = v.ref.liquidVbar;
            if (!(v.ref.liquidMask & 1) || !fin(Hl) || !fin(Vl) || !(Vl > 0)) return false;
            double reconstructed = 0;
            for (int k = 0; k < NS; ++k) reconstructed += Y[k];
            if (reconstructed != gm) return false;
            for (int k = 0; k < NS; ++k) Y[k] /= gm;
            GasPartials g;
            if (!refActiveGasDerivatives(mod().gas, T, pg, V, Y, gm, c, g)) return false;
            const double Wk = mod().gas.W[c], Wl = mod().liquid.W[0];
            const double liquidUbar = Hl - pl * Vl;
            jac[0][0] = v.Vp * p;
            jac[0][1] = v.VT * T;
            jac[1][0] = v.Ep * p / scale;
            jac[1][1] = v.ET * T / scale;
            jac[0][2] = dLiquid * (Vl / Wl - g.vbar / Wk);
            jac[1][2] = dLiquid * (liquidUbar / Wl - g.ubar / Wk) / scale;
            jac[2][0] = p * (Vl - g.vbar) / (kR * T);
            jac[2][1] = -(Hl - g.hbar) / (kR * T);
            jac[2][2] = g.dpsi * dLiquid / Wk;
            bool ok = true;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ok = ok && fin(jac[i][j]);
            return ok;
        }
        return false;
    }
    RAMD_NOINLINE bool jacobianFresh(const R* x, double p, double T, R scale, const E& v, R (&jac)[3][3]) {
        if constexpr (RJ) {
            if (!(p > 0) || !(T > 0) || !(scale > 0)) return false;
            double pg = p, pl = p;
            bool vapourCoordinates = false;
            if constexpr (CAP) {
                const double gp = gasPressure(p), lp = liquidPressure(p);
                pg = gp == 0 ? p : gp;
                pl = lp == 0 ? p : lp;
                if (!(pg > 0) || !(pl > 0) || !fin(pg) || !fin(pl)) return false;
                vapourCoordinates = this->adaptive;
            }
            const int c = mod().condensable;
            double gm[NS], Y[NS], total = 0;
            for (int k = 0; k < NS; ++k) gm[k] = q[k];
            if (!(q[c] > 0)) return false;
            double dLiquid;
            if (vapourCoordinates) {
                if (!(x[2] <= 0) || !fin(x[2])) return false;
                gm[c] = m_exp(x[2]) * q[c];
                if (!(gm[c] > 0)) return false;
                dLiquid = -gm[c];
            } else {
                if (!(x[2] >= 0 && x[2] < 1)) return false;
                gm[c] -= x[2] * q[c];
                dLiquid = q[c];
            }
            for (int k = 0; k < NS; ++k) total += gm[k];
            if (!(total > 0)) return false;
            for (int k = 0; k < NS; ++k) Y[k] = gm[k] / total;
            S gas;
            ++n.phases;
            if (gasTP(T, pg, Y, gas, nullptr) != kOk) return false;
            double hbar[NS], ubar[NS], vbar[NS], xm[NS], W = 0;
            if (!refPhasePartials(mod().gas, T, Y, gas, hbar, ubar, vbar, xm, W)) return false;
            const double totalMoles = total / W;
            S liquid;
            ++n.phases;
            if (liquidTP(T, pl, liquid, nullptr) != kOk) return false;
            double dpsi;
            if (!refGasCompositionPartial(mod().gas, T, xm, totalMoles, gas, c, dpsi)) return false;
            const double Wk = mod().gas.W[c], Wl = mod().liquid.W[0];
            jac[0][0] = v.Vp * p;
            jac[0][1] = v.VT * T;
            jac[1][0] = v.Ep * p / scale;
            jac[1][1] = v.ET * T / scale;
            jac[0][2] = dLiquid * (1.0 / liquid.rho - vbar[c] / Wk);
            jac[1][2] = dLiquid * (liquid.e - ubar[c] / Wk) / scale;
            const double liquidHbar = liquid.h * Wl;
            jac[2][0] = p * (liquid.V - vbar[c]) / (kR * T);
            jac[2][1] = -(liquidHbar - hbar[c]) / (kR * T);
            jac[2][2] = dpsi * dLiquid / Wk;
            bool ok = true;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ok = ok && fin(jac[i][j]);
            return ok;
        }
        return false;
    }
    // Start of an active-set solve: x = [ln p, ln T, seed fraction] and the
    // energy scale (reactiveDeviceFlash.h:374-390).
    // False (CAP, vapour coordinates) where the start has no positive vapour.
    RD bool activeStart(R seed, bool restart, R restartT, R* x, R& scale) const {
        const R rho = density();
        const R gT = restart ? restartT : guessT;
        scale = hi(R(1e5) * rho, hi(ab(energy), rho * R(2000) * hi(gT, Tmin)));
        x[0] = m_log(clamp(guessP, pmin, pmax));
        x[1] = m_log(clamp(gT, Tmin, Tmax));
        const R previousFraction = guessLiquid / q[mod().condensable];
        x[2] = seed < 0 ? clamp(previousFraction, R(P::fractionEdge), R(1) - R(P::fractionEdge)) : seed;
        if constexpr (CAP) {
            if (this->adaptive) {  // ln(vapour fraction); seed -3 (or almost no vapour) starts from a trace (:383-388)
                double vapour = seed < 0 ? 1 - previousFraction : 1 - seed;
                if (seed == R(-3) || (seed < 0 && vapour < 1e-6)) {
                    double other = 0;
                    for (int k = 0; k < NS; ++k)
                        if (k != mod().condensable) other += q[k];
                    vapour = other > 0 ? lo(1e-6, .01 * (other / q[mod().condensable])) : 1e-8;
                }
                if (!(vapour > 0 && fin(vapour))) return false;
                x[2] = m_log(lo(1., vapour));
            }
        }
        return true;
    }
    // Damped Newton on the one-liquid active set (reactiveDeviceFlash.h:391-448)
    // from x.  polish > 0 takes up to that many further Newton steps once the
    // tolerances hold, while the residual is above 1e-3 of them (the FP64
    // refinement of a single-precision candidate).  In single precision the
    // converged point is returned without the equilibrium sound speed.
    // Converged point: the equilibrium sound speed from the Jacobian's response
    // to a volume/energy perturbation (reactiveDeviceFlash.h:400-409).
    RAMD_NOINLINE bool activeFinish(const R* x, R scale, const R* f, E& value) {
        const R rho = density();
        R jac[3][3];
        if constexpr (A && NS == 1 && !P::single) {
            // mixed refinement: the exact (analytic) Jacobian at the converged
            // point, where the reference differences the residual (its c_eq
            // carries the differencing error, a few 1e-7)
            analyticJacobian(value, scale, jac);
        } else if constexpr (A && !P::single) {
            if (!analyticJacobianFinite(value, scale, jac) && !jacobian(x, scale, f, jac)) return false;
        } else {
            if (!referenceJacobian(x, scale, value, jac) && !jacobian(x, scale, f, jac)) return false;
        }
        R rhs[3] = {-R(1) / rho, value.p / (rho * scale), R(0)}, response[3];
        if (!solveN<3>(jac, rhs, response)) return false;
        const R squared = value.p * response[0];
        if (!fin(squared) || squared <= 0) return false;
        if constexpr (E::kOutputs) {  // the FP64 flashes only
            value.ceq = m_sqrt(squared);
            return value.ceq <= value.cf * (R(1) + R(2e-3));
        } else {
            return false;
        }
    }
    RAMD_NOINLINE bool activeNewton(R* x, R scale, E& value, int polish) {
        stalled = false;
        R f[3];
        if (!residual(x, scale, f, &value)) return false;
        bool polishing = false;  // converged already; a failed polishing step keeps the point
        for (int it = 0; it < 90; ++it) {
            const R chemical = hi(R(0), ab(f[2]));
            R jac[3][3];
            if (ab(f[0]) <= vtol && ab(f[1]) <= etol && chemical <= mutol) {
                if constexpr (E::kOutputs) {
                    value.volRes = ab(f[0]);
                    value.enRes = ab(f[1]);
                    value.chemRes = chemical;
                    value.iterations = it;
                }
                if (P::single) return true;
                const bool tight = ab(f[0]) <= R(1e-3) * vtol && ab(f[1]) <= R(1e-3) * etol &&
                                   chemical <= R(1e-3) * mutol;
                if (polish > 0 && !tight) {
                    --polish;
                    polishing = true;
                } else {
                    return activeFinish(x, scale, f, value);
                }
            }
            if constexpr (A && NS == 1) {
                analyticJacobian(value, scale, jac);
            } else if constexpr (A) {
                if (!analyticJacobianFinite(value, scale, jac) && !jacobian(x, scale, f, jac))
                    return polishing && activeFinish(x, scale, f, value);
            } else {
                if (!referenceJacobian(x, scale, value, jac) && !jacobian(x, scale, f, jac))
                    return polishing && activeFinish(x, scale, f, value);
            }
            R rhs[3] = {-f[0], -f[1], -f[2]}, step[3];
            if (!solveN<3>(jac, rhs, step)) return polishing && activeFinish(x, scale, f, value);
            R limiter = hi(R(1), hi(ab(step[0]) / R(.7), ab(step[1]) / R(.18)));
            bool vapourCoordinates = false;
            if constexpr (CAP) vapourCoordinates = this->adaptive;
            limiter = hi(limiter, ab(step[2]) / (vapourCoordinates ? R(2) : R(.3)));
            for (int j = 0; j < 3; ++j) step[j] /= limiter;
            bool accepted = false;
            const R previous = hi(hi(hi(R(0), ab(f[0])), ab(f[1])), ab(f[2]));
            if constexpr (WAVE) {
                const auto outsideAt = [vapourCoordinates](R y) {
                    return vapourCoordinates ? (!fin(y) || y > 0) : !(y >= 0 && y < 1);
                };
                // the loop evaluates the indices from `first` on: 0, or the
                // end of the rejected prefix of the partition coordinate
                // (bisected as below); lane j tries index first + j, and
                // every lane evaluates (lanes past the loop's end repeat its
                // last index and are not counted)
                int first = 0;
                if (outsideAt(x[2] + step[2])) {
                    int lower = 0, upper = 30;
                    while (upper - lower > 1) {
                        const int mid = (lower + upper) / 2;
                        if (outsideAt(x[2] + R(::ldexp(1.0, -mid)) * step[2])) lower = mid;
                        else upper = mid;
                    }
                    n.residuals += uint32_t(upper);
                    first = upper;
                }
                if (first < 30) {
                    const int lane = int(threadIdx.x & 63u);
                    const int index = first + lane < 30 ? first + lane : 29;
                    const R fraction = R(::ldexp(1.0, -index));
                    const Counts before = n;
                    R tx[3], tf[3];
                    for (int j = 0; j < 3; ++j) tx[j] = x[j] + fraction * step[j];
                    E trial;
                    zeroEval(trial);
                    const bool evaluated = residual(tx, scale, tf, &trial);
                    const bool ok = first + lane < 30 && evaluated &&
                                    hi(hi(hi(R(0), ab(tf[0])), ab(tf[1])), ab(tf[2])) <
                                        previous * (R(1) - R(1e-4) * fraction);
                    const unsigned long long won = __ballot(ok);
                    waveCounts(before, won ? __ffsll(won) - 1 : 29 - first);
                    if (won) {
                        const int w = __ffsll(won) - 1;
                        for (int j = 0; j < 3; ++j) {
                            x[j] = waveBroadcast(tx[j], w);
                            f[j] = waveBroadcast(tf[j], w);
                        }
                        value = waveBroadcast(trial, w);
                        accepted = true;
                    }
                }
            } else {
            R fraction = 1;
            for (int index = 0; index < 30; ++index, fraction *= R(.5)) {
                R tx[3], tf[3];
                for (int j = 0; j < 3; ++j) tx[j] = x[j] + fraction * step[j];
                // rejected partition coordinates (outside [0, 1), or positive or
                // non-finite in vapour coordinates)
                const auto outside = [vapourCoordinates](R y) {
                    return vapourCoordinates ? (!fin(y) || y > 0) : !(y >= 0 && y < 1);
                };
                if (outside(tx[2])) {
                    // rejected fractions form a prefix; bisect to its end (:419-438)
                    int lower = index, upper = 30;
                    while (upper - lower > 1) {
                        const int mid = (lower + upper) / 2;
                        const R y = x[2] + R(::ldexp(1.0, -mid)) * step[2];
                        if (outside(y)) lower = mid; else upper = mid;
                    }
                    n.residuals += uint32_t(upper - index);
                    if (upper == 30) break;
                    index = upper;
                    fraction = R(::ldexp(1.0, -upper));
                    for (int j = 0; j < 3; ++j) tx[j] = x[j] + fraction * step[j];
                }
                if constexpr (P::single)  // FP32 search: no movement now or at any smaller fraction
                    if (tx[0] == x[0] && tx[1] == x[1] && tx[2] == x[2]) break;
                E trial;
                if (residual(tx, scale, tf, &trial) &&
                    hi(hi(hi(R(0), ab(tf[0])), ab(tf[1])), ab(tf[2])) < previous * (R(1) - R(1e-4) * fraction)) {
                    for (int j = 0; j < 3; ++j) { x[j] = tx[j]; f[j] = tf[j]; }
                    value = trial;
                    accepted = true;
                    break;
                }
            }
            }
            if (!accepted) {
                stalled = P::single && previous <= R(P::handOver);
                return polishing && activeFinish(x, scale, f, value);
            }
        }
        return polishing && activeFinish(x, scale, f, value);
    }
    RAMD_NOINLINE bool activeFlash(R seed, E& value, bool restart, R restartT) {
        ++n.candidates;
        R x[3], scale;
        if (!activeStart(seed, restart, restartT, x, scale)) return false;
        return activeNewton(x, scale, value, 0);
    }
    // Stability (reactiveDeviceFlash.h:449-485) for at most one liquid.
    // evaluated: value itself came from the evaluation the test repeats (an
    // active-set candidate: same liquid mass, p, T, virtual liquids), so its
    // chemical potentials are the comparison's, bit for bit
    RAMD_NOINLINE bool stable(E& value, bool evaluated = false) {
        const R p = value.p, T = value.T;
        if (value.gasMass > 0) {
            E comparison;
            R vapour = R(-1);  // the candidate's own gas inventory in vapour coordinates (:452)
            if constexpr (CAP)
                if (value.capExplicitGas) vapour = value.capGasCondensable;
            if (evaluated) comparison = value;
            else if (!evaluate(value.liqMass, p, T, comparison, true, vapour)) return false;
            if (!NL || q[mod().condensable] == 0) return true;
            const R affinity = comparison.muLiquid - comparison.muGas;
            if (value.liqMass > 0) {
                if (!fin(affinity) || ab(affinity) > mutol * R(2)) return false;
            } else if (affinity < -mutol) return false;
            return true;
        }
        if (!NL || !(value.liqMass > 0)) return false;
        S prop;
        R mu[1];
        if (phaseStatus(0, liquidPressure(p), T, nullptr, prop, mu) != kOk) return false;
        const R muL = mu[0] / (R(8314.46261815324) * T);
        // Tangent-plane distance of the pure vapour of the condensable (:474-485, n = 1).
        R mole[NS], Y[NS], sum = 0;
        for (int k = 0; k < NS; ++k) mole[k] = 0;
        mole[mod().condensable] = 1;
        for (int k = 0; k < NS; ++k) { Y[k] = mole[k] * R(mod().weights[k]); sum += Y[k]; }
        for (int k = 0; k < NS; ++k) Y[k] /= sum;
        S gas;
        const int st = phaseStatus(-1, gasPressure(p), T, Y, gas, nullptr);
        R tangent;
        if (st != kOk) tangent = unavailable(st) ? R(Real<R>::big) * R(2) : R(0) / R(0);
        else {
            const R reference = R(0) + mole[mod().condensable] * muL;
            R W = 0;
            for (int k = 0; k < NS; ++k) W += mole[k] * R(mod().weights[k]);
            tangent = (gas.h - T * gas.s) * W / (R(8314.46261815324) * T) - reference;
        }
        return tangent >= -mutol;
    }
    RD void consider(E& candidate, E& best, bool& have, bool evaluated = false) {  // :486-489
        if (stable(candidate, evaluated)) {
            ++n.stable;
            if (!have || candidate.entropyDensity > best.entropyDensity) { best = candidate; have = true; }
        } else {
            ++n.failures;
        }
    }
    RD R restartTemperature(int t) const {  // reactiveFlashSeeds.h:13-16
        const R lower = hi(Tmin, liquidTmin), upper = lo(Tmax, liquidTc * (R(1) - R(1e-10)));
        const R frac = t == 0 ? R(0.02) : t == 1 ? R(0.25) : t == 2 ? R(0.5) : R(0.75);
        return lower + (upper - lower) * frac;
    }
    RD bool restartRange() const {
        return lo(Tmax, liquidTc * (R(1) - R(1e-10))) > hi(Tmin, liquidTmin);
    }
    // Candidate search (reactiveDeviceFlash.h:490-543); in vapour
    // coordinates (CAP) without the stable-gas pruning, and with the
    // trace-vapour start after the five seeds.
    RAMD_NOINLINE bool search(E& best) {
        bool have = false;
        E candidate;
        bool gasStable = false;
        bool vapourCoordinates = false;
        if constexpr (CAP) vapourCoordinates = this->adaptive;
        const bool prunable = mod().stableGasPrune != 0 && !vapourCoordinates;
        if (frozen(R(0), candidate, prunable ? 4 : 0)) {
            const uint32_t before = n.stable;
            consider(candidate, best, have);
            gasStable = n.stable > before && candidate.gasMass > 0;
        } else {
            ++n.failures;
        }
        if (!NL) return have;
        const bool prune = prunable && gasStable;
        const int c = mod().condensable;
        R other = 0;
        for (int k = 0; k < NS; ++k)
            if (k != c) other += q[k];
        if (other == 0) {
            if (frozen(q[c], candidate, 0)) consider(candidate, best, have);
            else ++n.failures;
        }
        const R seeds[5] = {R(-1), R(.5), R(.95), R(.1), R(.9999)};
        if (!prune && q[c] > 0) {
            for (int j = 0; j < 5; ++j) {
                if (activeFlash(seeds[j], candidate, false, R(0))) consider(candidate, best, have);
                else ++n.failures;
            }
            if (vapourCoordinates) {  // the trace-vapour start (:532)
                if (activeFlash(R(-3), candidate, false, R(0))) consider(candidate, best, have);
                else ++n.failures;
            }
        }
        if (!have && q[c] > 0 && restartRange())
            for (int t = 0; t < 4; ++t) {
                const R restartT = restartTemperature(t);
                for (int j = 0; j < 5; ++j) {
                    if (activeFlash(seeds[j], candidate, true, restartT)) consider(candidate, best, have);
                    else ++n.failures;
                }
            }
        return have;
    }
    // :544-560.  CAP: with boundaryRecovery, or for a capillary equilibrium
    // cell (capillaryBoundary), a result whose liquid holds nearly all of the
    // condensable with gas present, or no result, runs the search again in
    // vapour coordinates and the higher entropy density wins.
    RD bool run(E& result, bool capillaryBoundary = false) {
        if (!fin(energy) || !(density() > 0)) return false;
        for (int k = 0; k < NS; ++k)
            if (!fin(q[k]) || q[k] < 0) return false;
        if constexpr (!CAP) {
            return search(result);
        } else {
            E reference;
            zeroEval(reference);
            const bool have = search(reference);
            if (!mod().boundaryRecovery && !capillaryBoundary) {
                if (have) result = reference;
                return have;
            }
            if (have) {
                const bool boundary =
                    NL && q[mod().condensable] > 0 && reference.gasMass > 0 && reference.liqMass / q[mod().condensable] > 1 - 1e-4;
                if (!boundary) {
                    result = reference;
                    return true;
                }
            }
            this->adaptive = true;
            E recovered;
            zeroEval(recovered);
            const bool found = search(recovered);
            this->adaptive = false;
            if (!found && !have) return false;
            result = have && (!found || reference.entropyDensity >= recovered.entropyDensity) ? reference : recovered;
            return true;
        }
    }

    // ---- single-precision search (R = float): the candidate search with the
    // converged candidates recorded instead of selected.  Stability and
    // entropy are judged here only to decide, as the reference does, whether
    // the restart temperatures run; the FP64 refinement judges them again.
    RD void note(int16_t kind, int16_t index, const R* x, bool dedupe) {
        CandidateSet& set = *record;
        if (dedupe)
            for (uint32_t i = 0; i < set.count; ++i) {
                const Candidate& o = set.c[i];
                if (o.kind == kind && ab(float(x[0]) - o.x[0]) <= 1e-4f && ab(float(x[1]) - o.x[1]) <= 1e-4f &&
                    ab(float(x[2]) - o.x[2]) <= 1e-4f)
                    return;
            }
        if (set.count == kMaxCandidates) {
            set.flags |= kOverflow;
            return;
        }
        Candidate& o = set.c[set.count++];
        o.x[0] = float(x[0]); o.x[1] = float(x[1]); o.x[2] = float(x[2]);
        o.kind = kind;
        o.index = index;
    }
    RAMD_NOINLINE void searchRecord(CandidateSet& set) {
        record = &set;
        set.count = 0;
        set.flags = 0;
        if (!fin(energy) || !(density() > 0)) return;
        for (int k = 0; k < NS; ++k)
            if (!fin(q[k]) || q[k] < 0) return;
        bool have = false;
        E candidate, best;
        zeroEval(best);
        bool gasStable = false;
        const bool prunable = mod().stableGasPrune != 0;
        const bool gasOk = frozen(R(0), candidate, prunable ? 4 : 0);
        if (gasOk || stalled) {
            const R x[3] = {m_log(candidate.p), m_log(candidate.T), R(0)};
            if (gasOk) {
                const uint32_t before = n.stable;
                consider(candidate, best, have);
                gasStable = n.stable > before && candidate.gasMass > 0;
            }
            note(kGasCandidate, 0, x, false);
        }
        if (!NL) return;
        const bool prune = prunable && gasStable;
        if (prune) set.flags |= kSeedsPruned;
        const int c = mod().condensable;
        R other = 0;
        for (int k = 0; k < NS; ++k)
            if (k != c) other += q[k];
        if (other == 0) {
            const bool ok = frozen(q[c], candidate, 0);
            if (ok || stalled) {
                const R x[3] = {m_log(candidate.p), m_log(candidate.T), R(0)};
                if (ok) consider(candidate, best, have);
                note(kLiquidCandidate, 0, x, false);
            }
        }
        const R seeds[5] = {R(-1), R(.5), R(.95), R(.1), R(.9999)};
        if (!prune && q[c] > 0)
            for (int j = 0; j < 5; ++j) {
                ++n.candidates;
                R x[3], scale;
                activeStart(seeds[j], false, R(0), x, scale);
                const bool ok = activeNewton(x, scale, candidate, 0);
                if (ok || stalled) {
                    if (ok) consider(candidate, best, have);
                    note(kSeedCandidate, int16_t(j), x, true);
                }
            }
        if (!have && q[c] > 0 && restartRange()) {
            set.flags |= kRestartsSearched;
            for (int t = 0; t < 4; ++t) {
                const R restartT = restartTemperature(t);
                for (int j = 0; j < 5; ++j) {
                    ++n.candidates;
                    R x[3], scale;
                    activeStart(seeds[j], true, restartT, x, scale);
                    const bool ok = activeNewton(x, scale, candidate, 0);
                    if (ok || stalled) {
                        if (ok) consider(candidate, best, have);
                        note(kRestartCandidate, int16_t(5 * t + j), x, true);
                    }
                }
            }
        }
    }

    // ---- FP64 refinement of a recorded search (R = double): each recorded
    // candidate is finished by the reference's own solver started from the
    // single-precision point (plus up to two polishing Newton steps for the
    // active-set candidates), then judged by the FP64 stability test and
    // maximum entropy in the reference's order.  Cells whose single-precision
    // decisions (seed pruning, whether restarts were needed) disagree with
    // FP64, or whose set overflowed, run the plain FP64 search.
    // Gas candidate of a scalar-recovery model: Newton in T at fixed density
    // from the FP32 temperature, to 1e-3 of the energy tolerance, with the
    // scalar recovery's acceptance test (reactiveDeviceFlash.h:276-299).  The
    // probes skip the gas-root test: the final evaluation at (p, T) takes the
    // gas root there, and its volume residual rejects a root of another density.
    RAMD_NOINLINE bool gasPolish(R T, E& result) {
        const R rho = density();
        T = clamp(T, Tmin, Tmax);
        R F, cv, p, cp;
        if (!gasProbe(T, F, cv, p, &cp, false)) return false;
        const R scale = hi(R(1e5) * rho, rho * cp * T);
        int it = 0;
        for (; it < 8 && ab(F) > R(1e-3) * etol * scale; ++it) {
            const R next = T - F / cv;
            if (!fin(next) || next < Tmin || next > Tmax) return false;
            T = next;
            if (!gasProbe(T, F, cv, p, nullptr, false)) return false;
        }
        if (ab(F) > etol * scale) return false;
        if (!evaluate(R(0), p, T, result, false)) return false;
        result.volRes = ab(result.volume - R(1));
        result.enRes = ab(result.energy - energy) / energyScale(result);
        result.iterations = it;
        return result.volRes <= vtol && result.enRes <= etol;
    }
    // A frozen candidate from its single-precision point: the gas candidate
    // of a scalar-recovery model by the polish above, any other by the
    // frozen solve started there.  When that fails, the candidate is solved
    // from the cell's own guess, as the reference starts.  (A failed polish
    // goes straight to that solve: on a 160^3 box, every cell whose polish
    // failed also failed the frozen solve from the single-precision point,
    // which cost about 700 phase evaluations each.)  deferred (when given):
    // that solve from the guess, which can take a thousand phase evaluations,
    // is left to the deferred refinement (*deferred = true, false returned),
    // where a wave shares each line search.
    RAMD_NOINLINE bool finishFrozen(R mass0, const Candidate& c, E& value, int stallLimit, bool* deferred = nullptr) {
        if (mass0 == 0 && mod().scalarRecovery) {
            if (gasPolish(m_exp(R(c.x[1])), value)) return true;
            if (deferred) { *deferred = true; return false; }
            return frozen(mass0, value, stallLimit);
        }
        const R p0 = guessP, T0 = guessT;
        guessP = m_exp(R(c.x[0]));
        guessT = m_exp(R(c.x[1]));
        bool ok = frozen(mass0, value, stallLimit, R(0), 2);
        guessP = p0;
        guessT = T0;
        if (!ok && deferred) { *deferred = true; return false; }
        if (!ok) ok = frozen(mass0, value, stallLimit);  // from the cell's own guess, as the reference starts
        return ok;
    }
    RAMD_NOINLINE bool finishActive(const Candidate& c, E& value) {
        ++n.candidates;
        const bool restart = c.kind == kRestartCandidate;
        const R seeds[5] = {R(-1), R(.5), R(.95), R(.1), R(.9999)};
        R x[3], scale;
        activeStart(seeds[c.index % 5], restart, restart ? restartTemperature(c.index / 5) : R(0), x, scale);
        x[0] = R(c.x[0]);
        x[1] = R(c.x[1]);
        x[2] = R(c.x[2]);
        return activeNewton(x, scale, value, 2);
    }
    // The winner of an FP64 search run inside the refinement (a deferred
    // cell, or the FP64 restarts) stops, like the reference's, at the first
    // iterate inside the tolerances; it is refined further from its own
    // point as every recorded candidate is (up to two more Newton steps
    // toward 1e-3 of the tolerances).  A polish that fails or lands on
    // another phase set keeps the searched result.
    RAMD_NOINLINE void polishBest(E& best) {
        const int c = mod().condensable;
        const R p0 = guessP, T0 = guessT;
        guessP = best.p;
        guessT = best.T;
        E value;
        bool ok;
        if (!(best.liqMass > 0)) {
            ok = frozen(R(0), value, 0, R(0), 2);
        } else if (!NL || best.liqMass >= q[c]) {
            ok = frozen(best.liqMass, value, 0, R(0), 2);
        } else {
            R x[3], scale;
            activeStart(R(.5), false, R(0), x, scale);
            x[0] = m_log(best.p);
            x[1] = m_log(best.T);
            x[2] = best.liqMass / q[c];
            ok = activeNewton(x, scale, value, 2);
        }
        guessP = p0;
        guessT = T0;
        if (ok && value.active == best.active) best = value;
    }
    RAMD_NOINLINE bool searchPolished(E& best) {
        if (!search(best)) return false;
        polishBest(best);
        return true;
    }
    // deferred (when given): instead of running the plain FP64 search or the
    // FP64 restart search inline, the refinement stops with *deferred = true;
    // the caller reruns such cells with deferred == nullptr in a launch of
    // their own, so the long searches do not hold up the other lanes of a wave.
    RAMD_NOINLINE bool refine(const CandidateSet& set, E& best, bool* deferred = nullptr) {
        if (!fin(energy) || !(density() > 0)) return false;
        for (int k = 0; k < NS; ++k)
            if (!fin(q[k]) || q[k] < 0) return false;
        if (set.flags & kOverflow) {
            if (deferred) { *deferred = true; return false; }
            return searchPolished(best);
        }
        bool have = false;
        E candidate;
        bool gasStable = false;
        const bool prunable = mod().stableGasPrune != 0;
        uint32_t i = 0;
        const bool recorded = i < set.count && set.c[i].kind == kGasCandidate;
        // With seed pruning the reference decides on its own all-gas Newton,
        // started from the cell's guess with stall limit 4
        // (reactiveDeviceFlash.h:515-523); a polish from the single-precision
        // point can converge where that solve crawls and fails, so the gas
        // candidate is that very solve here.
        // (Its polish, two more Newton steps once it has converged, does not
        // change whether it converges.)
        bool gasOk = false;
        if (prunable) gasOk = frozen(R(0), candidate, 4, R(0), 2);
        else if (recorded) gasOk = finishFrozen(R(0), set.c[i], candidate, 0, deferred);
        if (deferred && *deferred) return false;
        if (recorded) ++i;
        if (gasOk) {
            const uint32_t before = n.stable;
            consider(candidate, best, have);
            gasStable = n.stable > before && candidate.gasMass > 0;
        } else {
            ++n.failures;
        }
        if (!NL) return have;
        const bool prune = prunable && gasStable;
        if (!prune && (set.flags & kSeedsPruned)) {
            // the single-precision search pruned the seeds FP64 needs
            have = false;
            if (deferred) { *deferred = true; return false; }
            return searchPolished(best);
        }
        // FP64 prunes where the single-precision search did not: the
        // reference's search stops at the gas (and pure-liquid) candidate,
        // so the recorded seeds are skipped (have holds: the gas is stable)
        const int c = mod().condensable;
        if (i < set.count && set.c[i].kind == kLiquidCandidate) {
            if (finishFrozen(q[c], set.c[i], candidate, 0, deferred)) consider(candidate, best, have);
            else if (deferred && *deferred) return false;
            else ++n.failures;
            ++i;
        }
        if (prune) return have;
        for (; i < set.count && set.c[i].kind == kSeedCandidate; ++i) {
            if (finishActive(set.c[i], candidate)) consider(candidate, best, have, true);
            else ++n.failures;
        }
        if (!have && q[c] > 0 && restartRange()) {
            if (!(set.flags & kRestartsSearched)) {
                if (deferred) { *deferred = true; return false; }
                // the single-precision search found a stable candidate that FP64
                // rejects: search the restart temperatures in FP64
                const R seeds[5] = {R(-1), R(.5), R(.95), R(.1), R(.9999)};
                for (int t = 0; t < 4; ++t)
                    for (int j = 0; j < 5; ++j) {
                        if (activeFlash(seeds[j], candidate, true, restartTemperature(t)))
                            consider(candidate, best, have);
                        else ++n.failures;
                    }
                if (have) polishBest(best);
            } else {
                for (; i < set.count; ++i) {
                    if (finishActive(set.c[i], candidate)) consider(candidate, best, have, true);
                    else ++n.failures;
                }
            }
        }
        return have;
    }
};

}  // namespace ramd
#endif
