// SPDX-License-Identifier: GPL-3.0-or-later
// Adjoint of one SSPRK2 step's pieces, for gradients of a step's output with
// respect to the Peng-Robinson parameters and the reservoir (inlet) states.
//
// Transposed face flux: every stage computes rate_c = sum_f s_cf F_f A / V
// (reactiveTransportKernels.h:343-356, 502-526); given the adjoint of the
// rates, each cell gathers, over its six faces, the face-flux adjoint
// F_bar_f = sum over the face's cells of s_cf A / V rate_bar_c and the
// vector-Jacobian product of the HLL flux with respect to its own state (q,
// p, c_frozen), evaluated with forward-mode dual numbers.  Reservoir states
// collect the same products from their fixed-state faces.  No atomics: the
// result is deterministic.
//
// Flash backward by the implicit-function theorem at the converged state
// (the flash is not unrolled): with unknowns x = (ln p, ln T[, liquid mass])
// and residuals R = (V - 1, E - e[, mu_l/RT - mu_g/RT]) of the equilibrium the
// flash converged to (reactiveDeviceFlash.h:208-262, 327-337), outputs
// y = (p, c_frozen) and inputs (rho, e, ln a-scale, ln b-scale):
//   lambda = J^-T (dy/dx)^T y_bar,   in_bar = (dy/din)^T y_bar - (dR/din)^T lambda.
// The partial derivatives are central differences of the explicit phase
// evaluation at fixed x (no iteration inside: the cubic roots are polished to
// rounding), steps 1e-6 in ln p, ln T and (relative) rho, 1e-5 rho in the
// liquid mass (three-point one-sided next to its bounds), and the caller's
// relative step in the PR scales (four HEM handles built with a, b scaled by
// 1 +/- h).  One 3x3 solve per cell.
#include <hip/hip_runtime.h>
#include <cstring>
#include <memory>
#include <vector>

#include "hip_common.h"
#include "dual.h"
#include "flash_device.h"
#include "hem_internal.h"
#include "transport_internal.h"
#include "cartesian_kernels.h"
#include "reactive_amd.h"

namespace ramd {

// ---------------------------------------------------------------- face flux
// HLL face flux (per unit area) of the inviscid model, any arithmetic type.
template <int NS, int NV, class R>
__device__ inline void hllFlux(const R* qL, R pL, R cL, const R* qR, R pR, R cR, const double* n, double wave,
                               R* flux) {
    R rhoL = qL[0], rhoR = qR[0];
    for (int s = 1; s < NS; ++s) {
        rhoL = rhoL + qL[s];
        rhoR = rhoR + qR[s];
    }
    R unL = R(0.0), unR = R(0.0);
    for (int d = 0; d < 3; ++d) {
        unL = unL + qL[NS + d] / rhoL * n[d];
        unR = unR + qR[NS + d] / rhoR * n[d];
    }
    const R aL = wave * cL, aR = wave * cR;
    const R m1 = unL - aL < unR - aR ? unL - aL : unR - aR;
    const R m2 = unL + aL > unR + aR ? unL + aL : unR + aR;
    const R wl = 0.0 < m1 ? R(0.0) : m1, wr = 0.0 > m2 ? R(0.0) : m2;
    const R inv = 1.0 / (wr - wl);
    const R x = wr * (unL - wl) * inv, y = wl * (wr - unR) * inv;
    const R pm = (wr * pL - wl * pR) * inv, pe = (wr * pL * unL - wl * pR * unR) * inv;
    for (int v = 0; v < NV; ++v) {
        R f = x * qL[v] + y * qR[v];
        if (v >= NS && v < NS + 3) f = f + pm * n[v - NS];
        if (v == NS + 3) f = f + pe;
        flux[v] = f;
    }
}

// Adjoint of the rates with respect to each cell's own (q, p, c): one thread
// per cell of the box.  rateBar, qBar are SoA with the transport's stride.
template <int NS, int NV>
__global__ __launch_bounds__(256) void stageAdjointKernel(const BoxDev* __restrict__ boxp,
                                                          const double* __restrict__ q, size_t stride,
                                                          const double* __restrict__ pArr,
                                                          const double* __restrict__ cArr,
                                                          const double* __restrict__ rateBar,
                                                          double* __restrict__ qBar, double* __restrict__ pBar,
                                                          double* __restrict__ cBar) {
    using D = Dual<NV + 2>;
    const BoxDev& box = *boxp;
    const CartesianArgs& a = box.a;
    const long cells = long(box.cells);
    const long c = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const long nxy = long(a.nx) * a.ny;
    const int i = int(c % a.nx), j = int((c / a.nx) % a.ny), k = int(c / nxy);
    const double scale = a.area * a.invV;
    D own[NV], pOwn, cOwn;
    for (int v = 0; v < NV; ++v) own[v] = D::variable(q[size_t(v) * stride + c], v);
    pOwn = D::variable(pArr[c], NV);
    cOwn = D::variable(cArr[c], NV + 1);
    double bar[NV + 2];
    for (int t = 0; t < NV + 2; ++t) bar[t] = 0;
    for (int slot = 0; slot < 6; ++slot) {
        const int ax = slot >> 1, plus = slot & 1;
        const int coord = ax == 0 ? i : (ax == 1 ? j : k);
        const int ext = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
        const bool bnd = plus ? coord == ext - 1 : coord == 0;
        const long step = ax == 0 ? 1 : (ax == 1 ? long(a.nx) : nxy);
        const double* n = a.normal[slot][bnd ? 1 : 0];
        const bool selfLeft = bnd || plus;
        double fbar[NV];
        D other[NV], pOther, cOther;
        if (bnd) {
            for (int v = 0; v < NV; ++v) fbar[v] = -scale * rateBar[size_t(v) * stride + c];
            const int kind = boundaryKindOf(a, slot, i, j, k);
            if (kind >= 0) {  // reservoir: a constant here
                const long f = long(a.fixedBase) + kind;
                for (int v = 0; v < NV; ++v) other[v] = D(q[size_t(v) * stride + f]);
                pOther = D(pArr[f]);
                cOther = D(cArr[f]);
            } else {
                for (int v = 0; v < NV; ++v) other[v] = own[v];
                pOther = pOwn;
                cOther = cOwn;
                if (kind == -2) {  // slip wall: mirrored normal momentum
                    D normal(0.0);
                    for (int d = 0; d < 3; ++d) normal = normal + own[NS + d] * n[d];
                    for (int d = 0; d < 3; ++d) other[NS + d] = own[NS + d] - 2.0 * normal * n[d];
                }
            }
        } else {
            const long nb = plus ? c + step : c - step;
            const long owner = selfLeft ? c : nb, neighbour = selfLeft ? nb : c;
            for (int v = 0; v < NV; ++v)
                fbar[v] = scale * (rateBar[size_t(v) * stride + neighbour] - rateBar[size_t(v) * stride + owner]);
            for (int v = 0; v < NV; ++v) other[v] = D(q[size_t(v) * stride + nb]);
            pOther = D(pArr[nb]);
            cOther = D(cArr[nb]);
        }
        D flux[NV];
        if (selfLeft) hllFlux<NS, NV, D>(own, pOwn, cOwn, other, pOther, cOther, n, a.waveFactor, flux);
        else hllFlux<NS, NV, D>(other, pOther, cOther, own, pOwn, cOwn, n, a.waveFactor, flux);
        for (int v = 0; v < NV; ++v)
            for (int t = 0; t < NV + 2; ++t) bar[t] += fbar[v] * flux[v].d[t];
    }
    for (int v = 0; v < NV; ++v) qBar[size_t(v) * stride + c] = bar[v];
    pBar[c] = bar[NV];
    cBar[c] = bar[NV + 1];
}

// Reservoir adjoint, one thread per fixed-state face (listed in side-plane
// order): the face's product with respect to the reservoir state, per face;
// the host sums them per reservoir in list order.
template <int NS, int NV>
__global__ void fixedAdjointKernel(const BoxDev* __restrict__ boxp, const int4* __restrict__ list, int count,
                                   const double* __restrict__ q, size_t stride, const double* __restrict__ pArr,
                                   const double* __restrict__ cArr, const double* __restrict__ rateBar,
                                   double* __restrict__ out) {
    using D = Dual<NV + 2>;
    const int e = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= count) return;
    const CartesianArgs& a = boxp->a;
    const int4 face = list[e];  // cell, slot, reservoir, -
    const long c = face.x;
    const int slot = face.y;
    const long f = long(a.fixedBase) + face.z;
    const double* n = a.normal[slot][1];
    const double scale = a.area * a.invV;
    D own[NV], res[NV];
    for (int v = 0; v < NV; ++v) {
        own[v] = D(q[size_t(v) * stride + c]);
        res[v] = D::variable(q[size_t(v) * stride + f], v);
    }
    const D pRes = D::variable(pArr[f], NV), cRes = D::variable(cArr[f], NV + 1);
    D flux[NV];
    hllFlux<NS, NV, D>(own, D(pArr[c]), D(cArr[c]), res, pRes, cRes, n, a.waveFactor, flux);
    for (int t = 0; t < NV + 2; ++t) {
        double s = 0;
        for (int v = 0; v < NV; ++v) s += -scale * rateBar[size_t(v) * stride + c] * flux[v].d[t];
        out[size_t(e) * (NV + 2) + t] = s;
    }
}

// ---------------------------------------------------------------- flash
struct Residuals {
    double r[3], cf;
};

// Residuals and frozen sound speed of the equilibrium state x at fixed phase
// amounts (see the file comment); false if the evaluation fails.
template <int NL>
__device__ inline bool equilibrium(const HemModel<1>& model, double rho, double energy, int mode, const double* x,
                                   Residuals& out) {
    Counts n = {0, 0, 0, 0, 0, 0};
    Flash<1, NL> f(model, n);
    f.q[0] = rho;
    f.energy = energy;
    const double mass0 = mode == 2 ? x[2] : (mode == 1 ? rho : 0.0);
    Eval v;
    zeroEval(v);
    if (!f.evaluate(mass0, elem::exp(x[0]), elem::exp(x[1]), v, mode == 2)) return false;
    out.r[0] = v.volume - 1;
    out.r[1] = v.energy - energy;
    out.r[2] = mode == 2 ? v.muLiquid - v.muGas : 0.0;
    out.cf = v.cf;
    return ::isfinite(out.cf) && out.cf > 0;
}

// Solves A^T y = b for n <= 3 by Gaussian elimination with partial pivoting.
__device__ inline bool solveTransposed(int n, const double (&A)[3][3], const double* b, double* y) {
    double M[3][4];
    for (int r = 0; r < n; ++r) {
        for (int col = 0; col < n; ++col) M[r][col] = A[col][r];
        M[r][n] = b[r];
    }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (::fabs(M[r][col]) > ::fabs(M[piv][col])) piv = r;
        if (!(::fabs(M[piv][col]) > 0)) return false;
        if (piv != col)
            for (int t = 0; t <= n; ++t) {
                const double s = M[col][t];
                M[col][t] = M[piv][t];
                M[piv][t] = s;
            }
        for (int r = col + 1; r < n; ++r) {
            const double g = M[r][col] / M[col][col];
            for (int t = col; t <= n; ++t) M[r][t] -= g * M[col][t];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = M[r][n];
        for (int t = r + 1; t < n; ++t) s -= M[r][t] * y[t];
        y[r] = s / M[r][r];
    }
    for (int r = 0; r < n; ++r)
        if (!::isfinite(y[r])) return false;
    return true;
}

// One thread per cell; q and qBar cell-major (count x nv), thetaBar
// (count x 2: d/d ln a-scale, d/d ln b-scale).  models: base, a+, a-, b+, b-.
template <int NL>
__global__ __launch_bounds__(64) void flashAdjointKernel(const HemModel<1>* const* __restrict__ models,
                                                         double relStep, const double* __restrict__ q, int nv,
                                                         const ThermoOut* __restrict__ states,
                                                         const int* __restrict__ success, size_t count,
                                                         const double* __restrict__ pBar,
                                                         const double* __restrict__ cBar, double* __restrict__ qBar,
                                                         double* __restrict__ thetaBar,
                                                         unsigned long long* __restrict__ skipped) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= count) return;
    const double* qc = q + c * nv;
    double* out = qBar + c * nv;
    for (int v = 0; v < nv; ++v) out[v] = 0;
    thetaBar[2 * c] = thetaBar[2 * c + 1] = 0;
    if (success && !success[c]) return;  // a failed flash kept its seed: no dependence on the inputs
    const ThermoOut s = states[c];
    const double rho = qc[0], m0 = qc[1], m1 = qc[2], m2 = qc[3], E = qc[4];
    const double energy = E - .5 * (m0 * m0 + m1 * m1 + m2 * m2) / rho;
    const double ml = NL ? s.liquidMass[0] : 0.0;
    const int mode = NL && ml > 0 && s.gasMass > 0 ? 2 : (NL && ml > 0 ? 1 : 0);
    const int n = mode == 2 ? 3 : 2;
    const double x[3] = {elem::log(s.p), elem::log(s.T), ml};
    // directions 0..2: x; 3: rho; 4, 5: ln a-scale, ln b-scale (scaled models)
    auto eval = [&](int dir, double delta, Residuals& out) -> bool {
        double xs[3] = {x[0], x[1], x[2]};
        double r = rho;
        if (dir < 3) xs[dir] += delta;
        if (dir == 3) r += delta;
        return equilibrium<NL>(*models[0], r, energy, mode, xs, out);
    };
    Residuals base;
    bool ok = eval(0, 0.0, base);
    // derivative of (R, c_f) along one direction: central where the step keeps
    // inside the open range (lo, hi), else three-point one-sided (second order)
    auto derive = [&](int dir, double h, double at, double lo, double hi, double* dR, double& dc) -> bool {
        Residuals a, b;
        if (at - h > lo && at + h < hi) {
            if (!eval(dir, h, a) || !eval(dir, -h, b)) return false;
            for (int r = 0; r < n; ++r) dR[r] = (a.r[r] - b.r[r]) / (2 * h);
            dc = (a.cf - b.cf) / (2 * h);
        } else {
            const double sgn = at + 2 * h < hi ? 1.0 : -1.0;
            if (sgn < 0 && !(at - 2 * h > lo)) return false;
            if (!eval(dir, sgn * h, a) || !eval(dir, 2 * sgn * h, b)) return false;
            for (int r = 0; r < n; ++r) dR[r] = sgn * (-3 * base.r[r] + 4 * a.r[r] - b.r[r]) / (2 * h);
            dc = sgn * (-3 * base.cf + 4 * a.cf - b.cf) / (2 * h);
        }
        return true;
    };
    const double big = 1e300;
    double J[3][3] = {}, cx[3] = {};
    for (int d = 0; d < n && ok; ++d) {
        double col[3];
        // liquid mass: R is linear in it, c_f rational; keep 0 < m < rho
        ok = d < 2 ? derive(d, 1e-6, x[d], -big, big, col, cx[d]) : derive(2, 1e-5 * rho, ml, 0.0, rho, col, cx[2]);
        for (int r = 0; r < n; ++r) J[r][d] = col[r];
    }
    // input directions: rho (a two-phase cell keeps rho above its liquid mass), ln a-scale, ln b-scale
    double Rin[3][3] = {}, cin[3] = {};
    if (ok) {
        double col[3];
        ok = derive(3, 1e-6 * rho, rho, mode == 2 ? ml : -big, big, col, cin[0]);
        for (int r = 0; r < n; ++r) Rin[r][0] = col[r];
    }
    for (int sdir = 0; sdir < 2 && ok; ++sdir) {
        Residuals rp, rm;
        ok = equilibrium<NL>(*models[1 + 2 * sdir], rho, energy, mode, x, rp) &&
             equilibrium<NL>(*models[2 + 2 * sdir], rho, energy, mode, x, rm);
        for (int r = 0; r < n; ++r) Rin[r][1 + sdir] = (rp.r[r] - rm.r[r]) / (2 * relStep);
        cin[1 + sdir] = (rp.cf - rm.cf) / (2 * relStep);
    }
    double lambda[3] = {};
    if (ok) {
        // (dy/dx)^T y_bar with y = (p = exp(x0), c_frozen)
        double g[3];
        for (int d = 0; d < n; ++d) g[d] = cx[d] * cBar[c];
        g[0] += s.p * pBar[c];
        ok = solveTransposed(n, J, g, lambda);
    }
    if (!ok) {
        atomicAdd(skipped, 1ull);
        return;
    }
    // in_bar = (dy/din)^T y_bar - (dR/din)^T lambda; dR1/de = -1
    double inBar[3];
    for (int t = 0; t < 3; ++t) {
        double sum = cin[t] * cBar[c];
        for (int r = 0; r < n; ++r) sum -= Rin[r][t] * lambda[r];
        inBar[t] = sum;
    }
    const double eBar = lambda[1];
    // e = E - |m|^2 / (2 rho)
    const double kin = .5 * (m0 * m0 + m1 * m1 + m2 * m2) / (rho * rho);
    out[0] = inBar[0] + eBar * kin;
    out[1] = -eBar * m0 / rho;
    out[2] = -eBar * m1 / rho;
    out[3] = -eBar * m2 / rho;
    out[4] = eBar;
    thetaBar[2 * c] = inBar[1];
    thetaBar[2 * c + 1] = inBar[2];
}

template <class T>
struct DeviceBuffer {
    T* p = nullptr;
    explicit DeviceBuffer(size_t n) { RAMD_HIP(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T))); }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Neumaier sum in index order (deterministic).
inline double orderedSum(const std::vector<double>& x, size_t stride, size_t offset) {
    double s = 0, comp = 0;
    for (size_t i = offset; i < x.size(); i += stride) {
        const double t = s + x[i];
        comp += std::fabs(s) >= std::fabs(x[i]) ? (s - t) + x[i] : (x[i] - t) + s;
        s = t;
    }
    return s + comp;
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int reactive_amd_flash_adjoint(void* hem, void* const* scaled, double relStep, const double* q, int nv,
                               const ReactiveThermoState* states, const int* success, size_t count,
                               const double* pBar, const double* cBar, double* qBar, double* thetaBar,
                               uint64_t* skipped, char* error, size_t errorSize) {
    try {
        require(hem && scaled && q && states && pBar && cBar && qBar && thetaBar, "Null flash adjoint argument");
        HemHandle& h = *static_cast<HemHandle*>(hem);
        require(h.ns == 1 && nv == 5, "The flash adjoint covers one gas species (nv = 5)");
        require(relStep > 0 && relStep < 1e-2, "Invalid relative step of the PR scales");
        const HemModel<1>* models[5] = {static_cast<const HemModel<1>*>(h.model)};
        for (int i = 0; i < 4; ++i) {
            require(scaled[i], "Null scaled HEM handle");
            const HemHandle& o = *static_cast<const HemHandle*>(scaled[i]);
            require(o.ns == 1 && o.nl == h.nl && o.device == h.device, "Scaled HEM handles differ from the base");
            models[1 + i] = static_cast<const HemModel<1>*>(o.model);
        }
        RAMD_HIP(hipSetDevice(h.device));
        const size_t n = std::max<size_t>(count, 1);
        DeviceBuffer<const HemModel<1>*> dm(5);
        DeviceBuffer<double> dq(n * nv), dpb(n), dcb(n), dqb(n * nv), dtb(2 * n);
        DeviceBuffer<ThermoOut> ds(n);
        DeviceBuffer<int> dsucc(n);
        DeviceBuffer<unsigned long long> dskip(1);
        hipStream_t st = h.stream;
        RAMD_HIP(hipMemcpyAsync(dm.p, models, sizeof(models), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(dq.p, q, count * nv * sizeof(double), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(ds.p, states, count * sizeof(ThermoOut), hipMemcpyHostToDevice, st));
        if (success) RAMD_HIP(hipMemcpyAsync(dsucc.p, success, count * sizeof(int), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(dpb.p, pBar, count * sizeof(double), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(dcb.p, cBar, count * sizeof(double), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemsetAsync(dskip.p, 0, sizeof(unsigned long long), st));
        if (count) {
            const unsigned blocks = unsigned((count + 63) / 64);
            if (h.nl)
                hipLaunchKernelGGL(flashAdjointKernel<1>, dim3(blocks), dim3(64), 0, st, dm.p, relStep, dq.p, nv,
                                   ds.p, success ? dsucc.p : nullptr, count, dpb.p, dcb.p, dqb.p, dtb.p, dskip.p);
            else
                hipLaunchKernelGGL(flashAdjointKernel<0>, dim3(blocks), dim3(64), 0, st, dm.p, relStep, dq.p, nv,
                                   ds.p, success ? dsucc.p : nullptr, count, dpb.p, dcb.p, dqb.p, dtb.p, dskip.p);
            RAMD_HIP(hipGetLastError());
        }
        std::vector<double> tb(2 * count);
        unsigned long long skip = 0;
        RAMD_HIP(hipMemcpyAsync(qBar, dqb.p, count * nv * sizeof(double), hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipMemcpyAsync(tb.data(), dtb.p, tb.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipMemcpyAsync(&skip, dskip.p, sizeof(skip), hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipStreamSynchronize(st));
        thetaBar[0] = orderedSum(tb, 2, 0);
        thetaBar[1] = orderedSum(tb, 2, 1);
        if (skipped) *skipped = skip;
        return 0;
    } catch (const std::exception& e) {
        writeError(error, errorSize, e.what());
        return 1;
    }
}

int reactive_amd_transport_stage_adjoint(void* transport, const double* q, const ReactiveTransportState* states,
                                         const double* rateBar, double* qBar, double* pBar, double* cBar,
                                         double* fixedBar) {
    if (!transport) return -1;
    Transport& t = *static_cast<Transport*>(transport);
    std::unique_lock<std::mutex> lock(t.entry, std::try_to_lock);
    if (!lock.owns_lock()) return 2;
    try {
        require(q && states && rateBar && qBar && pBar && cBar, "Null stage adjoint argument");
        require(!t.viscous() && !t.capillary,
                "The stage adjoint covers the inviscid flux (viscosity, conductivity, WALE and capillary model off)");
        require(t.cartesian.valid, "The stage adjoint needs a structured-box transport mesh");
        require(t.scheme == kHll, "The stage adjoint covers the HLL flux");
        require(t.ns == 1 && t.nv == 5, "The stage adjoint covers one species (nv = 5)");
        require(t.cfg.fixed == 0 || fixedBar, "Null reservoir adjoint output");
        RAMD_HIP(hipSetDevice(t.device));
        const size_t cells = t.cfg.cells, all = t.all;
        const int nv = t.nv;
        // SoA q, p, c over cells + reservoirs (reservoirs from the handle's own state)
        std::vector<double> soa(all * nv), p(all), c(all), rb(all * nv, 0.0);
        for (size_t i = 0; i < cells; ++i) {
            for (int k = 0; k < nv; ++k) {
                soa[k * all + i] = q[i * nv + k];
                rb[k * all + i] = rateBar[i * nv + k];
            }
            p[i] = states[i].p;
            c[i] = states[i].sound;
        }
        const hipStream_t st = t.stream;
        DeviceBuffer<double> dq(all * nv), dp(all), dc(all), drb(all * nv), dqb(all * nv), dpb(all), dcb(all);
        for (int k = 0; k < nv; ++k) {
            RAMD_HIP(hipMemcpyAsync(dq.p + k * all, soa.data() + k * all, cells * sizeof(double),
                                    hipMemcpyHostToDevice, st));
            if (t.cfg.fixed)
                RAMD_HIP(hipMemcpyAsync(dq.p + k * all + cells, t.q + k * all + cells, t.cfg.fixed * sizeof(double),
                                        hipMemcpyDeviceToDevice, st));
        }
        std::vector<ReactiveTransportState> fixedStates(t.cfg.fixed);
        if (t.cfg.fixed)
            RAMD_HIP(hipMemcpyAsync(fixedStates.data(), t.state + cells, t.cfg.fixed * sizeof(ReactiveTransportState),
                                    hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipStreamSynchronize(st));
        for (size_t j = 0; j < t.cfg.fixed; ++j) {
            p[cells + j] = fixedStates[j].p;
            c[cells + j] = fixedStates[j].sound;
        }
        RAMD_HIP(hipMemcpyAsync(dp.p, p.data(), all * sizeof(double), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(dc.p, c.data(), all * sizeof(double), hipMemcpyHostToDevice, st));
        RAMD_HIP(hipMemcpyAsync(drb.p, rb.data(), all * nv * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((stageAdjointKernel<1, 5>), dim3(unsigned((cells + 255) / 256)), dim3(256), 0, st,
                           t.cartesian.box, dq.p, all, dp.p, dc.p, drb.p, dqb.p, dpb.p, dcb.p);
        RAMD_HIP(hipGetLastError());
        std::vector<double> qb(all * nv), pb(all), cb(all);
        RAMD_HIP(hipMemcpyAsync(qb.data(), dqb.p, all * nv * sizeof(double), hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipMemcpyAsync(pb.data(), dpb.p, all * sizeof(double), hipMemcpyDeviceToHost, st));
        RAMD_HIP(hipMemcpyAsync(cb.data(), dcb.p, all * sizeof(double), hipMemcpyDeviceToHost, st));
        // reservoir faces, in side-plane order
        std::vector<int4> list;
        if (t.cfg.fixed && t.cartesian.boundaryKind) {
            const CartesianInfo& ci = t.cartesian;
            std::vector<int32_t> kinds(size_t(boundaryPlanes(ci.nx, ci.ny, ci.nz)));
            RAMD_HIP(hipMemcpyAsync(kinds.data(), ci.boundaryKind, kinds.size() * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, st));
            RAMD_HIP(hipStreamSynchronize(st));
            for (int slot = 0; slot < 6; ++slot) {
                const int ax = slot >> 1;
                const int e0 = ax == 0 ? ci.ny : ci.nx, e1 = ax == 2 ? ci.ny : ci.nz;
                for (int w = 0; w < e1; ++w)
                    for (int u = 0; u < e0; ++u) {
                        int i = 0, j = 0, k = 0;
                        const int fixedCoord = (slot & 1) ? (ax == 0 ? ci.nx : ax == 1 ? ci.ny : ci.nz) - 1 : 0;
                        if (ax == 0) { i = fixedCoord; j = u; k = w; }
                        else if (ax == 1) { i = u; j = fixedCoord; k = w; }
                        else { i = u; j = w; k = fixedCoord; }
                        const int32_t kind = kinds[size_t(boundaryPlane(ci.nx, ci.ny, ci.nz, slot, i, j, k))];
                        if (kind >= 0) {
                            const long cell = long(i) + long(ci.nx) * (j + long(ci.ny) * k);
                            list.push_back(make_int4(int(cell), slot, kind, 0));
                        }
                    }
            }
        }
        std::vector<double> perFace(list.size() * (nv + 2));
        if (!list.empty()) {
            DeviceBuffer<int4> dl(list.size());
            DeviceBuffer<double> dout(perFace.size());
            RAMD_HIP(hipMemcpyAsync(dl.p, list.data(), list.size() * sizeof(int4), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL((fixedAdjointKernel<1, 5>), dim3(unsigned((list.size() + 63) / 64)), dim3(64), 0, st,
                               t.cartesian.box, dl.p, int(list.size()), dq.p, all, dp.p, dc.p, drb.p, dout.p);
            RAMD_HIP(hipGetLastError());
            RAMD_HIP(hipMemcpyAsync(perFace.data(), dout.p, perFace.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    st));
            RAMD_HIP(hipStreamSynchronize(st));
        }
        RAMD_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < cells; ++i) {
            for (int k = 0; k < nv; ++k) qBar[i * nv + k] = qb[k * all + i];
            pBar[i] = pb[i];
            cBar[i] = cb[i];
        }
        if (t.cfg.fixed) {
            std::vector<double> sum(t.cfg.fixed * (nv + 2), 0.0), comp(sum.size(), 0.0);
            for (size_t e = 0; e < list.size(); ++e)
                for (int m = 0; m < nv + 2; ++m) {
                    double& s = sum[size_t(list[e].z) * (nv + 2) + m];
                    double& cc = comp[size_t(list[e].z) * (nv + 2) + m];
                    const double x = perFace[e * (nv + 2) + m], tt = s + x;
                    cc += std::fabs(s) >= std::fabs(x) ? (s - tt) + x : (x - tt) + s;
                    s = tt;
                }
            for (size_t m = 0; m < sum.size(); ++m) fixedBar[m] = sum[m] + comp[m];
        }
        t.error[0] = 0;
        return 0;
    } catch (const std::exception& e) {
        std::snprintf(t.error, sizeof(t.error), "%s", e.what());
        return 1;
    }
}

}  // extern "C"
