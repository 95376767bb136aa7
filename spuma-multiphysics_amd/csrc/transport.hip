// SPDX-License-Identifier: GPL-3.0-or-later
// Conservative transport on gfx950: first-order HLL face flux, SSPRK2 update
// and CFL reduction behind the reactive_transport_* ABI
// (src/reactiveTransport/reactiveTransport.h:48-137).
//
// Data layout in HBM: conserved q is structure-of-arrays, q[k*(cells+fixed)+c]
// (the reference's qi(), reactiveTransportKernels.h:85), so every kernel
// streams each variable with unit stride.  Face geometry is compacted to
// 64-byte records (owner, resolved right cell, kind, normal, area, distance,
// ownerWeight), the cell
// face lists are a CSR of signed face ids in the reference's order (faces
// sorted by id within a cell, owner side negative), which fixes the FP64
// summation order of every cell rate.
//
// Kernels (one launch each, no host round trip between them):
//   primitives   rho = sum q_k, u = (rho u)/rho               (Cells)
//   gradient     per-cell Gauss velocity gradient [+ WALE nu_t] (Gradients, WaleViscosities)
//   hll          per-face HLL coefficients [+ viscous record]  (Faces)
//   advance      per (cell, variable) CSR gather + SSPRK2     (Rhs + Advance)
//   boundary     per-variable boundary flux, fixed 256-lane tree (BoundaryPartial/Finish)
//   faceSpeed / cellStep / min-reduction                       (FaceSpeeds, Step, Minimum)
// Structured Cartesian boxes take the fused stage kernel of cartesian.hip.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hip_common.h"
#include "transport_internal.h"
#include "flux_device.h"
#include "reactive_amd.h"
#include "reactive_transport.h"

namespace ramd {

static_assert(sizeof(ReactiveTransportFace) == 80 && sizeof(ReactiveTransportState) == 56, "transport ABI");
static_assert(sizeof(ReactiveTransportConfig) == 88 && sizeof(ReactiveTransportOptionsV2) == 88, "transport ABI");
static_assert(sizeof(ReactiveTransportOptionsV21) == 120 && sizeof(ReactiveTransportToken) == 24, "transport ABI");
static_assert(sizeof(ReactiveTransportPrimitive) == 32, "transport ABI");

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// ---------------------------------------------------------------- kernels
__global__ void primitivesKernel(const double* __restrict__ q, size_t all, int ns, double* __restrict__ prim) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= all) return;
    double rho = 0;
    for (int k = 0; k < ns; ++k) rho += q[size_t(k) * all + c];
    prim[c] = rho;
    for (int d = 0; d < 3; ++d) prim[size_t(d + 1) * all + c] = q[size_t(ns + d) * all + c] / rho;
}

// velocities() (reactiveTransportKernels.h:145-149): the owner and (mirrored
// for slip walls) right velocities and their normal components.
__device__ inline void normalVelocities(const FaceDev& f, const double* __restrict__ prim, size_t all, double* ul,
                                        double* ur, double& unL, double& unR) {
    double dot = 0;
    for (int d = 0; d < 3; ++d) {
        ul[d] = prim[size_t(d + 1) * all + f.owner];
        ur[d] = prim[size_t(d + 1) * all + f.right];
        dot += ur[d] * f.n[d];
    }
    if (f.kind == 1)
        for (int d = 0; d < 3; ++d) ur[d] -= 2 * dot * f.n[d];
    unL = 0;
    unR = 0;
    for (int d = 0; d < 3; ++d) {
        unL += ul[d] * f.n[d];
        unR += ur[d] * f.n[d];
    }
}

// ---- viscous stress, conduction and the WALE closure ----------------------
// Parameters of the viscous part of the face flux: the cell velocity
// gradients, the WALE eddy viscosities (null without the closure) and the
// constant viscosity and conductivity.
struct ViscousArgs {
    const double* gradient;
    const double* nut;
    double viscosity, conductivity;
};

// Gradients (reactiveTransportKernels.h:290-301): Gauss sum over the cell's
// faces in CSR order of the interpolated face velocity, ownerWeight on the
// owner side.  With the WALE closure (Wale) the same thread goes on to the
// cell's eddy viscosity (WaleViscosities, :303-313); a cell the closure
// rejects sets *error.
__device__ inline bool waleEddyViscosity(const double* g, double delta, double cw, double& nut);

template <bool Wale>
__global__ __launch_bounds__(256) void gradientKernel(const FaceDev* __restrict__ faces,
                                                      const uint32_t* __restrict__ row,
                                                      const int32_t* __restrict__ incidence,
                                                      const double* __restrict__ invV, const double* __restrict__ prim,
                                                      size_t all, size_t cells, double* __restrict__ gradient,
                                                      const double* __restrict__ filterWidth, double cw,
                                                      const ReactiveTransportState* __restrict__ state,
                                                      double* __restrict__ nut, unsigned* __restrict__ error) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    double g[9];
    for (int j = 0; j < 9; ++j) g[j] = 0;
    for (uint32_t e = row[c]; e < row[c + 1]; ++e) {
        const int32_t entry = incidence[e];
        const FaceDev f = faces[entry < 0 ? -entry - 1 : entry - 1];
        double ul[3], ur[3], unL, unR;
        normalVelocities(f, prim, all, ul, ur, unL, unR);
        const double scale = (entry < 0 ? 1.0 : -1.0) * f.area * invV[c];
        for (int i = 0; i < 3; ++i) {
            const double face = f.weight * ul[i] + (1 - f.weight) * ur[i];
            for (int d = 0; d < 3; ++d) g[3 * i + d] += face * f.n[d] * scale;
        }
    }
    for (int j = 0; j < 9; ++j) gradient[c * 9 + j] = g[j];
    if (Wale) {
        double value = 0;
        const bool ok = waleEddyViscosity(g, filterWidth[c], cw, value);
        if (!ok || !isfinite(value) || value < 0 || !isfinite(state[c].rho * value)) atomicOr(error, 1u);
        else nut[c] = value;
    }
}

// ReactiveWale::eddyViscosity (reactiveWale.h:93-167): the invariants of g
// scaled by max |g_ij| (S = symm(g), Sd = dev(symm(g g))), nu_t = (Cw
// delta)^2 (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)), the product
// Cw^2 delta^2 scale ratio assembled from mantissas and exponents
// (:47-80) so that no partial product over- or underflows.  frexp, ldexp
// and sqrt are exact or correctly rounded, so every step rounds as on the
// host.  delta > 0 is the cell's filter width.
__device__ inline bool waleEddyViscosity(const double* g, double delta, double cw, double& nut) {
    if (!isfinite(delta) || !(delta > 0) || !isfinite(cw) || cw < 0) return false;
    double scale = 0;
    for (int i = 0; i < 9; ++i) {
        if (!isfinite(g[i])) return false;
        const double m = g[i] < 0 ? -g[i] : g[i];
        if (m > scale) scale = m;
    }
    if (scale == 0 || cw == 0) {
        nut = 0;
        return true;
    }
    double s[9], sq[9];
    for (int i = 0; i < 9; ++i) s[i] = g[i] / scale;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double v = 0;
            for (int k = 0; k < 3; ++k) v += s[3 * i + k] * s[3 * k + j];
            sq[3 * i + j] = v;
        }
    const double trace = sq[0] + sq[4] + sq[8];
    double strain2 = 0, dev2 = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double strain = 0.5 * (s[3 * i + j] + s[3 * j + i]);
            double dev = 0.5 * (sq[3 * i + j] + sq[3 * j + i]);
            if (i == j) dev -= trace / 3.0;
            strain2 += strain * strain;
            dev2 += dev * dev;
        }
    if (!isfinite(strain2) || !isfinite(dev2)) return false;
    if (dev2 == 0) {
        nut = 0;
        return true;
    }
    const double strainPower = strain2 * strain2 * sqrt(strain2);
    const double root = sqrt(dev2);
    const double devPower = dev2 * root;
    const double denominator = strainPower + dev2 * sqrt(root);
    if (!(denominator > 0) || !isfinite(denominator) || !isfinite(devPower)) return false;
    const double ratio = devPower / denominator;
    if (ratio == 0) {
        nut = 0;
        return true;
    }
    int ec, ed, eg, er, em;
    const double mc = frexp(cw, &ec), md = frexp(delta, &ed), mg = frexp(scale, &eg), mr = frexp(ratio, &er);
    const double m = frexp(mc * mc * md * md * mg * mr, &em);
    nut = ldexp(m, 2 * ec + 2 * ed + eg + er + em);
    return isfinite(nut);
}

// The face's rho nu_t: owner and right densities times the eddy viscosities
// of the owner and (on internal faces) the neighbour, weighted by ownerWeight
// (faceNut, reactiveTransportKernels.h:80-83).
__device__ inline double faceRhoNut(const FaceDev& f, double rhoL, double rhoR, const double* __restrict__ nut) {
    return f.weight * rhoL * nut[f.owner] + (1 - f.weight) * rhoR * nut[f.kind == 0 ? f.right : f.owner];
}

// Viscous traction and the viscous + conductive energy flux of one face
// (Faces, reactiveTransportKernels.h:418-439): the interpolated gradient with
// its normal part replaced by the two-point difference, the stress
// mu (g + g^T - 2/3 div I) n (normal part only on slip walls), its work with
// the interpolated face velocity, and -k (T_R - T_L) / distance off slip walls.
__device__ inline double4 viscousWork(const FaceDev& f, const double* ul, const double* ur,
                                      const ReactiveTransportState& sl, const ReactiveTransportState& sr,
                                      const ViscousArgs& v, double energy0 = 0.0) {
    double t[3] = {0, 0, 0}, energy = energy0;
    if (v.viscosity > 0 || v.nut) {
        double g[9];
        for (int j = 0; j < 9; ++j) {
            g[j] = v.gradient[size_t(f.owner) * 9 + j];
            if (f.kind == 0) g[j] = f.weight * g[j] + (1 - f.weight) * v.gradient[size_t(f.right) * 9 + j];
        }
        for (int i = 0; i < 3; ++i) {
            double normal = 0;
            for (int d = 0; d < 3; ++d) normal += g[i * 3 + d] * f.n[d];
            const double target = (ur[i] - ul[i]) / ((f.kind == 1 ? 2.0 : 1.0) * f.distance);
            for (int d = 0; d < 3; ++d) g[i * 3 + d] += f.n[d] * (target - normal);
        }
        const double div = g[0] + g[4] + g[8];
        const double mu = v.viscosity + (v.nut ? faceRhoNut(f, sl.rho, sr.rho, v.nut) : 0.0);
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d)
                t[i] += mu * (g[3 * i + d] + g[3 * d + i] - (i == d ? 2 * div / 3 : 0.0)) * f.n[d];
        if (f.kind == 1) {
            double dot = 0;
            for (int d = 0; d < 3; ++d) dot += t[d] * f.n[d];
            for (int d = 0; d < 3; ++d) t[d] = dot * f.n[d];
        }
        for (int d = 0; d < 3; ++d) energy -= t[d] * (f.weight * ul[d] + (1 - f.weight) * ur[d]);
    }
    if (v.conductivity > 0 && f.kind != 1) energy -= v.conductivity * (sr.T - sl.T) / f.distance;
    return make_double4(t[0], t[1], t[2], energy);
}

// ---- resolved interface with surface tension (sigma > 0) -----------------
// The material colour (cells + fixed) is supplied by the caller; the
// geometry kernels below turn it into a per-cell interface normal, area
// density |grad colour| and curvature, and the face kernel then takes the
// contact-preserving capillary flux in place of HLL.
struct CapillaryArgs {
    const double* color;        // null: no capillary model
    const double* normal;       // cells x 3, pointing from colour 1 to colour 0
    const double* areaDensity;  // cells
    const double* curvature;    // cells
    unsigned* error;            // set by a rejected geometry or face
    double sigma;
    int species;                // the condensable; its liquid inventory is variable ns + 4
};

// Face colour (faceColor, reactiveUnstructuredInterface.h:33-42): the far
// side is the neighbour, the fixed state's colour, or the owner itself.
__device__ inline double interfaceFaceColor(const FaceDev& f, const double* __restrict__ color) {
    return f.weight * color[f.owner] + (1 - f.weight) * color[f.right];
}

// Face normal of the interface (faceNormal, reactiveUnstructuredInterface.h:
// 75-87): the interpolated cell normals (a boundary face takes the owner's on
// both sides), scaled to unit length when not zero.
__device__ inline void interfaceFaceNormal(const FaceDev& f, const double* __restrict__ normal, double* out) {
    const size_t l = size_t(f.owner), r = f.kind == 0 ? size_t(f.right) : l;
    double mag2 = 0;
    for (int d = 0; d < 3; ++d) {
        out[d] = f.weight * normal[3 * l + d] + (1 - f.weight) * normal[3 * r + d];
        mag2 += out[d] * out[d];
    }
    if (mag2 > 0) {
        const double inv = 1 / sqrt(mag2);
        for (int d = 0; d < 3; ++d) out[d] *= inv;
    }
}

// InterfaceGradient (reactiveTransportKernels.h:200-204 -> gradientCell,
// reactiveUnstructuredInterface.h:44-73): Gauss gradient of the colour with
// the cell's own colour taken from every face value (a uniform colour has
// none), its length as the interface area density, and the unit normal
// -g/|g| where |g| times the cell length cbrt(V) (length[c], taken on the
// host) exceeds epsilon, else zero.  A colour outside [0, 1] or a non-finite
// result sets *error.
__global__ __launch_bounds__(256) void interfaceGradientKernel(
    const FaceDev* __restrict__ faces, const uint32_t* __restrict__ row, const int32_t* __restrict__ incidence,
    const double* __restrict__ invV, const double* __restrict__ length, size_t cells, double epsilon,
    const double* __restrict__ color, double* __restrict__ gradient, double* __restrict__ normal,
    double* __restrict__ areaDensity, unsigned* __restrict__ error) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const double own = color[c];
    if (!(isfinite(own) && own >= 0 && own <= 1)) {
        atomicOr(error, 1u);
        return;
    }
    double g[3] = {0, 0, 0};
    for (uint32_t e = row[c]; e < row[c + 1]; ++e) {
        const int32_t entry = incidence[e];
        const FaceDev& f = faces[entry < 0 ? -entry - 1 : entry - 1];
        const double orientation = entry < 0 ? 1.0 : -1.0;
        const double cf = interfaceFaceColor(f, color);
        if (!(isfinite(cf) && cf >= 0 && cf <= 1)) {
            atomicOr(error, 1u);
            return;
        }
        for (int d = 0; d < 3; ++d) g[d] += (cf - own) * orientation * f.area * f.n[d];
    }
    double mag2 = 0;
    for (int d = 0; d < 3; ++d) {
        g[d] *= invV[c];
        mag2 += g[d] * g[d];
        if (!isfinite(g[d])) {
            atomicOr(error, 1u);
            return;
        }
        gradient[3 * c + d] = g[d];
    }
    const double mag = sqrt(mag2);
    if (!isfinite(mag) || mag < 0) {
        atomicOr(error, 1u);
        return;
    }
    areaDensity[c] = mag;
    const bool active = mag * length[c] > epsilon;
    for (int d = 0; d < 3; ++d) normal[3 * c + d] = active ? -g[d] / mag : 0.0;
}

// InterfaceCurvature (reactiveTransportKernels.h:205-210 -> curvatureCell,
// reactiveUnstructuredInterface.h:89-100): divergence of the face normals
// relative to the cell's own, and the surface energy sigma |grad colour|.
__global__ __launch_bounds__(256) void interfaceCurvatureKernel(
    const FaceDev* __restrict__ faces, const uint32_t* __restrict__ row, const int32_t* __restrict__ incidence,
    const double* __restrict__ invV, size_t cells, double sigma, const double* __restrict__ normal,
    const double* __restrict__ areaDensity, double* __restrict__ curvature, double* __restrict__ energy,
    unsigned* __restrict__ error) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    double sum = 0;
    for (uint32_t e = row[c]; e < row[c + 1]; ++e) {
        const int32_t entry = incidence[e];
        const FaceDev& f = faces[entry < 0 ? -entry - 1 : entry - 1];
        const double orientation = entry < 0 ? 1.0 : -1.0;
        double nf[3];
        interfaceFaceNormal(f, normal, nf);
        for (int d = 0; d < 3; ++d) sum += (nf[d] - normal[3 * c + d]) * orientation * f.area * f.n[d];
    }
    const double candidate = sum * invV[c];
    if (isfinite(candidate)) curvature[c] = candidate;
    else atomicOr(error, 1u);
    energy[c] = sigma * areaDensity[c];
}

// One side of a capillary face (State, reactiveBalancedCapillary.h:18-20).
struct CapillarySide {
    double rho, u[3], p, sound, energy, color;
};

__device__ inline bool capillarySideValid(const CapillarySide& s) {  // valid (:40-47)
    if (!isfinite(s.rho) || s.rho <= 0 || !isfinite(s.p) || s.p <= 0 || !isfinite(s.sound) || s.sound <= 0 ||
        !isfinite(s.energy) || !isfinite(s.color) || s.color < 0 || s.color > 1)
        return false;
    for (int d = 0; d < 3; ++d)
        if (!isfinite(s.u[d])) return false;
    return true;
}

// Capillary face flux: the face's interface geometry (faceGeometry,
// reactiveUnstructuredInterface.h:102-128: interpolated curvature, colour
// and area density, surface stress sigma |grad c| (I - n n)) and the
// contact-preserving flux (faceFlux, reactiveBalancedCapillary.h:55-142):
// the HLLC solve on pi = p - sigma kappa colour (HLL where the contact
// leaves the acoustic bounds), plus the shared capillary traction
// sigma kappa c_f n - C n and its work with the face velocity.  Writes the
// four face coefficients to w, the capillary momentum flux to capM and its
// energy flux to capE; false where the reference rejects the face.
__device__ inline bool capillaryFlux(const FaceDev& f, const CapillarySide& L, const CapillarySide& R,
                                     const CapillaryArgs& ca, double4& w, double* capM, double& capE) {
    const size_t l = size_t(f.owner), r = f.kind == 0 ? size_t(f.right) : l;
    const double curv = f.weight * ca.curvature[l] + (1 - f.weight) * ca.curvature[r];
    const double colorFace = interfaceFaceColor(f, ca.color);
    double ni[3];
    interfaceFaceNormal(f, ca.normal, ni);
    const double scale = ca.sigma * (f.weight * ca.areaDensity[l] + (1 - f.weight) * ca.areaDensity[r]);
    double stress[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) stress[3 * i + j] = scale * ((i == j ? 1.0 : 0.0) - ni[i] * ni[j]);
    if (!isfinite(curv) || !isfinite(colorFace) || colorFace < 0 || colorFace > 1 || !isfinite(scale)) return false;
    if (!capillarySideValid(L) || !capillarySideValid(R)) return false;
    double n2 = 0, unL = 0, unR = 0;
    for (int d = 0; d < 3; ++d) {
        if (!isfinite(f.n[d])) return false;
        n2 += f.n[d] * f.n[d];
        unL += L.u[d] * f.n[d];
        unR += R.u[d] * f.n[d];
    }
    if (!isfinite(n2) || ::fabs(n2 - 1) > 1e-10) return false;
    for (int j = 0; j < 9; ++j)
        if (!isfinite(stress[j])) return false;
    const double piL = L.p - ca.sigma * curv * L.color;
    const double piR = R.p - ca.sigma * curv * R.color;
    if (!isfinite(piL) || !isfinite(piR)) return false;
    const double sl = dmin(0, dmin(unL - L.sound, unR - R.sound));
    const double sr = dmax(0, dmax(unL + L.sound, unR + R.sound));
    const double den = L.rho * (sl - unL) - R.rho * (sr - unR);
    if (!isfinite(den) || den == 0 || !isfinite(sl) || !isfinite(sr) || sr <= sl) return false;
    const double star = (piR - piL + L.rho * unL * (sl - unL) - R.rho * unR * (sr - unR)) / den;
    double advectL = 0, advectR = 0, contact, mass, pMom, pEn;
    if (!isfinite(star) || star <= sl || star >= sr) {  // the contact outran the bounds: HLL
        const double inv = 1 / (sr - sl);
        contact = (sr * unL - sl * unR) * inv;
        advectL = sr * (unL - sl) * inv;
        advectR = sl * (sr - unR) * inv;
        mass = advectL * L.rho + advectR * R.rho;
        pMom = (sr * piL - sl * piR) * inv;
        pEn = (sr * piL * unL - sl * piR * unR) * inv;
    } else {
        const bool useLeft = star >= 0;
        const CapillarySide& s = useLeft ? L : R;
        const double un = useLeft ? unL : unR, wave = useLeft ? sl : sr, pi = useLeft ? piL : piR;
        const bool supersonic = (useLeft && sl >= 0) || (!useLeft && sr <= 0);
        contact = supersonic ? un : star;
        if (supersonic) {
            mass = s.rho * un;
            pMom = pi;
            pEn = pi * un;
        } else {
            const double rhoStar = s.rho * (wave - un) / (wave - star);
            const double piStar = pi + s.rho * (wave - un) * (star - un);
            const double energyStar = ((wave - un) * s.energy - pi * un + piStar * star) / (wave - star);
            mass = rhoStar * star;
            pMom = piStar + mass * (star - un);
            pEn = star * (energyStar + piStar) - mass * s.energy / s.rho;
        }
        if (useLeft) advectL = mass / L.rho;
        else advectR = mass / R.rho;
    }
    const double advective = advectL * L.energy + advectR * R.energy;
    double fu[3], average = 0;
    for (int d = 0; d < 3; ++d) {
        fu[d] = 0.5 * (L.u[d] + R.u[d]);
        average += fu[d] * f.n[d];
    }
    for (int d = 0; d < 3; ++d) fu[d] += (contact - average) * f.n[d];
    capE = 0;
    bool finiteMomentum = true;
    for (int i = 0; i < 3; ++i) {
        double traction = 0;
        for (int j = 0; j < 3; ++j) traction += stress[3 * i + j] * f.n[j];
        capM[i] = ca.sigma * curv * colorFace * f.n[i] - traction;
        const double momentum = advectL * L.rho * L.u[i] + advectR * R.rho * R.u[i] + pMom * f.n[i] + capM[i];
        finiteMomentum = finiteMomentum && isfinite(momentum) && isfinite(capM[i]);
        capE += fu[i] * capM[i];
    }
    const double total = advective + pEn + capE;
    if (!isfinite(mass) || !isfinite(pMom) || !isfinite(pEn) || !isfinite(capE) || !isfinite(total) ||
        !finiteMomentum)
        return false;
    w = make_double4(advectL, advectR, pMom, pEn);
    return true;
}

// Face coefficients (Faces, reactiveTransportKernels.h:343-356): HLL, or the
// HLLC option (flux_device.h), which also reads the total energies from q.
// A face the HLLC flux rejects keeps its HLL coefficients and is counted.
// With V the same thread adds the viscous record (visc[fi]); with the WALE
// closure a non-finite traction or energy sets *waleError.
template <bool V>
__global__ void hllKernel(const FaceDev* __restrict__ faces, size_t nf, const double* __restrict__ prim, size_t all,
                          const ReactiveTransportState* __restrict__ state, double waveFactor,
                          double4* __restrict__ work, int scheme, const double* __restrict__ q, int ns,
                          unsigned long long* __restrict__ failures, ViscousArgs va, double4* __restrict__ visc,
                          unsigned* __restrict__ waleError, CapillaryArgs ca) {
    const size_t fi = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (fi >= nf) return;
    const FaceDev f = faces[fi];
    double ul[3], ur[3], unL, unR;
    normalVelocities(f, prim, all, ul, ur, unL, unR);
    const double pL = state[f.owner].p, pR = state[f.right].p;
    const double aL = waveFactor * state[f.owner].sound, aR = waveFactor * state[f.right].sound;
    const double left = dmin(0, dmin(unL - aL, unR - aR)), right = dmax(0, dmax(unL + aL, unR + aR));
    const double invWave = 1 / (right - left);
    double4 w;
    w.x = right * (unL - left) * invWave;
    w.y = left * (right - unR) * invWave;
    w.z = (right * pL - left * pR) * invWave;
    w.w = (right * pL * unL - left * pR * unR) * invWave;
    if (scheme == kHllc && !ca.color &&
        !hllcCoefficients(state[f.owner].rho, ul, pL, aL, q[size_t(ns + 3) * all + f.owner], state[f.right].rho, ur,
                          pR, aR, q[size_t(ns + 3) * all + f.right], f.n, w))
        atomicAdd(failures, 1ull);
    // capillary model (reactiveTransportKernels.h:357-417): its flux replaces
    // the coefficients, its energy flux opens the face's energy record and its
    // momentum flux is taken from the traction after the viscous part (:493)
    double capM[3] = {0, 0, 0}, energy0 = 0;
    if (ca.color) {
        const CapillarySide L{state[f.owner].rho, {ul[0], ul[1], ul[2]}, pL, aL,
                              q[size_t(ns + 3) * all + f.owner], ca.color[f.owner]};
        const CapillarySide R{state[f.right].rho, {ur[0], ur[1], ur[2]}, pR, aR,
                              q[size_t(ns + 3) * all + f.right], ca.color[f.right]};
        double capE = 0;
        if (!capillaryFlux(f, L, R, ca, w, capM, capE)) {
            atomicOr(ca.error, 1u);
            return;
        }
        energy0 += capE;
    }
    work[fi] = w;
    if (V) {
        double4 v = viscousWork(f, ul, ur, state[f.owner], state[f.right], va, energy0);
        if (ca.color) {
            // the liquid inventory of the condensable on both sides (:478-486)
            const size_t sk = size_t(ca.species), lk = size_t(ns + 4);
            const double slq = q[sk * all + f.owner], srq = q[sk * all + f.right];
            const double ml = q[lk * all + f.owner], mr = q[lk * all + f.right];
            if (!isfinite(slq) || !isfinite(srq) || !isfinite(ml) || !isfinite(mr) || slq < 0 || srq < 0 || ml < 0 ||
                mr < 0 || ml > slq || mr > srq)
                atomicOr(ca.error, 1u);
            v.x -= capM[0];
            v.y -= capM[1];
            v.z -= capM[2];
        }
        if (va.nut && !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w))) atomicOr(waleError, 2u);
        visc[fi] = v;
    }
}

// faceFlux (reactiveTransportKernels.h:180-191); visc (null for the
// inviscid model) holds the face's traction and energy flux.
__device__ inline double faceFlux(const FaceDev& f, const double4& w, const double4* __restrict__ visc, uint32_t fi,
                                  const double* __restrict__ q, size_t all, int ns, int k) {
    const double ql = q[size_t(k) * all + f.owner];
    double qr = q[size_t(k) * all + f.right];
    if (f.kind == 1 && k >= ns && k < ns + 3) {
        double normal = 0;
        for (int d = 0; d < 3; ++d) normal += q[size_t(ns + d) * all + f.right] * f.n[d];
        qr -= 2 * normal * f.n[k - ns];
    }
    const double4 v = visc ? visc[fi] : make_double4(0, 0, 0, 0);
    double flux = w.x * ql + w.y * qr;
    if (k >= ns && k < ns + 3) flux += (w.z - 0.0) * f.n[k - ns];
    if (k == ns + 3) flux += w.w;
    if (k < ns) flux += 0.0;
    else if (k < ns + 3) flux -= k == ns ? v.x : (k == ns + 1 ? v.y : v.z);
    else if (k == ns + 3) flux += v.w;
    return flux * f.area;
}

// Rhs::value (reactiveTransportKernels.h:502-526) and Advance (:551-560).
// mode 0: stage 0 -> out = q + dt L; 1: stage 1 -> out = .5 out + .5 (q + dt L);
// 2: rhs only -> out = L (cell-major index of a cells x nv array).
__global__ void advanceKernel(const FaceDev* __restrict__ faces, const double4* __restrict__ work,
                              const double4* __restrict__ visc, const uint32_t* __restrict__ row,
                              const int32_t* __restrict__ incidence,
                              const double* __restrict__ invV, const double* __restrict__ q, size_t cells, size_t all,
                              int ns, int nv, double dt, int mode, double* __restrict__ out) {
    const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cells * size_t(nv)) return;
    const size_t c = j % cells;
    const int k = int(j / cells);
    double rate = 0;
    const double inv = invV[c];
    for (uint32_t e = row[c]; e < row[c + 1]; ++e) {
        const int32_t entry = incidence[e];
        const uint32_t fi = uint32_t(entry < 0 ? -entry - 1 : entry - 1);
        const double sign = entry < 0 ? -1 : 1;
        rate += sign * faceFlux(faces[fi], work[fi], visc, fi, q, all, ns, k) * inv;
    }
    const size_t index = size_t(k) * all + c;
    if (mode == 0) out[index] = q[index] + dt * rate;
    else if (mode == 1) out[index] = .5 * out[index] + .5 * (q[index] + dt * rate);
    else out[c * nv + k] = rate;
}

// BoundaryPartial: block g = (variable k, tile) sums its lanes' strided faces,
// then the 256-lane tree of reduceGroups (reactiveTransport.cpp:70-81, 233-246).
__global__ __launch_bounds__(256) void boundaryPartialKernel(const FaceDev* __restrict__ faces,
                                                             const double4* __restrict__ work,
                                                             const double4* __restrict__ visc,
                                                             const uint32_t* __restrict__ boundary, size_t nb,
                                                             size_t partitions, const double* __restrict__ q,
                                                             size_t all, int ns, double* __restrict__ partial) {
    __shared__ double values[256];
    const size_t g = blockIdx.x;
    const int k = int(g / partitions);
    const size_t tile = g % partitions;
    double sum = 0;
    for (size_t j = tile * 256 + threadIdx.x; j < nb; j += partitions * 256) {
        const uint32_t fi = boundary[j];
        sum += faceFlux(faces[fi], work[fi], visc, fi, q, all, ns, k);
    }
    values[threadIdx.x] = sum;
    __syncthreads();
    for (unsigned stride = 128; stride; stride /= 2) {
        if (threadIdx.x < stride) values[threadIdx.x] = values[threadIdx.x] + values[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[g] = values[0];
}

__global__ __launch_bounds__(256) void boundaryFinishKernel(const double* __restrict__ partial, size_t partitions,
                                                            double* __restrict__ rate) {
    __shared__ double values[256];
    const int k = blockIdx.x;
    values[threadIdx.x] = threadIdx.x < partitions ? partial[size_t(k) * partitions + threadIdx.x] : 0;
    __syncthreads();
    for (unsigned stride = 128; stride; stride /= 2) {
        if (threadIdx.x < stride) values[threadIdx.x] = values[threadIdx.x] + values[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) rate[k] = values[0];
}

// FaceSpeeds (reactiveTransportKernels.h:315-342): wave speed plus twice the
// face diffusion number D / distance, D the largest of 4 mu / (3 rho) and
// k / (rho cv) over the two sides (mu with the WALE rho nu_t when nut is set;
// no species diffusion or scalar closure).
__global__ void faceSpeedKernel(const FaceDev* __restrict__ faces, size_t nf, const double* __restrict__ prim,
                                size_t all, const ReactiveTransportState* __restrict__ state, double waveFactor,
                                double viscosity, double conductivity, const double* __restrict__ nut,
                                double* __restrict__ speed) {
    const size_t fi = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (fi >= nf) return;
    const FaceDev f = faces[fi];
    double ul[3], ur[3], unL, unR;
    normalVelocities(f, prim, all, ul, ur, unL, unR);
    const ReactiveTransportState& sl = state[f.owner];
    const ReactiveTransportState& sr = state[f.right];
    const double aL = waveFactor * sl.sound, aR = waveFactor * sr.sound;
    const double left = dmin(0, dmin(unL - aL, unR - aR)), right = dmax(0, dmax(unL + aL, unR + aR));
    const double invWave = 1 / (right - left);
    const double muFace = viscosity + (nut ? faceRhoNut(f, sl.rho, sr.rho, nut) : 0.0);
    const double D = 0.0 + dmax(dmax(4 * muFace / (3 * sl.rho), 4 * muFace / (3 * sr.rho)),
                                dmax(conductivity / (sl.rho * sl.cv), conductivity / (sr.rho * sr.cv)));
    double s = dmax(::fabs(unL) + aL, ::fabs(unR) + aR) + 2 * (D + 0.0) / f.distance;
    if (!isfinite(unL) || !isfinite(unR) || !isfinite(s) || !isfinite(invWave) || right <= left) s = -1;
    speed[fi] = s;
}

// The capillary wave step's inputs (areaDensity null: no capillary model);
// length[c] = cbrt(V_c), taken on the host.
struct CapillaryStepArgs {
    const double* areaDensity;
    const double* length;
    const ReactiveTransportState* state;
    double sigma, cfl, epsilon;
};

// Step (reactiveTransportKernels.h:529-550) + the absorbing validity map of
// Minimum (reactiveTransport.cpp:222-232): invalid -> -1; block minimum.
__global__ __launch_bounds__(256) void cellStepKernel(const FaceDev* __restrict__ faces,
                                                      const double* __restrict__ speed,
                                                      const uint32_t* __restrict__ row,
                                                      const int32_t* __restrict__ incidence,
                                                      const double* __restrict__ invV, size_t cells, double cfl,
                                                      double maximum, double* __restrict__ blockMin,
                                                      CapillaryStepArgs cs) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double v = __builtin_huge_val();
    if (c < cells) {
        double denominator = 0, step;
        bool bad = false;
        for (uint32_t e = row[c]; e < row[c + 1]; ++e) {
            const int32_t entry = incidence[e];
            const uint32_t fi = uint32_t(entry < 0 ? -entry - 1 : entry - 1);
            const double s = speed[fi];
            if (!isfinite(s) || s < 0) { bad = true; break; }
            denominator += faces[fi].area * s;
        }
        if (bad) step = -1;
        else {
            step = denominator > 0 ? dmin(maximum, cfl / (invV[c] * denominator)) : maximum;
            if (cs.areaDensity && cs.sigma > 0) {  // capillary wave step (reactiveTransportKernels.h:540-547)
                constexpr double pi = 3.141592653589793238462643383279502884;
                const double volume = 1 / invV[c];
                if (cs.areaDensity[c] * cs.length[c] > cs.epsilon) {
                    const double capDt = cs.cfl * sqrt(cs.state[c].rho * volume / (pi * cs.sigma));
                    step = dmin(step, capDt);
                }
            }
            if (!isfinite(denominator) || denominator < 0) step = -1;
        }
        v = isfinite(step) && step > 0 ? step : -1;
    }
    __shared__ double lane[4];
    v = waveMin(v);
    if ((threadIdx.x & 63) == 0) lane[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lane[0];
        for (int i = 1; i < int(blockDim.x >> 6); ++i) m = lane[i] < m ? lane[i] : m;
        blockMin[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(256) void minReduceKernel(const double* __restrict__ in, size_t n,
                                                       double* __restrict__ out) {
    double v = __builtin_huge_val();
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) v = in[i] < v ? in[i] : v;
    __shared__ double lane[4];
    v = waveMin(v);
    if ((threadIdx.x & 63) == 0) lane[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lane[0];
        for (int i = 1; i < int(blockDim.x >> 6); ++i) m = lane[i] < m ? lane[i] : m;
        *out = m;
    }
}

// Cell-major host layout <-> SoA device layout (Layout, reactiveTransportKernels.h:569-576).
__global__ void packKernel(const double* __restrict__ aos, size_t cells, int nv, size_t stride, size_t offset,
                           double* __restrict__ soa) {
    const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cells * size_t(nv)) return;
    const size_t c = j / nv, k = j % nv;
    soa[k * stride + offset + c] = aos[j];
}
__global__ void unpackKernel(const double* __restrict__ soa, size_t cells, int nv, size_t stride,
                             double* __restrict__ aos) {
    const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cells * size_t(nv)) return;
    const size_t c = j / nv, k = j % nv;
    aos[j] = soa[k * stride + c];
}
__global__ void primitiveUploadKernel(const ReactiveTransportPrimitive* __restrict__ in, size_t cells, size_t all,
                                      double* __restrict__ prim) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    prim[c] = in[c].rho;
    for (int d = 0; d < 3; ++d) prim[size_t(d + 1) * all + c] = in[c].u[d];
}

// HLL coefficients of the boundary faces only, primitives rebuilt from q
// (Cells + Faces restricted to the boundary list, same arithmetic).
__global__ void boundaryHllKernel(const FaceDev* __restrict__ faces, const uint32_t* __restrict__ boundary, size_t nb,
                                  const double* __restrict__ q, size_t all, int ns, const double* __restrict__ pArr,
                                  const double* __restrict__ cArr, const double* __restrict__ rArr, double waveFactor,
                                  double4* __restrict__ work, int scheme, unsigned long long* __restrict__ failures) {
    const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const uint32_t fi = boundary[j];
    const FaceDev f = faces[fi];
    double rhoL = 0, rhoR = 0;
    for (int k = 0; k < ns; ++k) {
        rhoL += q[size_t(k) * all + f.owner];
        rhoR += q[size_t(k) * all + f.right];
    }
    double ul[3], ur[3], dot = 0;
    for (int d = 0; d < 3; ++d) {
        ul[d] = q[size_t(ns + d) * all + f.owner] / rhoL;
        ur[d] = q[size_t(ns + d) * all + f.right] / rhoR;
        dot += ur[d] * f.n[d];
    }
    if (f.kind == 1)
        for (int d = 0; d < 3; ++d) ur[d] -= 2 * dot * f.n[d];
    double unL = 0, unR = 0;
    for (int d = 0; d < 3; ++d) {
        unL += ul[d] * f.n[d];
        unR += ur[d] * f.n[d];
    }
    const double pL = pArr[f.owner], pR = pArr[f.right];
    const double aL = waveFactor * cArr[f.owner], aR = waveFactor * cArr[f.right];
    const double left = dmin(0, dmin(unL - aL, unR - aR)), right = dmax(0, dmax(unL + aL, unR + aR));
    const double invWave = 1 / (right - left);
    double4 w;
    w.x = right * (unL - left) * invWave;
    w.y = left * (right - unR) * invWave;
    w.z = (right * pL - left * pR) * invWave;
    w.w = (right * pL * unL - left * pR * unR) * invWave;
    if (scheme == kHllc &&
        !hllcCoefficients(rArr[f.owner], ul, pL, aL, q[size_t(ns + 3) * all + f.owner], rArr[f.right], ur, pR, aR,
                          q[size_t(ns + 3) * all + f.right], f.n, w))
        atomicAdd(failures, 1ull);
    work[fi] = w;
}

inline unsigned gridFor(size_t n, unsigned block) { return unsigned((n + block - 1) / block); }

// ---------------------------------------------------------------- handle
Transport::Transport(int backend, const ReactiveTransportConfig& c, const double* volumes,
                     const ReactiveTransportFace* hostFaces, const double* fixedQ,
                     const ReactiveTransportState* fixedStates, const ReactiveTransportOptionsV2* options,
                     const ReactiveTransportOptionsV21* options21)
    : cfg(c) {
    require(backend == 0 || backend == 1, "Unknown transport backend");
    require(backend == 1, "backend=0 (serial host operators) is not part of the MI355X transport build; use backend=1");
    int devices = 0;
    RAMD_HIP(hipGetDeviceCount(&devices));
    require(devices > 0, "No HIP device");
    device = currentDevice();
    if (options) {
        require(options->abiVersion == 1 && options->structBytes == sizeof(*options) && options->bridgeCells > 0,
                "Invalid transport options ABI");
        require(options->physicalModelHash[64] == 0 && std::strlen(options->physicalModelHash) == 64,
                "Invalid transport model hash");
        physicalHash = options->physicalModelHash;
        bridgeCells = options->bridgeCells;
        require(!options->recomputeGas, "Recompute requires ideal-gas diffusive transport");
    }
    if (options21) {
        cost.blockThreads = options21->blockThreads;
        cost.pinnedBytes = 0;
    }
    require(cfg.cells && cfg.species && cfg.faces, "Empty transport mesh/model");
    require(cfg.cells <= size_t(INT64_MAX) && cfg.faces < size_t(INT64_MAX) && cfg.fixed <= size_t(INT64_MAX),
            "Transport index overflow");
    require(cfg.mechanical ? cfg.variables == cfg.species + 6
                           : (cfg.variables >= cfg.species + 4 && cfg.variables <= cfg.species + 6),
            "Wrong transport conserved-variable count");
    require(std::isfinite(cfg.maxBytes) && cfg.maxBytes > 0, "Invalid transport memory budget");
    require(std::isfinite(cfg.waveFactor) && cfg.waveFactor >= 1, "Invalid transport wave factor");
    require(std::isfinite(cfg.viscosity) && std::isfinite(cfg.conductivity) && std::isfinite(cfg.diffusivity) &&
                cfg.viscosity >= 0 && cfg.conductivity >= 0 && cfg.diffusivity >= 0,
            "Invalid transport coefficients");
    require(cfg.diffusivity == 0,
            "Diffusive species transport (diffusivity > 0) is not provided by the MI355X build");
    require(!cfg.mechanical, "The mechanical-equilibrium closure is not provided by the MI355X build");
    require(volumes && hostFaces, "Null transport geometry");
    require(cfg.cells <= std::numeric_limits<size_t>::max() - cfg.fixed, "Fixed-state index overflow");
    all = cfg.cells + cfg.fixed;
    require(all < (size_t(1) << 31) && cfg.faces < (size_t(1) << 31),
            "This build indexes cells and faces with 32-bit integers (< 2^31)");
    nv = int(cfg.variables);
    ns = int(cfg.species);
    std::vector<uint32_t> row(cfg.cells + 1, 0);
    std::vector<uint32_t> boundary;
    for (size_t c2 = 0; c2 < cfg.cells; ++c2)
        require(std::isfinite(volumes[c2]) && volumes[c2] > 0, "Invalid transport volume");
    std::vector<FaceDev> faces(cfg.faces);
    for (size_t i = 0; i < cfg.faces; ++i) {
        const auto& f = hostFaces[i];
        require(f.owner >= 0 && size_t(f.owner) < cfg.cells && f.neighbour >= -1 &&
                    (f.neighbour < 0 || size_t(f.neighbour) < cfg.cells),
                "Invalid transport face cell");
        require(f.kind >= 0 && f.kind <= 3 && (f.kind == 0) == (f.neighbour >= 0), "Invalid transport face kind");
        require(f.kind != 3 || (f.fixed >= 0 && size_t(f.fixed) < cfg.fixed && !cfg.mechanical),
                "Invalid fixed boundary state index");
        double norm = 0;
        for (double component : f.normal) norm += component * component;
        require(std::isfinite(norm) && std::abs(norm - 1) < 1e-8, "Invalid transport unit normal");
        require(std::isfinite(f.area) && std::isfinite(f.distance) && std::isfinite(f.ownerWeight) && f.area > 0 &&
                    f.distance > 0 && f.ownerWeight >= 0 && f.ownerWeight <= 1,
                "Invalid transport face geometry");
        ++row[f.owner + 1];
        if (f.neighbour >= 0) ++row[f.neighbour + 1];
        else boundary.push_back(uint32_t(i));
        FaceDev& d = faces[i];
        d.owner = int32_t(f.owner);
        d.right = int32_t(f.neighbour >= 0 ? f.neighbour : (f.kind == 3 ? int64_t(cfg.cells) + f.fixed : f.owner));
        d.kind = f.kind;
        d.pad = 0;
        for (int k = 0; k < 3; ++k) d.n[k] = f.normal[k];
        d.area = f.area;
        d.distance = f.distance;
        d.weight = f.ownerWeight;
    }
    for (size_t c2 = 0; c2 < cfg.cells; ++c2) row[c2 + 1] += row[c2];
    std::vector<int32_t> incidence(row.back());
    std::vector<uint32_t> next(row);
    for (size_t i = 0; i < cfg.faces; ++i) {
        incidence[next[hostFaces[i].owner]++] = -int32_t(i) - 1;
        if (hostFaces[i].neighbour >= 0) incidence[next[hostFaces[i].neighbour]++] = int32_t(i) + 1;
    }
    nBoundary = boundary.size();
    partitions = std::min(size_t(256), std::max(size_t(1), (nBoundary + 255) / 256));
    profile.boundaryPartitions = partitions;
    bridgeCells = std::min(bridgeCells, std::max(cfg.cells, cfg.fixed));
    RAMD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    this->faces = allocate<FaceDev>(cfg.faces);
    invV = allocate<double>(cfg.cells);
    rowDev = allocate<uint32_t>(row.size());
    incidenceDev = allocate<int32_t>(incidence.size());
    boundaryDev = allocate<uint32_t>(std::max<size_t>(1, nBoundary));
    q = allocate<double>(all * nv);
    initial = allocate<double>(all * nv);
    state = allocate<ReactiveTransportState>(all);
    prim = allocate<double>(4 * all);
    work = allocate<double4>(cfg.faces);
    faceSpeed = allocate<double>(cfg.faces);
    blockMin = allocate<double>(gridFor(cfg.cells, 256) + 1);
    boundaryPartial = allocate<double>(nv * partitions);
    boundaryRate = allocate<double>(nv);
    staging = allocate<double>(std::max(all, cfg.cells) * nv);
    fluxFailures = allocate<unsigned long long>(1);
    RAMD_HIP(hipMemsetAsync(fluxFailures, 0, sizeof(unsigned long long), stream));
    if (cfg.viscosity > 0) gradient = allocate<double>(cfg.cells * 9);
    if (cfg.viscosity > 0 || cfg.conductivity > 0) visc = allocate<double4>(cfg.faces);
    deviceProfile.faceWorkspaceBytes = cfg.faces * (sizeof(double4) + sizeof(double)) + (visc ? cfg.faces * 32 : 0);
    std::vector<double> inverse(cfg.cells);
    for (size_t c2 = 0; c2 < cfg.cells; ++c2) {
        inverse[c2] = 1 / volumes[c2];
        require(std::isfinite(inverse[c2]), "Invalid inverse volume");
    }
    upload(this->faces, faces.data(), faces.size());
    upload(invV, inverse.data(), inverse.size());
    upload(rowDev, row.data(), row.size());
    upload(incidenceDev, incidence.data(), incidence.size());
    if (nBoundary) upload(boundaryDev, boundary.data(), nBoundary);
    // the structured-box kernels carry the inviscid flux only
    if (!viscous()) cartesian = detectCartesian(hostFaces, volumes, row, incidence);
    if (cfg.fixed) {
        require(fixedQ, "Null packed source");
        validateStates(fixedStates, cfg.fixed);
        upload(staging, fixedQ, cfg.fixed * nv);
        launchPack(staging, cfg.fixed, q, cfg.cells);
        RAMD_HIP(hipMemcpyAsync(initial, q, all * nv * sizeof(double), hipMemcpyDeviceToDevice, stream));
        upload(state + cfg.cells, fixedStates, cfg.fixed);
        std::vector<double> fixedPrim(4 * cfg.fixed);
        for (size_t c2 = 0; c2 < cfg.fixed; ++c2) {
            double rho = 0;
            for (int k = 0; k < ns; ++k) rho += fixedQ[c2 * nv + k];
            require(std::isfinite(rho) && rho > 0, "Invalid fixed density");
            fixedPrim[c2] = rho;
            for (int d = 0; d < 3; ++d) {
                fixedPrim[(d + 1) * cfg.fixed + c2] = fixedQ[c2 * nv + ns + d] / rho;
                require(std::isfinite(fixedPrim[(d + 1) * cfg.fixed + c2]), "Invalid fixed velocity");
            }
        }
        for (int d = 0; d < 4; ++d) upload(prim + d * all + cfg.cells, fixedPrim.data() + d * cfg.fixed, cfg.fixed);
    }
    finish();
}

Transport::~Transport() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocations) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

void* Transport::allocateBytes(size_t bytes) {
    if (!bytes) return nullptr;
    require(double(stats.allocatedBytes) + double(bytes) <= cfg.maxBytes,
            "Transport exceeds maxDeviceMemoryGB/workspace budget");
    void* p = nullptr;
    RAMD_HIP(hipMalloc(&p, bytes));
    allocations.push_back(p);
    stats.allocatedBytes += bytes;
    peakBytes = std::max<uint64_t>(peakBytes, stats.allocatedBytes);
    return p;
}

void Transport::uploadBytes(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    require(src && dst, "Null transport input buffer");
    RAMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    stats.uploadedBytes += bytes;
    ++memcpyCalls;
}
void Transport::downloadBytes(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    require(src && dst, "Null transport output buffer");
    RAMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    stats.downloadedBytes += bytes;
    ++memcpyCalls;
}
void Transport::finish() {
    ++synchronizations;
    RAMD_HIP(hipStreamSynchronize(stream));
}
void Transport::launched() { ++stats.kernelLaunches; }

void Transport::validateStates(const ReactiveTransportState* s, size_t count) {  // reactiveTransport.cpp:537-545
    require(!count || s, "Null transport state");
    for (size_t c2 = 0; c2 < count; ++c2)
        require(std::isfinite(s[c2].p) && s[c2].p > 0 && std::isfinite(s[c2].T) && s[c2].T > 0 &&
                    std::isfinite(s[c2].rho) && s[c2].rho > 0 && std::isfinite(s[c2].sound) && s[c2].sound > 0 &&
                    std::isfinite(s[c2].cv) && s[c2].cv > 0 && std::isfinite(s[c2].gasMass) && s[c2].gasMass >= 0 &&
                    std::isfinite(s[c2].dilatation),
                "Invalid recovered transport state");
}

void Transport::launchPack(const double* aos, size_t cells, double* soa, size_t offset) {
    hipLaunchKernelGGL(packKernel, dim3(gridFor(cells * nv, 256)), dim3(256), 0, stream, aos, cells, nv, all, offset,
                       soa);
    RAMD_HIP(hipGetLastError());
    launched();
    ++cost.layoutLaunches;
    ++cost.layoutTiles;
}
void Transport::launchUnpack(const double* soa, double* hostOut) {
    hipLaunchKernelGGL(unpackKernel, dim3(gridFor(cfg.cells * nv, 256)), dim3(256), 0, stream, soa, cfg.cells, nv,
                       all, staging);
    RAMD_HIP(hipGetLastError());
    launched();
    ++cost.layoutLaunches;
    ++cost.layoutTiles;
    downloadBytes(hostOut, staging, cfg.cells * nv * sizeof(double));
}

void Transport::uploadQ(const double* hostQ) {
    require(hostQ, "Null packed source");
    upload(staging, hostQ, cfg.cells * nv);
    launchPack(staging, cfg.cells, q, 0);
    ++profile.conservedUploads;
    profile.conservedUploadBytes += cfg.cells * nv * sizeof(double);
}

void Transport::uploadState(const ReactiveTransportState* states) {
    require(states, "Null transport state");
    if (stateShadowValid && stateShadow.size() == cfg.cells &&
        std::memcmp(stateShadow.data(), states, cfg.cells * sizeof(ReactiveTransportState)) == 0)
        return;
    stateShadowValid = false;
    validateStates(states, cfg.cells);
    upload(state, states, cfg.cells);
    profile.stateUploadBytes += cfg.cells * sizeof(ReactiveTransportState);
    stateShadow.assign(states, states + cfg.cells);
    stateShadowValid = true;
}

void Transport::primitives() {
    hipLaunchKernelGGL(primitivesKernel, dim3(gridFor(all, 256)), dim3(256), 0, stream, q, all, ns, prim);
    RAMD_HIP(hipGetLastError());
    launched();
}

// prepareResident's order (reactiveTransport.cpp:1379-1393): states, Cells,
// then the WALE viscosities or, for the flux of a viscous model, the velocity
// gradients, then the face coefficients or the CFL face speeds.
void Transport::prepareResident(const ReactiveTransportState* states, bool transport) {
    require(!capillary || colorReady, "Capillary geometry must be uploaded for this stage");
    uploadState(states);
    primitives();
    if (wale) buildWale();
    else if (transport && cfg.viscosity > 0) launchGradients(false);
    if (transport) launchFaces();
    else launchFaceSpeeds();
}

void Transport::launchFaces() {
    const ViscousArgs va{gradient, wale ? nut : nullptr, cfg.viscosity, cfg.conductivity};
    CapillaryArgs ca{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, 0};
    if (capillary) {
        ca = CapillaryArgs{capColor, capNormal, capArea, capCurvature, capError, capSigma, capSpecies};
        RAMD_HIP(hipMemsetAsync(capError, 0, sizeof(unsigned), stream));
    }
    if (wale) RAMD_HIP(hipMemsetAsync(waleError, 0, sizeof(unsigned), stream));
    if (faceRecords())
        hipLaunchKernelGGL(hllKernel<true>, dim3(gridFor(cfg.faces, 256)), dim3(256), 0, stream, faces, cfg.faces,
                           prim, all, state, cfg.waveFactor, work, scheme, q, ns, fluxFailures, va, visc, waleError,
                           ca);
    else
        hipLaunchKernelGGL(hllKernel<false>, dim3(gridFor(cfg.faces, 256)), dim3(256), 0, stream, faces, cfg.faces,
                           prim, all, state, cfg.waveFactor, work, scheme, q, ns, fluxFailures, va, visc, waleError,
                           ca);
    RAMD_HIP(hipGetLastError());
    launched();
    ++deviceProfile.transportFaceLaunches;
    if (capillary) {  // reactiveTransport.cpp:1405-1408
        unsigned status = 0;
        RAMD_HIP(hipMemcpyAsync(&status, capError, sizeof(status), hipMemcpyDeviceToHost, stream));
        finish();
        require(!status, "Invalid capillary face flux");
        capillaryProfile.faceFluxBuilds += cfg.faces;
    }
    if (wale) {
        unsigned status = 0;
        RAMD_HIP(hipMemcpyAsync(&status, waleError, sizeof(status), hipMemcpyDeviceToHost, stream));
        finish();
        require(!status, "Nonfinite WALE traction/energy flux; stage remains uncommitted");
    }
}

void Transport::launchGradients(bool withWale) {
    const unsigned blocks = gridFor(cfg.cells, 256);
    if (withWale)
        hipLaunchKernelGGL(gradientKernel<true>, dim3(blocks), dim3(256), 0, stream, faces, rowDev, incidenceDev, invV,
                           prim, all, cfg.cells, gradient, filterWidth, waleCw, state, nut, waleError);
    else
        hipLaunchKernelGGL(gradientKernel<false>, dim3(blocks), dim3(256), 0, stream, faces, rowDev, incidenceDev,
                           invV, prim, all, cfg.cells, gradient, filterWidth, waleCw, state, nut, waleError);
    RAMD_HIP(hipGetLastError());
    launched();
}

// buildWale (reactiveTransport.cpp:841-849): gradients and eddy viscosities
// in one pass, then the closure's status.
void Transport::buildWale() {
    RAMD_HIP(hipMemsetAsync(waleError, 0, sizeof(unsigned), stream));
    launchGradients(true);
    ++waleProfile.gradientBuilds;
    ++waleProfile.viscosityBuilds;
    waleProfile.cellsEvaluated += cfg.cells;
    unsigned status = 0;
    RAMD_HIP(hipMemcpyAsync(&status, waleError, sizeof(status), hipMemcpyDeviceToHost, stream));
    finish();
    require(!status, "Invalid WALE gradient/filter width/eddy viscosity");
}

// installWale (reactiveTransport.cpp:636-652): once, before any upload or
// step.  The filter width cbrt(1 / (1 / V)) of each cell is fixed by the
// geometry and taken on the host, as the reference's filter does.
void Transport::installWale(const ReactiveWaleOptionsV1& options) {
    require(options.abiVersion == 1 && options.structBytes == sizeof(options), "Invalid WALE options ABI");
    require(!wale && !highestVersion && !stats.stages && !stats.stepQueries && profile.conservedUploads == 0 &&
                profile.stateUploadBytes == 0,
            "WALE is immutable and must be installed before stepping");
    require(!stepAttached, "WALE must be installed before a resident step is created on this transport");
    require(!cfg.mechanical, "WALE mechanical-environment closure is not implemented");
    require(std::isfinite(options.Cw) && options.Cw >= 0, "Invalid WALE coefficient");
    const uint64_t before = stats.allocatedBytes;
    std::vector<double> inverse(cfg.cells), width(cfg.cells);
    downloadBytes(inverse.data(), invV, cfg.cells * sizeof(double));
    finish();
    for (size_t c = 0; c < cfg.cells; ++c) {
        const double volume = 1 / inverse[c];
        width[c] = std::isfinite(volume) && volume > 0 ? std::cbrt(volume) : std::numeric_limits<double>::quiet_NaN();
    }
    if (!gradient) gradient = allocate<double>(cfg.cells * 9);
    if (!visc) visc = allocate<double4>(cfg.faces);
    nut = allocate<double>(cfg.cells);
    filterWidth = allocate<double>(cfg.cells);
    waleError = allocate<unsigned>(1);
    upload(filterWidth, width.data(), cfg.cells);
    RAMD_HIP(hipMemsetAsync(nut, 0, cfg.cells * sizeof(double), stream));
    finish();
    waleCw = options.Cw;
    wale = true;
    cartesian.valid = false;
    waleProfile.workspaceBytes = stats.allocatedBytes - before;
}

// setCapillary (reactiveTransport.cpp:859-885): once, before any stage.
// The cell length cbrt(V) that gates the interface normal and the capillary
// wave step is taken on the host from the uploaded inverse volumes, as the
// reference's host arithmetic takes it.
void Transport::installCapillary(const ReactiveCapillaryOptionsV1& o) {
    require(o.abiVersion == 1 && o.structBytes == sizeof(o) && std::isfinite(o.sigma) && o.sigma > 0 &&
                std::isfinite(o.capillaryCfl) && o.capillaryCfl > 0 && o.capillaryCfl <= 1 &&
                std::isfinite(o.geometryEpsilon) && o.geometryEpsilon >= 0 && o.geometryEpsilon < 1,
            "Invalid capillary options ABI/model");
    require(!capillary && !attemptOpen && !stats.stages && !highestVersion,
            "Capillary model must be installed once before stepping");
    require(!cfg.mechanical && size_t(nv) == cfg.species + 5 && o.condensableSpecies >= 0 &&
                size_t(o.condensableSpecies) < cfg.species,
            "Capillary model needs one mapped conserved liquid inventory");
    require(!stepAttached, "The capillary model must be installed before a resident step is created on this transport");
    std::vector<double> inverse(cfg.cells), length(cfg.cells);
    downloadBytes(inverse.data(), invV, cfg.cells * sizeof(double));
    finish();
    for (size_t c = 0; c < cfg.cells; ++c) length[c] = std::cbrt(1 / inverse[c]);
    capColor = allocate<double>(all);
    capGradient = allocate<double>(cfg.cells * 3);
    capNormal = allocate<double>(cfg.cells * 3);
    capArea = allocate<double>(cfg.cells);
    capCurvature = allocate<double>(cfg.cells);
    capEnergy = allocate<double>(cfg.cells);
    capLength = allocate<double>(cfg.cells);
    capError = allocate<unsigned>(1);
    if (!visc) visc = allocate<double4>(cfg.faces);
    for (double* a : {capGradient, capNormal}) RAMD_HIP(hipMemsetAsync(a, 0, cfg.cells * 3 * sizeof(double), stream));
    for (double* a : {capArea, capCurvature, capEnergy})
        RAMD_HIP(hipMemsetAsync(a, 0, cfg.cells * sizeof(double), stream));
    RAMD_HIP(hipMemsetAsync(capColor, 0, all * sizeof(double), stream));
    upload(capLength, length.data(), cfg.cells);
    finish();
    capSigma = o.sigma;
    capCfl = o.capillaryCfl;
    capEpsilon = o.geometryEpsilon;
    capSpecies = int(o.condensableSpecies);
    capillary = true;
    cartesian.valid = false;
    capillaryProfile.capillaryWorkspaceBytes = sizeof(double) * (all + 9 * cfg.cells) + sizeof(uint32_t);
}

// capillaryGeometry (reactiveTransport.cpp:886-962), diffuse geometry: the
// colours are checked and uploaded, the gradient and curvature kernels run,
// and the surface energy (and, when asked, curvature and normals) come back.
// The geometry is a pure function of the colours, so a request with the same
// colour bytes (and no normals asked) returns the kept outputs.
void Transport::capillaryGeometry(const double* cellColor, const double* fixedColor, double* energy,
                                  double* curvature, double* normal) {
    require(capillary && cellColor && energy && (!cfg.fixed || fixedColor),
            "Missing capillary model, color or surface-energy output");
    for (size_t c = 0; c < cfg.cells; ++c)
        require(std::isfinite(cellColor[c]) && cellColor[c] >= 0 && cellColor[c] <= 1, "Invalid material color");
    for (size_t c = 0; c < cfg.fixed; ++c)
        require(std::isfinite(fixedColor[c]) && fixedColor[c] >= 0 && fixedColor[c] <= 1,
                "Invalid fixed material color");
    auto same = [](const std::vector<double>& kept, const double* a, size_t n) {
        return kept.size() == n && (n == 0 || std::memcmp(kept.data(), a, n * sizeof(double)) == 0);
    };
    if (!normal && colorReady && geometryShadowValid && (!curvature || !curvatureShadow.empty()) &&
        same(colorShadow, cellColor, cfg.cells) && same(fixedColorShadow, fixedColor, cfg.fixed)) {
        std::copy(energyShadow.begin(), energyShadow.end(), energy);
        if (curvature) std::copy(curvatureShadow.begin(), curvatureShadow.end(), curvature);
        return;
    }
    geometryShadowValid = false;
    colorReady = false;
    RAMD_HIP(hipMemsetAsync(capError, 0, sizeof(unsigned), stream));
    upload(capColor, cellColor, cfg.cells);
    if (cfg.fixed) upload(capColor + cfg.cells, fixedColor, cfg.fixed);
    const unsigned blocks = gridFor(cfg.cells, 256);
    hipLaunchKernelGGL(interfaceGradientKernel, dim3(blocks), dim3(256), 0, stream, faces, rowDev, incidenceDev, invV,
                       capLength, cfg.cells, capEpsilon, capColor, capGradient, capNormal, capArea, capError);
    RAMD_HIP(hipGetLastError());
    launched();
    hipLaunchKernelGGL(interfaceCurvatureKernel, dim3(blocks), dim3(256), 0, stream, faces, rowDev, incidenceDev,
                       invV, cfg.cells, capSigma, capNormal, capArea, capCurvature, capEnergy, capError);
    RAMD_HIP(hipGetLastError());
    launched();
    downloadBytes(energy, capEnergy, cfg.cells * sizeof(double));
    if (curvature) downloadBytes(curvature, capCurvature, cfg.cells * sizeof(double));
    if (normal) downloadBytes(normal, capNormal, cfg.cells * 3 * sizeof(double));
    unsigned status = 0;
    downloadBytes(&status, capError, sizeof(status));
    finish();
    require(!status, "Invalid capillary geometry");
    colorReady = true;
    ++capillaryProfile.geometryBuilds;
    colorShadow.assign(cellColor, cellColor + cfg.cells);
    fixedColorShadow.assign(fixedColor ? fixedColor : cellColor, fixedColor ? fixedColor + cfg.fixed : cellColor);
    energyShadow.assign(energy, energy + cfg.cells);
    if (curvature) curvatureShadow.assign(curvature, curvature + cfg.cells);
    else curvatureShadow.clear();
    geometryShadowValid = true;
    capillaryProfile.geometryCells += cfg.cells;
    capillaryProfile.geometryKernels += 2;
    capillaryProfile.colorUploadBytes += (cfg.cells + cfg.fixed) * sizeof(double);
    capillaryProfile.geometryDownloadBytes +=
        cfg.cells * sizeof(double) * (1 + (curvature ? 1 : 0) + (normal ? 3 : 0));
}

void Transport::launchFaceSpeeds() {
    hipLaunchKernelGGL(faceSpeedKernel, dim3(gridFor(cfg.faces, 256)), dim3(256), 0, stream, faces, cfg.faces, prim,
                       all, state, cfg.waveFactor, cfg.viscosity, cfg.conductivity, wale ? nut : nullptr, faceSpeed);
    RAMD_HIP(hipGetLastError());
    launched();
    ++deviceProfile.cflFaceLaunches;
}

void Transport::gather() {
    hipLaunchKernelGGL(boundaryPartialKernel, dim3(unsigned(nv * partitions)), dim3(256), 0, stream, faces, work,
                       visc, boundaryDev, nBoundary, partitions, q, all, ns, boundaryPartial);
    RAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(boundaryFinishKernel, dim3(unsigned(nv)), dim3(256), 0, stream, boundaryPartial, partitions,
                       boundaryRate);
    RAMD_HIP(hipGetLastError());
    launched();
    launched();
}

void Transport::boundaryRatesFrom(const double* qv, const double* p, const double* c, const double* r, double* out,
                                  unsigned long long* failures, hipStream_t s, size_t stride) {
    if (nBoundary) {
        hipLaunchKernelGGL(boundaryHllKernel, dim3(gridFor(nBoundary, 256)), dim3(256), 0, s, faces, boundaryDev,
                           nBoundary, qv, stride, ns, p, c, r, cfg.waveFactor, work, scheme, failures);
        RAMD_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(boundaryPartialKernel, dim3(unsigned(nv * partitions)), dim3(256), 0, s, faces, work,
                       static_cast<const double4*>(nullptr), boundaryDev, nBoundary, partitions, qv, stride, ns,
                       boundaryPartial);
    RAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(boundaryFinishKernel, dim3(unsigned(nv)), dim3(256), 0, s, boundaryPartial, partitions, out);
    RAMD_HIP(hipGetLastError());
}

void Transport::stableStep(double cfl, double maximum, double* dt) {  // reactiveTransport.cpp:1445-1456
    require(dt && std::isfinite(cfl) && cfl > 0 && cfl <= .5 && std::isfinite(maximum) && maximum > 0,
            "Invalid transport CFL controls");
    const unsigned blocks = gridFor(cfg.cells, 256);
    const CapillaryStepArgs cs{capillary ? capArea : nullptr, capLength, state, capSigma, capCfl, capEpsilon};
    hipLaunchKernelGGL(cellStepKernel, dim3(blocks), dim3(256), 0, stream, faces, faceSpeed, rowDev, incidenceDev, invV,
                       cfg.cells, cfl, maximum, blockMin, cs);
    RAMD_HIP(hipGetLastError());
    hipLaunchKernelGGL(minReduceKernel, dim3(1), dim3(256), 0, stream, blockMin, size_t(blocks), blockMin + blocks);
    RAMD_HIP(hipGetLastError());
    launched();
    launched();
    downloadBytes(dt, blockMin + blocks, sizeof(double));
    finish();
    require(std::isfinite(*dt) && *dt > 0, "Invalid wave/diffusion timestep");
    ++stats.stepQueries;
}

void Transport::checkFluxFailures() {
    unsigned long long failed = 0;
    RAMD_HIP(hipMemcpyAsync(&failed, fluxFailures, sizeof(failed), hipMemcpyDeviceToHost, stream));
    RAMD_HIP(hipMemsetAsync(fluxFailures, 0, sizeof(failed), stream));
    finish();
    require(failed == 0, "HLLC face flux failed");  // the capillary branch's failCapillary()
}

void Transport::advance(double dt, int stage, double* hostQ, double* boundary) {
    if (cartesian.valid) {
        cartesianAdvance(dt, stage);
    } else {
        hipLaunchKernelGGL(advanceKernel, dim3(gridFor(cfg.cells * nv, 256)), dim3(256), 0, stream, faces, work,
                           visc, rowDev, incidenceDev, invV, q, cfg.cells, all, ns, nv, dt, stage, initial);
        RAMD_HIP(hipGetLastError());
        launched();
    }
    if (scheme == kHllc) checkFluxFailures();
    std::swap(q, initial);
    if (hostQ) {
        launchUnpack(q, hostQ);
        profile.conservedDownloadBytes += cfg.cells * nv * sizeof(double);
    }
    downloadBytes(boundary, boundaryRate, nv * sizeof(double));
    finish();
    haveInitial = stage == 0;
    stageDt = dt;
    ++stats.stages;
}

void Transport::rhs(double* hostRhs, double* boundary) {
    double* out = rhsBuffer ? rhsBuffer : (rhsBuffer = allocate<double>(cfg.cells * nv));
    hipLaunchKernelGGL(advanceKernel, dim3(gridFor(cfg.cells * nv, 256)), dim3(256), 0, stream, faces, work, visc,
                       rowDev, incidenceDev, invV, q, cfg.cells, all, ns, nv, 0.0, 2, out);
    RAMD_HIP(hipGetLastError());
    launched();
    if (scheme == kHllc) checkFluxFailures();
    downloadBytes(hostRhs, out, cfg.cells * nv * sizeof(double));
    downloadBytes(boundary, boundaryRate, nv * sizeof(double));
    finish();
}

void Transport::uploadPrimitives(const ReactiveTransportPrimitive* primitive, const ReactiveTransportState* states) {
    require(primitive, "Null transport primitives");  // reactiveTransport.cpp:850-858
    for (size_t c2 = 0; c2 < cfg.cells; ++c2) {
        require(std::isfinite(primitive[c2].rho) && primitive[c2].rho > 0, "Invalid primitive density");
        for (double u : primitive[c2].u) require(std::isfinite(u), "Invalid primitive velocity");
    }
    uploadState(states);
    upload(reinterpret_cast<double*>(staging), reinterpret_cast<const double*>(primitive), cfg.cells * 4);
    hipLaunchKernelGGL(primitiveUploadKernel, dim3(gridFor(cfg.cells, 256)), dim3(256), 0, stream,
                       reinterpret_cast<const ReactiveTransportPrimitive*>(staging), cfg.cells, all, prim);
    RAMD_HIP(hipGetLastError());
    launched();
    profile.primitiveUploadBytes += cfg.cells * sizeof(ReactiveTransportPrimitive);
}

template <class F>
int protect(void* handle, F f) {  // reactiveTransport.cpp:1471-1491
    if (!handle) return -1;
    auto& t = *static_cast<Transport*>(handle);
    std::unique_lock<std::mutex> use(t.entry, std::defer_lock);
    if (!use.try_lock()) return 2;
    try {
        RAMD_HIP(hipSetDevice(t.device));
        f(t);
        t.error[0] = 0;
        return 0;
    } catch (const std::exception& ex) {
        try { (void)hipStreamSynchronize(t.stream); } catch (...) {}
        t.haveInitial = false;
        t.haveResident = false;
        ++t.cost.failedCalls;
        std::snprintf(t.error, sizeof(t.error), "%s", ex.what());
        return 1;
    }
}

int stub(void* handle, const char* name, const char* what) {
    return protect(handle, [&](Transport&) {
        throw std::runtime_error(std::string(name) + ": " + what + " is not provided by the MI355X transport build");
    });
}

}  // namespace ramd

using namespace ramd;

extern "C" {

int reactive_amd_transport_set_flux(void* handle, int scheme) {
    if (!handle) return -1;
    Transport& t = *static_cast<Transport*>(handle);
    std::unique_lock<std::mutex> lock(t.entry, std::try_to_lock);
    if (!lock.owns_lock()) return 2;
    if (scheme != kHll && scheme != kHllc) {
        std::snprintf(t.error, sizeof(t.error), "Unknown face flux scheme %d", scheme);
        return 1;
    }
    t.scheme = scheme;
    t.error[0] = 0;
    return 0;
}

void* reactive_transport_create(int backend, const ReactiveTransportConfig* cfg, const double* volumes,
                                const ReactiveTransportFace* faces, const double* q, const ReactiveTransportState* states,
                                const double* gasY, const double* gasH, char* error, size_t size) {
    try {
        require(cfg, "Null transport configuration");
        require(!cfg->diffusivity || (gasY && gasH), "Missing fixed gas properties");
        return new Transport(backend, *cfg, volumes, faces, q, states);
    } catch (const std::exception& ex) {
        writeError(error, size, ex.what());
        return nullptr;
    }
}

void* reactive_transport_create_v2(int backend, const ReactiveTransportConfig* cfg,
                                   const ReactiveTransportOptionsV2* options, const double* volume,
                                   const ReactiveTransportFace* faces, const double* q,
                                   const ReactiveTransportState* states, const double* y, const double* h, char* error,
                                   size_t size) {
    (void)y; (void)h;
    try {
        require(cfg && options, "Missing v2 configuration");
        return new Transport(backend, *cfg, volume, faces, q, states, options);
    } catch (const std::exception& ex) {
        writeError(error, size, ex.what());
        return nullptr;
    }
}

void* reactive_transport_create_v21(int backend, const ReactiveTransportConfig* cfg,
                                    const ReactiveTransportOptionsV21* opt, const double* volume,
                                    const ReactiveTransportFace* faces, const double* q,
                                    const ReactiveTransportState* state, const double* y, const double* h, char* error,
                                    size_t errorSize) {
    (void)y; (void)h;
    try {  // reactiveTransport.cpp:1837-1852
        require(cfg && opt && opt->abiVersion == 1 && opt->structBytes == sizeof(*opt), "Invalid v2.1 staging options");
        require(opt->slots == 1, "This version supports one event-owned staging slot");
        require(opt->blockThreads == 32 || opt->blockThreads == 64 || opt->blockThreads == 128 ||
                    opt->blockThreads == 256,
                "Unsupported transport block size");
        const size_t cellBytes = cfg->variables * sizeof(double);
        require(cellBytes > 0 && opt->slotBytes >= cellBytes, "Staging budget cannot hold one cell");
        require(opt->pinnedBudgetBytes >= opt->slotBytes, "Pinned staging budget smaller than slot budget");
        const size_t cells = opt->bridgeCellsOverride ? opt->bridgeCellsOverride : opt->slotBytes / cellBytes;
        require(cells > 0 && cells * cellBytes <= opt->slotBytes, "Bridge override exceeds byte budget");
        ReactiveTransportOptionsV2 old{};
        old.abiVersion = 1;
        old.structBytes = sizeof(old);
        old.bridgeCells = cells;
        old.recomputeGas = opt->recomputeGas;
        std::memcpy(old.physicalModelHash, opt->physicalModelHash, 65);
        return new Transport(backend, *cfg, volume, faces, q, state, &old, opt);
    } catch (const std::exception& ex) {
        writeError(error, errorSize, ex.what());
        return nullptr;
    }
}

void reactive_transport_destroy(void* t) { delete static_cast<Transport*>(t); }
const char* reactive_transport_error(void* t) { return t ? static_cast<Transport*>(t)->error : "Null transport"; }
int reactive_transport_is_cuda(void* t) { return t ? 1 : 0; }

int reactive_transport_stats(void* t, ReactiveTransportStats* result) {
    return protect(t, [&](Transport& x) {
        require(result, "Null transport statistics");
        *result = x.stats;
    });
}

int reactive_transport_stable_step(void* t, const double* q, const ReactiveTransportState* state, double cfl,
                                   double maximum, double* dt) {
    return protect(t, [&](Transport& x) {
        require(dt && std::isfinite(cfl) && cfl > 0 && cfl <= .5 && std::isfinite(maximum) && maximum > 0,
                "Invalid transport CFL controls");
        x.haveResident = false;
        x.uploadQ(q);
        x.prepareResident(state, false);
        x.stableStep(cfl, maximum, dt);
    });
}

int reactive_transport_stable_step_primitives(void* t, const ReactiveTransportPrimitive* primitive,
                                              const ReactiveTransportState* state, double cfl, double maximum,
                                              double* dt) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1647-1666
        x.uploadPrimitives(primitive, state);
        if (x.wale) x.buildWale();
        x.launchFaceSpeeds();
        x.stableStep(cfl, maximum, dt);
    });
}

int reactive_transport_rhs(void* t, const double* q, const ReactiveTransportState* states, const double* gasY,
                           const double* gasH, double* rhs, double* boundary) {
    (void)gasY; (void)gasH;
    return protect(t, [&](Transport& x) {
        require(rhs && boundary, "Null transport output buffer");
        x.haveInitial = false;
        x.haveResident = false;
        x.uploadQ(q);
        x.prepareResident(states, true);
        x.gather();
        x.rhs(rhs, boundary);
    });
}

int reactive_transport_stage(void* t, double* q, const ReactiveTransportState* states, const double* gasY,
                             const double* gasH, double dt, int stage, double* boundary) {
    (void)gasY; (void)gasH;
    return protect(t, [&](Transport& x) {
        require(std::isfinite(dt) && dt > 0 && (stage == 0 || stage == 1), "Invalid RK stage");
        require(stage == 0 || (x.haveInitial && x.stageDt == dt), "RK stage 1 requires matching stage 0");
        require(q && boundary, "Null transport output buffer");
        x.haveResident = false;
        x.uploadQ(q);
        x.prepareResident(states, true);
        x.gather();
        x.advance(dt, stage, q, boundary);
    });
}

int reactive_transport_upload_conserved(void* t, const double* q, uint64_t version) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1556-1564
        require(version, "Conserved version must be nonzero");
        if (x.haveResident && version == x.residentVersion) return;
        require(version > x.highestVersion, "Conserved upload version must increase");
        x.haveInitial = false;
        x.uploadQ(q);
        x.finish();
        x.residentVersion = x.highestVersion = version;
        x.haveResident = true;
    });
}

int reactive_transport_stage_resident(void* t, const ReactiveTransportState* states, const double* gasY,
                                      const double* gasH, double dt, int stage, uint64_t inputVersion,
                                      uint64_t outputVersion, double* q, double* boundary) {
    (void)gasY; (void)gasH;
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1709-1719
        require(x.haveResident && x.residentVersion == inputVersion, "Stale resident conserved state");
        require(outputVersion > x.highestVersion, "Output conserved version must increase");
        require(std::isfinite(dt) && dt > 0 && (stage == 0 || stage == 1), "Invalid RK stage");
        require(stage == 0 || (x.haveInitial && x.stageDt == dt), "RK stage 1 requires matching stage 0");
        require(boundary, "Null RK output buffer");
        x.prepareResident(states, true);
        x.gather();
        x.advance(dt, stage, q, boundary);
        x.residentVersion = x.highestVersion = outputVersion;
        ++x.profile.residentStages;
    });
}

int reactive_transport_begin_attempt(void* handle, const char* hash, uint64_t id) {
    return protect(handle, [&](Transport& t) {  // reactiveTransport.cpp:1749-1759
        require(hash && t.physicalHash == hash, "Transport physical model hash mismatch");
        require(!t.attemptOpen && id > t.lastAttempt, "Stale/open transport attempt");
        t.attemptId = t.lastAttempt = id;
        t.tokenStage = 0;
        t.attemptOpen = true;
        t.haveResident = t.haveInitial = false;
    });
}

int reactive_transport_end_attempt(void* handle, uint64_t id, int commit) {
    return protect(handle, [&](Transport& t) {  // reactiveTransport.cpp:1760-1771
        require((commit == 0 || commit == 1) && t.attemptOpen && id == t.attemptId, "Wrong transport attempt completion");
        require(!commit || t.tokenStage == 2, "Cannot commit incomplete transport RK");
        t.finish();
        t.attemptOpen = false;
        t.haveInitial = false;
        if (!commit) t.haveResident = false;
    });
}

int reactive_transport_advance_resident_v2(void* handle, ReactiveTransportToken input, ReactiveTransportToken output,
                                           const ReactiveTransportState* state, const ReactiveGasPartition* partition,
                                           const double* y, const double* h, double dt, double* boundary) {
    (void)partition; (void)y; (void)h;
    return protect(handle, [&](Transport& t) {  // reactiveTransport.cpp:1772-1783
        require(t.attemptOpen && input.attemptId == t.attemptId && output.attemptId == input.attemptId,
                "Wrong RK attempt");
        require(input.stageId == t.tokenStage && input.stageId < 2 && output.stageId == input.stageId + 1,
                "Stale RK stage token");
        require(t.haveResident && input.contentVersion == t.residentVersion &&
                    output.contentVersion > t.highestVersion,
                "Stale RK content token");
        require(std::isfinite(dt) && dt > 0 && boundary, "Invalid resident RK inputs");
        require(input.stageId == 0 || (t.haveInitial && t.stageDt == dt), "Missing matching RK stage 0");
        t.prepareResident(state, true);
        t.gather();
        t.advance(dt, int(input.stageId), nullptr, boundary);
        t.tokenStage = output.stageId;
        t.highestVersion = t.residentVersion = output.contentVersion;
        ++t.profile.residentStages;
    });
}

int reactive_transport_download_conserved(void* handle, ReactiveTransportToken token, double* output) {
    return protect(handle, [&](Transport& t) {  // reactiveTransport.cpp:1784-1788
        require(t.attemptOpen && token.attemptId == t.attemptId && token.stageId == t.tokenStage && t.haveResident &&
                    token.contentVersion == t.residentVersion,
                "Stale explicit download token");
        require(output, "Null unpack output");
        t.launchUnpack(t.q, output);
        t.profile.conservedDownloadBytes += t.cfg.cells * t.nv * sizeof(double);
        t.finish();
    });
}

int reactive_transport_memory_v2(void* handle, ReactiveTransportMemoryV2* out) {
    return protect(handle, [&](Transport& t) {
        require(out, "Missing memory output");
        const auto& c = t.cfg;
        *out = {t.stats.allocatedBytes, t.peakBytes, t.bridgeCells * c.variables * sizeof(double),
                2 * (c.cells + c.fixed) * c.variables * sizeof(double), 0,
                t.rhsBuffer ? c.cells * c.variables * sizeof(double) : 0, t.deviceCopies, t.synchronizations};
    });
}

int reactive_transport_profile(void* t, ReactiveTransportProfile* result) {
    return protect(t, [&](Transport& x) {
        require(result, "Null transport profile");
        *result = x.profile;
    });
}

int reactive_transport_device_profile(void* t, ReactiveTransportDeviceProfile* result) {
    return protect(t, [&](Transport& x) {
        require(result, "Null device profile");
        *result = x.deviceProfile;
    });
}

int reactive_transport_profile_v21(void* handle, ReactiveTransportProfileV21* out) {
    return protect(handle, [&](Transport& t) {
        require(out, "Missing transport v2.1 profile");
        auto result = t.cost;
        result.memcpyCalls = t.memcpyCalls;
        result.payloadBytes = t.stats.uploadedBytes + t.stats.downloadedBytes + t.deviceCopies;
        result.slotCells = t.bridgeCells;
        result.slots = 1;
        result.deviceStagingBytes = std::max(t.all, t.cfg.cells) * t.cfg.variables * sizeof(double);
        *out = result;
    });
}

// ---- exported stubs (physics that is off on this path) -------------------
int reactive_transport_set_gas_thermo(void* t, const ReactiveGasThermoSpecies*, size_t, const ReactiveGasThermoRegion*,
                                      size_t, const int64_t*, size_t) {
    return stub(t, "reactive_transport_set_gas_thermo", "device NASA gas thermo (diffusive transport)");
}
int reactive_transport_stage_resident_gas(void* t, const ReactiveTransportState*, const ReactiveGasPartition*, double,
                                          int, uint64_t, uint64_t, double*, double*) {
    return stub(t, "reactive_transport_stage_resident_gas", "diffusive gas transport");
}
int reactive_transport_rhs_gas(void* t, const double*, const ReactiveTransportState*, const ReactiveGasPartition*,
                               double*, double*) {
    return stub(t, "reactive_transport_rhs_gas", "diffusive gas transport");
}
int reactive_transport_gas_properties_resident(void* t, const ReactiveTransportState*, const ReactiveGasPartition*,
                                               double*, double*) {
    return stub(t, "reactive_transport_gas_properties_resident", "diffusive gas transport");
}
int reactive_transport_set_capillary_v1(void* t, const ReactiveCapillaryOptionsV1* options) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1566-1568
        require(options, "Null capillary options");
        x.installCapillary(*options);
    });
}
int reactive_transport_capillary_geometry_v1(void* t, const double* color, const double* fixedColor,
                                             double* surfaceEnergy, double* curvature, double* normalXYZ) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1591-1594
        x.capillaryGeometry(color, fixedColor, surfaceEnergy, curvature, normalXYZ);
    });
}
int reactive_transport_capillary_profile_v1(void* t, ReactiveCapillaryProfileV1* profile) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1620-1624
        require(profile && profile->abiVersion == 1 && profile->structBytes == sizeof(*profile),
                "Invalid capillary profile ABI");
        *profile = x.capillaryProfile;
    });
}
int reactive_transport_replace_stage_state_v1(void* t, ReactiveTransportToken before, uint64_t version,
                                              const double* q) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1636-1645
        // the post-RK1 state of an open attempt is replaced (after the driver's
        // coupled flash); the stage-0 snapshot that RK2 needs is kept
        require(x.attemptOpen && x.tokenStage == 1 && before.attemptId == x.attemptId && before.stageId == 1 &&
                    x.haveResident && before.contentVersion == x.residentVersion && x.haveInitial &&
                    version > x.highestVersion && q,
                "Invalid post-RK1 capillary/flash state replacement");
        x.uploadQ(q);
        x.finish();
        x.residentVersion = x.highestVersion = version;
    });
}
int reactive_transport_capillary_resident_begin_v1(void* t, size_t, const double*, const double*, const double*,
                                                   const void*, const double*, double, int, int*) {
    return stub(t, "reactive_transport_capillary_resident_begin_v1", "resident capillary closure");
}
int reactive_transport_capillary_resident_select_v1(void* t, int, void*, int*) {
    return stub(t, "reactive_transport_capillary_resident_select_v1", "resident capillary closure");
}
int reactive_transport_capillary_resident_update_v1(void* t, size_t, double*, int*) {
    return stub(t, "reactive_transport_capillary_resident_update_v1", "resident capillary closure");
}
int reactive_transport_capillary_resident_relax_v1(void* t) {
    return stub(t, "reactive_transport_capillary_resident_relax_v1", "resident capillary closure");
}
int reactive_transport_capillary_resident_commit_v1(void* t, void*, double*, double*, double*) {
    return stub(t, "reactive_transport_capillary_resident_commit_v1", "resident capillary closure");
}
int reactive_transport_install_cartesian_v1(void* t, const void*) {
    return stub(t, "reactive_transport_install_cartesian_v1", "Cartesian implicit-interface geometry");
}
int reactive_transport_cartesian_info_v1(void* t, void*) {
    return stub(t, "reactive_transport_cartesian_info_v1", "Cartesian implicit-interface geometry");
}
int reactive_transport_set_implicit_v1(void* t, const void*) {
    return stub(t, "reactive_transport_set_implicit_v1", "implicit capillary geometry");
}
int reactive_transport_implicit_coefficients_v1(void* t, double*, const double*, size_t, size_t*) {
    return stub(t, "reactive_transport_implicit_coefficients_v1", "implicit capillary geometry");
}
int reactive_transport_geometric_diagnostic_v1(void* t, const void*, const double*, const ReactiveTransportState*,
                                               const double*, const void*, const void*, double*, double*, void*) {
    return stub(t, "reactive_transport_geometric_diagnostic_v1", "geometric capillary diagnostic");
}
int reactive_transport_set_wale_v1(void* t, const ReactiveWaleOptionsV1* options) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1668-1670
        require(options, "Null WALE options");
        x.installWale(*options);
    });
}
int reactive_transport_wale_profile_v1(void* t, ReactiveWaleProfileV1* profile) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1699-1702
        require(profile && profile->abiVersion == 1 && profile->structBytes == sizeof(*profile),
                "Invalid WALE profile ABI");
        *profile = x.waleProfile;
    });
}
int reactive_transport_wale_primitives_v1(void* t, const ReactiveTransportPrimitive* primitive,
                                          const ReactiveTransportState* state, double* nut) {
    return protect(t, [&](Transport& x) {  // reactiveTransport.cpp:1703-1708
        require(x.wale && nut, "WALE not selected or missing output");
        x.uploadPrimitives(primitive, state);
        x.buildWale();
        x.downloadBytes(nut, x.nut, x.cfg.cells * sizeof(double));
        x.finish();
    });
}
int reactive_transport_set_wale_scalars_v1(void* t, const void*) {
    return stub(t, "reactive_transport_set_wale_scalars_v1", "WALE scalar closure");
}
int reactive_transport_wale_scalar_profile_v1(void* t, void*) {
    return stub(t, "reactive_transport_wale_scalar_profile_v1", "WALE scalar closure");
}
int reactive_transport_wale_scalar_fields_v1(void* t, const double*, const double*, const double*, const double*) {
    return stub(t, "reactive_transport_wale_scalar_fields_v1", "WALE scalar closure");
}
int reactive_transport_set_wale_pr_model_v1(void* t, const void*) {
    return stub(t, "reactive_transport_set_wale_pr_model_v1", "WALE PR properties");
}
int reactive_transport_wale_pr_properties_v1(void* t, const void*, const double*, size_t, const void*) {
    return stub(t, "reactive_transport_wale_pr_properties_v1", "WALE PR properties");
}
int reactive_transport_wale_pr_properties_v2(void* t, const void*, const double*, size_t, const void*, int) {
    return stub(t, "reactive_transport_wale_pr_properties_v2", "WALE PR properties");
}
int reactive_transport_wale_pr_profile_v1(void* t, void*) {
    return stub(t, "reactive_transport_wale_pr_profile_v1", "WALE PR properties");
}
void* reactive_closure_scalar_create_v1(size_t, char* error, size_t errorSize) {
    writeError(error, errorSize,
               "reactive_closure_scalar_create_v1: the scalar closure candidate is not provided by the MI355X build");
    return nullptr;
}
void reactive_closure_scalar_destroy_v1(void*) {}
int reactive_closure_scalar_run_v1(void*, const void*, size_t, void*, void*, char* error, size_t errorSize) {
    writeError(error, errorSize,
               "reactive_closure_scalar_run_v1: the scalar closure candidate is not provided by the MI355X build");
    return 1;
}

}
