// SPDX-License-Identifier: GPL-3.0-or-later
// HLLC face coefficients for the transport kernels: the reference's
// contact-preserving capillary flux (src/reactiveInterface/
// reactiveBalancedCapillary.h:55-142) with sigma = curvature = colorFace = 0 and
// no surface stress, so the modified pressure pi equals p.  Its coefficients
// replace the HLL ones exactly as the capillary branch of Faces does
// (reactiveTransportKernels.h:358-417).  Same operation order as the
// reference, so results are bitwise equal to its host build.
#ifndef REACTIVE_AMD_FLUX_DEVICE_H
#define REACTIVE_AMD_FLUX_DEVICE_H
#include <hip/hip_runtime.h>

namespace ramd {

enum FluxScheme : int { kHll = 0, kHllc = 1 };

__device__ inline bool fluxFinite(double x) {
    return x == x && x <= 1.7976931348623157e308 && x >= -1.7976931348623157e308;
}

// w = (advectL, advectR, pressureMomentum, pressureEnergy).  rho are the
// recovered-state densities, u the (slip-mirrored) side velocities, a the
// wave-factor-scaled sound speeds, E the total-energy densities.  Returns
// false where the reference's faceFlux fails; w is then left unchanged.
__device__ inline bool hllcCoefficients(double rhoL, const double* uL, double pL, double aL, double EL, double rhoR,
                                        const double* uR, double pR, double aR, double ER, const double* n,
                                        double4& w) {
    if (!fluxFinite(rhoL) || rhoL <= 0 || !fluxFinite(pL) || pL <= 0 || !fluxFinite(aL) || aL <= 0 ||
        !fluxFinite(EL))
        return false;
    if (!fluxFinite(rhoR) || rhoR <= 0 || !fluxFinite(pR) || pR <= 0 || !fluxFinite(aR) || aR <= 0 ||
        !fluxFinite(ER))
        return false;
    for (int d = 0; d < 3; ++d)
        if (!fluxFinite(uL[d]) || !fluxFinite(uR[d])) return false;
    double n2 = 0, unL = 0, unR = 0;
    for (int d = 0; d < 3; ++d) {
        if (!fluxFinite(n[d])) return false;
        n2 += n[d] * n[d];
        unL += uL[d] * n[d];
        unR += uR[d] * n[d];
    }
    if (!fluxFinite(n2) || ::fabs(n2 - 1) > 1e-10) return false;
    const double m1 = unL - aL < unR - aR ? unL - aL : unR - aR;
    const double m2 = unL + aL > unR + aR ? unL + aL : unR + aR;
    const double sl = 0 < m1 ? 0 : m1, sr = 0 > m2 ? 0 : m2;
    const double den = rhoL * (sl - unL) - rhoR * (sr - unR);
    if (!fluxFinite(den) || den == 0 || !fluxFinite(sl) || !fluxFinite(sr) || sr <= sl) return false;
    const double star = (pR - pL + rhoL * unL * (sl - unL) - rhoR * unR * (sr - unR)) / den;
    double4 out = make_double4(0, 0, 0, 0);
    double mass, contact;
    if (!fluxFinite(star) || star <= sl || star >= sr) {  // HLL fallback for waves outside the bounds
        const double inv = 1 / (sr - sl);
        contact = (sr * unL - sl * unR) * inv;
        out.x = sr * (unL - sl) * inv;
        out.y = sl * (sr - unR) * inv;
        mass = out.x * rhoL + out.y * rhoR;
        out.z = (sr * pL - sl * pR) * inv;
        out.w = (sr * pL * unL - sl * pR * unR) * inv;
    } else {
        const bool useLeft = star >= 0;
        const double rho = useLeft ? rhoL : rhoR, un = useLeft ? unL : unR, wave = useLeft ? sl : sr;
        const double pi = useLeft ? pL : pR, E = useLeft ? EL : ER;
        const bool supersonic = (useLeft && sl >= 0) || (!useLeft && sr <= 0);
        contact = supersonic ? un : star;
        if (supersonic) {
            mass = rho * un;
            out.z = pi;
            out.w = pi * un;
        } else {  // star state of the side the contact leaves behind
            const double rhoStar = rho * (wave - un) / (wave - star);
            const double piStar = pi + rho * (wave - un) * (star - un);
            const double energyStar = ((wave - un) * E - pi * un + piStar * star) / (wave - star);
            mass = rhoStar * star;
            out.z = piStar + mass * (star - un);
            out.w = star * (energyStar + piStar) - mass * E / rho;
        }
        if (useLeft) out.x = mass / rhoL;
        else out.y = mass / rhoR;
    }
    // the reference rejects a face whose assembled flux is not finite
    const double advective = out.x * EL + out.y * ER;
    double faceU[3], average = 0, capillaryEnergy = 0;
    for (int d = 0; d < 3; ++d) {
        faceU[d] = 0.5 * (uL[d] + uR[d]);
        average += faceU[d] * n[d];
    }
    for (int d = 0; d < 3; ++d) faceU[d] += (contact - average) * n[d];
    for (int d = 0; d < 3; ++d) capillaryEnergy += faceU[d] * 0.0;
    const double total = advective + out.w + capillaryEnergy;
    if (!fluxFinite(mass) || !fluxFinite(out.z) || !fluxFinite(out.w) || !fluxFinite(capillaryEnergy) ||
        !fluxFinite(total))
        return false;
    for (int d = 0; d < 3; ++d)
        if (!fluxFinite(out.x * rhoL * uL[d] + out.y * rhoR * uR[d] + out.z * n[d] + 0.0)) return false;
    w = out;
    return true;
}

}  // namespace ramd
#endif
