// SPDX-License-Identifier: GPL-3.0-or-later
// CPU ORACLE — test infrastructure only (see oracle.h).  Serial restatement
// of the reference's transport operators for backend 0:
// Cells / Gradients / WaleViscosities / Faces / Rhs / Advance / FaceSpeeds /
// Step and the Minimum and BoundaryPartial/BoundaryFinish reductions
// (src/reactiveTransport/reactiveTransportKernels.h, reactiveTransport.cpp).
// Scope: constant viscosity and conductivity, the WALE stress closure
// (reactiveWale.h; no scalar closure), diffusivity = 0, no mechanical
// closure, no capillary model; other configurations are rejected at create.
// Face flux: HLL (reactiveTransportKernels.h:343-356), or the HLLC option
// (scheme 1): the capillary branch's contact-preserving flux
// (reactiveBalancedCapillary.h:55-142) with sigma = curvature = 0 and no
// surface stress, whose coefficients replace the HLL ones
// (reactiveTransportKernels.h:358-417); viscous traction and work and the
// conductive heat flux follow (:418-439).
#include "oracle.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {
namespace {

inline double mn(double a, double b) { return a < b ? a : b; }
inline double mx(double a, double b) { return a > b ? a : b; }
void require(bool ok, const char* message) { if (!ok) throw std::runtime_error(message); }

struct Prim { double rho, u[3]; };
struct Work { double advectL, advectR, pressureMomentum, pressureEnergy; };
// Viscous traction and the viscous + conductive energy flux of a face
// (FaceWork::traction/energy, reactiveTransportKernels.h:418-439).
struct Visc { double traction[3], energy; };
struct Primitive { double rho, u[3]; };
static_assert(sizeof(Primitive) == 32, "ReactiveTransportPrimitive");

inline bool finite(double x) { return x == x && x <= 1.7976931348623157e308 && x >= -1.7976931348623157e308; }

// WALE kinematic eddy viscosity of one cell (ReactiveWale::eddyViscosity,
// reactiveWale.h:93-167): the invariants of the velocity gradient g
// (row-major, g[3i+j] = d u_i / d x_j) scaled by max |g_ij|, S = symm(g),
// Sd = dev(symm(g g)), then nu_t = (Cw delta)^2 (Sd:Sd)^(3/2) /
// ((S:S)^(5/2) + (Sd:Sd)^(5/4)) with the product Cw^2 delta^2 scale ratio
// assembled from mantissas and exponents (:47-80).  False where the
// reference fails.
bool waleEddyViscosity(const double* g, double delta, double cw, double& nut) {
    nut = std::nan("");
    if (!std::isfinite(delta) || delta < 0 || !std::isfinite(cw) || cw < 0) return false;
    double scale = 0;
    for (int i = 0; i < 9; ++i) {
        if (!std::isfinite(g[i])) return false;
        const double m = g[i] < 0 ? -g[i] : g[i];
        if (m > scale) scale = m;
    }
    double strain2 = 0, dev2 = 0;
    if (scale != 0) {
        double s[9], sq[9] = {};
        for (int i = 0; i < 9; ++i) s[i] = g[i] / scale;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) sq[3 * i + j] += s[3 * i + k] * s[3 * k + j];
        const double trace = sq[0] + sq[4] + sq[8];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double strain = 0.5 * (s[3 * i + j] + s[3 * j + i]);
                double dev = 0.5 * (sq[3 * i + j] + sq[3 * j + i]);
                if (i == j) dev -= trace / 3.0;
                strain2 += strain * strain;
                dev2 += dev * dev;
            }
        if (!std::isfinite(strain2) || !std::isfinite(dev2)) return false;
    }
    if (scale == 0 || cw == 0 || delta == 0 || dev2 == 0) {
        nut = 0;
        return true;
    }
    const double strainPower = strain2 * strain2 * std::sqrt(strain2);
    const double root = std::sqrt(dev2);
    const double devPower = dev2 * root;
    const double denominator = strainPower + dev2 * std::sqrt(root);
    if (!(denominator > 0) || !std::isfinite(denominator) || !std::isfinite(devPower)) return false;
    const double ratio = devPower / denominator;
    if (ratio == 0) {
        nut = 0;
        return true;
    }
    int ec = 0, ed = 0, eg = 0, er = 0, em = 0;
    const double mc = std::frexp(cw, &ec), md = std::frexp(delta, &ed);
    const double mg = std::frexp(scale, &eg), mr = std::frexp(ratio, &er);
    const double m = std::frexp(mc * mc * md * md * mg * mr, &em);
    nut = std::ldexp(m, 2 * ec + 2 * ed + eg + er + em);
    if (!std::isfinite(nut)) {
        nut = std::nan("");
        return false;
    }
    return true;
}

// HLLC coefficients for one face (reactiveBalancedCapillary.h:55-142, sigma =
// curvature = colorFace = 0, zero surface stress, so pi = p).  Inputs: side
// densities, velocities, pressures, wave speeds a = waveFactor c, total
// energies and the unit normal.  False where the reference's faceFlux fails.
bool hllcFace(double rhoL, const double* uL, double pL, double aL, double EL, double rhoR, const double* uR,
              double pR, double aR, double ER, const double* n, Work& w) {
    // State validity (:40-46; color = 0)
    if (!finite(rhoL) || rhoL <= 0 || !finite(pL) || pL <= 0 || !finite(aL) || aL <= 0 || !finite(EL)) return false;
    if (!finite(rhoR) || rhoR <= 0 || !finite(pR) || pR <= 0 || !finite(aR) || aR <= 0 || !finite(ER)) return false;
    for (int d = 0; d < 3; ++d)
        if (!finite(uL[d]) || !finite(uR[d])) return false;
    double n2 = 0, unL = 0, unR = 0;
    for (int d = 0; d < 3; ++d) {
        if (!finite(n[d])) return false;
        n2 += n[d] * n[d];
        unL += uL[d] * n[d];
        unR += uR[d] * n[d];
    }
    if (!finite(n2) || std::fabs(n2 - 1) > 1e-10) return false;
    const double sl = mn(0, mn(unL - aL, unR - aR)), sr = mx(0, mx(unL + aL, unR + aR));
    const double den = rhoL * (sl - unL) - rhoR * (sr - unR);
    if (!finite(den) || den == 0 || !finite(sl) || !finite(sr) || sr <= sl) return false;
    const double star = (pR - pL + rhoL * unL * (sl - unL) - rhoR * unR * (sr - unR)) / den;
    Work out{0, 0, 0, 0};
    double mass, contact;
    if (!finite(star) || star <= sl || star >= sr) {  // HLL fallback (:79-89)
        const double inv = 1 / (sr - sl);
        contact = (sr * unL - sl * unR) * inv;
        out.advectL = sr * (unL - sl) * inv;
        out.advectR = sl * (sr - unR) * inv;
        mass = out.advectL * rhoL + out.advectR * rhoR;
        out.pressureMomentum = (sr * pL - sl * pR) * inv;
        out.pressureEnergy = (sr * pL * unL - sl * pR * unR) * inv;
    } else {
        const bool useLeft = star >= 0;
        const double rho = useLeft ? rhoL : rhoR, un = useLeft ? unL : unR, wave = useLeft ? sl : sr;
        const double pi = useLeft ? pL : pR, E = useLeft ? EL : ER;
        const bool supersonic = (useLeft && sl >= 0) || (!useLeft && sr <= 0);
        contact = supersonic ? un : star;
        if (supersonic) {  // :96-101
            mass = rho * un;
            out.pressureMomentum = pi;
            out.pressureEnergy = pi * un;
        } else {  // star state (:103-110)
            const double rhoStar = rho * (wave - un) / (wave - star);
            const double piStar = pi + rho * (wave - un) * (star - un);
            const double energyStar = ((wave - un) * E - pi * un + piStar * star) / (wave - star);
            mass = rhoStar * star;
            out.pressureMomentum = piStar + mass * (star - un);
            out.pressureEnergy = star * (energyStar + piStar) - mass * E / rho;
        }
        if (useLeft) out.advectL = mass / rhoL;
        else out.advectR = mass / rhoR;
    }
    // finiteness of the assembled flux (:114-140): zero capillary terms
    const double advective = out.advectL * EL + out.advectR * ER;
    double faceU[3], average = 0, capillaryEnergy = 0;
    for (int d = 0; d < 3; ++d) {
        faceU[d] = 0.5 * (uL[d] + uR[d]);
        average += faceU[d] * n[d];
    }
    for (int d = 0; d < 3; ++d) faceU[d] += (contact - average) * n[d];
    for (int d = 0; d < 3; ++d) capillaryEnergy += faceU[d] * 0.0;
    const double total = advective + out.pressureEnergy + capillaryEnergy;
    if (!finite(mass) || !finite(out.pressureMomentum) || !finite(out.pressureEnergy) || !finite(capillaryEnergy) ||
        !finite(total))
        return false;
    for (int d = 0; d < 3; ++d) {
        const double momentum = out.advectL * rhoL * uL[d] + out.advectR * rhoR * uR[d] + out.pressureMomentum * n[d] + 0.0;
        if (!finite(momentum)) return false;
    }
    w = out;
    return true;
}

struct Transport {
    Config cfg;
    std::vector<Face> faces;
    std::vector<double> inverseVolume;
    std::vector<size_t> row, boundary;
    std::vector<int64_t> incidence;
    std::vector<double> q, initial, rhs;  // SoA: qi(c,k) = k*(cells+fixed)+c
    std::vector<TState> state;
    std::vector<Prim> prim;
    std::vector<Work> work;
    std::vector<Visc> visc;
    std::vector<double> gradient, nut;  // gradient[c*9 + 3i + j] = d u_i / d x_j
    std::vector<double> faceSpeed, step;
    size_t partitions = 1;
    int scheme = 0;  // 0 HLL, 1 HLLC
    bool wale = false, used = false;
    double waleCw = 0;
    bool haveInitial = false;
    double stageDt = 0;
    std::string error;

    size_t all() const { return cfg.cells + cfg.fixed; }
    size_t qi(size_t c, size_t k) const { return k * all() + c; }

    // Constructor checks and CSR incidence (reactiveTransport.cpp:410-488).
    Transport(const Config& c, const double* volume, const Face* f, const double* fixedQ, const TState* fixedStates)
        : cfg(c) {
        require(cfg.cells && cfg.species && cfg.faces, "Empty transport mesh/model");
        require(cfg.variables >= cfg.species + 4 && cfg.variables <= cfg.species + 6, "Wrong transport conserved-variable count");
        require(std::isfinite(cfg.waveFactor) && cfg.waveFactor >= 1, "Invalid transport wave factor");
        require(std::isfinite(cfg.viscosity) && std::isfinite(cfg.conductivity) && std::isfinite(cfg.diffusivity) &&
                    cfg.viscosity >= 0 && cfg.conductivity >= 0 && cfg.diffusivity >= 0,
                "Invalid transport coefficients");
        require(cfg.diffusivity == 0 && !cfg.mechanical,
                "Oracle covers the non-diffusive, non-mechanical transport only");
        require(volume && f, "Null transport geometry");
        row.assign(cfg.cells + 1, 0);
        for (size_t i = 0; i < cfg.cells; ++i) require(std::isfinite(volume[i]) && volume[i] > 0, "Invalid transport volume");
        for (size_t i = 0; i < cfg.faces; ++i) {
            const Face& x = f[i];
            require(x.owner >= 0 && size_t(x.owner) < cfg.cells && x.neighbour >= -1 &&
                        (x.neighbour < 0 || size_t(x.neighbour) < cfg.cells), "Invalid transport face cell");
            require(x.kind >= 0 && x.kind <= 3 && (x.kind == 0) == (x.neighbour >= 0), "Invalid transport face kind");
            require(x.kind != 3 || (x.fixed >= 0 && size_t(x.fixed) < cfg.fixed), "Invalid fixed boundary state index");
            double norm = 0;
            for (double component : x.normal) norm += component * component;
            require(std::isfinite(norm) && std::abs(norm - 1) < 1e-8, "Invalid transport unit normal");
            require(std::isfinite(x.area) && std::isfinite(x.distance) && std::isfinite(x.ownerWeight) && x.area > 0 &&
                        x.distance > 0 && x.ownerWeight >= 0 && x.ownerWeight <= 1, "Invalid transport face geometry");
            ++row[x.owner + 1];
            if (x.neighbour >= 0) ++row[x.neighbour + 1];
            else boundary.push_back(i);
        }
        for (size_t i = 0; i < cfg.cells; ++i) row[i + 1] += row[i];
        incidence.resize(row.back());
        std::vector<size_t> next(row);
        for (size_t i = 0; i < cfg.faces; ++i) {
            incidence[next[f[i].owner]++] = -int64_t(i) - 1;
            if (f[i].neighbour >= 0) incidence[next[f[i].neighbour]++] = int64_t(i) + 1;
        }
        faces.assign(f, f + cfg.faces);
        inverseVolume.resize(cfg.cells);
        for (size_t i = 0; i < cfg.cells; ++i) inverseVolume[i] = 1 / volume[i];
        q.assign(all() * cfg.variables, 0);
        initial.assign(all() * cfg.variables, 0);
        state.assign(all(), TState{});
        prim.assign(all(), Prim{});
        work.assign(cfg.faces, Work{});
        visc.assign(cfg.faces, Visc{});
        if (cfg.viscosity > 0) gradient.assign(cfg.cells * 9, 0);
        faceSpeed.assign(cfg.faces, 0);
        step.assign(cfg.cells, 0);
        partitions = std::min(size_t(256), std::max(size_t(1), (boundary.size() + 255) / 256));
        for (size_t i = 0; i < cfg.fixed; ++i)
            for (size_t k = 0; k < cfg.variables; ++k)
                q[qi(cfg.cells + i, k)] = initial[qi(cfg.cells + i, k)] = fixedQ[i * cfg.variables + k];
        require(!cfg.fixed || fixedStates, "Null transport state");
        validateStates(fixedStates, cfg.fixed);
        for (size_t i = 0; i < cfg.fixed; ++i) state[cfg.cells + i] = fixedStates[i];
        for (size_t i = 0; i < cfg.fixed; ++i) {  // fixed primitives (reactiveTransport.cpp:492-499)
            Prim& p = prim[cfg.cells + i];
            for (size_t k = 0; k < cfg.species; ++k) p.rho += fixedQ[i * cfg.variables + k];
            require(std::isfinite(p.rho) && p.rho > 0, "Invalid fixed density");
            for (int d = 0; d < 3; ++d) {
                p.u[d] = fixedQ[i * cfg.variables + cfg.species + d] / p.rho;
                require(std::isfinite(p.u[d]), "Invalid fixed velocity");
            }
        }
    }
    static void validateStates(const TState* s, size_t n) {  // reactiveTransport.cpp:537-545
        for (size_t c = 0; c < n; ++c)
            require(std::isfinite(s[c].p) && s[c].p > 0 && std::isfinite(s[c].T) && s[c].T > 0 && std::isfinite(s[c].rho) &&
                        s[c].rho > 0 && std::isfinite(s[c].sound) && s[c].sound > 0 && std::isfinite(s[c].cv) && s[c].cv > 0 &&
                        std::isfinite(s[c].gasMass) && s[c].gasMass >= 0 && std::isfinite(s[c].dilatation),
                    "Invalid recovered transport state");
    }
    size_t rightCell(const Face& f) const {  // reactiveTransportKernels.h:134-136
        return f.neighbour >= 0 ? size_t(f.neighbour) : (f.kind == 3 ? cfg.cells + size_t(f.fixed) : size_t(f.owner));
    }
    double rightValue(const Face& f, size_t k) const {  // :137-144
        const size_t r = rightCell(f);
        double value = q[qi(r, k)];
        if (f.kind == 1 && k >= cfg.species && k < cfg.species + 3) {
            double normal = 0;
            for (int d = 0; d < 3; ++d) normal += q[qi(r, cfg.species + d)] * f.normal[d];
            value -= 2 * normal * f.normal[k - cfg.species];
        }
        return value;
    }
    void velocities(const Face& f, double* ul, double* ur) const {  // :145-149
        const Prim& l = prim[f.owner];
        const Prim& r = prim[rightCell(f)];
        double dot = 0;
        for (int d = 0; d < 3; ++d) { ul[d] = l.u[d]; ur[d] = r.u[d]; dot += r.u[d] * f.normal[d]; }
        if (f.kind == 1)
            for (int d = 0; d < 3; ++d) ur[d] -= 2 * dot * f.normal[d];
    }
    double faceFlux(size_t fi, size_t k) const {  // :180-191
        const Face& f = faces[fi];
        const Work& w = work[fi];
        const Visc& v = visc[fi];
        const double ql = q[qi(f.owner, k)], qr = rightValue(f, k);
        double flux = w.advectL * ql + w.advectR * qr;
        if (k >= cfg.species && k < cfg.species + 3) flux += (w.pressureMomentum - 0.0) * f.normal[k - cfg.species];
        if (k == cfg.species + 3) flux += w.pressureEnergy;
        if (k < cfg.species) flux += 0.0;                        // species diffusion, D = 0
        else if (k < cfg.species + 3) flux -= v.traction[k - cfg.species];
        else if (k == cfg.species + 3) flux += v.energy;
        return flux * f.area;
    }
    // faceNut (reactiveTransportKernels.h:80-83) weighted by density: the
    // face's rho nu_t, owner side by ownerWeight.
    double faceRhoNut(const Face& f, const TState& sl, const TState& sr) const {
        const double right = nut[f.neighbour >= 0 ? size_t(f.neighbour) : size_t(f.owner)];
        return f.ownerWeight * sl.rho * nut[f.owner] + (1 - f.ownerWeight) * sr.rho * right;
    }
    void gradients() {  // Gradients, :290-301
        for (size_t c = 0; c < cfg.cells; ++c) {
            double g[9] = {};
            for (size_t j = row[c]; j < row[c + 1]; ++j) {
                const int64_t entry = incidence[j];
                const Face& f = faces[size_t(entry < 0 ? -entry - 1 : entry - 1)];
                double ul[3], ur[3];
                velocities(f, ul, ur);
                const double scale = (entry < 0 ? 1.0 : -1.0) * f.area * inverseVolume[c];
                for (int i = 0; i < 3; ++i)
                    for (int d = 0; d < 3; ++d)
                        g[3 * i + d] += (f.ownerWeight * ul[i] + (1 - f.ownerWeight) * ur[i]) * f.normal[d] * scale;
            }
            for (int j = 0; j < 9; ++j) gradient[c * 9 + j] = g[j];
        }
    }
    void buildWale() {  // buildWale, reactiveTransport.cpp:841-849: Gradients + WaleViscosities (:303-313)
        gradients();
        bool failed = false;
        for (size_t c = 0; c < cfg.cells; ++c) {
            const double volume = 1 / inverseVolume[c];
            double delta = std::nan(""), value = 0;
            bool ok = std::isfinite(volume) && volume > 0;  // filterWidthFromCellVolume (reactiveWale.h:171-182)
            if (ok) {
                delta = std::cbrt(volume);
                ok = std::isfinite(delta) && delta > 0;
            }
            ok = ok && waleEddyViscosity(&gradient[c * 9], delta, waleCw, value);
            if (!ok || !std::isfinite(value) || value < 0 || !std::isfinite(state[c].rho * value)) {
                failed = true;
                continue;
            }
            nut[c] = value;
        }
        require(!failed, "Invalid WALE gradient/filter width/eddy viscosity");
    }
    // The viscous and conductive part of Faces (:418-439) for one face.
    Visc viscousWork(const Face& f, const double* ul, const double* ur, const TState& sl, const TState& sr) const {
        Visc w{{0, 0, 0}, 0};
        if (cfg.viscosity > 0 || wale) {
            const size_t l = size_t(f.owner), r = rightCell(f);
            double g[9];
            for (int j = 0; j < 9; ++j) {
                g[j] = gradient[l * 9 + j];
                if (f.neighbour >= 0) g[j] = f.ownerWeight * g[j] + (1 - f.ownerWeight) * gradient[r * 9 + j];
            }
            for (int i = 0; i < 3; ++i) {  // replace the normal derivative by the two-point one
                double normal = 0;
                for (int d = 0; d < 3; ++d) normal += g[i * 3 + d] * f.normal[d];
                const double target = (ur[i] - ul[i]) / ((f.kind == 1 ? 2 : 1) * f.distance);
                for (int d = 0; d < 3; ++d) g[i * 3 + d] += f.normal[d] * (target - normal);
            }
            const double div = g[0] + g[4] + g[8];
            const double mu = cfg.viscosity + (wale ? faceRhoNut(f, sl, sr) : 0);
            for (int i = 0; i < 3; ++i)
                for (int d = 0; d < 3; ++d)
                    w.traction[i] += mu * (g[3 * i + d] + g[3 * d + i] - (i == d ? 2 * div / 3 : 0)) * f.normal[d];
            if (f.kind == 1) {  // a slip wall carries the normal stress only
                double dot = 0;
                for (int d = 0; d < 3; ++d) dot += w.traction[d] * f.normal[d];
                for (int d = 0; d < 3; ++d) w.traction[d] = dot * f.normal[d];
            }
            for (int d = 0; d < 3; ++d) w.energy -= w.traction[d] * (f.ownerWeight * ul[d] + (1 - f.ownerWeight) * ur[d]);
        }
        if (cfg.conductivity > 0 && f.kind != 1) w.energy -= cfg.conductivity * (sr.T - sl.T) / f.distance;
        return w;
    }
    void cells() {  // Cells, :193-199
        for (size_t c = 0; c < all(); ++c) {
            Prim p{};
            for (size_t k = 0; k < cfg.species; ++k) p.rho += q[qi(c, k)];
            for (int d = 0; d < 3; ++d) p.u[d] = q[qi(c, cfg.species + d)] / p.rho;
            prim[c] = p;
        }
    }
    void facesHll() {  // Faces: HLL coefficients (:343-356), viscous and conductive terms (:418-439)
        bool waleFailed = false;
        for (size_t fi = 0; fi < cfg.faces; ++fi) {
            const Face& f = faces[fi];
            const TState& sl = state[f.owner];
            const TState& sr = state[rightCell(f)];
            double ul[3], ur[3], unL = 0, unR = 0;
            velocities(f, ul, ur);
            for (int d = 0; d < 3; ++d) { unL += ul[d] * f.normal[d]; unR += ur[d] * f.normal[d]; }
            const double aL = cfg.waveFactor * sl.sound, aR = cfg.waveFactor * sr.sound;
            const double left = mn(0, mn(unL - aL, unR - aR)), right = mx(0, mx(unL + aL, unR + aR));
            const double invWave = 1 / (right - left);
            Work w;
            w.advectL = right * (unL - left) * invWave;
            w.advectR = left * (right - unR) * invWave;
            w.pressureMomentum = (right * sl.p - left * sr.p) * invWave;
            w.pressureEnergy = (right * sl.p * unL - left * sr.p * unR) * invWave;
            if (scheme == 1) {
                const size_t e = cfg.species + 3;
                require(hllcFace(sl.rho, ul, sl.p, aL, q[qi(f.owner, e)], sr.rho, ur, sr.p, aR, rightValue(f, e),
                                 f.normal, w),
                        "HLLC face flux failed");
            }
            work[fi] = w;
            const Visc v = viscousWork(f, ul, ur, sl, sr);
            if (wale && (!std::isfinite(v.energy) || !std::isfinite(v.traction[0]) || !std::isfinite(v.traction[1]) ||
                         !std::isfinite(v.traction[2])))
                waleFailed = true;
            visc[fi] = v;
        }
        require(!waleFailed, "Nonfinite WALE traction/energy flux; stage remains uncommitted");
    }
    void faceSpeeds() {  // FaceSpeeds, :315-342 (no scalar closure)
        for (size_t fi = 0; fi < cfg.faces; ++fi) {
            const Face& f = faces[fi];
            const TState& sl = state[f.owner];
            const TState& sr = state[rightCell(f)];
            double ul[3], ur[3], unL = 0, unR = 0;
            velocities(f, ul, ur);
            for (int d = 0; d < 3; ++d) { unL += ul[d] * f.normal[d]; unR += ur[d] * f.normal[d]; }
            const double aL = cfg.waveFactor * sl.sound, aR = cfg.waveFactor * sr.sound;
            const double left = mn(0, mn(unL - aL, unR - aR)), right = mx(0, mx(unL + aL, unR + aR));
            const double invWave = 1 / (right - left);
            const double rhoNut = wale ? faceRhoNut(f, sl, sr) : 0;
            const double muFace = cfg.viscosity + rhoNut;
            const double D = cfg.diffusivity + mx(mx(4 * muFace / (3 * sl.rho), 4 * muFace / (3 * sr.rho)),
                                                  mx(cfg.conductivity / (sl.rho * sl.cv), cfg.conductivity / (sr.rho * sr.cv)));
            double speed = mx(std::fabs(unL) + aL, std::fabs(unR) + aR) + 2 * (D + 0) / f.distance;
            if (!std::isfinite(unL) || !std::isfinite(unR) || !std::isfinite(speed) || !std::isfinite(invWave) || right <= left)
                speed = -1;
            faceSpeed[fi] = speed;
        }
    }
    double rate(size_t c, size_t k) const {  // Rhs::value, :502-526
        double r = 0;
        for (size_t j = row[c]; j < row[c + 1]; ++j) {
            const int64_t entry = incidence[j];
            const size_t fi = size_t(entry < 0 ? -entry - 1 : entry - 1);
            const double sign = entry < 0 ? -1 : 1;
            r += sign * faceFlux(fi, k) * inverseVolume[c];
        }
        return r;
    }
    static double tree256(double* v) {  // reduceGroups combine order, reactiveTransport.cpp:70-81
        for (size_t stride = 128; stride; stride /= 2)
            for (size_t j = 0; j < stride; ++j) v[j] = v[j] + v[j + stride];
        return v[0];
    }
    void boundaryRates(double* out) const {  // BoundaryPartial/Finish, :233-246
        std::vector<double> partial(cfg.variables * partitions);
        double v[256];
        for (size_t g = 0; g < cfg.variables * partitions; ++g) {
            const size_t k = g / partitions, tile = g % partitions;
            for (size_t lane = 0; lane < 256; ++lane) {
                double sum = 0;
                for (size_t j = tile * 256 + lane; j < boundary.size(); j += partitions * 256) sum += faceFlux(boundary[j], k);
                v[lane] = sum;
            }
            partial[g] = tree256(v);
        }
        for (size_t k = 0; k < cfg.variables; ++k) {
            for (size_t lane = 0; lane < 256; ++lane) v[lane] = lane < partitions ? partial[k * partitions + lane] : 0;
            out[k] = tree256(v);
        }
    }
    // prepareResident's order (reactiveTransport.cpp:1379-1393): Cells, then the
    // WALE viscosities, or for the flux the velocity gradients of a viscous model.
    void prepare(bool transport) {
        if (wale) buildWale();
        else if (transport && cfg.viscosity > 0) gradients();
    }
    void upload(const double* hostQ, const TState* states) {
        used = true;
        require(hostQ, "Null packed source");
        for (size_t c = 0; c < cfg.cells; ++c)
            for (size_t k = 0; k < cfg.variables; ++k) q[qi(c, k)] = hostQ[c * cfg.variables + k];
        require(states, "Null transport state");
        validateStates(states, cfg.cells);
        for (size_t c = 0; c < cfg.cells; ++c) state[c] = states[c];
        cells();
    }
    double minimumStep(double cfl, double maximum) {  // Step + Minimum, :529-550 and reactiveTransport.cpp:1445-1456
        require(std::isfinite(cfl) && cfl > 0 && cfl <= .5 && std::isfinite(maximum) && maximum > 0, "Invalid transport CFL controls");
        for (size_t c = 0; c < cfg.cells; ++c) {
            double denominator = 0;
            bool bad = false;
            for (size_t j = row[c]; j < row[c + 1]; ++j) {
                const int64_t entry = incidence[j];
                const size_t fi = size_t(entry < 0 ? -entry - 1 : entry - 1);
                const double speed = faceSpeed[fi];
                if (!std::isfinite(speed) || speed < 0) { bad = true; break; }
                denominator += faces[fi].area * speed;
            }
            if (bad) { step[c] = -1; continue; }
            step[c] = denominator > 0 ? mn(maximum, cfl / (inverseVolume[c] * denominator)) : maximum;
            if (!std::isfinite(denominator) || denominator < 0) step[c] = -1;
        }
        double dt;
        if (cfg.cells == 1) dt = step[0];
        else {
            dt = HUGE_VAL;
            for (size_t c = 0; c < cfg.cells; ++c) dt = mn(dt, std::isfinite(step[c]) && step[c] > 0 ? step[c] : -1);
        }
        require(std::isfinite(dt) && dt > 0, "Invalid wave/diffusion timestep");
        return dt;
    }
    void advance(double dt, int stage) {  // Advance, :551-560, then swap (reactiveTransport.cpp:1461-1468)
        for (size_t j = 0; j < cfg.cells * cfg.variables; ++j) {
            const size_t c = j % cfg.cells, k = j / cfg.cells, index = qi(c, k);
            const double r = rate(c, k);
            if (stage == 0) initial[index] = q[index] + dt * r;
            else initial[index] = .5 * initial[index] + .5 * (q[index] + dt * r);
        }
        std::swap(q, initial);
        haveInitial = stage == 0;
        stageDt = dt;
    }
};

template <class F> int guard(void* h, F f) {
    if (!h) return -1;
    Transport& t = *static_cast<Transport*>(h);
    try {
        f(t);
        t.error.clear();
        return 0;
    } catch (const std::exception& e) {
        t.haveInitial = false;
        t.error = e.what();
        return 1;
    }
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

void* oracle_transport_create(const Config* cfg, const double* volume, const Face* faces, const double* fixedQ,
                              const TState* fixedStates, char* error, size_t size) {
    try {
        require(cfg, "Null transport configuration");
        return new Transport(*cfg, volume, faces, fixedQ, fixedStates);
    } catch (const std::exception& e) {
        if (error && size) std::snprintf(error, size, "%s", e.what());
        return nullptr;
    }
}
void oracle_transport_destroy(void* h) { delete static_cast<Transport*>(h); }
const char* oracle_transport_error(void* h) { return h ? static_cast<Transport*>(h)->error.c_str() : "Null transport"; }

int oracle_transport_rhs(void* h, const double* q, const TState* states, double* rhs, double* boundary) {
    return guard(h, [&](Transport& t) {
        t.haveInitial = false;
        t.upload(q, states);
        t.prepare(true);
        t.facesHll();
        for (size_t c = 0; c < t.cfg.cells; ++c)
            for (size_t k = 0; k < t.cfg.variables; ++k) rhs[c * t.cfg.variables + k] = t.rate(c, k);
        t.boundaryRates(boundary);
    });
}
int oracle_transport_stage(void* h, double* q, const TState* states, double dt, int stage, double* boundary) {
    return guard(h, [&](Transport& t) {
        require(std::isfinite(dt) && dt > 0 && (stage == 0 || stage == 1), "Invalid RK stage");
        require(stage == 0 || (t.haveInitial && t.stageDt == dt), "RK stage 1 requires matching stage 0");
        t.upload(q, states);
        t.prepare(true);
        t.facesHll();
        t.boundaryRates(boundary);
        t.advance(dt, stage);
        for (size_t c = 0; c < t.cfg.cells; ++c)
            for (size_t k = 0; k < t.cfg.variables; ++k) q[c * t.cfg.variables + k] = t.q[t.qi(c, k)];
    });
}
int oracle_transport_stable_step(void* h, const double* q, const TState* states, double cfl, double maximum, double* dt) {
    return guard(h, [&](Transport& t) {
        require(dt && std::isfinite(cfl) && cfl > 0 && cfl <= .5 && std::isfinite(maximum) && maximum > 0,
                "Invalid transport CFL controls");
        t.upload(q, states);
        t.prepare(false);
        t.faceSpeeds();
        *dt = t.minimumStep(cfl, maximum);
    });
}
// uploadPrimitives (reactiveTransport.cpp:850-858): cell primitives and states.
static void uploadPrimitives(Transport& t, const Primitive* primitive, const TState* states) {
    require(primitive, "Null transport primitives");
    for (size_t c = 0; c < t.cfg.cells; ++c) {
        require(std::isfinite(primitive[c].rho) && primitive[c].rho > 0, "Invalid primitive density");
        for (double u : primitive[c].u) require(std::isfinite(u), "Invalid primitive velocity");
    }
    require(states, "Null transport state");
    Transport::validateStates(states, t.cfg.cells);
    t.used = true;
    for (size_t c = 0; c < t.cfg.cells; ++c) {
        t.state[c] = states[c];
        t.prim[c].rho = primitive[c].rho;
        for (int d = 0; d < 3; ++d) t.prim[c].u[d] = primitive[c].u[d];
    }
}
int oracle_transport_stable_step_primitives(void* h, const Primitive* primitive, const TState* states, double cfl,
                                            double maximum, double* dt) {
    return guard(h, [&](Transport& t) {  // reactiveTransport.cpp:1647-1666
        uploadPrimitives(t, primitive, states);
        if (t.wale) t.buildWale();
        t.faceSpeeds();
        require(dt, "Null transport output buffer");
        *dt = t.minimumStep(cfl, maximum);
    });
}
// installWale (reactiveTransport.cpp:636-652): once, before any upload.
int oracle_transport_set_wale(void* h, double cw) {
    return guard(h, [&](Transport& t) {
        require(!t.wale && !t.used, "WALE is immutable and must be installed before stepping");
        require(std::isfinite(cw) && cw >= 0, "Invalid WALE coefficient");
        t.gradient.assign(t.cfg.cells * 9, 0);
        t.nut.assign(t.cfg.cells, 0);
        t.waleCw = cw;
        t.wale = true;
    });
}
int oracle_transport_wale_primitives(void* h, const Primitive* primitive, const TState* states, double* nut) {
    return guard(h, [&](Transport& t) {  // reactiveTransport.cpp:1703-1708
        require(t.wale && nut, "WALE not selected or missing output");
        uploadPrimitives(t, primitive, states);
        t.buildWale();
        for (size_t c = 0; c < t.cfg.cells; ++c) nut[c] = t.nut[c];
    });
}
int oracle_transport_set_scheme(void* h, int scheme) {
    return guard(h, [&](Transport& t) {
        require(scheme == 0 || scheme == 1, "Unknown face flux scheme");
        t.scheme = scheme;
    });
}
// One HLLC face: left/right = {rho, u0, u1, u2, p, a, E}; out = {advectL,
// advectR, pressureMomentum, pressureEnergy}.  Returns 1, or 0 on failure.
int oracle_hllc_face(const double* left, const double* right, const double* normal, double* out) {
    Work w{};
    if (!hllcFace(left[0], left + 1, left[4], left[5], left[6], right[0], right + 1, right[4], right[5], right[6],
                  normal, w))
        return 0;
    out[0] = w.advectL; out[1] = w.advectR; out[2] = w.pressureMomentum; out[3] = w.pressureEnergy;
    return 1;
}
// The velocity gradients (cells x 9) of the last WALE build or viscous flux.
int oracle_transport_gradient(void* h, double* out) {
    return guard(h, [&](Transport& t) {
        require(!t.gradient.empty() && out, "No velocity gradient on this transport");
        for (size_t j = 0; j < t.gradient.size(); ++j) out[j] = t.gradient[j];
    });
}
// One WALE eddy viscosity: 1 and nut, or 0 where the reference fails.
int oracle_wale_eddy_viscosity(const double* gradient, double delta, double cw, double* nut) {
    return waleEddyViscosity(gradient, delta, cw, *nut) ? 1 : 0;
}
// Conservation totals sum_c q_ck V_c in serial cell order (ReactiveFoam.C:953-958).
void oracle_totals(size_t cells, size_t variables, const double* q, const double* volume, double* out) {
    for (size_t k = 0; k < variables; ++k) out[k] = 0;
    for (size_t c = 0; c < cells; ++c)
        for (size_t k = 0; k < variables; ++k) out[k] += q[c * variables + k] * volume[c];
}

}
