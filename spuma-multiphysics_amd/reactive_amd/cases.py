"""Synthetic Cartesian N2O cases for the HLL + SSPRK2 + HEM-flash step.

Geometry: an N^3 box of cubic cells (edge h), cell c = i + N (j + N k).
Faces follow OpenFOAM's upper-triangular order: internal faces by owner,
then neighbour (+x, +y, +z); then the six boundary patches (-x, +x, -y, +y,
-z, +z) in cell order, all of kind 2 (extrapolate / zero gradient,
reactiveTransportKernels.h:134-136).

State: per cell T ~ U(Tlo, Thi), rho log-uniform, u_d ~ N(0, sigma_u).  The
internal energy is the isochoric (T, V) equilibrium value: single-phase PR
outside the saturation dome, the lever rule between saturated liquid and
vapour inside it.  Saturation comes from equal fugacity of the PR cubic's two
roots, tabulated on a fine temperature grid and interpolated.  This is input
synthesis only; the flash under test recovers (T, p, phase split) from the
resulting conserved state.
"""
import math
import numpy as np

from . import abi, mechanism

R = mechanism.GAS_CONSTANT
OMEGA_A = 4.5723552892138218e-1
OMEGA_B = 7.77960739038885e-2
SQ2 = math.sqrt(2.0)


# ---------------------------------------------------------------- geometry
def cartesian_faces(n, h):
    """Face array (abi.FACE_DTYPE) and per-cell volumes for an n^3 box."""
    nc = n ** 3
    idx = np.arange(nc, dtype=np.int64).reshape(n, n, n)  # [k, j, i]
    own, nei, axis = [], [], []
    # internal faces, grouped by owner then neighbour (+x < +y < +z in index)
    ox = idx[:, :, :-1].ravel(); oy = idx[:, :-1, :].ravel(); oz = idx[:-1, :, :].ravel()
    owners = np.concatenate([ox, oy, oz])
    neighbours = np.concatenate([ox + 1, oy + n, oz + n * n])
    axes = np.concatenate([np.zeros(ox.size, np.int8), np.ones(oy.size, np.int8), np.full(oz.size, 2, np.int8)])
    order = np.lexsort((neighbours, owners))
    own.append(owners[order]); nei.append(neighbours[order]); axis.append(axes[order])
    internal = owners.size
    bsign = []
    for ax, (sl_lo, sl_hi) in enumerate([((slice(None), slice(None), 0), (slice(None), slice(None), n - 1)),
                                          ((slice(None), 0, slice(None)), (slice(None), n - 1, slice(None))),
                                          ((0, slice(None), slice(None)), (n - 1, slice(None), slice(None)))]):
        for sign, sl in ((-1.0, sl_lo), (1.0, sl_hi)):
            cells = np.sort(idx[sl].ravel())
            own.append(cells); nei.append(np.full(cells.size, -1, np.int64))
            axis.append(np.full(cells.size, ax, np.int8)); bsign.append(np.full(cells.size, sign))
    own = np.concatenate(own); nei = np.concatenate(nei); axis = np.concatenate(axis)
    nf = own.size
    faces = np.zeros(nf, dtype=abi.FACE_DTYPE)
    faces["owner"] = own
    faces["neighbour"] = nei
    faces["fixed"] = -1
    normal = np.zeros((nf, 3))
    normal[np.arange(nf), axis] = 1.0
    normal[internal:] *= np.concatenate(bsign)[:, None]
    faces["normal"] = normal
    faces["area"] = h * h
    faces["distance"] = np.where(nei >= 0, h, 0.5 * h)
    faces["ownerWeight"] = np.where(nei >= 0, 0.5, 1.0)
    faces["kind"] = np.where(nei >= 0, 0, 2)
    volume = np.full(nc, h ** 3)
    return faces, volume


# ------------------------------------------------------- PR thermodynamics
class PureFluid:
    """Pure-species Peng-Robinson + NASA9 in numpy (textbook PR; same
    constants as the flash model image)."""

    def __init__(self, name="N2O", a_scale=1.0, b_scale=1.0):
        d = mechanism.SPECIES[name]
        self.W = d["W"]; self.a = d["a"] * a_scale; self.b = d["b"] * b_scale
        self.kappa = mechanism.kappa(d["omega"])
        self.Tc = self.a * OMEGA_B / (self.b * OMEGA_A * R)
        self.ranges = np.array(d["ranges"]); self.coef = np.array(d["nasa9"])

    def alpha(self, T):
        s = np.sqrt(T / self.Tc)
        f = 1.0 + self.kappa * (1.0 - s)
        df = -self.kappa / (2.0 * self.Tc * s)
        return self.a * f * f, self.a * 2.0 * f * df  # a*alpha, d(a alpha)/dT

    def h_ideal(self, T):
        """Molar ideal-gas enthalpy from the NASA9 fits (J/kmol)."""
        T = np.asarray(T, float)
        r = np.clip(np.searchsorted(self.ranges[1:-1], T, side="right"), 0, len(self.coef) - 1)
        a = self.coef[r].T
        hRT = -a[0] / T ** 2 + a[1] * np.log(T) / T + a[2] + a[3] * T / 2 + a[4] * T ** 2 / 3 \
            + a[5] * T ** 3 / 4 + a[6] * T ** 4 / 5 + a[7] / T
        return hRT * R * T

    def pressure(self, T, V):
        aa, _ = self.alpha(T)
        return R * T / (V - self.b) - aa / (V * V + 2 * self.b * V - self.b ** 2)

    def energy(self, T, rho):
        """Specific internal energy e(T, rho) [J/kg]."""
        V = self.W / rho
        aa, daa = self.alpha(T)
        L = np.log((V + (1 + SQ2) * self.b) / (V + (1 - SQ2) * self.b)) / (2 * SQ2 * self.b)
        return (self.h_ideal(T) - R * T + L * (T * daa - aa)) / self.W

    def roots(self, T, p):
        """Real cubic roots in V above the covolume (ascending)."""
        aa, _ = self.alpha(T)
        b = self.b; RT = R * T
        # V^3 + (b - RT/p) V^2 + (aa/p - 3b^2 - 2 RT b/p) V + (b^3 + RT b^2/p - aa b/p) = 0
        c = [1.0, b - RT / p, aa / p - 3 * b * b - 2 * RT * b / p, b ** 3 + RT * b * b / p - aa * b / p]
        r = np.roots(c)
        r = np.sort(r[np.abs(r.imag) < 1e-12 * np.abs(r.real).max()].real)
        return r[r > b]

    def lnphi(self, T, p, V):
        aa, _ = self.alpha(T)
        Z = p * V / (R * T); B = self.b * p / (R * T); A = aa * p / (R * T) ** 2
        return Z - 1 - np.log(Z - B) - A / (2 * SQ2 * B) * np.log((Z + (1 + SQ2) * B) / (Z + (1 - SQ2) * B))

    def saturation(self, T):
        """(p_sat, V_liquid, V_vapour) at subcritical T by equal fugacity."""
        # bracket in ln p on the three-root band
        lo, hi = math.log(1.0), math.log(R * T / self.b)
        ps = np.exp(np.linspace(lo, math.log(self.a / self.b ** 2), 400))
        prev = None
        bracket = None
        for p in ps:
            r = self.roots(T, p)
            if r.size >= 3:
                g = self.lnphi(T, p, r[0]) - self.lnphi(T, p, r[-1])
                if prev is not None and prev[1] * g <= 0:
                    bracket = (prev[0], p)
                    break
                prev = (p, g)
            elif prev is not None:
                break
        if bracket is None:
            return None
        a_, b_ = math.log(bracket[0]), math.log(bracket[1])
        ga = prev[1]
        for _ in range(80):
            m = 0.5 * (a_ + b_)
            r = self.roots(T, math.exp(m))
            if r.size < 3:
                break
            g = self.lnphi(T, math.exp(m), r[0]) - self.lnphi(T, math.exp(m), r[-1])
            if (g > 0) == (ga > 0):  # same sign as the lower end: the root lies above m
                a_, ga = m, g
            else:
                b_ = m
        p = math.exp(0.5 * (a_ + b_))
        r = self.roots(T, p)
        return p, r[0], r[-1]


class MixtureFluid:
    """Single-phase Peng-Robinson gas mixture (van der Waals mixing with
    a_ij = sqrt(a_i a_j)) and NASA9 ideal parts, in numpy: the internal
    energy and pressure at (T, rho, Y) that seed a synthetic mixture state."""

    def __init__(self, names=("N2O", "N2")):
        self.species = [PureFluid(n) for n in names]
        self.W = np.array([f.W for f in self.species])

    def _mix(self, T, Y):
        Y = np.atleast_2d(Y)
        inv = Y / self.W
        W = 1.0 / inv.sum(1)
        x = inv * W[:, None]
        q, dq = [], []
        for f in self.species:
            sq = np.sqrt(T / f.Tc)
            g = 1.0 + f.kappa * (1.0 - sq)
            q.append(np.abs(g))
            dq.append(-np.sign(g) * f.kappa / (2.0 * f.Tc * sq))
        aa = np.zeros_like(T)
        da = np.zeros_like(T)
        for i, fi in enumerate(self.species):
            for j, fj in enumerate(self.species):
                w = x[:, i] * x[:, j] * math.sqrt(fi.a * fj.a)
                aa += w * q[i] * q[j]
                da += w * (dq[i] * q[j] + q[i] * dq[j])
        b = sum(x[:, k] * f.b for k, f in enumerate(self.species))
        return x, W, aa, da, b

    def energy(self, T, rho, Y):
        T = np.asarray(T, float)
        x, W, aa, da, b = self._mix(T, Y)
        V = W / rho
        L = np.log((V + (1 + SQ2) * b) / (V + (1 - SQ2) * b)) / (2 * SQ2 * b)
        h0 = sum(x[:, k] * f.h_ideal(T) for k, f in enumerate(self.species))
        return (h0 - R * T + L * (T * da - aa)) / W

    def pressure(self, T, rho, Y):
        T = np.asarray(T, float)
        x, W, aa, da, b = self._mix(T, Y)
        V = W / rho
        return R * T / (V - b) - aa / (V * V + 2 * b * V - b * b)


_SAT_CACHE = {}


def saturation_table(fluid, tlo=148.0, thi=None, n=1200):
    key = (fluid.W, tlo, thi, n)
    if key not in _SAT_CACHE:
        thi = thi if thi is not None else fluid.Tc - 0.05
        Ts, rows = [], []
        for T in np.linspace(tlo, thi, n):
            s = fluid.saturation(T)
            if s is not None:
                Ts.append(T); rows.append(s)
        rows = np.array(rows)
        _SAT_CACHE[key] = (np.array(Ts), rows[:, 0], rows[:, 1], rows[:, 2])
    return _SAT_CACHE[key]


def equilibrium_states(T, rho, fluid=None):
    """Isochoric equilibrium of pure fluid at (T, rho): returns (e, p,
    liquid mass density, two-phase flag)."""
    fluid = fluid or PureFluid("N2O")
    T = np.asarray(T, float); rho = np.asarray(rho, float)
    Ts, Ps, Vl, Vv = saturation_table(fluid)
    V = fluid.W / rho
    sub = T < Ts[-1]
    vl = np.interp(T, Ts, Vl); vv = np.interp(T, Ts, Vv); ps = np.exp(np.interp(T, Ts, np.log(Ps)))
    dome = sub & (V > vl) & (V < vv)
    e = fluid.energy(T, rho)
    p = fluid.pressure(T, V)
    liquid = np.where(V <= vl, rho, 0.0)
    if dome.any():
        x = (V[dome] - vl[dome]) / (vv[dome] - vl[dome])  # vapour mass fraction
        el = fluid.energy(T[dome], fluid.W / vl[dome]); ev = fluid.energy(T[dome], fluid.W / vv[dome])
        e[dome] = x * ev + (1 - x) * el
        p[dome] = ps[dome]
        liquid[dome] = (1 - x) * rho[dome]
    return e, p, liquid, dome


# ------------------------------------------------------------------ cases
def pure_n2o_cube(n, seed=0, edge=0.08, tlo=180.0, thi=300.0, rlo=1.0, rhi=1000.0, sigma_u=20.0):
    """Seeded pure-N2O cube: (faces, volume, q[nc, 5] cell-major, seeds[nc]
    abi.STATE_DTYPE, info).  q = [rho, rho u, rho (e + |u|^2 / 2)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = n ** 3
    T = rng.uniform(tlo, thi, nc)
    rho = np.exp(rng.uniform(math.log(rlo), math.log(rhi), nc))
    u = rng.normal(0.0, sigma_u, (nc, 3))
    e, p, liquid, dome = equilibrium_states(T, rho)
    q = np.empty((nc, 5))
    q[:, 0] = rho
    q[:, 1:4] = rho[:, None] * u
    q[:, 4] = rho * (e + 0.5 * (u * u).sum(1))
    seeds = np.zeros(nc, dtype=abi.STATE_DTYPE)
    seeds["p"] = p; seeds["T"] = T; seeds["rho"] = rho; seeds["e"] = e
    seeds["liquidMass"][:, 0] = liquid
    seeds["gasMass"] = rho - liquid
    seeds["activeLiquids"] = (liquid > 0).astype(np.int32)
    faces, volume = cartesian_faces(n, edge / n)
    return faces, volume, q, seeds, {"two_phase_fraction": float(dome.mean()), "seed": seed, "n": n}


def riemann_cube(n, edge=0.016, rho_left=60.0, rho_right=5.0, T=300.0):
    """Pure-N2O gas at rest in an n^3 box, dense on the x < edge / 2 side:
    (faces, volume, q, seeds).  Its first stage raises the wave speeds at the
    interface, so a step at CFL 0.5 fails the post-stage-0 recheck."""
    nc = n ** 3
    i = np.arange(nc) % n
    rho = np.where(i < n // 2, rho_left, rho_right).astype(float)
    Tc = np.full(nc, T)
    e, p, liquid, dome = equilibrium_states(Tc, rho)
    q = np.zeros((nc, 5))
    q[:, 0] = rho
    q[:, 4] = rho * e
    seeds = np.zeros(nc, dtype=abi.STATE_DTYPE)
    seeds["p"] = p; seeds["T"] = Tc; seeds["rho"] = rho; seeds["e"] = e
    seeds["liquidMass"][:, 0] = liquid
    seeds["gasMass"] = rho - liquid
    seeds["activeLiquids"] = (liquid > 0).astype(np.int32)
    faces, volume = cartesian_faces(n, edge / n)
    return faces, volume, q, seeds


def jet_cube(n, seed=0, edge=None, spacing=0.001, radius=0.0025, rho_jet=880.0, T_jet=293.15, speed=30.0,
             slip_sides=()):
    """pure_n2o_cube with two liquid-N2O jets entering through fixed-state
    (kind 3) boundary patches: jet 0 through the -x wall along +x, jet 1
    through the -y wall along +y, both centred at mid-height of the box (90
    degrees apart, axes meeting at the box centre), circular of `radius`.
    Each jet is a reservoir state (rho_jet, T_jet, |u| = speed along its
    axis).  `slip_sides` lists boundary sides (0..5 = -x +x -y +y -z +z)
    whose remaining faces become slip walls (kind 1).  The reservoir density
    is 880 kg/m^3: PR puts 293.15 K, 900 kg/m^3 at 14.87 MPa (inside pmax =
    50 MPa), but the reference flash finds no stable state for that liquid
    above about 885 kg/m^3; 880 kg/m^3 (12.97 MPa) is the closest round value
    it accepts.  Returns (faces,
    volume, q, seeds, fixed_q[2, 5], fixed_seeds[2], info); the fixed states'
    p, c_frozen come from a flash of fixed_q seeded by fixed_seeds."""
    edge = spacing * n if edge is None else edge
    faces, volume, q, seeds, info = pure_n2o_cube(n, seed=seed, edge=edge)
    h = edge / n
    own = faces["owner"]
    i, j, k = own % n, (own // n) % n, own // (n * n)
    normal = faces["normal"]
    bnd = faces["neighbour"] < 0
    side = np.full(len(faces), -1)
    for s_, (ax, sign) in enumerate([(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)]):
        side[bnd & (normal[:, ax] == sign)] = s_
    centre = 0.5 * edge
    yz = ((j + .5) * h - centre) ** 2 + ((k + .5) * h - centre) ** 2
    xz = ((i + .5) * h - centre) ** 2 + ((k + .5) * h - centre) ** 2
    jet0 = (side == 0) & (yz <= radius ** 2)
    jet1 = (side == 2) & (xz <= radius ** 2)
    for s_ in slip_sides:
        faces["kind"][(side == s_) & ~jet0 & ~jet1] = 1
    faces["kind"][jet0 | jet1] = 3
    faces["fixed"][jet0] = 0
    faces["fixed"][jet1] = 1
    fluid = PureFluid("N2O")
    e_jet = fluid.energy(T_jet, rho_jet)
    p_jet = fluid.pressure(T_jet, fluid.W / rho_jet)
    fixed_q = np.zeros((2, 5))
    for f_, axis in ((0, 0), (1, 1)):
        u = np.zeros(3)
        u[axis] = speed
        fixed_q[f_, 0] = rho_jet
        fixed_q[f_, 1:4] = rho_jet * u
        fixed_q[f_, 4] = rho_jet * (e_jet + 0.5 * speed ** 2)
    fixed_seeds = np.zeros(2, dtype=abi.STATE_DTYPE)
    fixed_seeds["p"] = p_jet; fixed_seeds["T"] = T_jet; fixed_seeds["rho"] = rho_jet; fixed_seeds["e"] = e_jet
    fixed_seeds["liquidMass"][:, 0] = rho_jet
    fixed_seeds["activeLiquids"] = 1
    info.update(jet_faces=(int(jet0.sum()), int(jet1.sum())), jet_pressure=float(p_jet))
    return faces, volume, q, seeds, fixed_q, fixed_seeds, info


def multi_case(count, seed=0, edges=(16, 24, 32, 40, 48), spacing=0.002):
    """Uneven batch of independent pure-N2O cubes: edge lengths (cells) drawn
    uniformly from `edges`, cell size `spacing` (2 mm, the 40^3 / 80 mm case),
    each cube seeded like pure_n2o_cube with its own seed.  Returns a list of
    pure_n2o_cube tuples."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice(np.asarray(edges), size=count)
    return [pure_n2o_cube(int(n), seed=seed * 1000 + i, edge=spacing * int(n)) for i, n in enumerate(sizes)]


def binary_cube(n, seed=0, edge=0.08, ymax=0.3, tlo=150.0, thi=320.0, rlo=0.5, rhi=1000.0, sigma_u=20.0):
    """Seeded binary N2O/N2 cube: Y_N2 ~ U(0, ymax), T ~ U(tlo, thi), rho
    log-uniform, u_d ~ N(0, sigma_u).  The internal energy is the
    single-phase PR mixture value at (T, rho, Y): the flash then finds the
    equilibrium (gas mixture + liquid N2O) with that energy.  Returns (faces,
    volume, q[nc, 6] = [rho_N2O, rho_N2, rho u, rho E], seeds, info)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nc = n ** 3
    yn2 = rng.uniform(0.0, ymax, nc)
    T = rng.uniform(tlo, thi, nc)
    rho = np.exp(rng.uniform(math.log(rlo), math.log(rhi), nc))
    u = rng.normal(0.0, sigma_u, (nc, 3))
    Y = np.stack([1 - yn2, yn2], 1)
    mix = MixtureFluid(("N2O", "N2"))
    e = mix.energy(T, rho, Y)
    pr = mix.pressure(T, rho, Y)
    q = np.empty((nc, 6))
    q[:, 0] = rho * Y[:, 0]
    q[:, 1] = rho * Y[:, 1]
    q[:, 2:5] = rho[:, None] * u
    q[:, 5] = rho * (e + 0.5 * (u * u).sum(1))
    seeds = np.zeros(nc, dtype=abi.STATE_DTYPE)
    seeds["p"] = np.clip(pr, 2e3, 4e7)
    seeds["T"] = T
    seeds["rho"] = rho
    seeds["e"] = e
    seeds["gasMass"] = rho
    faces, volume = cartesian_faces(n, edge / n)
    return faces, volume, q, seeds, {"seed": seed, "n": n}


def internal_energy(q, ns):
    """e_v = rhoE - |rho u|^2 / (2 rho) (ReactiveFoam.C:727-728)."""
    rho = q[:, :ns].sum(1)
    m = q[:, ns:ns + 3]
    return q[:, ns + 3] - .5 * (m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1] + m[:, 2] * m[:, 2]) / rho


def transport_states(states):
    """ReactiveTransportState from flash states (ReactiveFoam.C:419-420)."""
    out = np.zeros(states.shape, dtype=abi.TSTATE_DTYPE)
    for name, src in (("p", "p"), ("T", "T"), ("rho", "rho"), ("cv", "cv"), ("sound", "soundFrozen"),
                      ("gasMass", "gasMass")):
        out[name] = states[src]
    return out
