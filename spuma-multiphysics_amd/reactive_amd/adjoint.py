"""Gradients of one FP64 SSPRK2 step with respect to the Peng-Robinson
parameters and the jet inlet state, from hand-written adjoint kernels.

The step is the resident step's sequence on a box with two fixed-state liquid
jets (cases.jet_cube): flash #0 of q0 (step-start states), stage 0, flash #1
seeded from the step-start states, stage 1, flash #2 from the same seeds
(src/reactiveFoam/ReactiveFoam.C:961-1019), every piece through the library's
C ABI (reactive_gpu_hem_run_v1, reactive_transport_stage).  Its parameters are
theta = (a-scale, b-scale, inlet T, inlet |u|): the PR a and b of N2O (gas and
liquid tables) are multiplied by the scales, and each reservoir's conserved
state is (rho_in, rho_in u_in, rho_in (e(T_in, rho_in) + u_in^2 / 2)) with
e from the same scaled PR; the reservoirs' p and c_frozen are the flash of
that state.  dt is held at the CFL step of the base parameters, so the step
is a smooth function of theta.  The flashes run with tolerances 1e-3 as tight
as the reference's (vtol = etol = 1e-12, mutol = 1e-10): the step then sits
on its converged roots, which is what the implicit-function-theorem backward
differentiates and what a finite difference of the step can resolve.

Backward, in reverse order: flash #2 (implicit-function theorem,
reactive_amd_flash_adjoint), stage 1 (transposed HLL rates,
reactive_amd_transport_stage_adjoint), flash #1, stage 0, flash #0, the
reservoirs' flash, and the reservoir states' dependence on theta (central
differences of the explicit numpy PR energy).  StepPressure is the
torch.autograd.Function; JetStep holds the case and runs both passes.
"""
import ctypes as C

import numpy as np

from . import abi, cases, lib as rlib, model

TIGHT = {"vtol": 1e-12, "etol": 1e-12, "mutol": 1e-10}
THETA0 = (1.0, 1.0, 293.15, 30.0)


def _check(rc, lib, handle=None, err=None):
    if rc:
        msg = err.value.decode() if err is not None else lib.reactive_transport_error(handle).decode()
        raise RuntimeError(msg)


class JetStep:
    """One FP64 step of an n^3 jet box as a function of theta."""

    def __init__(self, n, seed=0, rel_step=1e-6, cfl=0.4, rho_jet=880.0):
        self.lib = rlib.load()
        self.n = n
        faces, volume, q, seeds, fixed_q, fixed_seeds, info = cases.jet_cube(n, seed=seed, rho_jet=rho_jet)
        self.faces, self.volume, self.q0, self.seeds = faces, volume, q, seeds
        self.fixed_seeds, self.rho_jet = fixed_seeds, rho_jet
        self.axes = (0, 1)  # jet 0 along +x, jet 1 along +y
        self.rel_step = rel_step
        self.cfl = cfl
        self.dt = None
        self.cells = len(q)
        # flash-consistent seeds at the base parameters; a cell the flash
        # fails on takes the conserved state and seeds of a converged cell
        hem = self.hem(self.image(1.0, 1.0), self.cells)
        try:
            states, ok = self.flash(hem, self.q0, self.seeds)
        finally:
            self.lib.reactive_gpu_hem_destroy_v1(hem)
        good, bad = np.where(ok == 1)[0], np.where(ok == 0)[0]
        if bad.size:
            self.q0[bad] = self.q0[good[0]]
            states[bad] = states[good[0]]
        self.seeds = states

    # ---- pieces
    def image(self, a_scale, b_scale):
        return model.model_bytes(model.build_model(limits=TIGHT, a_scale=a_scale, b_scale=b_scale))

    def hem(self, image, capacity):
        err = C.create_string_buffer(1024)
        h = self.lib.reactive_gpu_hem_create_v1(image, len(image), capacity, err, 1024)
        if not h:
            raise RuntimeError(err.value.decode())
        return h

    def flash(self, hem, q, seeds):
        err = C.create_string_buffer(1024)
        states = seeds.copy()
        ok = np.zeros(len(q), np.int32)
        prof = abi.HemProfile.fresh()
        _check(self.lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q[:, :1])),
                                                rlib.ptr(cases.internal_energy(q, 1)), len(q), rlib.ptr(states),
                                                rlib.ptr(ok), C.byref(prof), err, 1024), self.lib, err=err)
        return states, ok

    def reservoir_q(self, theta):
        """Conserved reservoir states (2 x 5) from the scaled numpy PR."""
        sa, sb, T_in, u_in = theta
        fluid = cases.PureFluid("N2O", a_scale=sa, b_scale=sb)
        e = float(fluid.energy(np.array([T_in]), self.rho_jet)[0])
        out = np.zeros((2, 5))
        for j, axis in enumerate(self.axes):
            out[j, 0] = self.rho_jet
            out[j, 1 + axis] = self.rho_jet * u_in
            out[j, 4] = self.rho_jet * (e + 0.5 * u_in * u_in)
        return out

    def transport(self, fixed_q, fixed_states):
        err = C.create_string_buffer(1024)
        cfg = abi.Config(cells=self.cells, species=1, variables=5, faces=len(self.faces), fixed=len(fixed_q),
                         viscosity=0, conductivity=0, diffusivity=0, waveFactor=1.1, maxBytes=64e9, mechanical=0)
        t = self.lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(self.volume), rlib.ptr(self.faces),
                                               rlib.ptr(fixed_q), rlib.ptr(cases.transport_states(fixed_states)),
                                               None, None, err, 1024)
        if not t:
            raise RuntimeError(err.value.decode())
        return t

    def stage(self, t, q, states, which):
        b = np.zeros(5)
        _check(self.lib.reactive_transport_stage(t, rlib.ptr(q), rlib.ptr(cases.transport_states(states)), None,
                                                 None, self.dt, which, rlib.ptr(b)), self.lib, t)

    # ---- forward
    def forward(self, theta):
        """p after the step (cells) and the tape the backward reads."""
        theta = tuple(float(x) for x in theta)
        image = self.image(theta[0], theta[1])
        hem = self.hem(image, self.cells)
        try:
            fq = self.reservoir_q(theta)
            seeds_in = self.fixed_seeds.copy()
            seeds_in["T"] = theta[2]
            s_in, ok_in = self.flash(hem, fq, seeds_in)
            if not ok_in.all():
                raise RuntimeError("reservoir state does not flash")
            t = self.transport(fq, s_in)
            try:
                s0, ok0 = self.flash(hem, self.q0, self.seeds)
                if self.dt is None:
                    dt = C.c_double()
                    _check(self.lib.reactive_transport_stable_step(
                        t, rlib.ptr(self.q0), rlib.ptr(cases.transport_states(s0)), self.cfl, 1.0, C.byref(dt)),
                        self.lib, t)
                    self.dt = dt.value
                q1 = self.q0.copy()
                self.stage(t, q1, s0, 0)
                s1, ok1 = self.flash(hem, q1, s0)
                q2 = q1.copy()
                self.stage(t, q2, s1, 1)
                s2, ok2 = self.flash(hem, q2, s0)
            finally:
                self.lib.reactive_transport_destroy(t)
        finally:
            self.lib.reactive_gpu_hem_destroy_v1(hem)
        tape = dict(theta=theta, fq=fq, s_in=s_in, s0=s0, ok0=ok0, q1=q1, s1=s1, ok1=ok1, q2=q2, s2=s2, ok2=ok2)
        return s2["p"].copy(), tape

    # ---- backward
    def flash_adjoint(self, hem, scaled, q, states, ok, pbar, cbar):
        err = C.create_string_buffer(1024)
        qbar = np.zeros_like(q)
        tbar = np.zeros(2)
        skipped = C.c_uint64()
        handles = (C.c_void_p * 4)(*scaled)
        _check(self.lib.reactive_amd_flash_adjoint(hem, handles, self.rel_step, rlib.ptr(np.ascontiguousarray(q)), 5,
                                                   rlib.ptr(states), rlib.ptr(ok.astype(np.int32)), len(q),
                                                   rlib.ptr(np.ascontiguousarray(pbar, dtype=np.float64)),
                                                   rlib.ptr(np.ascontiguousarray(cbar, dtype=np.float64)),
                                                   rlib.ptr(qbar), rlib.ptr(tbar), C.byref(skipped), err, 1024),
               self.lib, err=err)
        self.skipped += skipped.value
        return qbar, tbar

    def stage_adjoint(self, t, q, states, rate_bar):
        qbar = np.zeros_like(q)
        pbar = np.zeros(len(q))
        cbar = np.zeros(len(q))
        fixed = np.zeros((2, 7))
        _check(self.lib.reactive_amd_transport_stage_adjoint(t, rlib.ptr(np.ascontiguousarray(q)),
                                                             rlib.ptr(cases.transport_states(states)),
                                                             rlib.ptr(np.ascontiguousarray(rate_bar)),
                                                             rlib.ptr(qbar), rlib.ptr(pbar), rlib.ptr(cbar),
                                                             rlib.ptr(fixed)), self.lib, t)
        return qbar, pbar, cbar, fixed

    def backward(self, tape, p2bar):
        """d(sum p2bar * p2)/d theta (4)."""
        theta = tape["theta"]
        sa, sb = theta[0], theta[1]
        h = self.rel_step
        self.skipped = 0
        hem = self.hem(self.image(sa, sb), self.cells)
        scaled = [self.hem(self.image(sa * (1 + h), sb), self.cells), self.hem(self.image(sa * (1 - h), sb), self.cells),
                  self.hem(self.image(sa, sb * (1 + h)), self.cells), self.hem(self.image(sa, sb * (1 - h)), self.cells)]
        t = self.transport(tape["fq"], tape["s_in"])
        dt = self.dt
        zero = np.zeros(self.cells)
        # a cell whose flash #1 or #2 fails keeps its step-start state, so its
        # output is flash #0's: those adjoints pass to flash #0
        fail1, fail2 = tape["ok1"] == 0, tape["ok2"] == 0
        try:
            # flash #2: y_bar = (p_bar, 0)
            q2bar, tb = self.flash_adjoint(hem, scaled, tape["q2"], tape["s2"], tape["ok2"], p2bar, zero)
            lnbar = tb.copy()
            # stage 1 (q2 = q0/2 + (q1 + dt L(q1, s1))/2): the adjoint reads no
            # state the forward stage saved
            qb, pb1, cb1, fix1 = self.stage_adjoint(t, tape["q1"], tape["s1"], 0.5 * dt * q2bar)
            q1bar = 0.5 * q2bar + qb
            # flash #1: y_bar = (p_bar, c_bar) of the stage-1 states
            qb, tb = self.flash_adjoint(hem, scaled, tape["q1"], tape["s1"], tape["ok1"], pb1, cb1)
            q1bar += qb
            lnbar += tb
            # stage 0 (q1 = q0 + dt L(q0, s0)); q0 is an input constant
            _, pb0, cb0, fix0 = self.stage_adjoint(t, self.q0, tape["s0"], dt * q1bar)
            pb0 = pb0 + np.where(fail1, pb1, 0.0) + np.where(fail2, p2bar, 0.0)
            cb0 = cb0 + np.where(fail1, cb1, 0.0)
            # flash #0 of the step-start states
            _, tb = self.flash_adjoint(hem, scaled, self.q0, tape["s0"], tape["ok0"], pb0, cb0)
            lnbar += tb
            # reservoirs: fixed-state adjoints (q[5], p, c_frozen) from both stages
            fix = fix0 + fix1
            qb_in, tb = self.flash_adjoint(hem, scaled, tape["fq"], tape["s_in"], np.ones(2, np.int32), fix[:, 5],
                                           fix[:, 6])
            lnbar += tb
            fq_bar = fix[:, :5] + qb_in
        finally:
            self.lib.reactive_transport_destroy(t)
            for x in [hem] + scaled:
                self.lib.reactive_gpu_hem_destroy_v1(x)
        # reservoir conserved states as explicit functions of theta
        grad = np.array([lnbar[0] / sa, lnbar[1] / sb, 0.0, 0.0])
        for i in range(4):
            step = 1e-6 * abs(theta[i])
            up = list(theta); dn = list(theta)
            up[i] += step; dn[i] -= step
            dq = (self.reservoir_q(up) - self.reservoir_q(dn)) / (2 * step)
            grad[i] += float((fq_bar * dq).sum())
        return grad


# The library (and the HIP runtime it links) is loaded before torch: torch
# bundles an older HSA runtime, and the system HIP runtime does not load
# into a process that already holds it.
rlib.load()
try:
    import torch

    class StepPressure(torch.autograd.Function):
        """p after one step of `job` (a JetStep) as a differentiable function
        of theta (float64 tensor of 4); the backward runs the adjoint kernels."""

        @staticmethod
        def forward(ctx, theta, job):
            p2, tape = job.forward(theta.detach().cpu().numpy())
            ctx.job, ctx.tape = job, tape
            return torch.from_numpy(p2).to(theta.device)

        @staticmethod
        def backward(ctx, grad):
            g = ctx.job.backward(ctx.tape, grad.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g).to(grad.device), None
except ImportError:  # torch is the autograd front end only
    StepPressure = None
