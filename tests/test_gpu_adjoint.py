"""Adjoint of one FP64 SSPRK2 step (reactive_amd/adjoint.py): gradients of
the mean squared pressure error against a target step (PR a x 1.02, b x 0.99,
jet inlet T + 2 K) with respect to (a-scale, b-scale, inlet T, inlet |u|),
from the hand-written adjoint kernels behind a torch.autograd.Function,
against central finite differences of the same step (relative step 1e-6).

Pieces are pinned separately: the transposed HLL rates against finite
differences of reactive_transport_rhs, the flash backward against finite
differences of the flash."""
import numpy as np
import pytest

from reactive_amd import adjoint, cases, lib as rlib

torch = pytest.importorskip("torch")  # imported by reactive_amd.adjoint after the HIP library

pytestmark = pytest.mark.gpu

THETA0 = np.array([1.0, 1.0, 293.15, 30.0])
TARGET = np.array([1.02, 0.99, 295.15, 30.0])


def loss_and_grad(job, target, theta):
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    p2 = adjoint.StepPressure.apply(th, job)
    pt = torch.from_numpy(target)
    loss = ((p2 - pt) ** 2).mean()
    loss.backward()
    return float(loss), th.grad.numpy().copy()


def loss_difference(pu, pd, target):
    """mean((pu - t)^2) - mean((pd - t)^2), summed per cell as
    (pu - pd)(pu + pd - 2t): cells the perturbation leaves unchanged add
    exactly zero, so the difference carries no rounding of the whole loss."""
    return float(((pu - pd) * (pu + pd - 2 * target)).mean())


def phase_flips(tu, td):
    """Cells whose flash outcome (success, phase set) differs between two forwards."""
    n = 0
    for k in ("0", "1", "2"):
        n += int(((tu["ok" + k] != td["ok" + k]) | (tu["s" + k]["activeLiquids"] != td["s" + k]["activeLiquids"])).sum())
    return n


@pytest.mark.parametrize("n", [10, 80])
def test_step_gradient_matches_central_differences(n):
    """n = 80: the 80^3 differentiation workload.  A central difference is only
    meaningful where no flash changes its outcome between theta +- h; the step
    is taken from 1e-6, 3e-7, 1e-7 (relative), the first without such a flip."""
    job = adjoint.JetStep(n, seed=41)
    target, _ = job.forward(TARGET)
    job.dt = None  # dt from the CFL of the base parameters
    loss, grad = loss_and_grad(job, target, THETA0)
    print(f"\n{n}^3 jets: dt {job.dt:.4e}, loss {loss:.6e}, adjoint gradient {grad}, skipped cells {job.skipped}")
    fd = np.zeros(4)
    for i in range(4):
        for rel_h in (1e-6, 3e-7, 1e-7):
            h = rel_h * abs(THETA0[i])
            up, dn = THETA0.copy(), THETA0.copy()
            up[i] += h
            dn[i] -= h
            pu, tu = job.forward(up)
            pd, td = job.forward(dn)
            flips = phase_flips(tu, td)
            if flips == 0:
                break
            print(f"theta[{i}]: {flips} flash outcome(s) change within +-{rel_h:g}; smaller step")
        assert flips == 0
        fd[i] = loss_difference(pu, pd, target) / (2 * h)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("central differences:", fd, "relative difference:", rel)
    assert job.skipped == 0
    assert np.all(rel <= 1e-6), rel


def oracle_forward(job, theta, twin=False):
    """The same step on the oracle (the reference's CPU arithmetic with glibc,
    or its twin build): p after the step."""
    from helpers import OracleTransport, config, oracle_flash
    image = job.image(theta[0], theta[1])
    fq = job.reservoir_q(theta)
    seeds_in = job.fixed_seeds.copy()
    seeds_in["T"] = theta[2]
    s_in, ok_in, _ = oracle_flash(image, fq[:, :1], cases.internal_energy(fq, 1), seeds_in, 1, twin=twin)
    assert ok_in.all()
    cfg = config(job.cells, len(job.faces), fixed=2)
    ora = OracleTransport(cfg, job.volume, job.faces, fq, cases.transport_states(s_in))
    q = job.q0.copy()
    s0, _, _ = oracle_flash(image, q[:, :1], cases.internal_energy(q, 1), job.seeds, twin=twin)
    ora.stage(q, cases.transport_states(s0), job.dt, 0)
    s1, _, _ = oracle_flash(image, q[:, :1], cases.internal_energy(q, 1), s0, twin=twin)
    ora.stage(q, cases.transport_states(s1), job.dt, 1)
    s2, _, _ = oracle_flash(image, q[:, :1], cases.internal_energy(q, 1), s0, twin=twin)
    ora.close()
    return s2["p"].copy()


def test_gradient_matches_differences_of_the_reference_cpu_step():
    """The device adjoint against central differences of the reference's own
    CPU arithmetic (the glibc oracle, bitwise the compiled reference) on a 6^3
    jet box; the twin oracle's forward equals the device forward bit for bit."""
    job = adjoint.JetStep(6, seed=43)
    target, _ = job.forward(TARGET)
    job.dt = None
    _, grad = loss_and_grad(job, target, THETA0)
    p_dev, _ = job.forward(THETA0)
    assert np.array_equal(oracle_forward(job, THETA0, twin=True), p_dev)
    fd = np.zeros(4)
    for i in range(4):
        h = 1e-6 * abs(THETA0[i])
        up, dn = THETA0.copy(), THETA0.copy()
        up[i] += h
        dn[i] -= h
        fd[i] = loss_difference(oracle_forward(job, up), oracle_forward(job, dn), target) / (2 * h)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("\nadjoint", grad, "CPU reference differences", fd, "relative", rel)
    assert np.all(rel <= 1e-6), rel


def test_transposed_rates_match_rhs_differences():
    """rate_bar . dL/dq, dL/dp, dL/dc of the transposed stage against central
    differences of reactive_transport_rhs on a jet box with slip walls."""
    lib = rlib.load()
    job = adjoint.JetStep(6, seed=5)
    faces, volume, q, seeds, fq, fseeds, _ = cases.jet_cube(6, seed=5, slip_sides=(4, 5))
    image = job.image(1.0, 1.0)
    hem = job.hem(image, len(q))
    states, ok = job.flash(hem, q, seeds)
    good, bad = np.where(ok == 1)[0], np.where(ok == 0)[0]
    q[bad] = q[good[0]]
    states[bad] = states[good[0]]
    seeds_in = fseeds.copy()
    s_in, ok_in = job.flash(hem, fq, seeds_in)
    lib.reactive_gpu_hem_destroy_v1(hem)
    job.faces, job.volume, job.cells = faces, volume, len(q)
    t = job.transport(fq, s_in)
    st = cases.transport_states(states)
    rng = np.random.default_rng(3)
    rbar = rng.standard_normal(q.shape)

    def rhs(qq, ss):
        r = np.zeros_like(qq)
        b = np.zeros(5)
        assert lib.reactive_transport_rhs(t, rlib.ptr(qq), rlib.ptr(ss), None, None, rlib.ptr(r), rlib.ptr(b)) == 0
        return r

    qbar, pbar, cbar, fixed = job.stage_adjoint(t, q, states, rbar)
    for cell in rng.choice(len(q), 12, replace=False):
        for k in range(5):
            # the rates carry p / dx ~ 1e10: steps well above their rounding
            h = 1e-6 * max(abs(q[cell, k]), 100 * abs(q[cell, 0]))
            qp, qm = q.copy(), q.copy()
            qp[cell, k] += h
            qm[cell, k] -= h
            fdv = float((rbar * (rhs(qp, st) - rhs(qm, st))).sum()) / (2 * h)
            assert abs(qbar[cell, k] - fdv) <= 1e-6 * abs(fdv), (cell, k, qbar[cell, k], fdv)
        for name, bar in (("p", pbar), ("sound", cbar)):
            # the flux is linear in p: a large step is exact and keeps clear of rounding
            h = (1e-3 if name == "p" else 1e-5) * abs(st[name][cell])
            sp, sm = st.copy(), st.copy()
            sp[name][cell] += h
            sm[name][cell] -= h
            fdv = float((rbar * (rhs(q, sp) - rhs(q, sm))).sum()) / (2 * h)
            assert abs(bar[cell] - fdv) <= 1e-6 * abs(fdv) + 1e-12, (cell, name, bar[cell], fdv)
    assert np.abs(fixed).sum() > 0
    lib.reactive_transport_destroy(t)


def test_flash_backward_matches_flash_differences():
    """p_bar . dp/dq + c_bar . dc/dq of the implicit-function-theorem
    backward against central differences of the tight flash."""
    lib = rlib.load()
    job = adjoint.JetStep(6, seed=9)
    faces, volume, q, seeds, _, _, _ = cases.jet_cube(6, seed=9)
    image = job.image(1.0, 1.0)
    hem = job.hem(image, len(q))
    h = job.rel_step
    scaled = [job.hem(job.image(1 + h, 1), len(q)), job.hem(job.image(1 - h, 1), len(q)),
              job.hem(job.image(1, 1 + h), len(q)), job.hem(job.image(1, 1 - h), len(q))]
    states, ok = job.flash(hem, q, seeds)
    rng = np.random.default_rng(4)
    pbar, cbar = rng.standard_normal(len(q)), rng.standard_normal(len(q))
    job.skipped = 0
    qbar, tbar = job.flash_adjoint(hem, scaled, q, states, ok, pbar, cbar)
    assert job.skipped == 0
    checked = {0: 0, 1: 0, 2: 0}
    for cell in np.where(ok == 1)[0][:60]:
        s = states[cell]
        mode = 2 if s["liquidMass"][0] > 0 and s["gasMass"] > 0 else (1 if s["liquidMass"][0] > 0 else 0)
        for k in (0, 4):
            step = 1e-6 * abs(q[cell, k])
            qp, qm = q[cell:cell + 1].copy(), q[cell:cell + 1].copy()
            qp[0, k] += step
            qm[0, k] -= step
            sp, okp = job.flash(hem, qp, states[cell:cell + 1])
            sm, okm = job.flash(hem, qm, states[cell:cell + 1])
            if not (okp[0] and okm[0]) or sp["activeLiquids"][0] != sm["activeLiquids"][0]:
                continue
            fdv = (pbar[cell] * (sp["p"][0] - sm["p"][0]) + cbar[cell] * (sp["soundFrozen"][0] - sm["soundFrozen"][0])) / (2 * step)
            assert abs(qbar[cell, k] - fdv) <= 1e-4 * abs(fdv) + 1e-9 * abs(pbar[cell] * s["p"] / q[cell, k]), \
                (cell, k, mode, qbar[cell, k], fdv)
            checked[mode] += 1
    print("cells checked by phase configuration (gas, liquid, two-phase):", checked)
    assert checked[0] > 0 and checked[2] > 0
    for x in [hem] + scaled:
        lib.reactive_gpu_hem_destroy_v1(x)
