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


def loss_only(job, target, theta):
    p2, _ = job.forward(theta)
    return float(((p2 - target) ** 2).mean())


@pytest.mark.parametrize("n", [10])
def test_step_gradient_matches_central_differences(n):
    job = adjoint.JetStep(n, seed=41)
    target, _ = job.forward(TARGET)
    job.dt = None  # dt from the CFL of the base parameters
    loss, grad = loss_and_grad(job, target, THETA0)
    print(f"\n{n}^3 jets: dt {job.dt:.4e}, loss {loss:.6e}, adjoint gradient {grad}, skipped cells {job.skipped}")
    fd = np.zeros(4)
    for i in range(4):
        h = 1e-6 * abs(THETA0[i])
        up, dn = THETA0.copy(), THETA0.copy()
        up[i] += h
        dn[i] -= h
        fd[i] = (loss_only(job, target, up) - loss_only(job, target, dn)) / (2 * h)
    rel = np.abs(grad - fd) / np.abs(fd)
    print("central differences:", fd, "relative difference:", rel)
    assert job.skipped == 0
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
