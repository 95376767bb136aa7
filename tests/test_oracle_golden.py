"""The oracle restatement against the reference's own outputs (fixtures of
tests/golden/, written by make_golden.py from the compiled reference):
bit for bit, including the flash work counters."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, OracleTransport, config, oracle, oracle_flash, REF_PATH


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("part", ["cube", "edge"])
def test_flash_bitwise_vs_reference(part):
    g = load("flash_n2o.npz")
    mb = g["model"].tobytes()
    assert oracle().oracle_model_supported(C.create_string_buffer(mb, len(mb))) == 1
    out, success, counters = oracle_flash(mb, g[f"{part}_q"], g[f"{part}_energy"], g[f"{part}_seeds"])
    assert np.array_equal(success, g[f"{part}_success"])
    assert out.tobytes() == g[f"{part}_out"].tobytes()
    assert np.array_equal(counters, g[f"{part}_counters"])


def test_binary_flash_bitwise_vs_reference():
    """Binary N2O/N2 gas with liquid N2O (the two-species flash)."""
    g = load("flash_binary.npz")
    mb = g["model"].tobytes()
    assert oracle().oracle_model_supported(C.create_string_buffer(mb, len(mb))) == 1
    out, success, counters = oracle_flash(mb, g["q"], g["energy"], g["seeds"])
    assert np.array_equal(success, g["success"])
    assert out.tobytes() == g["out"].tobytes()
    assert np.array_equal(counters, g["counters"])


@pytest.mark.parametrize("name", ["cube6", "graph"])
def test_transport_bitwise_vs_reference(name):
    g = load("transport.npz")
    d = {k[len(name) + 1:]: g[k] for k in g.files if k.startswith(name + "_")}
    fq = d["fixed_q"] if len(d["fixed_q"]) else None
    fs = d["fixed_states"] if len(d["fixed_states"]) else None
    cfg = config(len(d["q"]), len(d["faces"]), fixed=0 if fq is None else len(fq))
    ora = OracleTransport(cfg, d["volume"], d["faces"], fq, fs)
    rhs, b = ora.rhs(d["q"], d["states"])
    assert np.array_equal(rhs, d["rhs"]) and np.array_equal(b, d["boundary"])
    rc, dt = ora.stable_step(d["q"], d["states"], 0.4, 1.0)
    assert rc == 0 and dt == float(d["dt"])
    q = d["q"].copy()
    assert np.array_equal(ora.stage(q, d["states"], dt, 0), d["b0"]) and np.array_equal(q, d["q1"])
    assert np.array_equal(ora.stage(q, d["states"], dt, 1), d["b1"]) and np.array_equal(q, d["q2"])
    ora.close()


@pytest.mark.skipif(not os.path.exists(REF_PATH), reason="oracle/_ref is built only where the reference is mounted")
def test_flash_fresh_inputs_vs_reference():
    """Fresh seeded cube, not in the fixtures: oracle == compiled reference."""
    from helpers import cube_case
    lib = C.CDLL(REF_PATH)
    V = C.c_void_p
    lib.ref_flash_batch.argtypes = [V, C.c_size_t, V, V, V, C.c_int, V, V, V]
    case = cube_case(6, seed=99)
    q, e, seeds, mb = case["q"][:, :1].copy(), case["energy"], case["seeds"], case["model"]
    out = seeds.copy()
    success = np.zeros(len(e), np.int32)
    lib.ref_flash_batch(C.create_string_buffer(mb, len(mb)), len(e), q.ctypes.data, e.ctypes.data,
                        seeds.ctypes.data, 1, out.ctypes.data, success.ctypes.data, None)
    mine, msuccess, _ = oracle_flash(mb, q, e, seeds)
    assert np.array_equal(success, msuccess) and out.tobytes() == mine.tobytes()


def oracle_frozen(mb, q, e, guesses, twin=False):
    """The oracle's frozen-inventory recovery, cell by cell: (states, success, counters[9])."""
    from helpers import p
    lib = oracle(twin)
    img = C.create_string_buffer(mb, len(mb))
    n = len(e)
    out = guesses.copy()
    success = np.zeros(n, np.int32)
    counters = np.zeros((n, 9), np.uint64)
    q = np.ascontiguousarray(q)
    for c in range(n):
        o = guesses[c:c + 1].copy()
        success[c] = lib.oracle_flash_frozen(img, p(q[c]), C.c_double(e[c]), p(guesses[c:c + 1]), p(o),
                                             p(counters[c]))
        if success[c]:
            out[c] = o[0]
    return out, success, counters


@pytest.mark.parametrize("part", ["seeded", "converged", "edge", "binary"])
def test_frozen_recovery_bitwise_vs_reference(part):
    """oracle_flash_frozen against the reference's recoverCapillary(...,
    equilibrium = 0) (reactiveDeviceFlash.h:572-590) on seeded, converged,
    edge-class and binary guesses: states, success flags and counters."""
    g = load("flash_frozen.npz")
    mb = (g["model_binary"] if part == "binary" else g["model"]).tobytes()
    out, success, counters = oracle_frozen(mb, g[f"{part}_q"], g[f"{part}_energy"], g[f"{part}_guess"])
    assert np.array_equal(success, g[f"{part}_success"])
    assert 0 < success.sum() < len(success) or part != "edge"
    assert out.tobytes() == g[f"{part}_out"].tobytes()
    assert np.array_equal(counters, g[f"{part}_counters"])


@pytest.mark.parametrize("part", ["cube", "edge", "binary"])
def test_stable_gas_prune_bitwise_vs_reference(part):
    """The opt-in stable-gas pruning (stableGasPrune = 1,
    reactiveDeviceFlash.h:509-523): the oracle's flash with the pruning model
    equals the reference's, and pruning changes the work (fewer candidates)."""
    g = load("flash_prune.npz")
    base = load("flash_binary.npz" if part == "binary" else "flash_n2o.npz")
    key = "" if part == "binary" else f"{part}_"
    mb = (g["model_binary"] if part == "binary" else g["model"]).tobytes()
    assert oracle().oracle_model_supported(C.create_string_buffer(mb, len(mb))) == 1
    out, success, counters = oracle_flash(mb, base[f"{key}q"], base[f"{key}energy"], base[f"{key}seeds"])
    assert np.array_equal(success, g[f"{part}_success"])
    assert out.tobytes() == g[f"{part}_out"].tobytes()
    assert np.array_equal(counters, g[f"{part}_counters"])
    assert counters[:, 2].sum() < base[f"{key}counters"][:, 2].sum()


@pytest.mark.parametrize("part", ["cube", "edge", "binary", "cube12"])
def test_analytic_jacobian_bitwise_vs_reference(part):
    """The reference's analytic active-set Jacobian (analyticJacobian = 1,
    reactiveDeviceFlashJacobian.h build/buildCached, dispatched at
    reactiveDeviceFlash.h:338-352): the oracle's flash with the analytic
    model equals the reference's (states, success flags and every counter,
    the analytic-Jacobian count included), and the mode changes the work
    (fewer phase evaluations than the finite-difference model)."""
    g = load("flash_analytic.npz")
    if part == "cube12":
        q, e, seeds, base = g["cube12_q"], g["cube12_energy"], g["cube12_seeds"], None
    else:
        base = load("flash_binary.npz" if part == "binary" else "flash_n2o.npz")
        key = "" if part == "binary" else f"{part}_"
        q, e, seeds = base[f"{key}q"], base[f"{key}energy"], base[f"{key}seeds"]
    mb = (g["model_binary"] if part == "binary" else g["model"]).tobytes()
    assert oracle().oracle_model_supported(C.create_string_buffer(mb, len(mb))) == 1
    out, success, counters = oracle_flash(mb, q, e, seeds)
    assert np.array_equal(success, g[f"{part}_success"])
    assert out.tobytes() == g[f"{part}_out"].tobytes()
    assert np.array_equal(counters, g[f"{part}_counters"])
    assert counters[:, 5].sum() > 0 and counters[:, 6].sum() == 0
    if base is not None:
        assert counters[:, 0].sum() < base[f"{key}counters"][:, 0].sum()
