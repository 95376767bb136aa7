"""Replay of captured HEM batches and restart from schema-3 checkpoints on
the device (reactive_amd.checkpoint).

The captures under tests/golden/ were written by the reference's own capture
tool (tools/hem_capture.cpp) around a seeded cube's equilibrium-flash batch;
the reference's compiled flash outputs for the same batch are stored beside
them.  Replayed through reactive_gpu_hem_run_v1, as tools/replay_hem_precision.py
does, the device must repeat itself bit for bit, equal the oracle's twin build
bit for bit and meet the reference (both the glibc oracle and the stored
outputs) at the FP64 parity gate.  A resident step restarted from a checkpoint
written after two steps must equal the uninterrupted run."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import GOLDEN, assert_reference_parity, config, cube_case, oracle_flash
from reactive_amd import abi, checkpoint, lib as rlib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["capture_n2o", "capture_n2o_n2"])
def test_replay_reference_capture(name):
    lib = rlib.load()
    cap = checkpoint.read_capture(os.path.join(GOLDEN, f"{name}.bin"))
    states, success, profile, seconds = checkpoint.replay_capture(lib, cap, repeats=2)
    again, success2, _, _ = checkpoint.replay_capture(lib, cap, repeats=1)
    assert states.tobytes() == again.tobytes() and np.array_equal(success, success2)
    print(f"{name}: {len(success)} cells, {int(success.sum())} recovered, kernel {seconds} s", flush=True)
    twin, tok, tcnt = oracle_flash(cap["model"], cap["q"], cap["energy"], cap["states"], twin=True)
    assert np.array_equal(success, tok)
    assert states.tobytes() == twin.tobytes()
    assert profile.phaseEvaluations == int(tcnt[:, 0].sum())
    ref, rok, _ = oracle_flash(cap["model"], cap["q"], cap["energy"], cap["states"])
    assert_reference_parity(states, success, ref, rok)
    stored = np.load(os.path.join(GOLDEN, f"{name}_out.npz"))
    assert np.array_equal(stored["success"], rok) and stored["out"].tobytes() == ref.tobytes()
    assert_reference_parity(states, success, stored["out"], stored["success"])


def test_checkpoint_restart_is_exact(tmp_path):
    """Two steps, a checkpoint of q and the states (with the step's history),
    then one more step from the checkpoint on a new handle: equal, bit for
    bit, to the third step of the uninterrupted run."""
    lib = rlib.load()
    case = cube_case(10, seed=91)
    mb = case["model"]
    err = C.create_string_buffer(1024)

    def handle():
        hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(case["q"]), err, 1024)
        cfg = config(len(case["q"]), len(case["faces"]))
        tr = lib.reactive_transport_create(1, C.byref(cfg), rlib.ptr(case["volume"]), rlib.ptr(case["faces"]),
                                           None, None, None, None, err, 1024)
        step = lib.reactive_amd_step_create(tr, hem, err, 1024)
        assert step, err.value
        return hem, tr, step

    def read(step):
        q = np.zeros_like(case["q"])
        s = np.zeros(len(q), dtype=abi.STATE_DTYPE)
        assert lib.reactive_amd_step_read(step, rlib.ptr(q), rlib.ptr(s)) == 0
        return q, s

    def close(hem, tr, step):
        lib.reactive_amd_step_destroy(step)
        lib.reactive_transport_destroy(tr)
        lib.reactive_gpu_hem_destroy_v1(hem)

    # uninterrupted: three steps
    hem, tr, step = handle()
    assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
    rep = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, 3, 0, C.byref(rep)) == 0, lib.reactive_amd_step_error(step)
    q3, s3 = read(step)
    close(hem, tr, step)
    # two steps, checkpoint, restart, one step
    hem, tr, step = handle()
    assert lib.reactive_amd_step_load(step, rlib.ptr(case["q"]), rlib.ptr(case["states"])) == 0
    rep2 = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, 2, 0, C.byref(rep2)) == 0, lib.reactive_amd_step_error(step)
    q2, s2 = read(step)
    close(hem, tr, step)
    history = dict(steps=int(rep2.steps), retries=0, time=1e-06, lastDt=rep2.dt, velocityScale=1.0,
                   energyScale=1.0, initial=np.zeros(5), boundary=np.array(list(rep2.boundaryIntegral)[:5]))
    d = checkpoint.write_checkpoint(str(tmp_path), q2, s2, history, identity="cube10")
    q, s, h = checkpoint.read_checkpoint(d, len(q2), 5)
    assert h["steps"] == 2 and h["lastDt"] == rep2.dt
    hem, tr, step = handle()
    assert lib.reactive_amd_step_load(step, rlib.ptr(q), rlib.ptr(s)) == 0
    rep1 = abi.StepReport.fresh()
    assert lib.reactive_amd_step_run(step, 0.4, 1.0, 1, 0, C.byref(rep1)) == 0, lib.reactive_amd_step_error(step)
    qr, sr = read(step)
    close(hem, tr, step)
    assert np.array_equal(qr, q3)
    assert sr.tobytes() == s3.tobytes()
