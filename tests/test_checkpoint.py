"""Schema-3 checkpoints and HEM captures (reactive_amd.checkpoint), on the CPU.

The capture fixtures under tests/golden/ were written by the reference's own
capture tool (tools/hem_capture.cpp, compiled from its source; see
tests/golden/make_capture.py): the reader must return exactly the batch the
tool was given, and the writer must reproduce the tool's bytes.  The
checkpoint reader and writer follow reactiveCheckpoint.H:34-85 (no reference
checkpoint file exists here: the writer needs OpenFOAM): round trips, the
commit marker's manifest, and each of the reference's rejections."""
import os
import struct

import numpy as np
import pytest

from helpers import GOLDEN
from reactive_amd import abi, cases, checkpoint

import importlib.util

_spec = importlib.util.spec_from_file_location("make_capture", os.path.join(GOLDEN, "make_capture.py"))
make_capture = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_capture)


@pytest.mark.parametrize("name", sorted(make_capture.SETS))
def test_capture_written_by_the_reference_tool(name, tmp_path):
    path = os.path.join(GOLDEN, f"{name}.bin")
    cap = checkpoint.read_capture(path)
    mb, q, e, seeds = make_capture.inputs(name)
    assert cap["model"] == mb
    assert cap["species"] == q.shape[1] and cap["capillary"] is None
    assert np.array_equal(cap["q"], q) and np.array_equal(cap["energy"], e)
    assert cap["states"].tobytes() == seeds.tobytes()
    mine = tmp_path / "again.bin"
    checkpoint.write_capture(str(mine), mb, q, e, seeds)
    assert mine.read_bytes() == open(path, "rb").read()


def test_capture_rejects_other_layouts(tmp_path):
    data = bytearray(open(os.path.join(GOLDEN, "capture_n2o.bin"), "rb").read())
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(data[:-8]))
    with pytest.raises(checkpoint.CheckpointError, match="missing or trailing"):
        checkpoint.read_capture(str(bad))
    data2 = bytearray(data)
    struct.pack_into("<Q", data2, 0, 0x1234)
    bad.write_bytes(bytes(data2))
    with pytest.raises(checkpoint.CheckpointError, match="not a HEM capture"):
        checkpoint.read_capture(str(bad))


def _state(n=6, seed=4):
    faces, volume, q, seeds, _ = cases.pure_n2o_cube(n, seed=seed)
    history = dict(steps=17, retries=2, time=1e-05, lastDt=3.5e-08, velocityScale=1.0, energyScale=2.5e5,
                   initial=np.arange(5.0) + 0.25, boundary=-np.arange(5.0) * 1e-3)
    return q, seeds, history


def test_checkpoint_round_trip(tmp_path):
    q, states, history = _state()
    d = checkpoint.write_checkpoint(str(tmp_path), q, states, history, identity="model-hash")
    assert os.path.basename(d) == "1e-05"
    marker = open(os.path.join(d, checkpoint.MARKER)).read().splitlines()
    assert marker[0] == "ReactiveFoam checkpoint 3"
    assert [line.split(" ", 1)[1] for line in marker[1:]] == ['"reactiveCheckpointInfo"', '"reactiveState.bin"',
                                                               '"reactiveStateIdentity"']
    q2, s2, h2 = checkpoint.read_checkpoint(d, len(q), q.shape[1])
    assert np.array_equal(q2, q) and s2.tobytes() == states.tobytes()
    for k in ("steps", "retries", "time", "lastDt", "velocityScale", "energyScale"):
        assert h2[k] == history[k]
    assert np.array_equal(h2["initial"], history["initial"]) and np.array_equal(h2["boundary"], history["boundary"])
    # the state file: header, history, baseline, boundary, q cell-major, states
    raw = open(os.path.join(d, checkpoint.STATE_FILE), "rb").read()
    assert struct.unpack_from("<6Q", raw) == (0x524643484B505433, 0x0102030405060708, 8, len(q), 5, 176)
    assert struct.unpack_from("<2Q4d", raw, 48) == (17, 2, 1e-05, 3.5e-08, 1.0, 2.5e5)
    assert np.array_equal(np.frombuffer(raw, "<f8", q.size, 96 + 80).reshape(q.shape), q)
    assert len(raw) == 96 + 80 + q.size * 8 + len(q) * 176


def test_checkpoint_rejections(tmp_path):
    q, states, history = _state()
    d = checkpoint.write_checkpoint(str(tmp_path / "a"), q, states, history, identity="x")
    state = os.path.join(d, checkpoint.STATE_FILE)
    with pytest.raises(checkpoint.CheckpointError, match="mismatch"):
        checkpoint.read_checkpoint(d, len(q) + 1, 5)  # another mesh
    raw = bytearray(open(state, "rb").read())
    raw[-1] ^= 1  # a changed byte: checksum mismatch
    open(state, "wb").write(bytes(raw))
    with pytest.raises(checkpoint.CheckpointError, match="checksum"):
        checkpoint.read_checkpoint(d)
    raw[-1] ^= 1
    open(state, "wb").write(bytes(raw) + b"\0" * 8)  # trailing data, recommitted
    checkpoint.commit_marker(d)
    with pytest.raises(checkpoint.CheckpointError, match="trailing"):
        checkpoint.read_checkpoint(d)
    struct.pack_into("<Q", raw, 0, 0x1)  # another magic
    open(state, "wb").write(bytes(raw))
    checkpoint.commit_marker(d)
    with pytest.raises(checkpoint.CheckpointError, match="layout"):
        checkpoint.read_checkpoint(d)
    e = checkpoint.write_checkpoint(str(tmp_path / "b"), q, states, history, identity="x")
    os.rename(e, os.path.join(str(tmp_path / "b"), "2e-05"))  # the directory is not the state's time
    with pytest.raises(checkpoint.CheckpointError, match="time"):
        checkpoint.read_checkpoint(os.path.join(str(tmp_path / "b"), "2e-05"))
    f = checkpoint.write_checkpoint(str(tmp_path / "c"), q, states, history, identity="x")
    os.remove(os.path.join(f, checkpoint.IDENTITY_FILES[0]))  # a listed file missing
    with pytest.raises(checkpoint.CheckpointError, match="checksum"):
        checkpoint.read_checkpoint(f)
    checkpoint.commit_marker(f)  # now the identity file is not listed at all
    with pytest.raises(checkpoint.CheckpointError, match="manifest"):
        checkpoint.read_checkpoint(f)
    os.remove(os.path.join(f, checkpoint.MARKER))
    with pytest.raises(checkpoint.CheckpointError, match="marker"):
        checkpoint.read_checkpoint(f)
    bad = dict(history, lastDt=-1.0)
    g = checkpoint.write_checkpoint(str(tmp_path / "d"), q, states, bad, identity="x")
    with pytest.raises(checkpoint.CheckpointError, match="time/history"):
        checkpoint.read_checkpoint(g)


def test_capture_from_checkpoint(tmp_path):
    q, states, history = _state(n=10, seed=6)
    d = checkpoint.write_checkpoint(str(tmp_path), q, states, history, identity="x")
    from reactive_amd import model
    mb = model.model_bytes(model.build_model())
    cap = checkpoint.capture_from_checkpoint(d, mb, 1)
    assert np.array_equal(cap["q"], q[:, :1]) and np.array_equal(cap["energy"], cases.internal_energy(q, 1))
    assert cap["states"].tobytes() == states.tobytes() and len(cap["energy"]) == 1000
    liquid = checkpoint.capture_from_checkpoint(d, mb, 1, select="liquid")
    assert np.array_equal(liquid["cells"], np.flatnonzero(states["liquidMass"][:, 0] > 0))
    out = tmp_path / "batch.bin"
    checkpoint.write_capture(str(out), cap["model"], cap["q"], cap["energy"], cap["states"])
    again = checkpoint.read_capture(str(out))
    assert np.array_equal(again["q"], cap["q"]) and again["states"].tobytes() == cap["states"].tobytes()
