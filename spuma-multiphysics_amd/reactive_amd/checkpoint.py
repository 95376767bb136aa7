"""ReactiveFoam state checkpoints (schema 3) and HEM batch captures.

Two binary formats of the reference, read and written here so that saved
solver states and captured flash batches replay on the MI355X library:

* The schema-3 checkpoint (src/reactiveFoam/reactiveCheckpoint.H:34-85): a
  time directory holding ``reactiveState.bin`` (header of six uint64 --
  magic 0x524643484b505433, endian tag 0x0102030405060708, scalar bytes,
  cells, variables, state bytes -- then the history: steps, retries (uint64),
  time, lastDt, velocityScale, energyScale (float64), the conservation
  baseline and boundary integral (variables float64 each), q cell-major
  (cells x variables float64) and one 176-byte ReactiveThermoState per cell)
  and the commit marker ``reactiveCheckpointComplete``: the line "ReactiveFoam
  checkpoint 3", then one line per regular file of the directory, in sorted
  order, with its SHA-256 and quoted name.  A directory is accepted only when
  the marker lists reactiveState.bin and an identity file, every listed file
  exists with that hash, and there are more than two entries
  (checkpoint::verify, :57-71).
* The HEM capture (tools/hem_capture.cpp:34-37): eight uint64 -- magic
  0x48454d4341503031, version 1, model bytes, cells, species, state bytes,
  capillary-record bytes, capillary flag -- then the 47,448-byte model image,
  q (cells x species float64), the internal energy density (cells float64),
  the seed states (cells x 176 bytes) and, with the flag, one 24-byte
  capillary input per cell.  tools/hem_checkpoint_batch.py builds such a
  batch from a checkpoint; tools/replay_hem_precision.py replays it.

`replay_capture` runs a capture through the library's drop-in
reactive_gpu_hem_run_v1 (no capillary records) or run_v2, as the reference's
replay tool does.
"""
import ctypes as C
import hashlib
import math
import os
import shlex
import struct

import numpy as np

from . import abi

STATE_MAGIC = 0x524643484B505433
ENDIAN_TAG = 0x0102030405060708
CAPTURE_MAGIC = 0x48454D4341503031
STATE_BYTES = 176
MODEL_BYTES = 47448
MARKER = "reactiveCheckpointComplete"
MARKER_SIGNATURE = "ReactiveFoam checkpoint 3"
STATE_FILE = "reactiveState.bin"
IDENTITY_FILES = ("reactiveStateIdentity", "reactiveStateIdentity.gz")
CAPILLARY_DTYPE = np.dtype([("color", "<f8"), ("pressureJump", "<f8"), ("equilibrium", "<i4"), ("pad", "<i4")])
assert CAPILLARY_DTYPE.itemsize == 24


class CheckpointError(RuntimeError):
    """A checkpoint or capture that does not match its format."""


def _require(ok, message):
    if not ok:
        raise CheckpointError(message)


def time_name(time):
    """OpenFOAM's default time-directory name (general format, 6 digits),
    which the reference requires the directory to carry (:74-75)."""
    return format(time, ".6g")


def _sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


# ---------------------------------------------------------------- checkpoint
def write_state(path, q, states, history):
    """reactiveState.bin (checkpoint::save, :34-47, non-mechanical states).
    history: dict with steps, retries, time, lastDt, velocityScale,
    energyScale, initial[nv], boundary[nv]."""
    q = np.ascontiguousarray(q, dtype="<f8")
    states = np.ascontiguousarray(states)
    _require(states.dtype.itemsize == STATE_BYTES, "states must be 176-byte ReactiveThermoState records")
    cells, nv = q.shape
    _require(len(states) == cells, "one state per cell")
    initial = np.asarray(history["initial"], dtype="<f8")
    boundary = np.asarray(history["boundary"], dtype="<f8")
    _require(initial.shape == (nv,) and boundary.shape == (nv,), "conservation baseline and boundary need nv values")
    with open(path, "wb") as out:
        out.write(struct.pack("<6Q", STATE_MAGIC, ENDIAN_TAG, 8, cells, nv, STATE_BYTES))
        out.write(struct.pack("<2Q", int(history.get("steps", 0)), int(history.get("retries", 0))))
        out.write(struct.pack("<4d", float(history["time"]), float(history.get("lastDt", 0.0)),
                              float(history.get("velocityScale", 1.0)), float(history.get("energyScale", 1.0))))
        out.write(initial.tobytes())
        out.write(boundary.tobytes())
        out.write(q.tobytes())
        out.write(states.tobytes())


def read_state(path, cells=None, nv=None):
    """reactiveState.bin: (q[cells, nv], states[cells], history), with the
    reference's layout and history checks (checkpoint::load, :72-85).  cells
    and nv, when given, must match (the mesh and closure of the reader)."""
    data = open(path, "rb").read()
    _require(len(data) >= 48, "Incomplete checkpoint state")
    magic, endian, scalar, n, v, state_bytes = struct.unpack_from("<6Q", data)
    _require(magic == STATE_MAGIC and endian == ENDIAN_TAG and scalar == 8 and state_bytes == STATE_BYTES and
             (cells is None or n == cells) and (nv is None or v == nv),
             "Checkpoint layout/mesh/closure mismatch; explicit migration required")
    need = 48 + 16 + 32 + 2 * v * 8 + n * v * 8 + n * STATE_BYTES
    _require(len(data) >= need, "Incomplete checkpoint state")
    _require(len(data) == need, "Checkpoint state has unexpected trailing data")
    steps, retries = struct.unpack_from("<2Q", data, 48)
    time, last_dt, velocity_scale, energy_scale = struct.unpack_from("<4d", data, 64)
    _require(math.isfinite(time) and math.isfinite(last_dt) and last_dt >= 0 and math.isfinite(velocity_scale) and
             velocity_scale > 0 and math.isfinite(energy_scale) and energy_scale > 0,
             "Invalid checkpoint time/history")
    offset = 96
    initial = np.frombuffer(data, "<f8", v, offset).copy()
    offset += v * 8
    boundary = np.frombuffer(data, "<f8", v, offset).copy()
    offset += v * 8
    _require(np.isfinite(initial).all(), "Invalid checkpoint conservation baseline")
    _require(np.isfinite(boundary).all(), "Invalid checkpoint boundary integral")
    q = np.frombuffer(data, "<f8", n * v, offset).reshape(n, v).copy()
    offset += n * v * 8
    _require(np.isfinite(q).all(), "Checkpoint field/binary conserved state mismatch")
    states = np.frombuffer(data, abi.STATE_DTYPE, n, offset).copy()
    history = dict(steps=steps, retries=retries, time=time, lastDt=last_dt, velocityScale=velocity_scale,
                   energyScale=energy_scale, initial=initial, boundary=boundary)
    return q, states, history


def commit_marker(directory):
    """Writes the commit marker over every regular file of the directory,
    sorted (Transaction::commit, reactiveCheckpoint.H:108-120)."""
    names = sorted(f for f in os.listdir(directory) if f != MARKER and os.path.isfile(os.path.join(directory, f)))
    lines = [MARKER_SIGNATURE]
    for name in names:
        lines.append(f"{_sha256(os.path.join(directory, name))} \"{name}\"")
    with open(os.path.join(directory, MARKER), "w") as out:
        out.write("\n".join(lines) + "\n")


def verify(directory):
    """checkpoint::verify (:57-71): the marker's signature, every listed file
    present with its SHA-256, a state file and an identity file listed."""
    path = os.path.join(directory, MARKER)
    _require(os.path.isfile(path), "Incomplete checkpoint: missing/invalid complete marker")
    lines = open(path).read().splitlines()
    _require(bool(lines) and lines[0] == MARKER_SIGNATURE, "Incomplete checkpoint: missing/invalid complete marker")
    state = identity = False
    count = 0
    for line in lines[1:]:
        parts = shlex.split(line)
        _require(len(parts) == 2, "Invalid checkpoint manifest entry")
        digest, name = parts
        _require(len(digest) == 64 and all(c in "0123456789abcdef" for c in digest) and "/" not in name and
                 name not in (".", ".."), "Invalid checkpoint manifest entry")
        file = os.path.join(directory, name)
        _require(os.path.isfile(file) and _sha256(file) == digest, f"Checkpoint checksum mismatch: {name}")
        state |= name == STATE_FILE
        identity |= name in IDENTITY_FILES
        count += 1
    _require(state and identity and count > 2, "Incomplete checkpoint manifest")


INFO_FILE = "reactiveCheckpointInfo"


def write_checkpoint(root, q, states, history, identity="", extra=None):
    """A committed schema-3 time directory root/<time name>: the state file,
    an identity file, a short text file naming the cells, variables and time
    (the reference's directories also hold the solver's field files; its
    verify() asks for more than two manifest entries), optional extra files
    {name: bytes}, and the marker.  Returns the directory."""
    directory = os.path.join(root, time_name(history["time"]))
    os.makedirs(directory, exist_ok=True)
    write_state(os.path.join(directory, STATE_FILE), q, states, history)
    with open(os.path.join(directory, IDENTITY_FILES[0]), "w") as out:
        out.write(identity)
    with open(os.path.join(directory, INFO_FILE), "w") as out:
        out.write(f"cells {len(q)}\nvariables {q.shape[1]}\ntime {history['time']!r}\n")
    for name, blob in (extra or {}).items():
        with open(os.path.join(directory, name), "wb") as out:
            out.write(blob)
    commit_marker(directory)
    return directory


def read_checkpoint(directory, cells=None, nv=None):
    """Verifies and reads a committed time directory: (q, states, history).
    The directory's name must be the time's (:74-75)."""
    verify(directory)
    q, states, history = read_state(os.path.join(directory, STATE_FILE), cells, nv)
    _require(time_name(history["time"]) == os.path.basename(os.path.normpath(directory)),
             "Invalid checkpoint time/history")
    return q, states, history


# ------------------------------------------------------------------- capture
def write_capture(path, model, q, energy, states, capillary=None):
    """A HEM batch in the capture layout (hem_capture.cpp:34-37)."""
    q = np.ascontiguousarray(q, dtype="<f8")
    count, ns = q.shape
    energy = np.ascontiguousarray(energy, dtype="<f8")
    states = np.ascontiguousarray(states)
    _require(len(model) == MODEL_BYTES and energy.shape == (count,) and len(states) == count and
             states.dtype.itemsize == STATE_BYTES, "capture inputs: model image, energy and states per cell")
    header = struct.pack("<8Q", CAPTURE_MAGIC, 1, len(model), count, ns, STATE_BYTES, CAPILLARY_DTYPE.itemsize,
                         0 if capillary is None else 1)
    with open(path, "xb") as out:
        out.write(header)
        out.write(bytes(model))
        out.write(q.tobytes())
        out.write(energy.tobytes())
        out.write(states.tobytes())
        if capillary is not None:
            cap = np.ascontiguousarray(capillary, dtype=CAPILLARY_DTYPE)
            _require(len(cap) == count, "one capillary record per cell")
            out.write(cap.tobytes())


def read_capture(path):
    """A HEM capture: dict(model, q[count, ns], energy, states, capillary or
    None), with the replay tool's checks (replay_hem_precision.py:29-37)."""
    data = open(path, "rb").read()
    _require(len(data) >= 64, "not a HEM capture")
    magic, version, model_bytes, count, ns, state_bytes, cap_bytes, has_cap = struct.unpack_from("<8Q", data)
    _require(magic == CAPTURE_MAGIC and version == 1, "not a HEM capture")
    _require(0 < count <= 1048576 and 1 <= ns <= 16 and model_bytes == MODEL_BYTES and state_bytes == STATE_BYTES and
             cap_bytes == CAPILLARY_DTYPE.itemsize and has_cap in (0, 1), "unsupported HEM capture layout")
    need = 64 + model_bytes + count * (ns * 8 + 8 + STATE_BYTES + (cap_bytes if has_cap else 0))
    _require(len(data) == need, "HEM capture has missing or trailing bytes")
    offset = 64
    model = data[offset:offset + model_bytes]
    offset += model_bytes
    q = np.frombuffer(data, "<f8", count * ns, offset).reshape(count, ns).copy()
    offset += count * ns * 8
    energy = np.frombuffer(data, "<f8", count, offset).copy()
    offset += count * 8
    states = np.frombuffer(data, abi.STATE_DTYPE, count, offset).copy()
    offset += count * STATE_BYTES
    capillary = None
    if has_cap:
        capillary = np.frombuffer(data, CAPILLARY_DTYPE, count, offset).copy()
        offset += count * CAPILLARY_DTYPE.itemsize
    _require(offset == len(data), "HEM capture has missing or trailing bytes")
    return dict(model=model, q=q, energy=energy, states=states, capillary=capillary, species=ns)


def capture_from_checkpoint(directory, model, species, select="all", limit=1048576, stride=1, offset=0):
    """The equilibrium-flash batch of a committed checkpoint
    (tools/hem_checkpoint_batch.py without an interface): species masses, the
    bulk internal energy rhoE - |rho u|^2 / (2 rho) as the driver forms it for
    the flash, the committed states as seeds.  select: all cells, those holding the condensable (species 0), or
    those with liquid.  Returns the read_capture dict."""
    from .cases import internal_energy
    q, states, history = read_checkpoint(directory)
    ns = species
    energy = internal_energy(q, ns)  # the driver's own expression (ReactiveFoam.C:727-728)
    mask = {"all": np.ones(len(q), bool), "condensable": q[:, 0] > 0,
            "liquid": states["liquidMass"][:, 0] > 0}[select]
    index = np.flatnonzero(mask)[::stride][offset:offset + limit]
    _require(1 <= len(index) <= 1048576, "a replayed batch holds 1 to 1048576 cells")
    return dict(model=bytes(model), q=np.ascontiguousarray(q[index, :ns]), energy=energy[index],
                states=states[index].copy(), capillary=None, species=ns, cells=index)


def replay_capture(lib, capture, repeats=1):
    """Runs a capture through reactive_gpu_hem_run_v1 (no capillary records)
    or run_v2 on a fresh handle, `repeats` times on identical inputs:
    (states, success, last profile, kernel seconds per repeat)."""
    model = capture["model"]
    count = len(capture["energy"])
    err = C.create_string_buffer(4096)
    handle = lib.reactive_gpu_hem_create_v1(model, len(model), count, err, len(err))
    if not handle:
        raise RuntimeError(err.value.decode())
    q = np.ascontiguousarray(capture["q"])
    energy = np.ascontiguousarray(capture["energy"])
    seconds = []
    try:
        for _ in range(repeats):
            states = capture["states"].copy()
            success = np.zeros(count, np.int32)
            profile = abi.HemProfile.fresh()
            if capture["capillary"] is None:
                rc = lib.reactive_gpu_hem_run_v1(handle, q.ctypes.data, energy.ctypes.data, count, states.ctypes.data,
                                                 success.ctypes.data, C.byref(profile), err, len(err))
            else:
                cap = np.ascontiguousarray(capture["capillary"])
                rc = lib.reactive_gpu_hem_run_v2(handle, q.ctypes.data, energy.ctypes.data, cap.ctypes.data, count,
                                                 states.ctypes.data, success.ctypes.data, C.byref(profile), err,
                                                 len(err))
            if rc:
                raise RuntimeError(err.value.decode())
            seconds.append(profile.kernelSeconds)
    finally:
        lib.reactive_gpu_hem_destroy_v1(handle)
    return states, success, profile, seconds
