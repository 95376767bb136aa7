"""The mixed-arithmetic equilibrium flash behind the drop-in HEM ABI
(reactive_gpu_hem_run_v1 on a handle set to REACTIVE_AMD_FLASH_MIXED, the
FP32 combined policy): against the plain FP64 batch of the same inputs and
both against the converged root, with the gates of the resident step's mixed
flash: no more than 0.2 % flips, T and p within 1e-10 of the root on every
cell whose phase set the root shares, at most 0.2 % of the converged cells on
another phase set than the root.  The FP32 search and the deferred FP64
queue must both have run, and two batches must agree bit for bit."""
import ctypes as C

import numpy as np
import pytest

from helpers import binary_case, cube_case, jet_case, oracle_flash
from test_gpu_mixed import one_step
from reactive_amd import abi, cases, lib as rlib, model
from reactive_amd.abi import FP64, MIXED

pytestmark = pytest.mark.gpu

TIGHT = {"vtol": 1e-12, "etol": 1e-12, "mutol": 1e-10}


def run(lib, mb, q, e, seeds, mode):
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), len(e), err, 1024)
    assert hem, err.value
    assert lib.reactive_amd_hem_set_arithmetic(hem, mode) == 0
    states = seeds.copy()
    ok = np.zeros(len(e), np.int32)
    prof = abi.HemProfile.fresh()
    rc = lib.reactive_gpu_hem_run_v1(hem, rlib.ptr(np.ascontiguousarray(q)), rlib.ptr(e), len(e), rlib.ptr(states),
                                     rlib.ptr(ok), C.byref(prof), err, 1024)
    assert rc == 0, err.value
    single, queued = C.c_uint64(), C.c_uint64()
    assert lib.reactive_amd_hem_mixed_report(hem, C.byref(single), C.byref(queued)) == 0
    lib.reactive_gpu_hem_destroy_v1(hem)
    return states, ok == 1, prof, int(single.value), int(queued.value)


@pytest.mark.parametrize("kind", ["n2o", "n2o-n2", "jets", "n2o-n2-prune"])
def test_hem_mixed_against_fp64_and_converged_root(kind):
    lib = rlib.load()
    if kind.startswith("n2o-n2"):
        case = binary_case(12, seed=8)
    elif kind == "jets":
        case = jet_case(16, seed=67)
    else:
        case = cube_case(16, seed=66)
    ns = case["ns"]
    species = ("N2O", "N2") if ns == 2 else ("N2O",)
    mb = case["model"]
    if kind.endswith("-prune"):
        mb = model.model_bytes(model.build_model(gas_species=species, stable_gas_prune=1))
    # the inputs of a step's second flash: the conserved state after one FP64
    # step, seeded with the states at the step's start
    qs, _, _ = one_step(lib, case, FP64)
    q = np.ascontiguousarray(qs[:, :ns])
    e = cases.internal_energy(qs, ns)
    seeds = case["states"]
    s64, ok64, p64, single64, _ = run(lib, mb, q, e, seeds, FP64)
    smx, okmx, pmx, single, queued = run(lib, mb, q, e, seeds, MIXED)
    again, okagain, _, _, _ = run(lib, mb, q, e, seeds, MIXED)
    assert again.tobytes() == smx.tobytes() and np.array_equal(okagain, okmx)
    assert single64 == 0 and single > 0
    assert pmx.phaseEvaluations > single and pmx.succeeded == int(okmx.sum())
    cells = len(e)
    both = ok64 & okmx
    flips = int((ok64 != okmx).sum() + (smx["activeLiquids"][both] != s64["activeLiquids"][both]).sum())
    tight = model.model_bytes(model.build_model(gas_species=species, limits=TIGHT,
                                                stable_gas_prune=1 if kind.endswith("-prune") else 0))
    root, rok, _ = oracle_flash(tight, q, e, seeds)
    rok = rok == 1
    use = okmx & rok & (smx["activeLiquids"] == root["activeLiquids"])
    dropped = int((okmx & rok & (smx["activeLiquids"] != root["activeLiquids"])).sum())
    err = {f: np.abs(smx[f][use] - root[f][use]) / np.abs(root[f][use]) for f in ("T", "p")}
    print(f"\n{kind}: cells {cells}, flips {flips}, FP64 failures {int((~ok64).sum())}, mixed failures "
          f"{int((~okmx).sum())}, deferred {queued}, other phase set than the root {dropped}, "
          f"max rel vs root T {err['T'].max():.2e} p {err['p'].max():.2e}; phases FP64 {p64.phaseEvaluations}, "
          f"mixed {pmx.phaseEvaluations} ({single} FP32)")
    assert flips <= max(2, cells // 500)
    assert dropped <= max(2, cells // 500)
    for f in ("T", "p"):
        assert err[f].max() <= 1e-10, (f, float(err[f].max()), int((err[f] > 1e-10).sum()))
    if kind == "n2o-n2":
        assert queued > 0


def test_hem_mixed_rules():
    lib = rlib.load()
    case = cube_case(4, seed=3)
    mb = case["model"]
    err = C.create_string_buffer(1024)
    hem = lib.reactive_gpu_hem_create_v1(mb, len(mb), 64, err, 1024)
    assert lib.reactive_amd_hem_set_arithmetic(None, MIXED) == -1
    assert lib.reactive_amd_hem_set_arithmetic(hem, 7) == 1
    assert lib.reactive_amd_hem_set_arithmetic(hem, MIXED) == 0
    # frozen-inventory batches run in FP64 whatever the arithmetic
    q = np.ascontiguousarray(case["q"][:, :1])
    e = cases.internal_energy(case["q"], 1)
    cap = np.zeros(len(e), dtype=abi.CAPILLARY_DTYPE)
    st = case["states"].copy()
    ok = np.zeros(len(e), np.int32)
    prof = abi.HemProfile.fresh()
    assert lib.reactive_gpu_hem_run_v2(hem, rlib.ptr(q), rlib.ptr(e), rlib.ptr(cap), len(e), rlib.ptr(st),
                                       rlib.ptr(ok), C.byref(prof), err, 1024) == 0, err.value
    single, queued = C.c_uint64(), C.c_uint64()
    assert lib.reactive_amd_hem_mixed_report(hem, C.byref(single), C.byref(queued)) == 0
    assert single.value == 0 and queued.value == 0
    lib.reactive_gpu_hem_destroy_v1(hem)
