"""The oracle's boundary search in vapour coordinates and its capillary
recoveries (oracle/oracle_flash.cpp) against the reference's own compiled
flash (tests/golden/flash_capillary.npz, written from oracle/_ref/libref.so):
states, success flags and every work counter bit for bit.  Sets: per-cell
colour, pressure jump and equilibrium flag on an N2O cube; boundaryRecovery =
1 binary N2O/N2 cubes with trace and ordinary N2 (finite-difference and
analytic Jacobian); capillary equilibrium cells with trace N2."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, oracle, oracle_capillary, oracle_flash

SETS = ["n2o", "trace", "trace_analytic", "binary", "trace_capillary"]


def load(name):
    z = np.load(os.path.join(GOLDEN, "flash_capillary.npz"))
    return {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}


def run(d, twin=False):
    mb = d["model"].tobytes()
    if "color" in d:
        return oracle_capillary(mb, d["q"], d["energy"], d["seeds"], d["color"], d["jump"], d["equilibrium"],
                                twin=twin)
    return oracle_flash(mb, d["q"], d["energy"], d["seeds"], twin=twin)


@pytest.mark.parametrize("name", SETS)
def test_capillary_oracle_matches_reference(name):
    d = load(name)
    out, ok, cnt = run(d)
    assert np.array_equal(ok, d["success"])
    assert out.tobytes() == d["out"].tobytes()
    # phases, residuals, candidates, stable, failures, analytic and fd Jacobians
    assert np.array_equal(cnt[:, :7], d["counters"][:, :7])


def test_boundary_search_runs():
    """The sets reach the second search: cells with more candidates than the
    five seeds of one search (the vapour-coordinate starts add six)."""
    for name in ("trace", "trace_capillary", "n2o"):
        d = load(name)
        extra = int((d["counters"][:, 2] > 5).sum())
        print(name, "cells with a second search or restarts:", extra)
        assert extra > 0


def test_trace_vapour_start():
    """Seed -3 of the vapour-coordinate search: a liquid cell with a trace of
    noncondensable gas converges to the state the reference finds."""
    d = load("trace_capillary")
    assert d["success"].sum() > 0.8 * len(d["success"])
