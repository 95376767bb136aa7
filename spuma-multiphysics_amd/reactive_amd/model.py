"""HEM model image: the 47,448-byte POD the flash ABI consumes.

The byte layout is the reference's ReactiveDeviceFlash::Model
(src/reactiveThermo/reactiveDeviceFlash.h:39-53, static_assert at :52-53:
size 47448, stableGasPrune at 45004, weights at 45008), built from
ReactiveDevicePR::Table<16,8> (reactiveDevicePR.h:76-99) and
ReactiveSolidThermo::Model (reactiveSolidThermo.h:21-26).  The tables are
filled the way reactiveDevicePRExport.h:24-123 fills them from Cantera:
NASA9 regions in file order, kappa(omega), a_ij = sqrt(a_i a_j); the model
header follows reactiveGpuHemModel.inc:3-18.
"""
import math
import numpy as np

from . import mechanism

MODEL_BYTES = 47448
MAX_SPECIES = 16
MAX_REGIONS = 8

REGION = np.dtype([("tmin", "<f8"), ("tmax", "<f8"), ("poly", "<i4"), ("pad", "<i4"),
                   ("coef", "<f8", (9,))])
SPECIES = np.dtype([("W", "<f8"), ("a", "<f8"), ("b", "<f8"), ("kappa", "<f8"),
                    ("nreg", "<i4"), ("pad", "<i4"), ("regions", REGION, (MAX_REGIONS,))])
TABLE = np.dtype([("ns", "<i4"), ("pad", "<i4"), ("refP", "<f8"),
                  ("species", SPECIES, (MAX_SPECIES,)), ("aBinary", "<f8", (MAX_SPECIES, MAX_SPECIES))])
SOLID = np.dtype([("enabled", "<i4"), ("n", "<i4"), ("Tmin", "<f8"), ("Tmax", "<f8"), ("Tref", "<f8"),
                  ("pref", "<f8"), ("pmax", "<f8"), ("molarVolume", "<f8"), ("molecularWeight", "<f8"),
                  ("hRef", "<f8"), ("sRef", "<f8"), ("T", "<f8", (64,)), ("cp", "<f8", (64,))])
MODEL = np.dtype([("phase", TABLE, (3,)), ("ns", "<i4"), ("nl", "<i4"), ("condensable", "<i4", (2,)),
                  ("boundaryRecovery", "<i4"), ("scalarRecovery", "<i4"), ("analyticJacobian", "<i4"),
                  ("stableGasPrune", "<i4"), ("weights", "<f8", (MAX_SPECIES,)),
                  ("liquidTmin", "<f8", (2,)), ("liquidTc", "<f8", (2,)),
                  ("Tmin", "<f8"), ("Tmax", "<f8"), ("pmin", "<f8"), ("pmax", "<f8"),
                  ("vtol", "<f8"), ("etol", "<f8"), ("mutol", "<f8"),
                  ("condensedKind", "<i4", (2,)), ("enforceSpeciesTemperatureBounds", "<i4"),
                  ("pad", "<i4"), ("solid", SOLID, (2,))])
assert REGION.itemsize == 96 and SPECIES.itemsize == 808 and TABLE.itemsize == 14992
assert SOLID.itemsize == 1104 and MODEL.itemsize == MODEL_BYTES
assert MODEL.fields["stableGasPrune"][1] == 45004 and MODEL.fields["weights"][1] == 45008


def _fill_table(table, names, a_scale=1.0, b_scale=1.0):
    table["ns"] = len(names)
    table["refP"] = mechanism.REFERENCE_PRESSURE
    for k, name in enumerate(names):
        data = mechanism.SPECIES[name]
        sp = table["species"][k]
        sp["W"] = data["W"]
        sp["a"] = data["a"] * a_scale
        sp["b"] = data["b"] * b_scale
        sp["kappa"] = mechanism.kappa(data["omega"])
        ranges = data["ranges"]
        sp["nreg"] = len(data["nasa9"])
        for r, coef in enumerate(data["nasa9"]):
            sp["regions"][r]["tmin"] = ranges[r]
            sp["regions"][r]["tmax"] = ranges[r + 1]
            sp["regions"][r]["poly"] = 9
            sp["regions"][r]["coef"][:] = coef
    for i, ni in enumerate(names):
        for j, nj in enumerate(names):
            table["aBinary"][i, j] = math.sqrt(mechanism.SPECIES[ni]["a"] * mechanism.SPECIES[nj]["a"]) * a_scale


def build_model(gas_species=("N2O",), liquid="N2O", analytic_jacobian=0, stable_gas_prune=0,
                scalar_recovery=1, limits=None, a_scale=1.0, b_scale=1.0, boundary_recovery=0):
    """Return the model image (numpy structured scalar) for a PR gas of
    `gas_species` with one pure liquid of `liquid` (None for no liquid).
    a_scale, b_scale multiply every species' PR a (and a_ij) and b;
    boundary_recovery = 1 selects the vapour-log boundary search
    (reactiveDeviceFlash.h:544-560)."""
    lim = dict(mechanism.LIMITS)
    if limits:
        lim.update(limits)
    m = np.zeros((), dtype=MODEL)
    _fill_table(m["phase"][0], list(gas_species), a_scale, b_scale)
    m["ns"] = len(gas_species)
    m["scalarRecovery"] = scalar_recovery
    m["analyticJacobian"] = analytic_jacobian
    m["stableGasPrune"] = stable_gas_prune
    m["boundaryRecovery"] = boundary_recovery
    for k, name in enumerate(gas_species):
        m["weights"][k] = mechanism.SPECIES[name]["W"]
    if liquid is not None:
        _fill_table(m["phase"][1], [liquid], a_scale, b_scale)
        m["nl"] = 1
        m["condensable"][0] = list(gas_species).index(liquid)
        m["liquidTmin"][0] = lim["liquidTmin"]
        m["liquidTc"][0] = lim["liquidTc"]
    for key in ("Tmin", "Tmax", "pmin", "pmax", "vtol", "etol", "mutol"):
        m[key] = lim[key]
    m["enforceSpeciesTemperatureBounds"] = lim["enforceSpeciesTemperatureBounds"]
    return m


def model_bytes(m):
    b = m.tobytes()
    assert len(b) == MODEL_BYTES
    return b
