"""Peng-Robinson + NASA9 species data for the N2O / N2 HEM configurations.

The numbers are the reference mechanism's own data
(examples/impinging-n2o-supercooled/cold-pr-148K.yaml): N2 at lines 52-111,
N2O at lines 172-231; PR a [Pa m^6/kmol^2], b [m^3/kmol] and the acentric
factor are the `equation-of-state` entries.  The HEM limits and tolerances are
the thermo configuration (cold-pr-148K-config.yaml:1-19).

Molecular weights follow Cantera's element weights, summed in element order
(N 14.007, O 15.999); the NASA9 reference pressure is one standard atmosphere.
Both values define every downstream result to the last bit.
"""

ELEMENT_WEIGHT = {"N": 14.007, "O": 15.999}
REFERENCE_PRESSURE = 101325.0
GAS_CONSTANT = 8314.46261815324  # J/(kmol K), reactiveDevicePR.h:31


def _weight(composition):
    w = 0.0
    for element in ("N", "O"):
        w += ELEMENT_WEIGHT[element] * composition.get(element, 0.0)
    return w


SPECIES = {
    "N2": {
        "composition": {"N": 2.0},
        "a": 148227.8749500853,
        "b": 0.02403707313054656,
        "omega": 0.0372,
        "ranges": [148.0, 182.34, 200.0, 1050.0, 3500.0],
        "nasa9": [
            [98.29709685176373, -3.927219570926414, 3.5657235265591836, -0.0005864923434821304,
             2.999170859129467e-06, -8.169508775555017e-09, 9.359447755361056e-12,
             -1028.6313331407114, 2.811951852573376],
            [0.0, 0.0, 3.517433744475874, -8.456160143098416e-05, 7.709235665872965e-07,
             -3.002427186776785e-09, 4.451086162653223e-12, -1045.881160054519, 3.01983695239548],
            [0.0, 0.0, 3.73100682, -0.0018315973, 4.32324662e-06, -3.04378151e-09,
             7.46071562e-13, -1062.87426, 2.16821198],
            [0.0, 0.0, 2.81166073, 0.00167067353, -6.79997428e-07, 1.32881379e-10,
             -1.02767442e-14, -869.811579, 6.6483805],
        ],
    },
    "N2O": {
        "composition": {"N": 2.0, "O": 1.0},
        "a": 417972.1865083438,
        "b": 0.02763390066995578,
        "omega": 0.1613,
        "ranges": [148.0, 182.34, 200.0, 1420.0, 3500.0],
        "nasa9": [
            [-13883.483314218653, 400.2145375936049, 0.08987442809300043, -0.002575847027809755,
             0.00015894592824637003, -5.833630345245842e-07, 6.961726174507204e-10,
             6995.331413063375, 24.693976510805115],
            [0.0, 0.0, 4.796459403683242, -0.03064976304265226, 0.000248176450083356,
             -7.193439839786354e-07, 7.637795272607834e-10, 8617.974553201218, 2.097742156376054],
            [0.0, 0.0, 2.51620681, 0.00916990584, -7.8942084e-06, 3.42198245e-09,
             -5.84582078e-13, 8720.2997, 9.7150932],
            [0.0, 0.0, 4.85543722, 0.0025805244, -9.33594196e-07, 1.54088463e-10,
             -9.24862861e-15, 8055.95827, -2.39060855],
        ],
    },
}
for _name, _data in SPECIES.items():
    _data["W"] = _weight(_data["composition"])

# cold-pr-148K-config.yaml:5-19
LIMITS = {
    "Tmin": 148.0, "Tmax": 3500.0, "pmin": 1000.0, "pmax": 50000000.0,
    "vtol": 1.0e-9, "etol": 1.0e-9, "mutol": 1.0e-7,
    "liquidTmin": 148.0, "liquidTc": 309.52,
    "enforceSpeciesTemperatureBounds": 1,
}


def kappa(omega):
    """PR kappa(omega) (reactiveDevicePRExport.h:89-92, Cantera 3.2)."""
    if omega <= 0.491:
        return 0.37464 + 1.54226 * omega - 0.26992 * omega * omega
    return 0.374642 + 1.487503 * omega - 0.164423 * omega * omega + 0.016666 * omega * omega * omega
