"""The shared elementary functions (spuma-multiphysics_amd/csrc/elementary.h),
exercised through the oracle's twin build, against 80-digit references
computed with Python's decimal module; the glibc build is the comparison."""
import decimal
import math
from decimal import Decimal as D

import numpy as np
import pytest

from helpers import oracle, p

decimal.getcontext().prec = 80


def lib(twin):
    return oracle(twin)


def _pi():
    def arctan_inv(x):
        x = D(x)
        s, t, n, sign = D(0), 1 / x, 1, 1
        while t / n > D(10) ** -78:
            s += sign * t / n
            t /= x * x
            n += 2
            sign = -sign
        return s
    return 16 * arctan_inv(5) - 4 * arctan_inv(239)


PI = _pi()


def dcos(x):
    x = D(x)
    twopi = 2 * PI
    x = x - twopi * (x / twopi).to_integral_value(rounding=decimal.ROUND_FLOOR)
    s, t, n = D(0), D(1), 0
    while abs(t) > D(10) ** -79:
        s += t
        t = -t * x * x / ((n + 1) * (n + 2))
        n += 2
    return s


def dacos(x):
    # Newton on cos(y) = x from the double estimate
    y = D(math.acos(x))
    for _ in range(4):
        c = dcos(y)
        s = (1 - c * c).sqrt()
        if s == 0:
            break
        y = y + (c - D(x)) / s
    return y


def ulps(got, exact):
    """|got - exact| in units of the last place of the double nearest exact."""
    e = float(exact)
    ulp = math.ulp(e) if e != 0 else math.ulp(0.0)
    return float(abs(D(got) - exact) / D(ulp))


CASES = {
    0: ("log", lambda r: np.exp(r.uniform(-40, 40, 600)), lambda x: D(x).ln(), 1.0),
    1: ("exp", lambda r: r.uniform(-40, 40, 600), lambda x: D(x).exp(), 1.0),
    2: ("pow(x, 1/3)", lambda r: np.concatenate([np.exp(r.uniform(-30, 30, 400)), np.exp(r.uniform(-744, 709, 400))]),
        lambda x: (D(1.0 / 3.0) * D(x).ln()).exp(), 1.0),
    3: ("acos", lambda r: r.uniform(-1, 1, 400), dacos, 2.0),
    4: ("cos", lambda r: r.uniform(0, 5.3, 400), dcos, 2.0),
}


@pytest.mark.parametrize("which", sorted(CASES))
def test_elementary_accuracy(which):
    name, gen, exact, bound = CASES[which]
    x = gen(np.random.default_rng(100 + which))
    if which == 3:
        x = np.concatenate([x, [-1.0, -0.5, 0.0, 0.5, 1.0, 0.9999999, -0.9999999]])
    twin, glibc = np.zeros_like(x), np.zeros_like(x)
    lib(True).oracle_elementary(which, len(x), p(x), p(twin))
    lib(False).oracle_elementary(which, len(x), p(x), p(glibc))
    err = np.array([ulps(t, exact(v)) for t, v in zip(twin, x)])
    assert err.max() <= bound, (name, err.max(), x[err.argmax()])
    same = float((twin == glibc).mean())
    # informational: the share of arguments on which the twin equals glibc
    print(f"{name}: max {err.max():.3f} ulp, equal to glibc on {100 * same:.1f} %")
    assert same > 0.5


def test_elementary_special_values():
    L = lib(True)
    x = np.array([0.0, math.inf, math.nan, -1.0, 1.0, 1e-310])
    y = np.zeros_like(x)
    L.oracle_elementary(0, len(x), p(x), p(y))
    assert y[0] == -math.inf and y[1] == math.inf and math.isnan(y[2]) and math.isnan(y[3]) and y[4] == 0.0
    assert abs(y[5] - math.log(1e-310)) <= 2 * math.ulp(math.log(1e-310))
    x = np.array([0.0, 800.0, -800.0, math.nan, 1.0])
    L.oracle_elementary(1, len(x), p(x), p(y[:5]))
    assert y[0] == 1.0 and y[1] == math.inf and y[2] == 0.0 and math.isnan(y[3]) and y[4] == math.e
