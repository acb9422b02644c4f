"""The integrator's exponential operator on the GPU against extended precision.

numerics.hpp:11-27 of the reference: exp(-x) and op(x) = (1 - exp(-x))/x, which
every Nose-Hoover chain element and the barostat take (thermostat.cpp:40-66,
md_driver.cpp:462-499). Three device forms: the barostat's (a Taylor series of
op for |x| <= 1/4, else a range-reduced exponential and a Newton reciprocal),
the chain sweeps' wide form (one branch-free formula for any x: op from the
series when the reduction leaves x as it is, else from expm1 = 2^k r p(r) +
2^k - 1) and their series form (|x| <= 1/4). Each is checked against numpy's
80-bit long double (exp, expm1) over the whole range the chains reach, so a
wrong coefficient, a dropped term or a wrong reduction constant shows as an
error of many ulp.
"""

from __future__ import annotations

import numpy as np
import pytest

from mdelast_amd import _native as nat

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
SERIES = 0.25   # the series form's range


def _points():
    rs = np.random.RandomState(7)
    small = rs.uniform(-SERIES, SERIES, 20000)
    tiny = rs.uniform(-1e-3, 1e-3, 2000)
    mid = rs.uniform(SERIES, 4.0, 20000) * rs.choice([-1.0, 1.0], 20000)
    big = np.exp(rs.uniform(np.log(4.0), np.log(700.0), 20000)) * rs.choice([-1.0, 1.0], 20000)
    edges = np.array([SERIES, -SERIES, np.nextafter(SERIES, 1.0), np.nextafter(-SERIES, -1.0),
                      0.5 * np.log(2.0), -0.5 * np.log(2.0), np.nextafter(0.5 * np.log(2.0), 1.0),
                      1e-300, -1e-300, 700.0, -700.0])
    return np.concatenate([small, tiny, mid, big, edges])


def _truth(x):
    xl = x.astype(np.longdouble)
    return np.exp(-xl), -np.expm1(-xl) / xl


@pytest.mark.parametrize("form", [0, 1])
def test_expx_against_long_double(form):
    x = _points()
    e, op = nat.expx_eval(x, form)
    e_true, op_true = _truth(x)
    rel_e = np.abs((e - e_true) / e_true).astype(np.float64)
    assert rel_e.max() <= 2.5 * EPS, (x[rel_e.argmax()], rel_e.max())
    err = np.abs(op - op_true).astype(np.float64)
    small = np.abs(x) <= SERIES
    assert (err[small] <= 4.0 * EPS * np.abs(op_true[small])).all()
    if form == 0:
        # beyond 1/4 the product (1 - e) * (1/x) carries e's rounding amplified
        # by e / (1 - e) (the reference's quotient carries the same), plus the
        # reciprocal and the product
        bound = EPS * (4.0 * np.abs(op_true) + 2.0 * e_true / np.abs(x.astype(np.longdouble)))
    else:
        # the wide form does not cancel: a few ulp everywhere
        bound = 6.0 * EPS * np.abs(op_true)
    bound = bound.astype(np.float64)
    assert (err <= bound).all(), (x[(err / bound).argmax()], float((err / bound).max()))


def test_series_form_in_its_range():
    x = np.concatenate([_points()[:22000], [0.0, SERIES, -SERIES]])
    e, op = nat.expx_eval(x, 2)
    e_true, op_true = _truth(np.where(x == 0.0, 1e-300, x))
    assert (np.abs(e - e_true) <= 2.5 * EPS * e_true).all()
    assert (np.abs(op - op_true) <= 4.0 * EPS * np.abs(op_true)).all()


def test_wide_form_at_zero_and_extremes():
    e, op = nat.expx_eval(np.array([0.0, -0.0, 800.0, -800.0]), 1)
    assert e[0] == 1.0 and op[0] == 1.0 and e[1] == 1.0 and op[1] == 1.0
    assert e[2] == 0.0 and np.isinf(e[3])
    assert abs(op[2] - 1.0 / 800.0) <= 4 * EPS / 800.0
    e, op = nat.expx_eval(np.array([800.0, -800.0]), 0)
    assert e[0] == 0.0 and np.isinf(e[1])
    assert abs(op[0] - 1.0 / 800.0) <= 4 * EPS / 800.0
