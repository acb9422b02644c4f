"""Known-answer tests of the quintic and B-spline tables (CPU).

The seven properties of the reference's native table test
(tools/test_poly_table.cpp:87-252) on the oracle's QuinticTable, the
product's coefficient records compared with the oracle's bit for bit, and the
interpolation checks of tests/test_tables.py:72-112.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import cells
import oracle as O
from mdelast_amd import md_driver_module as mdm
from mdelast_amd import units as U

EPS = 2.22e-16


def make(fun, h, n, kind=O.QUINTIC):
    x = np.arange(n + 1) * h
    f, f1, f2 = (np.array([g(v) for v in x]) for g in fun)
    return O.Table(kind, h, f, f1, f2)


def poly_fun(deg):
    def f(x):
        s = 0.0
        for m in range(deg, -1, -1):
            s = s * (x - 3) + 1.0 / (m + 1)
        return s

    def f1(x):
        s = 0.0
        for m in range(deg, 0, -1):
            s = s * (x - 3) + m / (m + 1.0)
        return s

    def f2(x):
        s = 0.0
        for m in range(deg, 1, -1):
            s = s * (x - 3) + m * (m - 1) / (m + 1.0)
        return s
    return f, f1, f2


SMOOTH = (lambda x: math.sin(3 * x) * math.exp(-x / 2) + 1 / (1 + x * x),
          lambda x: math.exp(-x / 2) * (3 * math.cos(3 * x) - 0.5 * math.sin(3 * x))
          - 2 * x / ((1 + x * x) ** 2),
          lambda x: math.exp(-x / 2) * (-8.75 * math.sin(3 * x) - 3 * math.cos(3 * x))
          + (6 * x * x - 2) / (1 + x * x) ** 3)


def max_err(T, fun, lo, hi, npts=4001):
    xs = np.linspace(lo, hi, npts)
    ref = np.array([[g(x) for g in fun] for x in xs])
    got = np.array([[T.value(x), T.d1(x), T.d2(x)] for x in xs])
    sc = np.maximum(np.abs(ref).max(0), 1e-300)
    return np.abs(got - ref).max(0) / sc


@pytest.mark.parametrize("h", [0.1, 0.0034])
@pytest.mark.parametrize("deg", range(6))
def test_polynomials_reproduced(h, deg):
    n = int(round(6.8 / h))
    fun = poly_fun(deg)
    T = make(fun, h, n)
    e = max_err(T, fun, 0.0, n * h)
    xs = np.linspace(0, n * h, 2001)
    fmax = max(abs(fun[0](x)) for x in xs)
    s1 = max(abs(fun[1](x)) for x in xs) or 1.0
    s2 = max(abs(fun[2](x)) for x in xs) or 1e-300
    df = 4 * EPS * fmax   # conditioning limit of any interpolant through rounded nodes
    tol = [1e-14, max(1e-13, 2 * 1.875 * df / h / s1), max(1e-13, 2 * 5.774 * df / h / h / s2)]
    for k in range(min(deg, 2) + 1):
        assert e[k] <= tol[k], (k, e[k], tol[k])


def test_convergence_orders():
    prev = None
    for n in (20, 40, 80, 160):
        e = max_err(make(SMOOTH, 4.0 / n, n), SMOOTH, 0.0, 4.0, 8001)
        if prev is not None:
            orders = np.log2(prev / e)
            assert np.all(np.abs(orders - np.array([6, 5, 4])) < 0.35), (n, orders)
        prev = e


def test_node_values():
    n, h = 2000, 0.0034
    T = make(SMOOTH, h, n)
    e = [0.0, 0.0, 0.0]
    for i in range(n + 1):
        x = i * h
        e[0] = max(e[0], abs(T.value(x) - SMOOTH[0](x)) / 2)
        e[1] = max(e[1], abs(T.d1(x) - SMOOTH[1](x)) / 4)
        e[2] = max(e[2], abs(T.d2(x) - SMOOTH[2](x)) / 12)
    assert e[0] < 1e-14 and e[1] < 1e-12 and e[2] < 1e-10


def _poly_deriv(c, k, t, h):
    """k-th derivative (in x) of the interval polynomial with coefficients c at t."""
    s = 0.0
    for m in range(5, k - 1, -1):
        fac = 1.0
        for j in range(k):
            fac *= (m - j)
        s = s * t + fac * c[m]
    return s / h ** k


@pytest.mark.parametrize("random_data", [False, True])
def test_c2_continuity(random_data):
    n, h = 400, 4.0 / 400
    if random_data:
        rng = np.random.RandomState(7)
        T = O.Table(O.QUINTIC, h, rng.uniform(-1, 1, n + 1), rng.uniform(-1, 1, n + 1) / h,
                    rng.uniform(-1, 1, n + 1) / (h * h))
    else:
        T = make(SMOOTH, h, n)
    c = T.coef()
    for k in range(3):
        L = np.array([_poly_deriv(c[i - 1], k, 1.0, h) for i in range(1, n)])
        R = np.array([_poly_deriv(c[i], k, 0.0, h) for i in range(1, n)])
        rel = np.abs(L - R).max() / max(np.abs(L).max(), np.abs(R).max())
        assert rel < (1e-12 if k == 2 else 1e-13), (k, rel)
    e, s = np.zeros(3), np.zeros(3)
    for i in range(1, n):
        x = i * h
        xl, xr = np.nextafter(x, -1.0), np.nextafter(x, 10.0)
        lv = np.array([T.value(xl), T.d1(xl), T.d2(xl)])
        rv = np.array([T.value(xr), T.d1(xr), T.d2(xr)])
        e = np.maximum(e, np.abs(lv - rv))
        s = np.maximum(s, np.abs(lv))
    # x moves by 2 ulp across the node, so f'' legitimately changes by f''' * 2 ulp;
    # with random data (|f'''| ~ 1/h^3) that alone is ~1e-11 of max|f''|
    assert np.all(e / s < (5e-11 if random_data else 1e-11)), e / s


def test_index_edge_cases():
    n, h = 2000, 0.0034
    xmax = n * h
    T = make(SMOOTH, h, n)
    assert abs(T.value(0.0) - SMOOTH[0](0.0)) < 1e-15
    assert T.value(-0.0) == T.value(0.0)
    assert abs(T.value(xmax) - SMOOTH[0](xmax)) < 1e-14
    assert abs(T.value(np.nextafter(xmax, 0)) - T.value(np.nextafter(xmax, 1e9))) < 1e-14
    assert abs(T.d2(np.nextafter(0.0, -1.0)) - T.d2(0.0)) < 1e-10
    c = T.coef()
    xe = xmax + 0.5 * h   # end polynomial continued
    assert abs(T.value(xe) - _poly_deriv(c[n - 1], 0, 1.5, h)) < 1e-14
    for i in range(0, n + 1, 7):
        for x in (np.nextafter(i * h, -1.0), i * h, np.nextafter(i * h, 1e9)):
            assert all(np.isfinite([T.value(x), T.d1(x), T.d2(x)]))
    for x in (-1e300, -1e10, 1e10, 3e9 * h, 1e300):
        T.value(x), T.d1(x), T.d2(x)   # no out-of-range read (the index is clamped)
    for x in (math.inf, -math.inf, math.nan):
        assert not np.isfinite(T.value(x))


def test_two_node_table():
    P = (lambda x: 1 + x * (2 + x * (-1 + x * (0.5 + x * (0.25 - 0.125 * x)))),
         lambda x: 2 + x * (-2 + x * (1.5 + x * (1 - 0.625 * x))),
         lambda x: -2 + x * (3 + x * (3 - 2.5 * x)))
    T = make(P, 1.0, 1)
    assert max_err(T, P, 0.0, 1.0).max() < 1e-14


# --- B-spline tables ------------------------------------------------------------
def test_bspline_reproduces_cubics():
    # the five-point end slopes are exact for polynomials of degree <= 4, so the
    # cubic B-spline interpolant of a cubic is the cubic itself
    P = lambda x: 0.3 - 1.2 * x + 0.7 * x * x - 0.05 * x ** 3
    n, h = 60, 0.1
    x = np.arange(n + 1) * h
    f = P(x)
    T = O.Table(O.BSPLINE, h, f, f, f)
    xs = np.linspace(0, n * h, 997)
    err = max(abs(T.value(v) - P(v)) for v in xs) / np.abs(f).max()
    assert err < 1e-12


def test_bspline_vs_quintic_on_zncu():
    # tests/test_tables.py:97-107: agreement of the two interpolations
    data = cells.setfl_data("ZnCu.eam.alloy")
    q = O.tables_from_setfl(data, "quintic")
    b = O.tables_from_setfl(data, "bspline")
    r = np.linspace(1.8, 6.7, 997) * U.angstrom
    for k in range(3):
        vq = np.array([[q[0][k].value(x), q[0][k].d1(x), q[0][k].d2(x)] for x in r])
        vb = np.array([[b[0][k].value(x), b[0][k].d1(x), b[0][k].d2(x)] for x in r])
        rel = np.abs(vq - vb).max(0) / np.abs(vq).max(0)
        assert np.all(rel < [1e-8, 1e-6, 1e-4]), (k, rel)
    for e in range(2):
        for fn in ("value", "d1", "d2"):
            a = np.array([getattr(q[1][e], fn)(x) for x in r])
            c = np.array([getattr(b[1][e], fn)(x) for x in r])
            assert np.abs(a - c).max() / np.abs(a).max() < 1e-4


def test_nan_argument():
    data = cells.setfl_data("ZnCu.eam.alloy")
    for interp in ("quintic", "bspline"):
        p, r, F = O.tables_from_setfl(data, interp)
        assert math.isnan(p[0].value(float("nan")))
        assert math.isnan(F[0].value(float("nan")))


# --- product coefficient records vs the oracle (host side, no GPU) ---------------
@pytest.mark.parametrize("potential", ["ZnCu.eam.alloy", "Cu01.eam.alloy.txt",
                                       "Mishin_updated-Ni-Al-Co-2013.eam.alloy"])
def test_product_quintic_records_bit_identical(potential):
    data = cells.setfl_data(potential)
    ot = O.tables_from_setfl(data, "quintic")
    for k, p in enumerate(data.pairs):
        t = mdm._Table("quintic", data.dr, p.phi, p.dphi, p.d2phi)
        assert np.array_equal(t.coefficients(), ot[0][k].coef())
    for k, e in enumerate(data.elements):
        t = mdm._Table("quintic", data.drho, e.F, e.dF, e.d2F)
        assert np.array_equal(t.coefficients(), ot[2][k].coef())


def test_product_bspline_records_match_oracle():
    data = cells.setfl_data("ZnCu.eam.alloy")
    ot = O.tables_from_setfl(data, "bspline")
    h = data.dr
    xs = np.linspace(0.5, data.rcut / U.angstrom, 3001) * U.angstrom
    for k, p in enumerate(data.pairs):
        rec = mdm._Table("bspline", h, p.phi, p.dphi, p.d2phi).coefficients()
        i = np.minimum((xs / h).astype(int), len(rec) - 1)
        t = xs / h - i
        for s, fn in enumerate(("value", "d1", "d2")):
            c = rec[i, 4 * s:4 * s + 4]
            got = c[:, 0] + t * (c[:, 1] + t * (c[:, 2] + t * c[:, 3]))
            ref = np.array([getattr(ot[0][k], fn)(x) for x in xs])
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
