"""The library's SETFL reader and quintic-spline node derivatives
(mdl_setfl_read, mdl_spline5_node_derivatives; include/mdelast_setfl.h) against
what pins them:

- polynomials of degree <= 5 are reproduced exactly by the degree-5
  interpolating spline, so their node derivatives are known in closed form (a
  wrong knot choice or end condition fails this);
- the reference's own path, numpy.loadtxt parsing and SciPy's
  UnivariateSpline(k=5, s=0) (get_potential.py:49-147): node values bit-identical,
  derivatives to round-off amplified by the grid spacing;
- the smoke cell's golden values (benchmarks/expected/smoke.json) through the
  oracle fed by this reader.
"""

from __future__ import annotations

import numpy as np
import pytest

import cells
import harness as H
from mdelast_amd import _native as nat
from mdelast_amd import setfl
from mdelast_amd import units as U

FILES = ["Cu01.eam.alloy.txt", "ZnCu.eam.alloy", "Mishin_updated-Ni-Al-Co-2013.eam.alloy"]


@pytest.mark.parametrize("grid", ["uniform", "graded"])
@pytest.mark.parametrize("deg", range(6))
def test_spline_reproduces_polynomials(grid, deg):
    x = np.linspace(0.0, 3.0, 41)
    if grid == "graded":
        x = 3.0 * (x / 3.0) ** 1.5
    c = np.random.RandomState(deg).normal(size=deg + 1)
    d1, d2 = setfl.node_derivatives(x, np.polyval(c, x))
    e1 = np.polyval(np.polyder(c), x) if deg >= 1 else 0 * x
    e2 = np.polyval(np.polyder(c, 2), x) if deg >= 2 else 0 * x
    assert np.abs(d1 - e1).max() <= 1e-11 * max(1.0, np.abs(e1).max())
    assert np.abs(d2 - e2).max() <= 1e-9 * max(1.0, np.abs(e2).max())


def test_spline_is_not_a_natural_or_clamped_one():
    """x^6 is not a degree-5 spline: the interpolant's node derivatives differ
    from the exact ones, and they equal SciPy's (same knots, same spline)."""
    x = np.linspace(0.0, 1.0, 30)
    y = x ** 6
    d1, d2 = setfl.node_derivatives(x, y)
    s1, s2 = setfl.node_derivatives_scipy(x, y)
    assert np.abs(d1 - 6 * x ** 5).max() > 1e-9
    assert np.abs(d1 - s1).max() < 1e-11 and np.abs(d2 - s2).max() < 1e-9


@pytest.mark.parametrize("name", FILES)
def test_reader_matches_the_reference_path(name):
    a = setfl.read_setfl(name)
    b = setfl.read_setfl_python(name)
    assert (a.Nrho, a.Nr, a.drho, a.dr, a.rcut, a.rhomax) == \
        (b.Nrho, b.Nr, b.drho, b.dr, b.rcut, b.rhomax)
    assert np.array_equal(a.r_grid, b.r_grid) and np.array_equal(a.rho_grid, b.rho_grid)
    worst = {}
    for ea, eb in zip(a.elements, b.elements):
        assert (ea.name, ea.atomic_number, ea.atomic_weight, ea.a_lattice_angstrom) == \
            (eb.name, eb.atomic_number, eb.atomic_weight, eb.a_lattice_angstrom)
        assert np.array_equal(ea.F, eb.F) and np.array_equal(ea.rho, eb.rho)
        for k in ("dF", "d2F", "drho", "d2rho"):
            x, y = getattr(ea, k), getattr(eb, k)
            worst[k] = max(worst.get(k, 0.0), np.abs(x - y).max() / np.abs(y).max())
    assert len(a.pairs) == len(b.pairs)
    for pa, pb in zip(a.pairs, b.pairs):
        assert (pa.i, pa.j) == (pb.i, pb.j) and np.array_equal(pa.phi, pb.phi)
        for k in ("dphi", "d2phi"):
            x, y = getattr(pa, k), getattr(pb, k)
            worst[k] = max(worst.get(k, 0.0), np.abs(x - y).max() / np.abs(y).max())
    print(name, {k: f"{v:.1e}" for k, v in worst.items()})
    # first derivatives agree to round-off; second derivatives carry it over h^2
    for k in ("dF", "drho", "dphi"):
        assert worst[k] < 1e-10, k
    for k in ("d2F", "d2rho", "d2phi"):
        assert worst[k] < 1e-6, k


def test_smoke_golden_through_the_library_reader():
    import json
    import os
    data = setfl.read_setfl("ZnCu.eam.alloy")
    cell = dict(cells.smoke_cell(), data=data)
    o = H.oracle_md(cell, H.md_setup(cell, static=True))
    s = o.scalars()
    ex = json.load(open(os.path.join(cells.GOLDEN, "smoke_expected.json")))["checks"]
    got = {"static.static_epot_eV_per_atom": s["epot_atom"] / U.ev,
           "static.static_pressure_GPa": s["pressure"] / 1e10}
    for k, v in got.items():
        e = ex[k]["expected_by_interpolation"]["quintic"]
        assert abs(v / e - 1) < 1e-12, (k, v, e)


def test_reader_errors(tmp_path):
    with pytest.raises(nat.MdlError, match="cannot open"):
        setfl.read_setfl(str(tmp_path / "missing.eam.alloy"))
    good = open(setfl.resolve_path("ZnCu.eam.alloy")).read().splitlines()
    p = tmp_path / "cut.eam.alloy"
    p.write_text("\n".join(good[:200]) + "\n")
    with pytest.raises(nat.MdlError, match="ends inside a table"):
        setfl.read_setfl(str(p))
    with pytest.raises(nat.MdlError, match="n >= 6"):
        setfl.node_derivatives(np.arange(5.0), np.arange(5.0))
    with pytest.raises(nat.MdlError, match="increase strictly"):
        setfl.node_derivatives(np.array([0, 1, 2, 2, 3, 4, 5.0]), np.zeros(7))
