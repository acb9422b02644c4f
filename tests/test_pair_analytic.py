"""Analytic pair potentials (Morse, Lennard-Jones; potential.hpp:29-75) on the CPU.

The oracle's restatement is pinned by what the mathematics fixes, not by
retyping its formula: the Lennard-Jones minimum at 2^(1/6) sigma with depth
epsilon + vcut and curvature 72 epsilon / r_min^2, V(sigma) = -vcut, V(rcut) = 0;
the Morse minimum at r0 with depth D and curvature 2 alpha^2 D, V = -3D/4 at
r0 + ln 2 / alpha; V' and V'' against central differences of V; and, for the MD
driver, the total pair energy and forces of a small cell against a numpy sum
over all minimum-image pairs and a finite difference of that sum.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import cells
import harness as H
from mdelast_amd import units as U

EPS_AR, SIG_AR = 0.0104 * U.ev, 3.405 * U.angstrom
D_CU, ALPHA_CU, R0_CU = 0.3429 * U.ev, 1.3588 / U.angstrom, 2.866 * U.angstrom


def _lj(rcut=3.5 * SIG_AR):
    import oracle as O
    return O.AnalyticPair(O.LJ, [EPS_AR, SIG_AR, rcut], rcut)


def _morse():
    import oracle as O
    rc = 10.0 / ALPHA_CU + R0_CU
    return O.AnalyticPair(O.MORSE, [D_CU, ALPHA_CU, R0_CU], rc)


def test_oracle_lj_closed_forms():
    t = _lj()
    rc = t.rcut
    vcut = 4 * EPS_AR * ((SIG_AR / rc) ** 12 - (SIG_AR / rc) ** 6)
    rmin = 2 ** (1 / 6) * SIG_AR
    assert t.value(rmin) == pytest.approx(-EPS_AR - vcut, rel=1e-14)
    assert abs(t.d1(rmin)) < 1e-13 * EPS_AR / SIG_AR
    assert t.d2(rmin) == pytest.approx(72 * EPS_AR / rmin ** 2, rel=1e-13)
    assert t.value(SIG_AR) == pytest.approx(-vcut, rel=1e-12)
    assert t.value(rc) == 0.0      # the shift is formed by the same products
    # V(r) > 0 inside sigma, < 0 outside; forces repulsive inside r_min
    assert t.value(0.9 * SIG_AR) > 0 > t.value(1.5 * SIG_AR)
    assert t.d1(0.95 * rmin) < 0 < t.d1(1.05 * rmin)


def test_oracle_morse_closed_forms():
    t = _morse()
    assert t.value(R0_CU) == pytest.approx(-D_CU, rel=1e-15)
    assert abs(t.d1(R0_CU)) < 1e-15 * D_CU * ALPHA_CU
    assert t.d2(R0_CU) == pytest.approx(2 * ALPHA_CU ** 2 * D_CU, rel=1e-14)
    assert t.value(R0_CU + math.log(2) / ALPHA_CU) == pytest.approx(-0.75 * D_CU, rel=1e-14)
    # large r: V -> -2 D exp(-alpha (r - r0))
    r = R0_CU + 30 / ALPHA_CU
    assert t.value(r) == pytest.approx(-2 * D_CU * math.exp(-30), rel=1e-12)


@pytest.mark.parametrize("which", ["lj", "morse"])
def test_oracle_analytic_derivatives_by_central_differences(which):
    t = _lj() if which == "lj" else _morse()
    lo = 0.9 * SIG_AR if which == "lj" else 0.8 * R0_CU
    for r in np.linspace(lo, 0.95 * t.rcut, 23):
        h = 2e-6 * r
        d1 = (t.value(r + h) - t.value(r - h)) / (2 * h)
        d2 = (t.d1(r + h) - t.d1(r - h)) / (2 * h)
        s1 = abs(t.d1(r)) + abs(t.value(r)) / r
        s2 = abs(t.d2(r)) + abs(t.d1(r)) / r
        assert abs(d1 - t.d1(r)) < 1e-6 * s1, r
        assert abs(d2 - t.d2(r)) < 1e-6 * s2, r


@pytest.mark.parametrize("which", ["lj", "morse"])
def test_host_calculate_matches_oracle(which):
    """md_driver_module.Morse / LennardJones.calculate (host, as the reference's
    bound calculate) against the oracle, and zero beyond the cut-off."""
    from mdelast_amd import md_driver_module as mdm
    if which == "lj":
        rc = 3.5 * SIG_AR
        p, t = mdm.LennardJones(rc, EPS_AR, SIG_AR), _lj(rc)
        assert p.vcut == pytest.approx(4 * EPS_AR * ((SIG_AR / rc) ** 12 - (SIG_AR / rc) ** 6),
                                       rel=1e-13)
        rs = np.linspace(0.85 * SIG_AR, rc, 50)
    else:
        p, t = mdm.Morse(10.0 / ALPHA_CU + R0_CU, D_CU, ALPHA_CU, R0_CU), _morse()
        rs = np.linspace(0.7 * R0_CU, p.rcut, 50)
    for r in rs:
        v = p.calculate(r)
        assert v.V == pytest.approx(t.value(r), rel=1e-14, abs=1e-30)
        assert v.dVdr == pytest.approx(t.d1(r), rel=1e-14, abs=1e-30)
        assert v.d2Vdr2 == pytest.approx(t.d2(r), rel=1e-14, abs=1e-30)
    z = p.calculate(p.rcut * 1.0001)
    assert (z.V, z.dVdr, z.d2Vdr2) == (0.0, 0.0, 0.0)


def _lj_spec(rc):
    return dict(form="lj", rcut=rc, params=[(EPS_AR, SIG_AR)])


def _brute_energy(pos, L, rc, t):
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    iu = np.triu_indices(len(pos), 1)
    rr = r[iu]
    rr = rr[rr <= rc]
    return sum(t.value(x) for x in rr)


@pytest.mark.parametrize("neighbors", [False, True])
def test_oracle_md_lj_energy_and_forces_brute_force(neighbors):
    """The oracle MD with an analytic pair potential and no EAM: total energy =
    sum over minimum-image pairs inside r_c; force on one atom = -dE/dx by a
    central difference of that sum."""
    rc = 2.5 * SIG_AR
    cell = cells.pair_cell(3, 5.4, _lj_spec(rc), [39.948], disp=0.15, disp_seed=3)
    setup = H.md_setup(cell, static=True)
    o = H.oracle_md(cell, setup, neighbors_list=neighbors, r_skin=1.2)
    t = _lj(rc)
    pos, L = cell["pos"], cell["box"][0]
    N = len(pos)
    e_brute = _brute_energy(pos, L, rc, t)
    so = o.scalars()
    assert so["epot_atom"] * N == pytest.approx(e_brute, rel=1e-12)
    assert so["epot_eam_atom"] == 0.0
    F = o.atoms("forces")
    for i, k in [(0, 0), (7, 1), (30, 2)]:
        h = 1e-5 * U.angstrom
        p1, p2 = pos.copy(), pos.copy()
        p1[i, k] += h
        p2[i, k] -= h
        fd = -(_brute_energy(p1, L, rc, t) - _brute_energy(p2, L, rc, t)) / (2 * h)
        assert F[i, k] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(F).max())
    # Newton's third law over the whole cell
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).max()


def test_material_helpers():
    from mdelast_amd import molecular_dynamics as md, get_potential as gp, \
        md_driver_module as mdm
    ar = md.get_Ar_LJ()
    p = ar.potential[0]
    assert isinstance(p, mdm.LennardJones) and ar.eam is None and ar.a_lattice is None
    assert p.rcut == pytest.approx(3.5 * SIG_AR) and p.epsilon == pytest.approx(EPS_AR)
    assert gp.get_Ar_LJ(rcut_fac=2.5).potential[0].rcut == pytest.approx(2.5 * SIG_AR)
    cu = md.get_Cu_Morse(a_factor=1.01)
    m = cu.potential[0]
    assert isinstance(m, mdm.Morse) and m.rcut == pytest.approx(10 / ALPHA_CU + R0_CU)
    assert cu.a_lattice == pytest.approx(1.01 * 3.615 * U.angstrom)
    # unavailable in the reference v1.2 (it raises before building the table)
    with pytest.raises(AttributeError):
        md.get_Cu_Table_Morse()


def test_analytic_handle_parameters():
    """The device parameter record of an analytic handle (no GPU needed)."""
    from mdelast_amd import md_driver_module as mdm
    rc = 3.0 * SIG_AR
    lj = mdm.LennardJones(rc, EPS_AR, SIG_AR)
    rec = lj.table.coefficients()
    assert rec.shape == (1, 4)
    assert rec[0, :2].tolist() == [EPS_AR, SIG_AR] and rec[0, 2] == lj.vcut
    mo = mdm.Morse(1e-7, D_CU, ALPHA_CU, R0_CU)
    assert mo.table.coefficients()[0, :3].tolist() == [D_CU, ALPHA_CU, R0_CU]
