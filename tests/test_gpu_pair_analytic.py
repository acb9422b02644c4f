"""Analytic pair potentials on the GPU (Morse, Lennard-Jones without EAM, FP64)
against the oracle's restatement of potential.hpp:29-75.

Tolerances as the tabulated FP64 path: 1e-12 scale-relative for the static
quantities (the device's exp differs from the host's by an ulp, divisions and
contractions by round-off); 1e-10 over short trajectories.
"""

from __future__ import annotations

import numpy as np
import pytest

import cells
import harness as H
from mdelast_amd import md_driver_module as mdm
from mdelast_amd import units as U

pytestmark = pytest.mark.gpu

EPS_AR, SIG_AR = 0.0104 * U.ev, 3.405 * U.angstrom
D_CU, ALPHA_CU, R0_CU = 0.3429 * U.ev, 1.3588 / U.angstrom, 2.866 * U.angstrom
RC_MORSE = 10.0 / ALPHA_CU + R0_CU


def _lj(rc=3.5 * SIG_AR):
    return dict(form="lj", rcut=rc, params=[(EPS_AR, SIG_AR)])


def _morse():
    return dict(form="morse", rcut=RC_MORSE, params=[(D_CU, ALPHA_CU, R0_CU)])


def _lj2():
    """Two species (Ar, Kr), Lorentz-Berthelot cross terms, one cut-off."""
    ek, sk = 0.014 * U.ev, 3.65 * U.angstrom
    return dict(form="lj", rcut=3.5 * sk,
                params=[(EPS_AR, SIG_AR), (np.sqrt(EPS_AR * ek), 0.5 * (SIG_AR + sk)), (ek, sk)])


CASES = {
    # all-pairs list (box < 3 list radii), solid argon, displaced
    "lj_solid_n5": lambda: cells.pair_cell(5, 5.26, _lj(), [39.948], disp=0.1, disp_seed=1),
    # cell-list path (box >= 3 list radii)
    "lj_solid_n9": lambda: cells.pair_cell(9, 5.26, _lj(), [39.948], disp=0.1, disp_seed=2),
    # the reference's LJ gas example: a = 8 A, 500 atoms
    "lj_gas_n5": lambda: cells.pair_cell(5, 8.0, _lj(), [39.948], disp=0.5, disp_seed=3),
    "morse_cu_n7": lambda: cells.pair_cell(7, 3.615, _morse(), [63.546], disp=0.08, disp_seed=4),
    "lj_two_species": lambda: cells.pair_cell(6, 5.5, _lj2(), [39.948, 83.798],
                                              type_fractions=[0.7, 0.3], seed=5, disp=0.1,
                                              disp_seed=5),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("neighbors", [False, True])
def test_static_vs_oracle(case, neighbors):
    cell = CASES[case]()
    setup = H.md_setup(cell, static=True)
    o = H.oracle_md(cell, setup, neighbors_list=neighbors, r_skin=1.2)
    g = H.gpu_md(cell, setup, neighbors_list=neighbors, r_skin=1.2)
    errs = H.compare_static(o, g, 1e-12)
    print(case, {k: f"{v:.1e}" for k, v in errs.items()})
    assert g.epot_eam_atom == 0.0


@pytest.mark.parametrize("case,ensemble", [("lj_gas_n5", "NVT"), ("lj_solid_n5", "NVE"),
                                           ("lj_solid_n5", "NPT"), ("morse_cu_n7", "NVT")])
def test_trajectory_vs_oracle(case, ensemble):
    cell = CASES[case]()
    setup = H.md_setup(cell, T=300.0, M=3, Mb=3, seed=0)
    dt = 5e-15 if case == "lj_gas_n5" else 2e-15
    kw = dict(ensemble=ensemble, dt=dt, neighbors_list=True, r_skin=1.5)
    o = H.oracle_md(cell, setup, **kw)
    g = H.gpu_md(cell, setup, **kw)
    worst = {}
    for step in range(40):
        if step % 10 == 0:
            o.calculate_neighbors()
            g.calculate_neighbors()
        o.integrate()
        g.integrate()
        so, sg = o.scalars(), H.gpu_scalars(g)
        for k in ("etot_cons", "temperature_kin", "epot_atom", "ekin_atom"):
            worst[k] = max(worst.get(k, 0), abs(sg[k] / so[k] - 1))
        xo, xg = o.atoms("coordinates"), g.coordinates_array()
        worst["coords"] = max(worst.get("coords", 0), float(np.abs(xo - xg).max()))
        vo, vg = o.atoms("velocities"), g.velocities_array()
        worst["vel"] = max(worst.get("vel", 0), float(np.abs(vo - vg).max() / np.abs(vo).max()))
        if ensemble == "NPT":
            worst["box"] = max(worst.get("box", 0), abs(sg["box0"] / so["box0"] - 1))
        assert sg["cycle"] == so["cycle"]
    print(case, ensemble, {k: f"{v:.1e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v < 1e-10, (k, v)


@pytest.mark.parametrize("which", ["lj", "morse"])
def test_table_eval_of_analytic_handles(which):
    """mdl_table_eval of an analytic handle (device formula) against the host
    calculate."""
    if which == "lj":
        p = mdm.LennardJones(3.5 * SIG_AR, EPS_AR, SIG_AR)
        rs = np.linspace(0.85 * SIG_AR, p.rcut, 301)
    else:
        p = mdm.Morse(RC_MORSE, D_CU, ALPHA_CU, R0_CU)
        rs = np.linspace(0.7 * R0_CU, p.rcut, 301)
    v, d1, d2 = p.table.eval(rs)
    host = np.array([[q.V, q.dVdr, q.d2Vdr2] for q in map(p.calculate, rs)])
    for k, a in enumerate((v, d1, d2)):
        sc = np.abs(host[:, k]).max()
        assert np.abs(a - host[:, k]).max() < 1e-14 * sc, k


def test_refused_configurations():
    cell = CASES["lj_solid_n5"]()
    setup = H.md_setup(cell, static=True)
    with pytest.raises(mdm.MDDriverError):
        H.gpu_md(cell, setup, precision="mixed")
    # an analytic pair potential beside EAM tables
    from mdelast_amd import get_potential
    pots, eams = get_potential.get_EAM_alloy_file("Cu01.eam.alloy.txt")[1:]
    lj = mdm.LennardJones(3.5 * SIG_AR, EPS_AR, SIG_AR)
    th = mdm.Thermostat(300.0, [1e-30], [0.0], 3.0)
    ba = mdm.Barostat(0.0, 1e-30, 0.0, mdm.Thermostat(300.0, [1e-30], [0.0], 1.0))
    with pytest.raises(mdm.MDDriverError, match="without EAM"):
        mdm.MDDriver(cell["pos"], setup["vel"], mdm.PotentialVector([lj]), mdm.EAMVector(eams),
                     0, 0, th, ba, mdm.Box(np.diag(cell["box"])), [1e-8], 3.0 * len(cell["pos"]),
                     [63.546 / U.Navogadro], [0] * len(cell["pos"]), 1e-15, False, 1.2, 1e3)


def test_ar_lj_gas_example(tmp_path, monkeypatch):
    """The reference's examples/ar_lj_gas/run.py set-up (500 atoms, a = 8 A, NVT
    at 300 K, dt = 5 fs, neighbour list every 10 steps with skin 1.5, g(r) every
    50 steps), shortened to 600 + 600 steps: the kinetic temperature averages
    near the thermostat's, the configurational one is finite and positive, and
    g(r) of the gas tends to 1 at large r."""
    from mdelast_amd import molecular_dynamics as md
    monkeypatch.setattr(mdm, "VERBOSE", False)
    run = md.MolecularDynamics(
        work_dir=str(tmp_path / "ar"), ensamble=mdm.MDDriver.NVT, material=md.get_Ar_LJ(),
        N_lattice_side=5, a_lattice=8.0 * U.angstrom, dt_global=5e-15, tau_thermostat=1e-13,
        tau_barostat=5e-13, temperature_initial=300.0, temperature_thermostat=300.0,
        pressure_barostat=0.0, dump_cycles=100000, slurm_plt=False, show_graphs=False,
        plot_graphs=True, fatal_conservation=1e3, cycles_thermalization=600, plot_freq=600,
        cycles_end=1200, neighbors=True, r_skin=1.5, cycles_neigbors=10, gr_stride=50,
        random_seed=0).run()
    T = run.series.temperature_kin.mean
    print("T_kin", T, "1/T_conf", run.series.temperature_conf_inverse.mean,
          "P GPa", run.series.pressure.mean / U.GPa)
    assert 240.0 < T < 360.0
    assert run.series.temperature_conf_inverse.mean > 0
    gr = run.gr_sum / max(run.gr_samples, 1)
    tail = gr[(run.rg > 12 * U.angstrom) & (run.rg < 19 * U.angstrom)]
    assert abs(tail.mean() - 1.0) < 0.1
