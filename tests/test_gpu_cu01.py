"""Pure Cu with the one-element loader (the reference's cu_300k example and the
CPU-runnable workload: 5^3 fcc cells = 500 atoms, a = 3.615 A, Cu01 EAM, FP64).

* static: a configuration displaced by 0.08 A, both interpolations, every
  static quantity against the oracle (1e-12; B-spline 1e-9), and the perfect
  lattice's Born tensor against Mishin et al. (2001), which also pins the oracle;
* dynamics: 40 NVT steps (chain of 3, 300 K, dt 1 fs) against the oracle,
  with the reference's default all-pairs loop and with the list (skin 1.2,
  rebuilt every 20 steps): every per-step observable within 1e-10;
* the product's loader: the Material of get_potential.get_EAM_file (node
  derivatives from the library, no SciPy) in the driver, against the oracle on
  the same node data (1e-12) and on SciPy's (the two derivative paths differ by
  at most 3e-8 relative in rho'');
* the caller: MolecularDynamics with that Material for a short NVT run.
"""

from __future__ import annotations

import numpy as np
import pytest

import harness as H
from mdelast_amd import get_potential, setfl
from mdelast_amd import units as U
from test_eam_file import CU01, cu01_cell

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_BSPLINE = 1e-9


@pytest.mark.parametrize("interp", ["quintic", "bspline"])
def test_static_displaced_vs_oracle(interp):
    cell = cu01_cell(disp=0.08, disp_seed=5)
    setup = H.md_setup(cell, static=True)
    o = H.oracle_md(cell, setup, interpolation=interp)
    g = H.gpu_md(cell, setup, interpolation=interp)
    errs = H.compare_static(o, g, TOL_FP64 if interp == "quintic" else TOL_BSPLINE)
    print(interp, {k: f"{v:.1e}" for k, v in errs.items()})


def test_perfect_lattice_born_tensor_is_mishin_2001():
    cell = cu01_cell()
    g = H.gpu_md(cell, H.md_setup(cell, static=True))
    s = H.gpu_scalars(g)
    CB = s["born_elasticity_tensor"] / 1e10
    assert abs(s["epot_atom"] / U.ev + 3.54) < 1e-4
    assert abs(np.mean(np.diag(CB)[:3]) - 169.9) < 0.1
    assert abs(np.mean([CB[0, 1], CB[0, 2], CB[1, 2]]) - 122.6) < 0.1
    assert abs(np.mean(np.diag(CB)[3:]) - 76.2) < 0.1


@pytest.mark.parametrize("neighbors", [False, True])
def test_nvt_trajectory_vs_oracle(neighbors):
    cell = cu01_cell()
    setup = H.md_setup(cell, T=300.0, M=3, seed=0)
    kw = dict(ensemble="NVT", dt=1e-15, neighbors_list=neighbors,
              r_skin=1.2 if neighbors else 1.0)
    worst, o, g = H.traj_compare(cell, setup, 40, rebuild_every=20, **kw)
    print(neighbors, {k: f"{v:.1e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v < 1e-10, (k, v)


def test_product_loader_on_the_device():
    m = get_potential.get_EAM_file(file_name=CU01)
    native = setfl.read_eam_file(CU01)
    cell = cu01_cell(disp=0.08, disp_seed=5, data=native)
    setup = H.md_setup(cell, static=True)
    from mdelast_amd import md_driver_module as mdm
    g = mdm.MDDriver(cell["pos"], setup["vel"], mdm.PotentialVector(m.potential),
                     mdm.EAMVector(m.eam), 0, 0, mdm.Thermostat(*setup["thermostat"]),
                     mdm.Barostat(*setup["barostat"], mdm.Thermostat(*setup["baro_thermostat"])),
                     mdm.Box(np.diag(cell["box"])), [1e-10], setup["Nf"],
                     [a / U.Navogadro for a in m.atomic_weight], [0] * len(cell["pos"]), 1e-15,
                     False, 1.0, 1e3)
    o = H.oracle_md(cell, setup)
    H.compare_static(o, g, TOL_FP64)
    scipy_cell = cu01_cell(disp=0.08, disp_seed=5)
    o2 = H.oracle_md(scipy_cell, setup)
    errs = H.compare_static(o2, g, 1.0)
    assert max(errs.values()) < 1e-7, errs


def test_molecular_dynamics_with_eam_file_material(tmp_path):
    from mdelast_amd.md_driver_module import MDDriver
    from mdelast_amd.molecular_dynamics import MolecularDynamics, get_EAM_file
    md = MolecularDynamics(
        work_dir=str(tmp_path), ensamble=MDDriver.NVT, material=get_EAM_file(file_name=CU01),
        N_lattice_side=5, dt_global=1e-15, tau_thermostat=5e-14, tau_barostat=5e-13,
        temperature_initial=300.0, temperature_thermostat=300.0,
        pressure_barostat=1e-10 * U.GPa, dump_cycles=10_000, plot_graphs=False,
        fatal_conservation=1e3, cycles_thermalization=1000, plot_freq=3000, cycles_end=4000,
        nchains_thermostat=3, neighbors=True, r_skin=1.2, cycles_neigbors=20, gr_stride=100,
        random_seed=0).run()
    s = md.series
    assert md.Nparticles == 500
    assert abs(s.temperature_kin.mean - 300.0) < 15.0
    assert abs(1.0 / s.temperature_conf_inverse.mean - 300.0) < 20.0
    assert float(np.max(s.etot_cons_err.data)) < 1e-4
    # near the zero-pressure lattice constant at 300 K the pressure is small
    assert abs(s.pressure.mean / U.GPa) < 2.0
