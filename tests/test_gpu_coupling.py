"""Configurational and average thermostat coupling (md_driver.cpp:512-531 of
the reference, bound at md_driver_module.cpp:186-191) on the device against the
oracle, whose coupling branches are pinned on the CPU in
test_thermostat_coupling.py. 40 steps of the 500-atom Zn06 cell, chain of 3,
NVT and NPT, with the list: every per-step observable, the coordinates,
velocities and chain momenta within 1e-10."""

from __future__ import annotations

import pytest

import cells
import harness as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ensemble", ["NVT", "NPT"])
@pytest.mark.parametrize("coupling", [1, 2])
def test_coupling_trajectory_vs_oracle(coupling, ensemble):
    cell = cells.zn06_cell(5)
    setup = H.md_setup(cell, T=300.0, M=3, Mb=3, seed=0)
    kw = dict(ensemble=ensemble, dt=2e-15, neighbors_list=True, r_skin=1.2, coupling=coupling)
    worst, o, g = H.traj_compare(cell, setup, 40, rebuild_every=20, **kw)
    print(coupling, ensemble, {k: f"{v:.1e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v < 1e-10, (k, v)
    # the coupling reaches the device: a kinetic run's chain differs
    gk = H.gpu_md(cell, setup, **dict(kw, coupling=0))
    gk.run(40, 1, 20)
    assert abs(gk.thermostat.p_eta[0] / g.thermostat.p_eta[0] - 1) > 1e-6
