"""The large workloads at their own sizes.

* 108,000-atom Ni-Al-Co (the throughput workload of bench.py: 30^3 fcc cells,
  Ni 75 / Al 20 / Co 5 at.%, a = 3.56 A, the library's SETFL reader), displaced
  by 0.07 A: every static quantity of the FP64 path against the oracle at 1e-12,
  of the mixed path at the mixed tolerances; then 200 mixed NVT steps at 600 K
  with the bench's settings, checked for momentum (the reference's FATAL bound
  is 1e-14; pair forces here are bit-antisymmetric), conservation, temperature
  and against the same 200 steps on the FP64 path;
* the 64-replica table (replicas.batched_replica_specs: edge 5-20 cells,
  500-32,000 atoms, 602,740 in all, Cu or Cu-Zn up to 10 at.% Zn, 100-900 K),
  20 batched steps in one launch per pass: every replica equal bit for bit to
  its own launches, property checks on all 64, and the smallest and largest
  replicas against the oracle (20 steps from the same state, then the static
  quantities of the state the batched run reached, mixed tolerances).
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
import harness as H
from mdelast_amd import get_potential, replicas as R, setfl
from mdelast_amd import units as U
from test_gpu_mixed import TOL as MIXED_TOL

pytestmark = pytest.mark.gpu


def _check_mixed(o, g):
    errs = H.compare_static(o, g, 1.0)
    print({k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if v > MIXED_TOL[k]}
    assert not bad, bad


@pytest.fixture(scope="module")
def nialco108k():
    import oracle as O
    w = bench.build_workload()
    rs = np.random.RandomState(11)
    L = w["box"]
    pos = w["pos"] + rs.normal(scale=0.07, size=w["pos"].shape) * U.angstrom
    pos = np.mod(pos + L / 2, L) - L / 2
    cell = dict(data=w["data"], pos=pos, types=w["types"].astype(int), box=np.full(3, L))
    setup = H.md_setup(cell, static=True)
    print("oracle: O(N^2) list and forces of 108,000 atoms", flush=True)
    o = O.MD(positions=pos, velocities=setup["vel"], types=w["types"], mass=w["mass"],
             box=[L] * 3, data=w["data"], ensemble=O.NVE, thermostat=setup["thermostat"],
             barostat=setup["barostat"], baro_thermostat=setup["baro_thermostat"],
             Nf=setup["Nf"], dt=1e-15, neighbors_list=True, r_skin=1.0)
    return w, cell, setup, o


def test_nialco108k_static_fp64_vs_oracle(nialco108k):
    w, cell, setup, o = nialco108k
    g = H.gpu_md(cell, setup, neighbors_list=True, r_skin=1.2)
    errs = H.compare_static(o, g, 1e-12)
    print({k: f"{v:.1e}" for k, v in errs.items()})


def test_nialco108k_static_mixed_vs_oracle(nialco108k):
    w, cell, setup, o = nialco108k
    _check_mixed(o, H.gpu_md(cell, setup, neighbors_list=True, r_skin=1.2, precision="mixed"))


def test_nialco108k_mixed_nvt_properties():
    w = bench.build_workload()
    a, keep_a = bench.make_driver(w, "mixed")
    b, keep_b = bench.make_driver(w, "fp64")
    ra = a.run(200, 1, 20)
    rb = b.run(200, 1, 20)
    assert np.all(np.isfinite(ra))
    m0 = w["mass"][0]
    ratio = (ra[:, 21:24] ** 2).sum(1) / (2 * m0 * ra[:, 4])
    print("max |P|^2/(2 m0 K):", ratio.max(), " max E_cons err:", ra[:, 7].max(),
          " T:", ra[-1, 11])
    assert ratio.max() < 1e-20
    assert ra[:, 7].max() < 2e-5
    assert 200.0 < ra[:, 11].min() and ra[:, 11].max() < 650.0
    # the mixed run tracks the FP64 run (chaos amplifies the 1e-7 table
    # differences slowly over 200 fs)
    for k, tol in ((6, 1e-6), (11, 1e-3), (1, 1e-6)):
        assert np.max(np.abs(ra[:, k] / rb[:, k] - 1)) < tol, k
    assert np.max(np.abs(ra[:, 7] - rb[:, 7])) < 2e-7
    assert a.neighbor_list_builds >= 10
    # both mixed passes scan the mid list (pairs within r_c + 0.15 A), refiltered
    # at every full build and when an atom may have moved 0.075 A since
    nb, nm = a.neighbor_list_builds, a.mid_list_builds
    print("full builds", nb, "mid filterings", nm)
    assert nb < nm < 200


def test_nialco108k_mixed_long_run():
    """20 000 mixed NVT steps (20 ps) of the 108 000-atom workload, as the bench
    runs it (graph chunks of 20 steps, the list rebuilt every 20): momentum at
    FP64 round-off, the conserved energy flat after the lattice start's first
    steps, the temperature at the thermostat's 600 K after the first 2 ps."""
    w = bench.build_workload()
    d, keep = bench.make_driver(w, "mixed")
    d.graph_steps = 20
    m0 = w["mass"][0]
    err0 = None
    for chunk in range(10):
        rec = d.run(2000, 100, 20)
        assert np.all(np.isfinite(rec))
        ratio = (rec[:, 21:24] ** 2).sum(1) / (2 * m0 * rec[:, 4])
        err = rec[:, 7].max()
        err0 = err if err0 is None else err0
        print(f"steps {2000 * (chunk + 1)}: max |P|^2/(2 m0 K) {ratio.max():.2e}, "
              f"max E_cons err {err:.3e}, T {rec[-1, 11]:.1f} K", flush=True)
        assert ratio.max() < 1e-20
        assert err < 2e-5 and err < 1.05 * err0
        # the lattice start pours kinetic energy into the potential first (T dips
        # to about 420 K); after 2 ps the thermostat holds 600 K
        lo = 200.0 if chunk == 0 else 560.0
        assert lo < rec[:, 11].min() and rec[:, 11].max() < 640.0


def _oracle_from_driver(d, data, T, q, p_eta):
    """Oracle with the driver's current state (NVT, the driver's chain)."""
    import oracle as O
    N = len(d.types)
    return O.MD(positions=d.positions_array(), velocities=d.velocities_array(),
                types=np.array(d.types), mass=np.array(d.mass), box=np.diag(d.box.h), data=data,
                ensemble=O.NVT, thermostat=(T, q, p_eta, 3.0 * (N - 1.0)),
                barostat=(0.0, 1.0, 0.0), baro_thermostat=(T, [1.0], [0.0], 1.0),
                Nf=3.0 * (N - 1.0), dt=1e-15, neighbors_list=True, r_skin=1.2)


def test_replica_table_batched():
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    data = setfl.read_setfl("ZnCu.eam.alloy")   # the node data of those tables
    masses = [e[2] / U.Navogadro for e in elements]
    specs = R.batched_replica_specs(64)
    N = [s["N"] for s in specs]
    assert sum(N) == 602740 and min(N) == 500 and max(N) == 32000
    a = [R.make_replica_driver(s, pots, eams, masses, lanes_per_atom=16) for s in specs]
    b = [R.make_replica_driver(s, pots, eams, masses, lanes_per_atom=16) for s in specs]
    pick = [int(np.argmin(N)), int(np.argmax(N))]
    start = {}
    for k in pick:
        th = a[k].thermostat
        start[k] = _oracle_from_driver(a[k], data, specs[k]["T"], np.array(th.q),
                                       np.array(th.p_eta))
    ra = R.run_replicas(a, 20, 1, 20, batched=True)
    rb = [d.run(20, 1, 20) for d in b]
    for k in range(64):
        assert np.array_equal(ra[k], rb[k]), k
    for k, r in enumerate(ra):
        assert r.shape == (20, 72) and np.all(np.isfinite(r)), k
        # the reference scales the check with the mass of type 0 (md_driver.cpp:620-627)
        m0 = masses[0]
        assert ((r[:, 21:24] ** 2).sum(1) / (2 * m0 * r[:, 4])).max() < 1e-14, k
        assert r[:, 7].max() < 1e-5, k
        T0 = specs[k]["T"]
        assert 0.2 * T0 < r[-1, 11] < 1.2 * T0, (k, r[-1, 11], T0)
    for k in pick:
        o = start[k]
        o.calculate_neighbors()
        o.integrate(20)
        so = o.scalars()
        g = H.gpu_scalars(a[k])
        print(N[k], "atoms: E_cons", g["etot_cons"] / so["etot_cons"] - 1, "T_kin",
              g["temperature_kin"] / so["temperature_kin"] - 1)
        assert abs(g["etot_cons"] / so["etot_cons"] - 1) < 1e-6
        assert abs(g["temperature_kin"] / so["temperature_kin"] - 1) < 1e-4
        assert np.abs(a[k].coordinates_array() - o.atoms("coordinates")).max() < 1e-8
        # the state the batched run reached: static quantities at mixed tolerances
        import oracle as O
        now = O.MD(positions=a[k].positions_array(), velocities=a[k].velocities_array(),
                   types=np.array(a[k].types), mass=np.array(masses),
                   box=np.diag(a[k].box.h), data=data, ensemble=O.NVE, Nf=3.0 * N[k],
                   dt=1e-15, neighbors_list=True, r_skin=1.0)
        _check_mixed(now, a[k])
