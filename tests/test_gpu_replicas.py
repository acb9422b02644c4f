"""Replicas advanced side by side (mdl_drivers_run) against the same replicas
advanced one at a time: the interleaving of launches across streams changes
nothing in any replica's arithmetic, so the records agree bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from mdelast_amd import get_potential, replicas as R
from mdelast_amd import units as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_side_by_side_equals_one_at_a_time(precision):
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=6, N=864, x_zn=0.0, T=800.0, place_seed=3, vel_seed=4),
             dict(edge=7, N=1372, x_zn=0.10, T=150.0, place_seed=5, vel_seed=6)]
    a = [R.make_replica_driver(s, pots, eams, masses, precision=precision) for s in specs]
    b = [R.make_replica_driver(s, pots, eams, masses, precision=precision) for s in specs]
    ra = R.run_replicas(a, 60, 1, 20)
    rb = [d.run(60, 1, 20) for d in b]
    for x, y in zip(ra, rb):
        assert x.shape == (60, 72)
        assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.forces_array(), db.forces_array())
    # temperatures are those of three different systems
    T = [r[-1][11] for r in ra]
    assert len(set(np.round(T, 6))) == 3


def test_side_by_side_self_decided_rebuilds():
    """Eight 32,000-atom replicas side by side with a thin skin, so that the list
    check decides rebuilds inside k_list on many steps, often on several
    streams at once. Their k_list grids share the device (each would fill it
    alone), and the records, forces and build counts equal the replicas run one
    at a time, bit for bit."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=20, N=32000, x_zn=0.05 * (k % 3), T=300.0 + 100.0 * k,
                  place_seed=10 + k, vel_seed=20 + k) for k in range(8)]

    def make():
        return [R.make_replica_driver(s, pots, eams, masses, precision="mixed",
                                      lanes_per_atom=16, skin=1.04) for s in specs]

    a, b = make(), make()
    builds0 = [d.neighbor_list_builds for d in a]
    ra = R.run_replicas(a, 60, 5, 0)
    rb = [d.run(60, 5, 0) for d in b]
    for x, y in zip(ra, rb):
        assert x.shape == (12, 72)
        assert np.array_equal(x, y)
    for da, db, b0 in zip(a, b, builds0):
        assert np.array_equal(da.forces_array(), db.forces_array())
        assert da.neighbor_list_builds == db.neighbor_list_builds
        assert da.neighbor_list_builds - b0 >= 2      # decided by the check, none forced


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_step_launches(precision):
    """Chunks of 20 steps replayed as captured graphs give the step-by-step
    launches' records and state bit for bit, with a sampled stride, forced
    rebuilds, a requested rebuild and a run whose length is not a multiple of
    the chunk."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=6, N=864, x_zn=0.0, T=800.0, place_seed=3, vel_seed=4)]
    a = [R.make_replica_driver(s, pots, eams, masses, precision=precision) for s in specs]
    b = [R.make_replica_driver(s, pots, eams, masses, precision=precision) for s in specs]
    for d in a:
        d.graph_steps = 20
    for da, db in zip(a, b):
        assert np.array_equal(da.run(45, 5, 20), db.run(45, 5, 20))
        da.calculate_neighbors()
        db.calculate_neighbors()
        assert np.array_equal(da.run(70, 5, 10), db.run(70, 5, 10))
    ra = R.run_replicas(a, 101, 1, 20)
    rb = R.run_replicas(b, 101, 1, 20)
    for x, y in zip(ra, rb):
        assert x.shape == (101, 72)
        assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.forces_array(), db.forces_array())
        assert np.array_equal(da.positions_array(), db.positions_array())


def test_joint_graph_equals_step_launches():
    """Replicas whose chunks are captured together as one graph (one launch
    per chunk for all of them) give the step-by-step results bit for bit,
    also after a driver of the set is replaced."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=6, N=864, x_zn=0.0, T=800.0, place_seed=3, vel_seed=4),
             dict(edge=5, N=500, x_zn=0.10, T=150.0, place_seed=5, vel_seed=6)]
    a = [R.make_replica_driver(s, pots, eams, masses, precision="mixed") for s in specs]
    b = [R.make_replica_driver(s, pots, eams, masses, precision="mixed") for s in specs]
    for d in a:
        d.graph_steps = 20
    for rep in range(2):
        ra = R.run_replicas(a, 61, 1, 20)
        rb = R.run_replicas(b, 61, 1, 20)
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
    # a new driver in place of one of the set: the cached graph must not be reused
    a[1] = R.make_replica_driver(specs[1], pots, eams, masses, precision="mixed")
    b[1] = R.make_replica_driver(specs[1], pots, eams, masses, precision="mixed")
    a[1].graph_steps = 20
    a[1].run(122, 1, 20)
    b[1].run(122, 1, 20)
    ra = R.run_replicas(a, 60, 5, 20)
    rb = R.run_replicas(b, 60, 5, 20)
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.forces_array(), db.forces_array())


def test_joint_graph_waits_for_every_stream():
    """A joint chunk launched on the first driver's stream starts after the
    steps still queued on the other drivers' streams: a small first replica
    and a large second one, stepped one by one up to the chunk boundary, then
    advanced by joint chunks, agree bit for bit with step launches."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=20, N=32000, x_zn=0.06, T=600.0, place_seed=3, vel_seed=4)]
    a = [R.make_replica_driver(s, pots, eams, masses, precision="mixed") for s in specs]
    b = [R.make_replica_driver(s, pots, eams, masses, precision="mixed") for s in specs]
    for d in a:
        d.graph_steps = 20
    for n in (7, 53, 60):
        ra = R.run_replicas(a, n, 1, 20)
        rb = R.run_replicas(b, n, 1, 20)
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.positions_array(), db.positions_array())


@pytest.mark.parametrize("lanes", [[16, 16, 16, 16], [8, 32, 0, 16]])
@pytest.mark.parametrize("skin", [1.2, 1.04])
def test_batched_launches_equal_one_at_a_time(skin, lanes):
    """Every pass of a step as one launch over all replicas (mdl_replicas_run;
    one launch per lanes-per-atom group) gives each replica's records, forces
    and positions bit for bit as its own launches do: with forced rebuilds
    (batched build launches), with a requested rebuild, and (thin skin) with
    rebuilds that the one-block list check decides by itself, over brute-force
    and cell lists; lanes per atom fixed or autotuned (0)."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=6, N=864, x_zn=0.0, T=800.0, place_seed=3, vel_seed=4),
             dict(edge=8, N=2048, x_zn=0.10, T=150.0, place_seed=5, vel_seed=6),
             dict(edge=5, N=500, x_zn=0.0, T=900.0, place_seed=7, vel_seed=8)]

    def make():
        return [R.make_replica_driver(s, pots, eams, masses, precision="mixed",
                                      lanes_per_atom=lp, skin=skin)
                for s, lp in zip(specs, lanes)]

    a, b = make(), make()
    builds0 = [d.neighbor_list_builds for d in a]
    ra = R.run_replicas(a, 60, 1, 20, batched=True)
    rb = [d.run(60, 1, 20) for d in b]
    for x, y in zip(ra, rb):
        assert x.shape == (60, 72)
        assert np.array_equal(x, y)
    for d in a + b:
        d.calculate_neighbors()
    ra = R.run_replicas(a, 45, 5, 0, batched=True)
    rb = [d.run(45, 5, 0) for d in b]
    for x, y in zip(ra, rb):
        assert x.shape == (9, 72)
        assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.forces_array(), db.forces_array())
        assert np.array_equal(da.positions_array(), db.positions_array())
        assert da.neighbor_list_builds == db.neighbor_list_builds
    if skin < 1.1:
        # rebuilds beyond the 3 forced and 1 requested ones came from the check
        assert all(d.neighbor_list_builds - b0 > 4 for d, b0 in zip(a, builds0))


@pytest.mark.parametrize("skin", [1.2, 1.04])
def test_batched_graph_replay_equals_batched_launches(skin):
    """Chunks of 20 batched steps replayed as one captured graph (the lead
    driver's graph_steps) give the batched launches' records, forces and
    positions bit for bit: sampled strides 5 and 4, forced builds every 10 and
    20 steps and none, a requested rebuild, rebuilds that the list check decides
    inside a replay (thin skin), runs that end inside a chunk, and a second set
    led by the same driver (tables sent again, graph captured again)."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    specs = [dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2),
             dict(edge=6, N=864, x_zn=0.0, T=800.0, place_seed=3, vel_seed=4),
             dict(edge=5, N=500, x_zn=0.10, T=900.0, place_seed=7, vel_seed=8)]

    def make():
        return [R.make_replica_driver(s, pots, eams, masses, precision="mixed",
                                      lanes_per_atom=lp, skin=skin)
                for s, lp in zip(specs, (16, 8, 16))]

    a, b = make(), make()
    a[0].graph_steps = 20
    builds0 = [d.neighbor_list_builds for d in a]
    for k, (n, stride, rebuild) in enumerate([(73, 5, 10), (47, 4, 20), (60, 5, 0)]):
        c0 = a[0].cycle
        ra = R.run_replicas(a, n, stride, rebuild, batched=True)
        rb = R.run_replicas(b, n, stride, rebuild, batched=True)
        for x, y in zip(ra, rb):
            assert x.shape[0] == sum((c0 + k) % stride == 0 for k in range(1, n + 1))
            assert np.array_equal(x, y)
        if k == 0:
            a[1].calculate_neighbors()
            b[1].calculate_neighbors()
    ra = R.run_replicas(a[:2], 40, 5, 10, batched=True)
    rb = R.run_replicas(b[:2], 40, 5, 10, batched=True)
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    for da, db in zip(a, b):
        assert np.array_equal(da.forces_array(), db.forces_array())
        assert np.array_equal(da.positions_array(), db.positions_array())
        assert da.neighbor_list_builds == db.neighbor_list_builds
    if skin < 1.1:
        # rebuilds beyond the forced (cycles 0-70 every 10, 80 and 100, 180-210
        # every 10 for the first two) and requested ones came from the check
        forced = (8 + 2 + 4, 8 + 2 + 4 + 1, 8 + 2)
        assert all(d.neighbor_list_builds - b0 > f for d, b0, f in zip(a, builds0, forced))


def test_batched_launches_refuse_mixed_settings():
    """A set that cannot share launches is refused before any driver moves."""
    elements, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    masses = [e[2] / U.Navogadro for e in elements]
    s = dict(edge=5, N=500, x_zn=0.06, T=300.0, place_seed=1, vel_seed=2)
    d16 = R.make_replica_driver(s, pots, eams, masses, lanes_per_atom=16)
    ref = R.make_replica_driver(s, pots, eams, masses, lanes_per_atom=16)
    dfp = R.make_replica_driver(s, pots, eams, masses, precision="fp64", lanes_per_atom=16)
    with pytest.raises(Exception, match="mixed precision"):
        R.run_replicas([d16, dfp], 5, 1, 0, batched=True)
    late = R.make_replica_driver(s, pots, eams, masses, lanes_per_atom=16)
    late.run(3, 1, 0)
    with pytest.raises(Exception, match="same cycle"):
        R.run_replicas([d16, late], 5, 1, 0, batched=True)
    # the refused calls left d16 untouched
    assert np.array_equal(d16.run(4, 1, 0), ref.run(4, 1, 0))
