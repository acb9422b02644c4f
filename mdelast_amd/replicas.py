"""Independent replicas advanced side by side on one device.

The reference estimates its error bars from replicas run as separate
single-core jobs (velocity and placement seeds; benchmarks/slurm_example.sh).
Here each replica is one driver with its own stream, and
``run_replicas`` advances all of them with one host call
(``mdl_drivers_run``): step k of every replica is launched before step k + 1
of any, so replicas of uneven size, composition and temperature overlap on the
device instead of queueing behind one another. With ``batched=True``
(``mdl_replicas_run``) each pass of a step is a single launch whose blocks
cover every replica, so many small systems fill the device together.

``batched_replica_specs`` draws the batched-replica workload: cubic FCC
supercells with edge 5 to 20 cells (500 to 32 000 atoms), each pure Cu or a
random Cu-Zn solution with up to 10 at.% Zn, temperature uniform in
100-900 K, NVT with dt = 1 fs.
"""

from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _native as nat
from . import md_driver_module as mdm
from . import units as U


def run_replicas(drivers: Sequence["mdm.MDDriver"], nsteps: int, sample_stride: int = 1,
                 rebuild_every: int = 0, records: bool = True,
                 batched: bool = False) -> List[np.ndarray]:
    """nsteps steps of every driver (each as MDDriver.run would do them); returns
    each driver's per-sample records (MDL_S_* layout).

    batched: every pass of a step is one launch over all replicas
    (``mdl_replicas_run``; mixed precision, all drivers at the same cycle; the
    pair passes one launch per lanes-per-atom group) instead of one launch per
    replica on its own stream (``mdl_drivers_run``). In a batched run the first
    driver's ``graph_steps`` (when set) replays chunks of that many batched steps
    as one captured graph, with the same results."""
    n = len(drivers)
    if n == 0:
        return []
    if len({id(d) for d in drivers}) != n:
        raise ValueError("a driver appears twice")
    handles = (C.c_void_p * n)(*[d.handle for d in drivers])
    bufs, ptrs = [], (nat._dp * n)()
    for i, d in enumerate(drivers):
        c0 = d.cycle
        nsamp = sum(1 for k in range(1, nsteps + 1) if (c0 + k) % sample_stride == 0)
        b = np.zeros((max(nsamp, 1), nat.NRECORD))
        bufs.append(b)
        ptrs[i] = nat.dptr(b) if records else None
    counts = (C.c_int64 * n)()
    for d in drivers:
        d._cache = None
    try:
        fn = nat.lib().mdl_replicas_run if batched else nat.lib().mdl_drivers_run
        st = fn(handles, n, int(nsteps), int(sample_stride), int(rebuild_every),
                ptrs if records else None, counts)
    finally:
        for d in drivers:
            d._cache = None
    if st != 0:
        raise mdm.MDDriverError(st, nat.lib().mdl_last_error().decode(errors="replace"))
    return [b[:counts[i]] for i, b in enumerate(bufs)]


def batched_replica_specs(n: int = 64, seed: int = 2026, edge=(5, 20), x_zn_max: float = 0.10,
                          T_range=(100.0, 900.0)):
    """The replica table: per replica its FCC edge (cells), Zn fraction (0 for
    the pure-Cu replicas, drawn for about half of them), temperature and seeds."""
    rs = np.random.RandomState(seed)
    specs = []
    for k in range(n):
        e = int(rs.randint(edge[0], edge[1] + 1))
        pure = bool(rs.rand() < 0.5)
        x = 0.0 if pure else float(rs.uniform(0.0, x_zn_max))
        T = float(rs.uniform(*T_range))
        specs.append(dict(edge=e, N=4 * e ** 3, x_zn=x, T=T, place_seed=int(rs.randint(2 ** 31)),
                          vel_seed=int(rs.randint(2 ** 31))))
    return specs


def fcc_positions(edge: int, a: float) -> np.ndarray:
    """Centred FCC supercell (molecular_dynamics.py:582-608), CGS."""
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    idx = np.indices((edge, edge, edge)).reshape(3, -1).T
    L = edge * a
    return ((idx[:, None, :] + 0.25 + basis[None, :, :]) * a).reshape(-1, 3) - L / 2.0


def make_replica_driver(spec, pots, eams, masses, *, a=3.6345 * U.angstrom, zn_type=0,
                        cu_type=1, dt=1e-15, M=10, tau_T=5e-14, precision="mixed",
                        lanes_per_atom=0, skin=1.2):
    """One NVT replica, set up as MolecularDynamics does (velocities at T with the
    centre-of-mass momentum removed and an exact rescale, chain momenta drawn
    from the chain masses)."""
    pos = fcc_positions(spec["edge"], a)
    N = len(pos)
    types = np.full(N, cu_type, dtype=np.uint32)
    nzn = int(round(spec["x_zn"] * N))
    if nzn:
        types[np.random.RandomState(spec["place_seed"]).choice(N, nzn, replace=False)] = zn_type
    m = np.asarray(masses)[types]
    rs = np.random.RandomState(spec["vel_seed"])
    T = spec["T"]
    sig = np.sqrt(U.k_boltz * T / m)
    v = np.stack([rs.normal(0.0, sig) for _ in range(3)], axis=1)
    v -= (v * m[:, None]).sum(0) / m.sum()
    v *= np.sqrt((3.0 * N - 3.0) * U.k_boltz * T / 2.0 / (0.5 * np.sum(m[:, None] * v ** 2)))
    Nf = 3.0 * (N - 1.0)
    kT = U.k_boltz * T
    q = np.ones(M) * tau_T ** 2 * kT
    q[0] *= Nf
    th = mdm.Thermostat(T, q, rs.normal(0.0, np.sqrt(q * kT)), Nf)
    ba = mdm.Barostat(0.0, 1.0, 0.0, mdm.Thermostat(T, [1.0], [0.0], 1.0))
    L = spec["edge"] * a
    return mdm.MDDriver(pos, v, mdm.PotentialVector(pots), mdm.EAMVector(eams), mdm.MDDriver.NVT,
                        mdm.MDDriver.THERMOSTAT_KINETIC, th, ba, mdm.Box(np.diag([L] * 3)), [1e-10],
                        Nf, masses, types, dt, True, skin, 1e3, precision=precision,
                        lanes_per_atom=lanes_per_atom)
