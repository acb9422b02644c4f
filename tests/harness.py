"""Build the oracle and the GPU driver from one specification.

The thermostat/barostat set-up follows MolecularDynamics.__init__
(molecular_dynamics.py:318, :505-543): one RandomState stream, draws in the
order velocities (x, y, z), thermostat p_eta, barostat-chain p_eta, p_epsilon.
"""

from __future__ import annotations

import numpy as np

import cells
from mdelast_amd import units as U

ENS = {"NVE": 0, "NVT": 1, "NPT": 2}


def md_setup(cell, *, T=300.0, tau_T=5e-14, tau_P=5e-13, P=1e-10 * U.GPa, M=3, Mb=3, seed=0,
             velocities="maxwell", static=False, tau_B=None):
    """Velocities and chain parameters as the reference's MolecularDynamics builds them.
    tau_B: time constant of the barostat's chain (default tau_P, as the reference)."""
    N = len(cell["pos"])
    Nf = 3.0 * (N - 1.0)
    if static:
        return dict(vel=np.zeros((N, 3)), Nf=3.0 * N,
                    thermostat=(300.0, [1e-30], [0.0], 3.0 * N), barostat=(0.0, 1e-30, 0.0),
                    baro_thermostat=(300.0, [1e-30], [0.0], 1.0))
    rs = np.random.RandomState(seed)
    m = np.array(cells.masses(cell["data"]))[cell["types"]]
    sig = np.sqrt(U.k_boltz * T / m)
    v = np.stack([rs.normal(0.0, sig), rs.normal(0.0, sig), rs.normal(0.0, sig)], axis=1)
    v -= (v * m[:, None]).sum(0) / m.sum()
    ekin = (3.0 * N - 3.0) * U.k_boltz * T / 2.0
    v *= np.sqrt(ekin / (0.5 * np.sum(m[:, None] * v ** 2)))
    kT = U.k_boltz * T
    q = np.ones(M) * tau_T ** 2 * kT
    q[0] *= Nf
    p_eta = rs.normal(0.0, np.sqrt(q * kT))
    qb = np.ones(Mb) * (tau_P if tau_B is None else tau_B) ** 2 * kT
    p_eta_b = rs.normal(0.0, np.sqrt(qb * kT))
    W = Nf * tau_P ** 2 * kT
    p_eps = rs.normal(0.0, np.sqrt(W * kT))
    return dict(vel=v, Nf=Nf, thermostat=(T, q, p_eta, Nf), barostat=(P, W, p_eps),
                baro_thermostat=(T, qb, p_eta_b, 1.0))


def oracle_md(cell, setup, *, interpolation="quintic", ensemble="NVE", dt=1e-15,
              neighbors_list=False, r_skin=1.0, fatal=1e3, coupling=0):
    import oracle as O
    if "pair" in cell:   # analytic pair potentials, no EAM
        spec = cell["pair"]
        kind = O.LJ if spec["form"] == "lj" else O.MORSE
        pair = [O.AnalyticPair(kind, list(p) + ([spec["rcut"]] if kind == O.LJ else []),
                               spec["rcut"]) for p in spec["params"]]
        return O.MD(positions=cell["pos"], velocities=setup["vel"], types=cell["types"],
                    mass=cells.masses(cell["data"]), box=cell["box"], tables=(pair, [], []),
                    has_eam=False, ensemble=ENS[ensemble], coupling=coupling,
                    thermostat=setup["thermostat"], barostat=setup["barostat"],
                    baro_thermostat=setup["baro_thermostat"], Nf=setup["Nf"], dt=dt,
                    neighbors_list=neighbors_list, r_skin=r_skin, fatal_conservation=fatal)
    return O.MD(positions=cell["pos"], velocities=setup["vel"], types=cell["types"],
                mass=cells.masses(cell["data"]), box=cell["box"], data=cell["data"],
                interpolation=interpolation, ensemble=ENS[ensemble], coupling=coupling,
                thermostat=setup["thermostat"], barostat=setup["barostat"],
                baro_thermostat=setup["baro_thermostat"], Nf=setup["Nf"], dt=dt,
                neighbors_list=neighbors_list, r_skin=r_skin, fatal_conservation=fatal)


_table_cache = {}


def gpu_tables(data, interpolation):
    from mdelast_amd import get_potential
    # keyed by the node-data object (kept alive in the entry, so its id is not
    # reused): one file read by two readers or derivative paths gives two entries
    key = (id(data), interpolation)
    if key not in _table_cache:
        _table_cache[key] = (data, get_potential.tables_from_setfl(data, interpolation))
    return _table_cache[key][1]


def gpu_md(cell, setup, *, interpolation="quintic", ensemble="NVE", dt=1e-15,
           neighbors_list=False, r_skin=1.0, fatal=1e3, coupling=0, lanes_per_atom=0,
           precision="fp64", rg=None):
    from mdelast_amd import md_driver_module as mdm
    if "pair" in cell:   # analytic pair potentials, no EAM
        spec = cell["pair"]
        cls = mdm.LennardJones if spec["form"] == "lj" else mdm.Morse
        pots, eams = [cls(spec["rcut"], *p) for p in spec["params"]], []
    else:
        pots, eams = gpu_tables(cell["data"], interpolation)
    thT, thq, thp, thNf = setup["thermostat"]
    bP, bW, bpe = setup["barostat"]
    bT, bq, bp, bNf = setup["baro_thermostat"]
    th = mdm.Thermostat(thT, thq, thp, thNf)
    ba = mdm.Barostat(bP, bW, bpe, mdm.Thermostat(bT, bq, bp, bNf))
    box = mdm.Box(np.diag(cell["box"]))
    if rg is None:
        rg = [(k + 0.5) * 0.01 * U.angstrom for k in range(10)]
    return mdm.MDDriver(cell["pos"], setup["vel"], mdm.PotentialVector(pots),
                        mdm.EAMVector(eams), ENS[ensemble], coupling, th, ba, box, rg,
                        setup["Nf"], cells.masses(cell["data"]),
                        [int(t) for t in cell["types"]], dt, neighbors_list, r_skin, fatal,
                        lanes_per_atom=lanes_per_atom, precision=precision)


def gpu_scalars(d):
    s = d.scalars_array()
    from mdelast_amd import _native as nat
    out = {k: s[v] for k, v in nat.S_INDEX.items()}
    out["born_stress_tensor"] = s[24:30]
    out["stress_tensor"] = s[30:36]
    out["born_elasticity_tensor"] = s[36:72].reshape(6, 6)
    return out


def rel_err(a, b, scale):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / scale)


def compare_static(o, g, tol):
    """Scale-relative differences of the static quantities (oracle o, GPU driver g).

    Scales: |E| for energies, max|F| for forces, the sums of |pair term| of the
    oracle (divided by V like the quantity) for virial, Born stress, Born
    elasticity and Laplacian. Returns a dict of errors; asserts each <= tol.
    """
    so, sg = o.scalars(), gpu_scalars(g)
    sc = o.scales()
    V = float(np.prod([so["box0"], so["box1"], so["box2"]]))
    errs = {}
    errs["epot_atom"] = abs(sg["epot_atom"] - so["epot_atom"]) / abs(so["epot_atom"])
    # no EAM: the embedding energy and rho are exactly zero on both sides
    errs["epot_eam_atom"] = abs(sg["epot_eam_atom"] - so["epot_eam_atom"]) / \
        (abs(so["epot_eam_atom"]) or 1.0)
    Fo, Fg = o.atoms("forces"), g.forces_array()
    errs["forces"] = rel_err(Fg, Fo, np.abs(Fo).max())
    errs["virial"] = abs(sg["virial"] - so["virial"]) / (sc["virial"] / 3.0)
    errs["born_stress"] = float(np.max(np.abs(sg["born_stress_tensor"] - so["born_stress_tensor"])
                                       / (sc["born_stress"] / V)))
    CBo, CBg = so["born_elasticity_tensor"], sg["born_elasticity_tensor"]
    up = np.triu_indices(6)
    errs["born_elasticity"] = float(np.max(np.abs(CBg - CBo)[up] / (sc["born_elasticity"][up] / V)))
    errs["lower_triangle_zero"] = float(np.abs(np.tril(CBg, -1)).max())
    errs["laplacian"] = abs(sg["total_laplacian"] - so["total_laplacian"]) / sc["laplacian"]
    errs["T_conf_inv"] = abs(sg["temperature_conf_inverse"] / so["temperature_conf_inverse"] - 1)
    errs["pressure"] = abs(sg["pressure"] - so["pressure"]) / (sc["virial"] / 3.0 / V)
    epo, epg = np.array(o.atoms("epot")), np.array(g.energies_pot)
    errs["energies_pot"] = rel_err(epg, epo, np.abs(epo).max())
    rho_o, rho_g = o.atoms("rho"), np.array(g.rho_eam)
    errs["rho"] = rel_err(rho_g, rho_o, np.abs(rho_o).max() or 1.0)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"above {tol}: {bad}"
    return errs


def traj_compare(cell, setup, nsteps, rebuild_every=20, **kw):
    """Step the oracle and the GPU driver side by side (integrate() one step at a
    time, a requested list rebuild every rebuild_every steps when the list is on)
    and return the worst relative differences of the per-step observables, the
    coordinates, velocities and chain momenta, with both drivers."""
    o = oracle_md(cell, setup, **kw)
    g = gpu_md(cell, setup, **kw)
    neighbors = kw.get("neighbors_list", False)
    ensemble = kw.get("ensemble", "NVE")
    worst = {}

    def upd(k, v):
        worst[k] = max(worst.get(k, 0.0), float(v))

    for step in range(nsteps):
        if neighbors and rebuild_every and step % rebuild_every == 0:
            o.calculate_neighbors()
            g.calculate_neighbors()
        o.integrate()
        g.integrate()
        so, sg = o.scalars(), gpu_scalars(g)
        for k in ("etot_cons", "temperature_kin", "epot_atom", "ekin_atom"):
            upd(k, abs(sg[k] / so[k] - 1))
        if so["temperature_conf_inverse"]:
            upd("T_conf_inv", abs(sg["temperature_conf_inverse"] / so["temperature_conf_inverse"]
                                  - 1))
        scale_p = abs(so["virial"]) + 2 * so["ekin_tot"] / 3
        V = so["box0"] * so["box1"] * so["box2"]
        upd("pressure", abs(sg["pressure"] - so["pressure"]) * V / scale_p)
        upd("coords", np.abs(o.atoms("coordinates") - g.coordinates_array()).max())
        vo, vg = o.atoms("velocities"), g.velocities_array()
        upd("vel", np.abs(vo - vg).max() / np.abs(vo).max())
        if ensemble == "NPT":
            upd("box", abs(sg["box0"] / so["box0"] - 1))
        assert sg["cycle"] == so["cycle"]
    if ensemble != "NVE":
        po, _ = o.chain(0)
        upd("p_eta", np.max(np.abs(np.array(g.thermostat.p_eta) - po) / np.abs(po).max()))
    if ensemble == "NPT":
        pb, _ = o.chain(1)
        upd("p_eta_baro", np.max(np.abs(np.array(g.barostat.thermostat.p_eta) - pb)
                                 / np.abs(pb).max()))
    return worst, o, g
