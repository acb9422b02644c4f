"""Test configurations shared by the CPU and GPU tests.

* smoke cell: benchmarks/smoke/run_smoke.py:102-129 of the reference (fcc 4^3,
  a = 3.6345 A, RandomState(2026) displacements of 0.08 A, 15 Zn);
* Ni-Al-Co cell: the reference's LAMMPS fixture (tests/golden/NiAlCo.data);
* fcc lattices as MolecularDynamics.build_fcc_lattice (molecular_dynamics.py:578-614).
"""

from __future__ import annotations

import json
import os

import numpy as np

from mdelast_amd import setfl, units as U

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_cache = {}


def setfl_data(name):
    """Node data through the reference's own derivative path (numpy parsing, SciPy
    k=5 spline), so that the oracle reproduces the reference's golden values to
    ~1e-14; the library's reader (setfl.read_setfl) is pinned against this one
    in test_setfl_native.py."""
    if name not in _cache:
        _cache[name] = setfl.read_setfl_python(name)
    return _cache[name]


def fcc_positions(nside, a):
    """Centred fcc supercell in A, atom order (i, j, k, basis) as the reference."""
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    pos = np.array([(np.array([i, j, k]) + 0.25 + b) * a
                    for i in range(nside) for j in range(nside)
                    for k in range(nside) for b in basis])
    return pos - nside * a / 2.0


def smoke_cell():
    data = setfl_data("ZnCu.eam.alloy")
    A0, ns = 3.6345, 4
    rs = np.random.RandomState(2026)
    pos = fcc_positions(ns, A0)
    pos = pos + rs.normal(scale=0.08, size=pos.shape)
    L = ns * A0
    pos = np.mod(pos + L / 2, L) - L / 2
    types = np.full(len(pos), data.names.index("Cu"))
    types[rs.choice(len(pos), size=15, replace=False)] = data.names.index("Zn")
    return dict(data=data, pos=pos * U.angstrom, types=types, box=np.full(3, L * U.angstrom))


def nialco_cell():
    data = setfl_data("Mishin_updated-Ni-Al-Co-2013.eam.alloy")
    lines = open(os.path.join(GOLDEN, "NiAlCo.data")).read().splitlines()
    lo, hi = (float(x) for x in next(l for l in lines if l.endswith("xlo xhi")).split()[:2])
    start = next(i for i, l in enumerate(lines) if l.startswith("Atoms")) + 2
    rows = np.array([[float(x) for x in l.split()] for l in lines[start:] if l.strip()])
    rows = rows[np.argsort(rows[:, 0])]
    L = hi - lo
    pos = np.mod(rows[:, 2:5] + L / 2, L) - L / 2
    return dict(data=data, pos=pos * U.angstrom, types=rows[:, 1].astype(int) - 1,
                box=np.full(3, L * U.angstrom))


def lammps_reference():
    ref = json.load(open(os.path.join(GOLDEN, "lammps_values.json")))["systems"]["NiAlCo"]
    F = np.loadtxt(os.path.join(GOLDEN, "NiAlCo_forces.dump"), skiprows=9)
    ref["forces"] = F[np.argsort(F[:, 0])][:, 1:]
    return ref


def alloy_cell(nside, a, potential, fractions, seed, disp=0.0, disp_seed=None):
    """fcc cell with species placed by RandomState(seed).permutation; fractions per species
    in the potential's order (the remainder goes to species 0)."""
    data = setfl_data(potential)
    pos = fcc_positions(nside, a)
    N = len(pos)
    counts = [int(round(f * N)) for f in fractions]
    counts[0] = N - sum(counts[1:])
    perm = np.random.RandomState(seed).permutation(N)
    types = np.zeros(N, dtype=int)
    k = 0
    for t, c in enumerate(counts):
        types[perm[k:k + c]] = t
        k += c
    if disp:
        pos = pos + np.random.RandomState(disp_seed).normal(scale=disp, size=pos.shape)
    L = nside * a
    pos = np.mod(pos + L / 2, L) - L / 2
    return dict(data=data, pos=pos * U.angstrom, types=types, box=np.full(3, L * U.angstrom))


def zn06_cell(nside=5):
    """Cu-Zn 6 at.% cell of sample_case/run.py:224-227 (Zn by RandomState(42).choice)."""
    data = setfl_data("ZnCu.eam.alloy")
    a = 3.6345
    pos = fcc_positions(nside, a)
    N = len(pos)
    nzn = int(round(0.06 * N))
    types = np.full(N, data.names.index("Cu"))
    types[np.random.RandomState(42).choice(N, size=nzn, replace=False)] = data.names.index("Zn")
    L = nside * a
    return dict(data=data, pos=pos * U.angstrom, types=types, box=np.full(3, L * U.angstrom))


def masses(data):
    return [e.atomic_weight / U.Navogadro for e in data.elements]


def maxwell_velocities(cell, T, seed):
    """Velocities as MolecularDynamics.set_initial_velocities (molecular_dynamics.py:616-642)."""
    m = np.array(masses(cell["data"]))[cell["types"]]
    rs = np.random.RandomState(seed)
    sig = np.sqrt(U.k_boltz * T / m)
    v = np.stack([rs.normal(0.0, sig), rs.normal(0.0, sig), rs.normal(0.0, sig)], axis=1)
    v -= (v * m[:, None]).sum(0) / m.sum()
    N = len(m)
    ekin = (3.0 * N - 3.0) * U.k_boltz * T / 2.0
    v *= np.sqrt(ekin / (0.5 * np.sum(m[:, None] * v ** 2)))
    return v


def pair_cell(nside, a, spec, atomic_weights, type_fractions=None, seed=0, disp=0.0,
              disp_seed=None):
    """fcc cell (lattice constant a in A) for analytic pair potentials.
    spec = dict(form="lj"|"morse", rcut=<cm>, params=[per slot: (epsilon, sigma) or
    (D, alpha, r0)]); type_fractions: fraction of each species (one species if None)."""
    from types import SimpleNamespace
    pos = fcc_positions(nside, a)
    rs = np.random.RandomState(seed)
    types = np.zeros(len(pos), dtype=int)
    if type_fractions is not None:
        types = rs.choice(len(type_fractions), size=len(pos), p=type_fractions)
    if disp:
        pos = pos + np.random.RandomState(disp_seed).normal(scale=disp, size=pos.shape)
    L = nside * a
    pos = np.mod(pos + L / 2, L) - L / 2
    data = SimpleNamespace(path=None,
                           elements=[SimpleNamespace(atomic_weight=w) for w in atomic_weights])
    return dict(data=data, pair=spec, pos=pos * U.angstrom, types=types,
                box=np.full(3, L * U.angstrom))
