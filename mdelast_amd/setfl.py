"""LAMMPS SETFL (eam/alloy) reader producing CGS node arrays.

This is the data side of the potential loader: it parses the file, converts
to CGS and computes the first and second derivatives at the table nodes. It
builds no driver objects (``get_potential`` does that), so the node arrays can
also be fed to other consumers.

Semantics follow src/python/get_potential.py:58-147 of the reference:

* header: 3 comment lines, then ``Nel name...``, then ``Nrho drho Nr dr rcut``;
  per element ``Z mass a lattice`` followed by Nrho values of F(rho) and Nr
  values of rho(r); then the pair blocks r*phi(r) in the order 11, 21, 22, 31,
  32, 33, ... (get_potential.py:122-146), stored at slot j(j+1)/2 + i, i <= j;
* CGS: F in erg (x eV), rho in cm^-3 (/ A^3), drho in cm^-3, dr in cm;
  phi = (r phi) / r * eV * A with phi[0] = phi[1] (get_potential.py:127-128);
* grids are ``numpy.linspace(0, d*(N-1), N)`` in CGS; the cutoff handed to the
  driver is the last grid node r_grid[-1], and rho_max = rho_grid[-1]
  (get_potential.py:94-95, :117, :134), not the header's rcut;
* node derivatives are those of the k=5 interpolating spline through the
  nodes, SciPy ``UnivariateSpline(k=5, s=0)`` (get_potential.py:49-55).

``read_setfl`` parses and differentiates in the library (``mdl_setfl_read``,
include/mdelast_setfl.h): no SciPy on the production path. The spline is the
same one SciPy builds (degree 5, interior knots at x[3..n-4]), solved from its
collocation equations in long double; ``read_setfl_python`` (numpy parsing,
SciPy derivatives) is kept as its cross-check.

The tokenizer reads whitespace-separated values rather than one per line, so
files with several values per line also load; for one-value-per-line files
(all three shipped potentials) the numbers are identical to ``numpy.loadtxt``.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import units as U

POTENTIAL_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "potentials")


def node_derivatives(x: np.ndarray, y: np.ndarray):
    """First and second derivatives at the nodes of the quintic interpolating
    spline (get_potential.py:49-55), computed in the library."""
    import ctypes as C
    from . import _native as nat
    x = np.ascontiguousarray(x, dtype=float)
    y = np.ascontiguousarray(y, dtype=float)
    d1, d2 = np.zeros_like(x), np.zeros_like(x)
    nat.check(nat.lib().mdl_spline5_node_derivatives(C.c_int64(len(x)), nat.dptr(x), nat.dptr(y),
                                                     nat.dptr(d1), nat.dptr(d2)))
    return d1, d2


def node_derivatives_scipy(x: np.ndarray, y: np.ndarray):
    """The same derivatives from SciPy, as get_potential.py:49-55 computes them
    (FITPACK k=5, s=0); the cross-check of ``node_derivatives``."""
    from scipy.interpolate import UnivariateSpline

    spl = UnivariateSpline(x=x, y=y, k=5, s=0)
    return spl.derivative(n=1)(x), spl.derivative(n=2)(x)


@dataclass
class SetflElement:
    name: str
    atomic_number: float
    atomic_weight: float
    a_lattice_angstrom: float
    F: np.ndarray = field(repr=False)
    dF: np.ndarray = field(repr=False)
    d2F: np.ndarray = field(repr=False)
    rho: np.ndarray = field(repr=False)
    drho: np.ndarray = field(repr=False)
    d2rho: np.ndarray = field(repr=False)


@dataclass
class SetflPair:
    i: int
    j: int
    phi: np.ndarray = field(repr=False)
    dphi: np.ndarray = field(repr=False)
    d2phi: np.ndarray = field(repr=False)


@dataclass
class SetflData:
    """CGS node data of one SETFL file. ``pairs`` is indexed by the driver slot."""

    path: str
    elements: List[SetflElement]
    pairs: List[SetflPair]
    Nrho: int
    drho: float
    Nr: int
    dr: float
    rcut: float          # r_grid[-1], the cutoff the driver uses
    rhomax: float        # rho_grid[-1]
    r_grid: np.ndarray = field(repr=False)
    rho_grid: np.ndarray = field(repr=False)
    rmin_fatal: float = 1e-2 * U.angstrom

    @property
    def names(self):
        return [e.name for e in self.elements]


def resolve_path(file_name: str) -> str:
    """Absolute paths are used as given; others are looked up in ``potentials/``."""
    if os.path.isabs(file_name):
        return file_name
    return os.path.join(POTENTIAL_DIR, file_name)


def read_setfl(file_name: str, with_derivatives: bool = True) -> SetflData:
    """Parse, convert and differentiate in the library (mdl_setfl_read)."""
    import ctypes as C
    from . import _native as nat
    path = resolve_path(file_name)
    L = nat.lib()
    h = C.c_void_p()
    nat.check(L.mdl_setfl_read(path.encode(), int(bool(with_derivatives)), C.byref(h)))
    try:
        nel, nrho, nr = C.c_int32(), C.c_int64(), C.c_int64()
        drho, dr, rcut, rhomax, rch = (C.c_double() for _ in range(5))
        nat.check(L.mdl_setfl_grid(h, C.byref(nel), C.byref(nrho), C.byref(drho), C.byref(nr),
                                   C.byref(dr), C.byref(rcut), C.byref(rhomax), C.byref(rch)))

        def nodes(which, index, n):
            f, d1, d2 = (np.zeros(n) for _ in range(3))
            nat.check(L.mdl_setfl_nodes(h, which, index, nat.dptr(f), nat.dptr(d1),
                                        nat.dptr(d2)))
            return (f, d1, d2) if with_derivatives else (f, None, None)

        elements = []
        for e in range(nel.value):
            name, lat = C.create_string_buffer(64), C.create_string_buffer(64)
            Z, mass, a = C.c_double(), C.c_double(), C.c_double()
            nat.check(L.mdl_setfl_element(h, e, name, 64, C.byref(Z), C.byref(mass), C.byref(a),
                                          lat, 64))
            F, dF, d2F = nodes(0, e, nrho.value)
            rho, drh, d2rh = nodes(1, e, nr.value)
            elements.append(SetflElement(name.value.decode(), Z.value, mass.value, a.value, F, dF,
                                         d2F, rho, drh, d2rh))
        n = nel.value
        pairs = []
        for j in range(n):
            for i in range(j + 1):
                phi, d1, d2 = nodes(2, j * (j + 1) // 2 + i, nr.value)
                pairs.append(SetflPair(i, j, phi, d1, d2))
    finally:
        L.mdl_setfl_destroy(h)
    r_grid = np.linspace(0.0, dr.value * (nr.value - 1.0), nr.value)
    rho_grid = np.linspace(0.0, drho.value * (nrho.value - 1.0), nrho.value)
    return SetflData(path=path, elements=elements, pairs=pairs, Nrho=nrho.value, drho=drho.value,
                     Nr=nr.value, dr=dr.value, rcut=rcut.value, rhomax=rhomax.value,
                     r_grid=r_grid, rho_grid=rho_grid)


def read_setfl_python(file_name: str, with_derivatives: bool = True,
                      derivatives=node_derivatives_scipy) -> SetflData:
    """numpy parsing and SciPy derivatives, as the reference reads the file: the
    cross-check of ``read_setfl``."""
    node_derivatives = derivatives
    path = resolve_path(file_name)
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    head = lines[3].split()
    nel = int(head[0])
    names = head[1:1 + nel]
    if len(names) != nel:
        raise ValueError(f"{path}: element line lists {len(names)} names for {nel} elements")
    p = lines[4].split()
    Nrho, drho_a, Nr, dr_a = int(p[0]), float(p[1]), int(p[2]), float(p[3])
    drho = drho_a / U.angstrom ** 3
    dr = dr_a * U.angstrom
    r_grid = np.linspace(0.0, dr * (Nr - 1.0), Nr)
    rho_grid = np.linspace(0.0, drho * (Nrho - 1.0), Nrho)

    # Per element: one header line (Z, mass, a, lattice) and Nrho + Nr values;
    # then Nel(Nel+1)/2 blocks of Nr values. Read value tokens line by line so that
    # the element header lines are recognised wherever the value blocks end.
    pos = 5
    elements = []

    def take(n):
        nonlocal pos
        out = []
        while len(out) < n:
            if pos >= len(lines):
                raise ValueError(f"{path}: file ends inside a table ({len(out)}/{n} values)")
            out.extend(float(t) for t in lines[pos].split())
            pos += 1
        if len(out) != n:
            raise ValueError(f"{path}: a table line crosses a block boundary")
        return np.array(out, dtype=float)

    for k in range(nel):
        zline = lines[pos].split()
        pos += 1
        Z, mass, a = (float(s) for s in zline[:3])
        F = take(Nrho) * U.ev
        rho = take(Nr) / U.angstrom ** 3
        if with_derivatives:
            drho_n, d2rho_n = node_derivatives(r_grid, rho)
            dF_n, d2F_n = node_derivatives(rho_grid, F)
        else:
            drho_n = d2rho_n = dF_n = d2F_n = None
        elements.append(SetflElement(names[k], Z, mass, a, F, dF_n, d2F_n, rho, drho_n, d2rho_n))

    npair = nel * (nel + 1) // 2
    pairs: List[SetflPair] = [None] * npair  # type: ignore[list-item]
    i = j = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(npair):
            phir = take(Nr)
            phi = phir / r_grid * U.ev * U.angstrom
            phi[0] = phi[1]
            if with_derivatives:
                d1, d2 = node_derivatives(r_grid, phi)
            else:
                d1 = d2 = None
            slot = j * (j + 1) // 2 + i
            pairs[slot] = SetflPair(i, j, phi, d1, d2)
            # blocks are 11, 21, 22, 31, 32, 33, ...: (i, j) = (0,0), (0,1), (1,1), (0,2), ...
            if i == j:
                j += 1
                i = 0
            else:
                i += 1
    return SetflData(path=path, elements=elements, pairs=pairs, Nrho=Nrho, drho=drho, Nr=Nr,
                     dr=dr, rcut=float(r_grid[-1]), rhomax=float(rho_grid[-1]), r_grid=r_grid,
                     rho_grid=rho_grid)
