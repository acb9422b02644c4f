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
    # "alloy": get_EAM_alloy_file's conversion; "single": get_EAM_file's (the
    # grids are formed in A and scaled afterwards, so the last bits differ)
    reader: str = "alloy"

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


# Leading entries of rho(r) and r*phi(r) in the Mendelev Cu FS file that hold the
# fill value 1e12 (get_potential.py:213-226); they take the first tabulated value.
_FS_FILL, _FS_NFILL = 1e12, 167


def read_eam_file(file_name: str, derivatives=node_derivatives) -> SetflData:
    """One-element EAM file as get_EAM_file reads it (get_potential.py:150-257,
    molecular_dynamics.py:123-234 of the reference).

    * line 6: ``Z mass a lattice``; line 5: ``Nrho drho Nr dr rcut``; every
      value after line 6 is table data, read as whitespace-separated tokens
      (``numpy.loadtxt(skiprows=6)``);
    * ``*.eam.alloy.txt`` (SETFL): Nrho values of F(rho), Nr of rho(r), Nr of
      r*phi(r), nothing else; ``*.eam.fs.txt`` (Finnis-Sinclair, five values a
      line): the same three blocks, Nrho and Nr multiples of 5, and the first
      167 entries of rho and of r*phi hold 1e12 and take entry 167; any other
      name is refused (the reference logs FATAL and exits);
    * unlike the alloy reader, the grids are formed in A and A^-3 and then
      scaled, and phi = (r phi) / r[A] * eV; phi[0] = phi[1];
    * node derivatives of the quintic interpolating spline on the CGS grids.

    Returns the node data as a one-element :class:`SetflData` (``reader`` =
    "single"); the file's element symbol is kept when line 4 names one."""
    path = resolve_path(file_name)
    if not (file_name.endswith("eam.alloy.txt") or file_name.endswith("eam.fs.txt")):
        raise ValueError(f"undefined potential format '{file_name}'")
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    if len(lines) < 7:
        raise ValueError(f"{path}: too short for an EAM file")
    Z, mass, a_lat = (float(s) for s in lines[5].split()[:-1][:3])
    p = [float(s) for s in lines[4].split()]
    if len(p) < 5:
        raise ValueError(f"{path}: line 5 must hold Nrho drho Nr dr rcut")
    Nrho, drho_a, Nr, dr_a = int(p[0]), p[1], int(p[2]), p[3]
    head = lines[3].split()
    name = head[1] if len(head) > 1 else "Cu"
    r_grid_a = np.linspace(0.0, dr_a * (Nr - 1.0), Nr)
    rho_grid_a = np.linspace(0.0, drho_a * (Nrho - 1.0), Nrho)
    values = np.array([float(t) for ln in lines[6:] for t in ln.split()], dtype=float)
    if file_name.endswith("eam.alloy.txt"):
        if any(len(ln.split()) > 1 for ln in lines[6:]) or len(values) != Nrho + 2 * Nr:
            raise ValueError(f"{path}: SETFL single-element file must hold Nrho + 2 Nr values "
                             "one per line")
        F = values[:Nrho].copy()
        rho = values[Nrho:Nrho + Nr].copy()
        phir = values[Nrho + Nr:].copy()
    else:
        rows = [ln.split() for ln in lines[6:] if ln.strip()]
        if any(len(r) != 5 for r in rows) or 5 * len(rows) != Nrho + 2 * Nr \
                or Nrho % 5 or Nr % 5:
            raise ValueError(f"{path}: FS file must hold Nrho + 2 Nr values, five a line, "
                             "with Nrho and Nr multiples of 5")
        F = values[:Nrho].copy()
        rho = values[Nrho:Nrho + Nr].copy()
        phir = values[Nrho + Nr:].copy()
        for arr, what in ((rho, "rho"), (phir, "r*phi")):
            if set(arr[:_FS_NFILL]) != {_FS_FILL}:
                raise ValueError(f"{path}: the first {_FS_NFILL} {what} values must be the "
                                 f"fill value {_FS_FILL:g}")
            arr[:_FS_NFILL] = arr[_FS_NFILL]
    F *= U.ev
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = phir / r_grid_a * U.ev
    phi[0] = phi[1]
    dr = dr_a * U.angstrom
    r_grid = r_grid_a * U.angstrom
    drho = drho_a / U.angstrom ** 3
    rho /= U.angstrom ** 3
    rho_grid = rho_grid_a / U.angstrom ** 3
    dphi, d2phi = derivatives(r_grid, phi)
    drh, d2rh = derivatives(r_grid, rho)
    dF, d2F = derivatives(rho_grid, F)
    el = SetflElement(name, Z, mass, a_lat, F, dF, d2F, rho, drh, d2rh)
    return SetflData(path=path, elements=[el], pairs=[SetflPair(0, 0, phi, dphi, d2phi)],
                     Nrho=Nrho, drho=drho, Nr=Nr, dr=dr, rcut=float(r_grid[-1]),
                     rhomax=float(rho_grid[-1]), r_grid=r_grid, rho_grid=rho_grid,
                     reader="single")
