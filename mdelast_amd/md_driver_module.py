"""Python mirror of the reference's pybind11 module ``md_driver_module``.

Same names, constructor signatures, attributes and enum values as
src/md_driver_module.cpp:16-224 of mdelast v1.2, backed by the C-ABI of
include/mdelast.h (libmdelast_hip.so). The MD state lives on the GPU; the
attributes are copies read back from it, as pybind11 returns copies.

Differences a caller can see:

* FATAL conditions raise :class:`MDDriverError` (the reference prints and calls
  ``exit(1)``); the message is the reference's FATAL text.
* ``Morse`` and ``LennardJones`` run in an ``MDDriver`` as pair potentials
  without EAM, in FP64 (the reference also allows them beside EAM tables).
* ``calculate_pair_correlation*`` count the pairs on the device (the same
  counts and normalisation as the reference's all-pairs loop).
* extra, optional: ``MDDriver(..., precision="fp64", lanes_per_atom=0)``,
  ``MDDriver.run(...)`` (many steps in one call, records returned as an
  array), ``*_array()`` getters returning numpy arrays, ``kernel_times()``.
"""

from __future__ import annotations

import ctypes as C
import enum
import math
import sys
from typing import Iterable, List

import numpy as np

from . import _native as nat
from . import units as _U

VERBOSE = True   # print the constructor banner and the every-100-cycles line, as the reference


class MDDriverError(nat.MdlError):
    """A FATAL condition of the driver (the reference exits the process)."""


# ---------------------------------------------------------------------------
# Vector3d (vector3d.hpp, md_driver_module.cpp:21-31)
# ---------------------------------------------------------------------------
class Vector3d:
    __slots__ = ("x", "y", "z")

    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def length2(self) -> float:
        return self.x * self.x + self.y * self.y + self.z * self.z

    def length(self) -> float:
        return math.sqrt(self.length2())

    def __repr__(self):
        return f"({self.x:f},{self.y:f},{self.z:f})"


def _vec_array(vs) -> np.ndarray:
    if isinstance(vs, np.ndarray):
        return np.ascontiguousarray(vs, dtype=float).reshape(-1, 3)
    return np.array([[v.x, v.y, v.z] for v in vs], dtype=float).reshape(-1, 3)


def _to_vectors(a: np.ndarray) -> List[Vector3d]:
    return [Vector3d(*row) for row in a.tolist()]


# ---------------------------------------------------------------------------
# Units (units.hpp, md_driver_module.cpp:196-211)
# ---------------------------------------------------------------------------
class Units:
    k_boltz = _U.k_boltz
    Navogadro = _U.Navogadro
    ev = _U.ev
    joule = _U.joule
    angstrom = _U.angstrom
    nm = _U.nm
    bar = _U.bar
    kbar = _U.kbar
    atm = _U.atm
    Pa = _U.Pa
    GPa = _U.GPa
    mus = _U.mus
    ns = _U.ns
    ps = _U.ps
    fs = _U.fs


# ---------------------------------------------------------------------------
# Potentials (potential.hpp, eam.hpp, poly_table.hpp; md_driver_module.cpp:36-88)
# ---------------------------------------------------------------------------
class PotentialValue:
    __slots__ = ("V", "dVdr", "d2Vdr2")

    def __init__(self, V, dVdr, d2Vdr2):
        self.V, self.dVdr, self.d2Vdr2 = float(V), float(dVdr), float(d2Vdr2)


_INTERP = {"quintic": nat.QUINTIC, "bspline": nat.BSPLINE}


class _Table:
    """One tabulated function on the grid x_k = k h (an mdl_table handle)."""

    def __init__(self, interpolation: str, h: float, f, f1, f2):
        if interpolation not in _INTERP:
            raise ValueError(f"FATAL: unknown table interpolation '{interpolation}' "
                             "(use 'quintic' or 'bspline')")
        arrs = [np.ascontiguousarray(a, dtype=float) for a in (f, f1, f2)]
        if not (len(arrs[0]) == len(arrs[1]) == len(arrs[2])):
            raise ValueError("FATAL: QuinticTable needs >= 2 nodes, equal array sizes, and h > 0")
        self.name = interpolation
        self.h = float(h)
        self.n_nodes = len(arrs[0])
        h_ = C.c_void_p()
        nat.check(nat.lib().mdl_table_create(_INTERP[interpolation], self.h, self.n_nodes,
                                             *(nat.dptr(a) for a in arrs), C.byref(h_)))
        self.handle = h_

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        lib_ = getattr(nat, "_lib", None) if nat is not None else None
        if getattr(self, "handle", None) and lib_ is not None:
            lib_.mdl_table_destroy(self.handle)
            self.handle = None

    def coefficients(self) -> np.ndarray:
        n = C.c_int64(0)
        s = C.c_int32(0)
        nat.check(nat.lib().mdl_table_coefficients(self.handle, None, C.byref(n), C.byref(s)))
        out = np.zeros((n.value, s.value))
        nat.check(nat.lib().mdl_table_coefficients(self.handle, nat.dptr(out), None, None))
        return out

    def eval(self, x):
        """value, d1, d2 at the points x (evaluated on the GPU)."""
        xs = np.ascontiguousarray(np.atleast_1d(x), dtype=float)
        v, d1, d2 = (np.zeros_like(xs) for _ in range(3))
        nat.check(nat.lib().mdl_table_eval(self.handle, len(xs), nat.dptr(xs), nat.dptr(v),
                                           nat.dptr(d1), nat.dptr(d2)))
        return v, d1, d2


class Potential:
    """Base of the pair potentials (potential.hpp:16-27)."""

    def __init__(self, rcut: float):
        self.rcut = float(rcut)

    def calculate(self, r: float) -> PotentialValue:
        raise NotImplementedError


def _pow6(s: float) -> float:
    """s^6 by binary powering, ((s s) s)^2, as boost::math::pow<6> forms it."""
    s3 = s * s * s
    return s3 * s3


class _AnalyticTable(_Table):
    """An analytic pair form as an mdl_table handle (mdl_table_create_analytic):
    the driver evaluates the formula on the GPU in place of a table lookup."""

    def __init__(self, kind: int, params):
        self.name = {nat.MORSE: "morse", nat.LJ: "lennard_jones"}[kind]
        self.h, self.n_nodes = 0.0, 0
        p = np.ascontiguousarray(params, dtype=float)
        h_ = C.c_void_p()
        nat.check(nat.lib().mdl_table_create_analytic(kind, nat.dptr(p), C.byref(h_)))
        self.handle = h_


class Morse(Potential):
    """Analytic Morse pair potential (potential.hpp:29-51). ``calculate`` evaluates
    on the host as the reference does; an ``MDDriver`` evaluates the same formula
    on the GPU (pair potentials without EAM, FP64)."""

    def __init__(self, rcut, D, alpha, r0):
        super().__init__(rcut)
        self.D, self.alpha, self.r0 = float(D), float(alpha), float(r0)
        self.rmin_fatal = 0.0
        self._table = None

    @property
    def table(self) -> _Table:
        if self._table is None:
            self._table = _AnalyticTable(nat.MORSE, [self.D, self.alpha, self.r0])
        return self._table

    def calculate(self, r):
        if r > self.rcut:
            return PotentialValue(0.0, 0.0, 0.0)
        e1 = math.exp(-self.alpha * (r - self.r0))
        e2 = e1 * e1
        return PotentialValue(self.D * (e2 - 2.0 * e1), -2.0 * self.D * self.alpha * (e2 - e1),
                              2.0 * self.alpha * self.alpha * self.D * (2 * e2 - e1))


class LennardJones(Potential):
    """Analytic Lennard-Jones pair potential, shifted to 0 at rcut (potential.hpp:53-75).
    ``calculate`` evaluates on the host as the reference does; an ``MDDriver``
    evaluates the same formula on the GPU (pair potentials without EAM, FP64)."""

    def __init__(self, rcut, epsilon, sigma):
        super().__init__(rcut)
        self.epsilon, self.sigma = float(epsilon), float(sigma)
        v6 = _pow6(self.sigma / self.rcut)
        self.vcut = 4.0 * self.epsilon * (v6 * v6 - v6)
        self.rmin_fatal = 0.0
        self._table = None

    @property
    def table(self) -> _Table:
        if self._table is None:
            self._table = _AnalyticTable(nat.LJ, [self.epsilon, self.sigma, self.rcut])
        return self._table

    def calculate(self, r):
        if r > self.rcut:
            return PotentialValue(0.0, 0.0, 0.0)
        v6 = _pow6(self.sigma / r)
        v12 = v6 * v6
        return PotentialValue(4.0 * self.epsilon * (v12 - v6) - self.vcut,
                              24.0 * self.epsilon * (v6 - 2.0 * v12) / r,
                              24.0 * self.epsilon * (26.0 * v12 - 7.0 * v6) / (r * r))


class TablePotential(Potential):
    """Tabulated pair potential phi(r) (potential.hpp:77-104)."""

    def __init__(self, rcut, rmin_fatal, dr_interp, V, dVdr, d2Vdr2, interpolation="quintic"):
        super().__init__(rcut)
        self.rmin_fatal = float(rmin_fatal)
        self.table = _Table(interpolation, dr_interp, V, dVdr, d2Vdr2)

    @property
    def interpolation(self) -> str:
        return self.table.name

    def calculate(self, r):
        r = float(r)
        if r < self.rmin_fatal:
            print(f"FATAL: PotentialTable dr <rmin_fatal {r:g} {self.rmin_fatal:g}")
        if r > self.rcut:
            return PotentialValue(0.0, 0.0, 0.0)
        v, d1, d2 = self.table.eval(r)
        return PotentialValue(v[0], d1[0], d2[0])

    def calculate_array(self, r):
        """(V, dV/dr, d2V/dr2) arrays at many r (extension)."""
        r = np.ascontiguousarray(r, dtype=float)
        v, d1, d2 = self.table.eval(r)
        out = r > self.rcut
        for a in (v, d1, d2):
            a[out] = 0.0
        return v, d1, d2


class EAM:
    """Density rho(r) and embedding F(rho) of one species (eam.hpp:10-71)."""

    def __init__(self, rcut, rhomax_fatal, dr_interp, drho_interp, rho, drhodr, d2rhodr2, F,
                 dFdrho, d2Fdrho2, interpolation="quintic"):
        self.rcut = float(rcut)
        self.rhomax_fatal = float(rhomax_fatal)
        self.table_rho = _Table(interpolation, dr_interp, rho, drhodr, d2rhodr2)
        self.table_F = _Table(interpolation, drho_interp, F, dFdrho, d2Fdrho2)

    @property
    def interpolation(self) -> str:
        return self.table_rho.name

    def _r(self, r, k):
        r = float(r)
        if r > self.rcut:
            return 0.0
        return float(self.table_rho.eval(r)[k][0])

    def rho(self, r):
        return self._r(r, 0)

    def drhodr(self, r):
        return self._r(r, 1)

    def d2rhodr2(self, r):
        return self._r(r, 2)

    def _F(self, rho, k, name):
        rho = float(rho)
        if rho > self.rhomax_fatal:
            raise MDDriverError(5, f"FATAL: EAM {name} rho >rhomax_fatal {rho:g} "
                                   f"{self.rhomax_fatal:g}")
        return float(self.table_F.eval(rho)[k][0])

    def F(self, rho):
        return self._F(rho, 0, "F")

    def dFdrho(self, rho):
        return self._F(rho, 1, "dFdrho")

    def d2Fdrho2(self, rho):
        return self._F(rho, 2, "d2Fdrho2")


class PotentialVector(list):
    """Opaque std::vector<shared_ptr<Potential>> (md_driver_module.cpp:87)."""


class EAMVector(list):
    """Opaque std::vector<shared_ptr<EAM>> (md_driver_module.cpp:88)."""


# ---------------------------------------------------------------------------
# Box, Thermostat, Barostat (md_driver.hpp:28-37, thermostat.hpp, barostat.hpp)
# ---------------------------------------------------------------------------
class Box:
    def __init__(self, h):
        self._h = np.array(h, dtype=float).reshape(3, 3)

    @property
    def h(self) -> np.ndarray:
        return self._h.copy()

    def volume(self) -> float:
        h = self._h
        # Eigen's 3x3 cofactor determinant (diagonal box: h00 * (h11 * h22))
        return float(h[0, 0] * (h[1, 1] * h[2, 2] - h[1, 2] * h[2, 1])
                     - h[0, 1] * (h[1, 0] * h[2, 2] - h[1, 2] * h[2, 0])
                     + h[0, 2] * (h[1, 0] * h[2, 1] - h[1, 1] * h[2, 0]))

    def length(self) -> float:
        return float(np.trace(self._h) / 3.0)


class Thermostat:
    def __init__(self, temperature, q, p_eta, Nf):
        self.temperature = float(temperature)
        self.q = [float(v) for v in q]
        self.p_eta = [float(v) for v in p_eta]
        self.eta = [0.0] * len(self.p_eta)
        self.nchain = len(self.p_eta)
        self.Nf = float(Nf)
        if self.nchain == 0 or len(self.q) != self.nchain:
            raise ValueError("Thermostat: q and p_eta need the same non-zero length")


class Barostat:
    def __init__(self, pressure, w, p_epsilon, thermostat: Thermostat):
        self.pressure = float(pressure)
        self.w = float(w)
        self.epsilon = 0.0
        self.p_epsilon = float(p_epsilon)
        self.thermostat = thermostat


class BarostatFlex:
    """Placeholder, as in the reference (barostat_flex.hpp): NPT_FLEX is rejected."""

    def __init__(self, pressure, w, pg, thermostat: Thermostat):
        self.pressure = float(pressure)
        self.w = float(w)
        self.pg = np.array(pg, dtype=float).reshape(3, 3)
        self.thermostat = thermostat


# ---------------------------------------------------------------------------
# MDDriver (md_driver.hpp:18-168; md_driver_module.cpp:93-191)
# ---------------------------------------------------------------------------
class _Ensamble(enum.IntEnum):
    NVE = 0
    NVT = 1
    NPT = 2
    NPT_FLEX = 3


class _ThermostatCoupling(enum.IntEnum):
    THERMOSTAT_KINETIC = 0
    THERMOSTAT_CONFIGURATIONAL = 1
    THERMOSTAT_AVERAGE = 2


_PRECISION = {"fp64": nat.FP64, "mixed": nat.MIXED}


class MDDriver:
    Ensamble = _Ensamble
    ThermostatCoupling = _ThermostatCoupling
    NVE = _Ensamble.NVE
    NVT = _Ensamble.NVT
    NPT = _Ensamble.NPT
    NPT_FLEX = _Ensamble.NPT_FLEX
    THERMOSTAT_KINETIC = _ThermostatCoupling.THERMOSTAT_KINETIC
    THERMOSTAT_CONFIGURATIONAL = _ThermostatCoupling.THERMOSTAT_CONFIGURATIONAL
    THERMOSTAT_AVERAGE = _ThermostatCoupling.THERMOSTAT_AVERAGE

    def __init__(self, positions, velocities, pot, eam, ensamble, thermostat_coupling,
                 thermostat: Thermostat, barostat: Barostat, box: Box, rg, Nf, mass, types,
                 dt_global, neighbors_list, r_skin, fatal_conservation, *,
                 precision: str = "fp64", lanes_per_atom: int = 0):
        self.handle = None
        pos = _vec_array(positions)
        vel = _vec_array(velocities)
        N = len(pos)
        if len(vel) != N:
            raise ValueError("positions and velocities differ in length")
        pots = list(pot)
        eams = list(eam)
        mass = np.ascontiguousarray(mass, dtype=float)
        ntypes = len(mass)
        types_a = np.ascontiguousarray(types, dtype=np.uint32)
        if len(types_a) != N:
            raise ValueError("types has a different length than positions")
        if len(pots) != ntypes * (ntypes + 1) // 2:
            raise ValueError(f"{len(pots)} pair potentials for {ntypes} species "
                             f"(need {ntypes * (ntypes + 1) // 2})")
        if eams and len(eams) != ntypes:
            raise ValueError(f"{len(eams)} EAM objects for {ntypes} species")
        for p in pots:
            if not isinstance(p, (TablePotential, Morse, LennardJones)):
                raise TypeError(f"{type(p).__name__} is not a pair potential")
        hb = box.h if isinstance(box, Box) else np.asarray(box, dtype=float)
        if np.any(hb != np.diag(np.diag(hb))):
            raise MDDriverError(12, "the box must be diagonal (md_driver.cpp:60)")
        if precision not in _PRECISION:
            raise ValueError(f"precision must be one of {tuple(_PRECISION)}")
        self._rg = [float(x) for x in rg]
        self._mass = [float(m) for m in mass]
        self._types = [int(t) for t in types_a]
        self._Nf = float(Nf)
        self._dt = float(dt_global)
        self._fatal = float(fatal_conservation)
        self._ensamble = _Ensamble(int(ensamble))
        self._coupling = _ThermostatCoupling(int(thermostat_coupling))
        self._thermostat_cfg = thermostat
        self._barostat_cfg = barostat
        self._box0 = hb.copy()
        self.N = N
        self._keep = keep = {}
        keep["pos"], keep["vel"], keep["types"], keep["mass"] = pos, vel, types_a, mass
        th, ba, bth = thermostat, barostat, barostat.thermostat
        keep["thq"] = np.ascontiguousarray(th.q, dtype=float)
        keep["thp"] = np.ascontiguousarray(th.p_eta, dtype=float)
        keep["bq"] = np.ascontiguousarray(bth.q, dtype=float)
        keep["bp"] = np.ascontiguousarray(bth.p_eta, dtype=float)
        keep["prc"] = np.array([p.rcut for p in pots], dtype=float)
        keep["prm"] = np.array([p.rmin_fatal for p in pots], dtype=float)
        keep["pair"] = (C.c_void_p * len(pots))(*[p.table.handle.value for p in pots])
        nt = max(1, len(eams))
        keep["rho"] = (C.c_void_p * nt)(*[e.table_rho.handle.value for e in eams])
        keep["F"] = (C.c_void_p * nt)(*[e.table_F.handle.value for e in eams])
        keep["erc"] = np.array([e.rcut for e in eams] or [0.0], dtype=float)
        keep["erm"] = np.array([e.rhomax_fatal for e in eams] or [0.0], dtype=float)
        self._tables = (pots, eams)      # the handles must outlive the driver
        cfg = nat.Config()
        cfg.N, cfg.ntypes = N, ntypes
        cfg.positions, cfg.velocities = nat.dptr(pos), nat.dptr(vel)
        cfg.types = types_a.ctypes.data_as(C.POINTER(C.c_uint32))
        cfg.mass = nat.dptr(mass)
        cfg.box = (C.c_double * 3)(*[float(hb[k, k]) for k in range(3)])
        cfg.ensemble, cfg.coupling = int(ensamble), int(thermostat_coupling)
        cfg.th_T, cfg.th_Nf, cfg.th_M = th.temperature, th.Nf, len(keep["thq"])
        cfg.th_q, cfg.th_p_eta = nat.dptr(keep["thq"]), nat.dptr(keep["thp"])
        cfg.bar_P, cfg.bar_W, cfg.bar_p_eps = ba.pressure, ba.w, ba.p_epsilon
        cfg.bth_T, cfg.bth_Nf, cfg.bth_M = bth.temperature, bth.Nf, len(keep["bq"])
        cfg.bth_q, cfg.bth_p_eta = nat.dptr(keep["bq"]), nat.dptr(keep["bp"])
        cfg.Nf, cfg.dt = float(Nf), float(dt_global)
        cfg.neighbors_list = int(bool(neighbors_list))
        cfg.r_skin = float(r_skin)
        cfg.fatal_conservation = float(fatal_conservation)
        cfg.pair = C.cast(keep["pair"], C.POINTER(C.c_void_p))
        cfg.pair_rcut, cfg.pair_rmin = nat.dptr(keep["prc"]), nat.dptr(keep["prm"])
        cfg.has_eam = int(len(eams) > 0)
        cfg.rho_tab = C.cast(keep["rho"], C.POINTER(C.c_void_p))
        cfg.F_tab = C.cast(keep["F"], C.POINTER(C.c_void_p))
        cfg.eam_rcut, cfg.eam_rhomax = nat.dptr(keep["erc"]), nat.dptr(keep["erm"])
        cfg.precision = _PRECISION[precision]
        cfg.lanes_per_atom = int(lanes_per_atom)
        self.precision = precision
        if VERBOSE:
            print("EAM object was given to MDDriver constructor" if eams else
                  "EAM object was not given to MDDriver constructor")
        h = C.c_void_p()
        st = nat.lib().mdl_driver_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise MDDriverError(st, nat.lib().mdl_last_error().decode(errors="replace"))
        self.handle = h
        self._sample = True
        self._cache = None
        self._gr = [0.0] * len(self._rg)
        self._gr_partial = [[0.0] * len(self._rg) for _ in range(ntypes * (ntypes + 1) // 2)]
        if VERBOSE:
            print(f"###############\nMDDriver constructor:\nN={N}\nNf={self._Nf:g}\n"
                  f"Ntypes={ntypes}\ndt_global={self._dt:g}\nEsamble: {self._ensamble.name}\n"
                  "###############")

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        lib_ = getattr(nat, "_lib", None) if nat is not None else None
        if getattr(self, "handle", None) and lib_ is not None:
            lib_.mdl_driver_destroy(self.handle)
            self.handle = None

    # -- calls -------------------------------------------------------------
    def _call(self, st):
        if st != 0:
            raise MDDriverError(st, nat.lib().mdl_last_error().decode(errors="replace"))

    def integrate(self):
        """One MD step (md_driver.cpp:563-601)."""
        self._cache = None
        try:
            self._call(nat.lib().mdl_driver_integrate(self.handle, 1))
        finally:
            self._cache = None
        if VERBOSE and (self.cycle - 1) % 100 == 0:
            self._print_line(self.cycle - 1)

    def calculate_neighbors(self):
        """Rebuild the neighbour list before the next force evaluation."""
        self._call(nat.lib().mdl_driver_calculate_neighbors(self.handle))

    def run(self, nsteps: int, sample_stride: int = 1, rebuild_every: int = 0,
            records: bool = True) -> np.ndarray:
        """nsteps steps in one call (the loop of MolecularDynamics.run without Python
        per step); returns one MDL_S_* record per sampled step (extension)."""
        cycle0 = self.cycle
        nsamp = sum(1 for k in range(1, nsteps + 1)
                    if (cycle0 + k) % sample_stride == 0) if records else 0
        buf = np.zeros((max(nsamp, 1), nat.NRECORD))
        n = C.c_int64(0)
        self._cache = None
        try:
            self._call(nat.lib().mdl_driver_run(self.handle, int(nsteps), int(sample_stride),
                                                int(rebuild_every),
                                                nat.dptr(buf) if records else None,
                                                C.byref(n)))
        finally:
            self._cache = None
        return buf[:n.value]

    @property
    def graph_steps(self) -> int:
        """Chunks of this many steps replayed as one captured graph in run() (0: off)."""
        return getattr(self, "_graph_steps", 0)

    @graph_steps.setter
    def graph_steps(self, steps: int):
        self._call(nat.lib().mdl_driver_set_graph_steps(self.handle, int(steps)))
        self._graph_steps = int(steps)

    def kernel_times(self):
        """Average ms of the density pass, force pass and whole step over the last run()."""
        out = np.zeros(3)
        self._call(nat.lib().mdl_driver_kernel_times(self.handle, nat.dptr(out)))
        return {"density_ms": out[0], "force_ms": out[1], "step_ms": out[2]}

    def _pair_correlation(self, mode):
        rg = np.asarray(self._rg, dtype=np.float64)
        gr = np.zeros(len(rg))
        gp = np.zeros((len(self._gr_partial), len(rg)))
        self._call(nat.lib().mdl_driver_pair_correlation(self.handle, nat.dptr(rg), len(rg),
                                                          mode, nat.dptr(gr), nat.dptr(gp)))
        if mode != 2:
            self._gr = gr.tolist()
        if mode != 0:
            self._gr_partial = gp.tolist()

    def calculate_pair_correlation(self):
        """g(r) of the current configuration into ``gr`` (md_driver.cpp:732-752)."""
        self._pair_correlation(0)

    def calculate_pair_correlations(self):
        """``gr`` and the per-species-pair ``gr_partial`` (md_driver.cpp:754-810)."""
        self._pair_correlation(1)

    def calculate_partial_pair_correlation(self):
        """``gr_partial`` only (md_driver.cpp:812-863)."""
        self._pair_correlation(2)

    # -- state -------------------------------------------------------------
    def _scalars(self) -> np.ndarray:
        if self._cache is None:
            out = np.zeros(nat.NSCALARS)
            self._call(nat.lib().mdl_driver_get_scalars(self.handle, nat.dptr(out)))
            self._cache = out
        return self._cache

    def scalars_array(self) -> np.ndarray:
        return self._scalars().copy()

    def _atoms(self, what: str) -> np.ndarray:
        n3 = what in ("positions", "coordinates", "velocities", "forces")
        out = np.zeros(3 * self.N if n3 else self.N)
        self._call(nat.lib().mdl_driver_get_atoms(self.handle, nat.A_INDEX[what], nat.dptr(out)))
        return out.reshape(self.N, 3) if n3 else out

    def positions_array(self):
        return self._atoms("positions")

    def coordinates_array(self):
        return self._atoms("coordinates")

    def velocities_array(self):
        return self._atoms("velocities")

    def forces_array(self):
        return self._atoms("forces")

    @property
    def positions(self):
        return _to_vectors(self._atoms("positions"))

    @property
    def coordinates(self):
        return _to_vectors(self._atoms("coordinates"))

    @property
    def velocities(self):
        return _to_vectors(self._atoms("velocities"))

    @property
    def forces(self):
        return _to_vectors(self._atoms("forces"))

    @property
    def rho_eam(self):
        return self._atoms("rho").tolist()

    @property
    def energies_pot(self):
        return self._atoms("epot").tolist()

    @property
    def energies_kin(self):
        return self._atoms("ekin").tolist()

    @property
    def ensamble(self):
        return self._ensamble

    @property
    def thermostat_coupling(self):
        return self._coupling

    def _chain(self, which, cfg: Thermostat) -> Thermostat:
        M = len(cfg.q)
        p, e = np.zeros(M), np.zeros(M)
        self._call(nat.lib().mdl_driver_get_chain(self.handle, which, nat.dptr(p), nat.dptr(e)))
        t = Thermostat(cfg.temperature, cfg.q, p.tolist(), cfg.Nf)
        t.eta = e.tolist()
        return t

    @property
    def thermostat(self) -> Thermostat:
        return self._chain(0, self._thermostat_cfg)

    @property
    def barostat(self) -> Barostat:
        s = self._scalars()
        b = Barostat(self._barostat_cfg.pressure, self._barostat_cfg.w,
                     s[nat.S_INDEX["bar_peps"]], self._chain(1, self._barostat_cfg.thermostat))
        b.epsilon = float(s[nat.S_INDEX["bar_eps"]])
        return b

    @property
    def box_initial(self) -> Box:
        return Box(self._box0)

    @property
    def box(self) -> Box:
        s = self._scalars()
        return Box(np.diag([s[nat.S_INDEX["box0"]], s[nat.S_INDEX["box1"]],
                            s[nat.S_INDEX["box2"]]]))

    def _s(self, key):
        return float(self._scalars()[nat.S_INDEX[key]])

    epot_eam_atom = property(lambda self: self._s("epot_eam_atom"))
    epot_atom = property(lambda self: self._s("epot_atom"))
    ekin_atom = property(lambda self: self._s("ekin_atom"))
    etot_atom = property(lambda self: self._s("etot_atom"))
    etot_cons = property(lambda self: self._s("etot_cons"))
    etot_cons_err = property(lambda self: self._s("etot_cons_err"))
    temperature_kin = property(lambda self: self._s("temperature_kin"))
    temperature_conf_inverse = property(lambda self: self._s("temperature_conf_inverse"))
    total_laplacian = property(lambda self: self._s("total_laplacian"))
    total_force2 = property(lambda self: self._s("total_force2"))
    pressure = property(lambda self: self._s("pressure"))
    time = property(lambda self: self._s("time"))

    @property
    def cycle(self) -> int:
        return int(self._scalars()[nat.S_INDEX["cycle"]])

    @property
    def born_stress_tensor(self) -> np.ndarray:
        return self._scalars()[nat.S_BORN_STRESS:nat.S_BORN_STRESS + 6].copy()

    @property
    def stress_tensor(self) -> np.ndarray:
        return self._scalars()[nat.S_STRESS:nat.S_STRESS + 6].copy()

    @property
    def born_elasticity_tensor(self) -> np.ndarray:
        return self._scalars()[nat.S_BORN_ELAST:nat.S_BORN_ELAST + 36].reshape(6, 6).copy()

    @property
    def dt_global(self):
        return self._dt

    @property
    def rg(self):
        return list(self._rg)

    @property
    def gr(self):
        return list(self._gr)

    @property
    def gr_partial(self):
        return [list(g) for g in self._gr_partial]

    @property
    def fatal_conservation(self):
        return self._fatal

    @property
    def sample_observables(self) -> bool:
        return self._sample

    @sample_observables.setter
    def sample_observables(self, flag):
        self._sample = bool(flag)
        self._call(nat.lib().mdl_driver_set_sample_observables(self.handle, int(self._sample)))

    @property
    def mass(self):
        return list(self._mass)

    @property
    def types(self):
        return list(self._types)

    @property
    def Nparticles_per_type(self):
        return np.bincount(self._types, minlength=len(self._mass)).tolist()

    @property
    def neighbor_list_builds(self) -> int:
        return int(self._scalars()[nat.S_NLIST_BUILDS])

    @property
    def neighbor_list_entries(self) -> int:
        return int(self._scalars()[nat.S_NLIST_ENTRIES])

    def pairs_within_rc(self) -> np.ndarray:
        """Per atom, the pairs within r_c at the last density pass (both directions;
        the pairs the force pass evaluates), a diagnostic without counterpart in
        the reference. FP64: the length of the compacted list the force pass walks;
        mixed: counted over the mid list with the passes' own r_c decision."""
        return self._atoms("pairs_rc").astype(np.int64)

    def mid_list_lengths(self) -> np.ndarray:
        """Per atom, the entries of the mid list the density pass scans (pairs within
        r_c + 0.3 A at its last filtering; r_c + 0.15 A in mixed precision, where
        the force pass scans it too), a diagnostic."""
        return self._atoms("pairs_mid").astype(np.int64)

    @property
    def lanes_per_atom_used(self) -> int:
        """Lanes per atom of the pair passes, as given or as the autotune chose."""
        return int(self._scalars()[nat.S_LANES_PER_ATOM])

    @property
    def mid_list_builds(self) -> int:
        """Filterings of the mid list that pass A scans (pairs within r_c + 0.3 A),
        a diagnostic without counterpart in the reference."""
        return int(self._scalars()[nat.S_NMID_BUILDS])

    @property
    def nhc_general_path_count(self):
        """Chain propagations that took at least one factor in the general form
        (range-reduced exp): those with a chain argument dt p/q beyond 1/4, or
        following one (diagnostic; no counterpart in the reference)."""
        return int(self._scalars()[nat.S_NHC_GENERAL])

    @property
    def warnings(self):
        s = self._scalars()
        return {"virial": int(s[nat.S_WARN_VIRIAL]), "rmin": int(s[nat.S_WARN_RMIN])}

    def _print_line(self, cycle):
        # md_driver.cpp:703-720
        s = self._scalars()
        th = self._thermostat_cfg
        V = self.box.volume()
        V0 = self.box_initial.volume()
        tci = s[nat.S_INDEX["temperature_conf_inverse"]]
        print(f"{cycle} t={s[nat.S_INDEX['time']]:1.3E} dt={self._dt:g}: "
              f"e={s[nat.S_INDEX['etot_atom']] / _U.ev:g} ep={s[nat.S_INDEX['epot_atom']] / _U.ev:g} "
              f"eEMB={s[nat.S_INDEX['epot_eam_atom']] / _U.ev:g} "
              f"ek={s[nat.S_INDEX['ekin_atom']] / _U.ev:g}[eV] "
              f"er0={s[nat.S_INDEX['etot_cons_err']]:1.1E} "
              f"Tk={s[nat.S_INDEX['temperature_kin']]:g}[K] "
              f"Tc={(1.0 / tci) if tci else float('inf'):g}[K] Tth={th.temperature:g}[K] "
              f"V/V0={V / V0:g} L={self.box.length() / _U.angstrom:g}[A] "
              f"P={s[nat.S_INDEX['pressure']] / _U.GPa:g} "
              f"Pb={self._barostat_cfg.pressure / _U.GPa:g}[GPa]", file=sys.stdout)
