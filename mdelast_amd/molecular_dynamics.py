"""The production caller of the MD step and its elasticity estimator.

Mirror of ``src/python/molecular_dynamics.py`` of mdelast v1.2 (its
``MolecularDynamics`` and ``Material`` classes) over this package's
``md_driver_module``:

- set-up as the reference (``molecular_dynamics.py:317-642``): keyword
  parameters, FCC lattice or explicit lattice, Maxwell-Boltzmann velocities with
  the centre-of-mass momentum removed and an exact rescale, Nose-Hoover chain
  masses Q_0 = N_f tau^2 kT, Q_j = tau^2 kT, barostat mass W = N_f tau_P^2 kT, and
  the random-number order v_x, v_y, v_z -> p_eta -> p_eta(barostat) -> p_eps;
- the step loop (``:1240-1285``): list rebuilds every ``cycles_neigbors`` steps,
  sampling every ``sample_stride`` steps, trajectory dumps, g(r) (and the partial
  g(r) of each species pair when there are several species) every ``gr_stride``
  steps and thermal averages every ``plot_freq`` steps after thermalisation. The loop runs on the device in
  stretches between those events (``MDDriver.run``), one Python call per stretch
  instead of one per step;
- ``perform_thermal_average`` (``:664-1170``): the stress-fluctuation estimator of
  the elastic constants, C_ij = <C^B_ij> - (<V>/kT)(<s_i s_j> - <s_i><s_j>) +
  kinetic term (+ bulk modulus term in NPT), with s = sigma^B in NVT and
  s = sigma^B - (N kT / V) delta in NPT; compressibility, bulk modulus and
  thermal expansion in NPT; heat capacities in NVE/NVT.

Not here: plots and pickle files. The series and thermal results go to
``series.npz`` with the keys of the reference's ``series.h5`` (``save_series``,
``series_io``), after every thermal average as the reference writes its file;
``save_json`` writes the thermal results alone.
"""

from __future__ import annotations

import json
import logging
import os
import time
from os import path

import numpy as np

from . import _native as nat
from .md_driver_module import (MDDriver, Box, Vector3d, Thermostat, Barostat, Units,
                               PotentialVector, EAMVector)

logger = logging.getLogger("MD")


# ---------------------------------------------------------------------------
# time series (averages.py, series.py of the reference: population statistics)
# ---------------------------------------------------------------------------
def get_mean_std(values):
    v = np.asarray(values, dtype=float)
    mean = np.average(v)
    return mean, np.sqrt(np.average((v - mean) ** 2))


def get_mean_std_cum(values):
    v = np.asarray(values, dtype=float)
    n = np.arange(1, len(v) + 1)
    mean_cum = np.cumsum(v) / n
    # (the variance from running sums can round to a tiny negative value)
    std_cum = np.sqrt(np.maximum(np.cumsum(v * v) / n - mean_cum ** 2, 0.0))
    return mean_cum, std_cum


class TimeSeries:
    """Scalar samples with statistics over the production window."""

    def __init__(self):
        self.data = []
        self.mean = self.std = self.std_mean = self.mean_cum = self.std_cum = None

    def append(self, x):
        self.data.append(x)

    def set_averages(self, cycles_thermalization=0, cycles_thermalization_cum=0):
        self.cycles_thermalization = cycles_thermalization
        self.mean, self.std = get_mean_std(self.data[cycles_thermalization:])
        self.std_mean = self.std / abs(self.mean) if self.mean != 0 else np.inf
        self.mean_cum, self.std_cum = get_mean_std_cum(self.data[cycles_thermalization_cum:])


class TimeSeriesVector:
    def __init__(self, N):
        self.N = N
        self.data = [TimeSeries() for _ in range(N)]

    def append(self, vec):
        for s, v in zip(self.data, vec):
            s.append(v)

    def get_time_array(self):
        return np.array([s.data for s in self.data], dtype=float)

    def set_averages(self, cut=0, cut_cum=0):
        for s in self.data:
            s.set_averages(cut, cut_cum)


class TimeSeriesMatrix:
    def __init__(self, N):
        self.N = N
        self.data = [[TimeSeries() for _ in range(N)] for _ in range(N)]

    def append(self, mat):
        for i in range(self.N):
            for j in range(self.N):
                self.data[i][j].append(mat[i][j])

    def get_time_array(self):
        return np.array([[s.data for s in row] for row in self.data], dtype=float)

    def set_averages(self, cut=0, cut_cum=0):
        for row in self.data:
            for s in row:
                s.set_averages(cut, cut_cum)


class ThermalResults:
    """Namespace of the aggregate properties after thermalisation."""


class Series:
    SCALARS = ("temperature_kin", "temperature_conf_inverse", "total_laplacian", "total_force2",
               "pressure", "etot_atom", "epot_atom", "epot_eam_atom", "ekin_atom", "etot_cons",
               "etot_cons_err", "V_V0", "L_L0")

    def __init__(self):
        self.time = []
        for name in self.SCALARS:
            setattr(self, name, TimeSeries())
        self.stress_tensor = TimeSeriesVector(6)
        self.born_stress_tensor = TimeSeriesVector(6)
        self.born_elasticity_tensor = TimeSeriesMatrix(6)
        self.thermal_results = ThermalResults()


# ---------------------------------------------------------------------------
class Material:
    """A single- or multi-species material (molecular_dynamics.py:48-103)."""

    def __init__(self, potential, atomic_weight, atom_name, eam=None, a_lattice=None,
                 lattice=None, atom_type=None, box_length=None):
        self.potential = list(potential) if isinstance(potential, (list, tuple, np.ndarray)) \
            else [potential]
        self.atomic_weight = list(atomic_weight) \
            if isinstance(atomic_weight, (list, tuple, np.ndarray)) else [atomic_weight]
        self.atom_name = atom_name
        if eam is None:
            self.eam = None
        else:
            self.eam = list(eam) if isinstance(eam, (list, tuple, np.ndarray)) else [eam]
        self.a_lattice = a_lattice
        self.lattice = lattice
        if lattice is None:
            self.atom_type = [0]
            self.box_length = None
        else:
            if atom_type is None or len(atom_type) != len(lattice):
                raise ValueError("an explicit lattice needs one atom_type per site")
            if box_length is None:
                raise ValueError("an explicit lattice needs box_length")
            self.atom_type = atom_type
            self.box_length = box_length


def get_EAM_file(*, file_name, a_factor=1.0):
    """One-element EAM file as a ``Material`` with quintic tables
    (molecular_dynamics.py:123-234); ``get_potential.get_EAM_file`` takes the
    interpolation as well."""
    from .get_potential import get_EAM_file as _get
    return _get(file_name=file_name, a_factor=a_factor)


def get_Cu_Morse(a_factor=1.0):
    """Copper as the analytic Morse pair potential (molecular_dynamics.py:259-275):
    D = 0.3429 eV, alpha = 1.3588 / A, r0 = 2.866 A, cut at 10/alpha + r0."""
    from .md_driver_module import Morse
    D = 0.3429 * Units.ev
    alpha = 1.3588 / Units.angstrom
    r0 = 2.866 * Units.angstrom
    return Material(potential=Morse(10.0 / alpha + r0, D, alpha, r0), atomic_weight=63.546,
                    atom_name="Cu", a_lattice=a_factor * 3.615 * Units.angstrom)


def get_Cu_Table_Morse():
    """Unavailable, as in mdelast v1.2: the reference reads ``rcut`` from
    ``Material.potential``, which is a list, and raises AttributeError before it
    builds the table (molecular_dynamics.py:237-241, docs/API_REFERENCE.md:39-40).
    The same error is raised here. A tabulated Morse potential is built with
    ``TablePotential`` from ``Morse.calculate`` at the nodes."""
    potential = get_Cu_Morse().potential
    raise AttributeError(f"'{type(potential).__name__}' object has no attribute 'rcut'")


def get_Ar_LJ(*, epsilon=None, sigma=None, rcut_fac=3.5):
    """Argon as the Lennard-Jones pair potential, epsilon = 0.0104 eV and
    sigma = 3.405 A by default, cut at rcut_fac sigma (molecular_dynamics.py:278-291)."""
    from .md_driver_module import LennardJones
    epsilon = 0.0104 * Units.ev if epsilon is None else epsilon
    sigma = 3.405 * Units.angstrom if sigma is None else sigma
    return Material(potential=LennardJones(rcut_fac * sigma, epsilon, sigma),
                    atomic_weight=39.948, atom_name="Ar")


def npt_to_nvt_material(npt) -> "Material":
    """The NVT stage's material after an NPT stage, handed over as the
    reference's two-stage calculation does (sample_case/run.py:251-283): the
    positions of the NPT run's last trajectory dump, scaled by the production
    mean (V/V0)^(1/3), in the production-mean box (the NPT stage's initial box
    times the same factor). ``npt`` is the finished NPT ``MolecularDynamics``;
    the result keeps its potentials, EAM tables, atomic weights and species
    names, and takes each atom's species from the dump."""
    import glob
    dumps = sorted(glob.glob(path.join(npt.dump_dir, "*.xyz")),
                   key=lambda p: int(path.splitext(path.basename(p))[0]))
    if not dumps:
        raise ValueError(f"no trajectory dump in {npt.dump_dir}")
    mat = npt.material
    species = {name: k for k, name in enumerate(mat.atom_name.split("-"))}
    scale = npt.series.V_V0.mean ** (1.0 / 3.0)
    positions, types = [], []
    with open(dumps[-1]) as f:
        lines = f.read().splitlines()
    for line in lines[2:]:
        fields = line.split()
        if len(fields) < 7:
            continue
        xyz = [float(v) * Units.angstrom * scale for v in fields[1:4]]
        positions.append(Vector3d(*xyz))
        types.append(species[fields[0]])
    if len(positions) != npt.Nparticles:
        raise ValueError(f"{dumps[-1]} holds {len(positions)} atoms, the run had "
                         f"{npt.Nparticles}")
    box = (npt.Lx * scale, npt.Ly * scale, npt.Lz * scale)
    return Material(potential=mat.potential, atomic_weight=mat.atomic_weight,
                    atom_name=mat.atom_name, eam=mat.eam, lattice=positions, atom_type=types,
                    box_length=box)


# MDL_S_* columns of the per-sample records returned by MDDriver.run
_R = nat.S_INDEX
_R_SB, _R_ST, _R_CB = nat.S_BORN_STRESS, nat.S_STRESS, nat.S_BORN_ELAST


class MolecularDynamics:
    """molecular_dynamics.py:282-1285. Keyword parameters as the reference:
    required ``material, temperature_initial, temperature_thermostat,
    tau_thermostat, pressure_barostat, tau_barostat, ensamble, dump_cycles,
    cycles_thermalization, plot_freq, cycles_end, dt_global``; optional
    ``nchains_thermostat, nchains_barostat, fatal_conservation, slurm_plt,
    show_graphs, plot_graphs, N_lattice_side, a_lattice, neighbors, r_skin,
    cycles_neigbors, gr_stride, initial_velocities, thermostat_coupling,
    random_seed, sample_stride``; and, as an extension, ``precision`` ("fp64" or
    "mixed") and ``lanes_per_atom`` for the driver."""

    KEYS_MUST = {"material", "temperature_initial", "temperature_thermostat", "tau_thermostat",
                 "pressure_barostat", "tau_barostat", "ensamble", "dump_cycles",
                 "cycles_thermalization", "plot_freq", "cycles_end", "dt_global"}
    KEYS_DEFAULT = {"nchains_thermostat", "nchains_barostat", "fatal_conservation", "slurm_plt",
                    "show_graphs", "plot_graphs", "N_lattice_side", "a_lattice", "neighbors",
                    "r_skin", "cycles_neigbors", "gr_stride", "initial_velocities",
                    "thermostat_coupling", "random_seed", "sample_stride", "precision",
                    "lanes_per_atom"}

    def __init__(self, *, work_dir=".", **kwargs):
        np.random.seed(kwargs.get("random_seed", 0))
        self.runtime_begin = time.time()
        missing = self.KEYS_MUST - kwargs.keys()
        if missing:
            raise ValueError(f"missing: {sorted(missing)}")
        extra = kwargs.keys() - (self.KEYS_MUST | self.KEYS_DEFAULT)
        if extra:
            raise ValueError(f"extra keys given: {sorted(extra)}")
        # defaults (molecular_dynamics.py:362-376)
        self.nchains_thermostat = 10
        self.nchains_barostat = 10
        self.fatal_conservation = 1e3
        self.slurm_plt = False
        self.show_graphs = False
        self.plot_graphs = True
        self.neighbors = False
        self.r_skin = 1000
        self.cycles_neigbors = 15
        self.gr_stride = 1
        self.sample_stride = 1
        self.thermostat_coupling = MDDriver.THERMOSTAT_KINETIC
        self.precision = "fp64"
        self.lanes_per_atom = 0
        for key, value in kwargs.items():
            if key == "ensamble" and isinstance(value, str):
                value = {"NVE": MDDriver.NVE, "NVT": MDDriver.NVT, "NPT": MDDriver.NPT}[value]
            if key == "thermostat_coupling" and isinstance(value, str):
                value = {"kinetic": MDDriver.THERMOSTAT_KINETIC,
                         "configurational": MDDriver.THERMOSTAT_CONFIGURATIONAL,
                         "average": MDDriver.THERMOSTAT_AVERAGE}[value]
            setattr(self, key, value)
        if int(self.sample_stride) != self.sample_stride or self.sample_stride < 1:
            raise ValueError(f"sample_stride must be a positive integer, got {self.sample_stride}")
        self.sample_stride = int(self.sample_stride)
        if self.sample_stride > 1:
            if self.thermostat_coupling != MDDriver.THERMOSTAT_KINETIC:
                raise ValueError("sample_stride > 1 requires kinetic thermostat coupling")
            for key in ("cycles_thermalization", "plot_freq"):
                if getattr(self, key) % self.sample_stride != 0:
                    raise ValueError(f"{key} must be a multiple of sample_stride")
        self.samples_thermalization = self.cycles_thermalization // self.sample_stride

        self.work_dir = path.abspath(work_dir)
        os.makedirs(self.work_dir, exist_ok=True)
        self.dump_dir = path.join(self.work_dir, "dump")
        os.makedirs(self.dump_dir, exist_ok=True)
        self.dump_counter = 0

        mat = self.material
        if mat.lattice is not None:
            self.initial_positions = mat.lattice
            self.Lx, self.Ly, self.Lz = mat.box_length
            self.Ntypes = list(mat.atom_type)
            self.Nparticles = len(self.initial_positions)
            self.box = Box(np.diag((self.Lx, self.Ly, self.Lz)))
            self.volume = self.box.volume()
            self.a_lattice = np.power(self.volume / self.Nparticles, 1.0 / 3.0)
            self.mass = np.array([mat.atomic_weight[a] for a in mat.atom_type]) / Units.Navogadro
        else:
            if mat.a_lattice is not None:
                self.a_lattice = mat.a_lattice
            elif "a_lattice" not in kwargs:
                raise ValueError("no lattice constant: give Material(a_lattice=...) or a_lattice")
            if not hasattr(self, "N_lattice_side"):
                raise ValueError("N_lattice_side is required for an FCC lattice")
            self.NsideX = self.NsideY = self.NsideZ = self.N_lattice_side
            self.Lx, self.Ly, self.Lz = (self.NsideX * self.a_lattice, self.NsideY * self.a_lattice,
                                         self.NsideZ * self.a_lattice)
            self.box = Box(np.diag((self.Lx, self.Ly, self.Lz)))
            self.volume = self.box.volume()
            self.build_fcc_lattice()
            self.Ntypes = [0] * self.Nparticles
            self.mass = np.array([mat.atomic_weight[0]] * self.Nparticles) / Units.Navogadro

        self.ndensity = self.Nparticles / self.volume
        self.Nf = 3.0 * (self.Nparticles - 1.0)
        self.thermalized = False
        drg = self.a_lattice / 500.0
        ng = int(min(self.Lx, self.Ly, self.Lz) / 2.0 / drg)
        self.rg = np.array([(k + 0.5) * drg for k in range(ng)])
        if getattr(self, "initial_velocities", None) is None:
            self.set_initial_velocities()
        elif len(self.initial_velocities) != self.Nparticles:
            raise ValueError("initial_velocities length does not match Nparticles")

        kT = Units.k_boltz * self.temperature_thermostat
        q = np.ones(self.nchains_thermostat) * self.tau_thermostat ** 2 * kT
        q[0] *= self.Nf
        thermostat = Thermostat(self.temperature_thermostat, q,
                                np.random.normal(0.0, np.sqrt(q * kT)), self.Nf)
        qb = np.ones(self.nchains_barostat) * self.tau_barostat ** 2 * kT
        thermostat_baro = Thermostat(self.temperature_thermostat, qb,
                                     np.random.normal(0.0, np.sqrt(qb * kT)), 1.0)
        w = self.Nf * self.tau_barostat ** 2 * kT
        barostat = Barostat(self.pressure_barostat, w, np.random.normal(0.0, np.sqrt(w * kT)),
                            thermostat_baro)
        self.driver = MDDriver(
            self.initial_positions, self.initial_velocities, PotentialVector(mat.potential),
            EAMVector([] if mat.eam is None else mat.eam), self.ensamble,
            self.thermostat_coupling, thermostat, barostat, self.box, self.rg, self.Nf,
            np.array(mat.atomic_weight) / Units.Navogadro, self.Ntypes, self.dt_global,
            self.neighbors, self.r_skin, self.fatal_conservation,
            precision=self.precision, lanes_per_atom=self.lanes_per_atom)
        self.series = Series()
        self.gr_sum = np.zeros(len(self.rg))
        self.gr_samples = 0
        # partial g(r) only for more than one species (molecular_dynamics.py:566-576)
        self.Ntypes_unique = len(set(int(t) for t in self.Ntypes))
        if self.Ntypes_unique > 1:
            g = self.driver.gr_partial
            self.gr_partial_sum = np.zeros((len(g), len(g[0])))
        else:
            self.gr_partial_sum = None
        self.gr_partial_samples = 0

    # -- set-up (molecular_dynamics.py:582-642) ------------------------------
    def build_fcc_lattice(self):
        basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
        offset = 0.25
        L = np.array([self.Lx, self.Ly, self.Lz])
        self.initial_positions = []
        for i in range(self.NsideX):
            for j in range(self.NsideY):
                for k in range(self.NsideZ):
                    for m in range(4):
                        xyz = -L / 2.0 + (np.array([i, j, k]) + offset + basis[m]) * self.a_lattice
                        self.initial_positions.append(Vector3d(*xyz))
        self.Nparticles = len(self.initial_positions)

    def set_initial_velocities(self):
        m = self.mass
        sigma = np.sqrt(Units.k_boltz * self.temperature_initial / m)
        v = [np.random.normal(0.0, sigma) for _ in range(3)]
        for c in range(3):
            v[c] = v[c] - np.sum(v[c] * m) / np.sum(m)
        ekin = (3.0 * self.Nparticles - 3.0) * Units.k_boltz * self.temperature_initial / 2.0
        scale = np.sqrt(ekin / (0.5 * np.sum(m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))))
        self.initial_velocities = [Vector3d(a * scale, b * scale, c * scale)
                                   for a, b, c in zip(*v)]

    # -- output ----------------------------------------------------------------
    def dump(self):
        """Trajectory frame as extended XYZ (molecular_dynamics.py:644-662)."""
        filename = path.join(self.dump_dir, f"{self.dump_counter}.xyz")
        names = self.material.atom_name.split("-")
        pos = self.driver.positions_array() / Units.angstrom
        vel = self.driver.velocities_array() / (Units.angstrom / Units.fs)
        head = (f"{self.Nparticles}\na= {self.a_lattice / Units.angstrom} "
                f"Properties=species:S:1:pos:R:3:vel:R:3 Time={self.driver.time / Units.ps}\n")
        body = "\n".join(f"{names[a]} {p[0]} {p[1]} {p[2]} {u[0]} {u[1]} {u[2]}"
                         for p, u, a in zip(pos, vel, self.Ntypes))
        with open(filename, "w") as f:
            print(head + body, file=f)
        self.dump_counter += 1

    def save_json(self, filename=None):
        filename = filename or path.join(self.work_dir, "thermal_results.json")
        out = {}
        for k, v in vars(self.series.thermal_results).items():
            out[k] = v.tolist() if isinstance(v, np.ndarray) else v
        with open(filename, "w") as f:
            json.dump(out, f, indent=1, default=float)
        return filename

    # -- samples ----------------------------------------------------------------
    def _append_records(self, rec):
        """append_sample (molecular_dynamics.py:1218-1238) for each record."""
        s = self.series
        V0 = self.driver.box_initial.volume()
        for r in rec:
            s.time.append(r[_R["time"]])
            s.temperature_kin.append(r[_R["temperature_kin"]])
            s.temperature_conf_inverse.append(r[_R["temperature_conf_inverse"]])
            s.total_laplacian.append(r[_R["total_laplacian"]])
            s.total_force2.append(r[_R["total_force2"]])
            s.pressure.append(r[_R["pressure"]])
            s.stress_tensor.append(r[_R_ST:_R_ST + 6])
            s.etot_atom.append(r[_R["etot_atom"]])
            s.epot_atom.append(r[_R["epot_atom"]])
            s.epot_eam_atom.append(r[_R["epot_eam_atom"]])
            s.ekin_atom.append(r[_R["ekin_atom"]])
            s.etot_cons.append(r[_R["etot_cons"]])
            s.etot_cons_err.append(r[_R["etot_cons_err"]])
            s.born_stress_tensor.append(r[_R_SB:_R_SB + 6])
            s.born_elasticity_tensor.append(r[_R_CB:_R_CB + 36].reshape(6, 6))
            V_V0 = r[_R["box0"]] * (r[_R["box1"]] * r[_R["box2"]]) / V0
            s.V_V0.append(V_V0)
            s.L_L0.append(V_V0 ** (1.0 / 3.0))

    # -- the loop (molecular_dynamics.py:1240-1285) ------------------------------
    def run(self):
        d = self.driver
        rebuild = self.cycles_neigbors if self.neighbors else 0
        while d.cycle <= self.cycles_end:
            c = d.cycle
            if c % self.dump_cycles == 0:
                self.dump()
            # run up to the next dump, thermal average or the end
            stop = min(self.cycles_end + 1, (c // self.dump_cycles + 1) * self.dump_cycles)
            t0 = self.cycles_thermalization
            if self.gr_stride < 1:
                raise ValueError(f"gr_stride must be >= 1, got {self.gr_stride}")
            for every in (self.plot_freq, self.gr_stride):
                # the first cycle after c (and after t0) where t0 + k * every acts
                k = max(1, (c + 1 - t0 + every - 1) // every)
                stop = min(stop, t0 + k * every)
            rec = d.run(stop - c, self.sample_stride, rebuild, records=True)
            self._append_records(rec)
            self.thermalized = d.cycle > self.cycles_thermalization
            if self.thermalized and (d.cycle - t0) % self.gr_stride == 0:
                # molecular_dynamics.py:1258-1272
                if self.gr_partial_sum is not None:
                    d.calculate_pair_correlations()
                    self.gr_sum += np.array(d.gr)
                    self.gr_samples += 1
                    self.gr_partial_sum += np.array(d.gr_partial)
                    self.gr_partial_samples += 1
                else:
                    d.calculate_pair_correlation()
                    self.gr_sum += np.array(d.gr)
                    self.gr_samples += 1
            if self.thermalized and (d.cycle - t0) % self.plot_freq == 0:
                self.perform_thermal_average()
                self.save_series()      # where the reference writes series.h5 (:1276-1279)
        self.runtime = time.time() - self.runtime_begin
        return self

    def save_series(self, filename=None):
        """The series and thermal results with the reference's series.h5 schema
        (molecular_dynamics.py:1182-1195), as ``series.npz`` in the work directory;
        ``series_io.summarize_run`` reads it as bench_lib.summarize_run reads
        series.h5."""
        from . import series_io
        self.series_path = series_io.save_series(
            self.series, filename or path.join(self.work_dir, "series.npz"))
        return self.series_path

    # -- averages (molecular_dynamics.py:664-1170) ---------------------------------
    def perform_thermal_average(self):
        s, tr, d = self.series, self.series.thermal_results, self.driver
        cut = self.samples_thermalization
        if not (self.thermalized and cut < len(s.time)):
            raise RuntimeError("thermal averages need production samples")
        tr.time = d.time
        tr.cycles_thermalization = self.cycles_thermalization
        tr.sample_stride = self.sample_stride
        for name in Series.SCALARS:
            getattr(s, name).set_averages(cut, cut)
        s.etot_cons_err.set_averages(0, 0)
        s.stress_tensor.set_averages(cut, cut)
        s.born_stress_tensor.set_averages(cut, cut)
        V0 = d.box_initial.volume()
        Vmean = s.V_V0.mean * V0
        s.particle_density = TimeSeries()
        s.particle_density.data = list(self.Nparticles / (np.array(s.V_V0.data) * V0))
        s.particle_density.set_averages(cut, cut)
        tr.gr = self.gr_sum / self.gr_samples if self.gr_samples > 0 else np.zeros_like(self.gr_sum)
        if self.gr_partial_sum is not None:
            tr.gr_partial = (self.gr_partial_sum / self.gr_partial_samples
                             if self.gr_partial_samples > 0
                             else np.zeros_like(self.gr_partial_sum))
        if self.ensamble in (MDDriver.NVT, MDDriver.NPT):
            Tth = d.thermostat.temperature
            tr.tk_tnh_err = abs(s.temperature_kin.mean / Tth - 1.0)
            tr.tc_tnh_err = abs(1.0 / s.temperature_conf_inverse.mean / Tth - 1.0)
        tr.pressure_ideal = Units.k_boltz * s.pressure.mean * self.ndensity
        tr.pressure_baro_err = abs(s.pressure.mean / d.barostat.pressure - 1.0) \
            if d.barostat.pressure != 0 else np.inf
        kT = Units.k_boltz * self.temperature_thermostat
        tr.bulk_modulus = 0.0
        npt = d.ensamble == MDDriver.NPT
        if npt:
            tr.compressibility = (s.V_V0.std * V0) ** 2 / (Vmean * kT)
            tr.bulk_modulus = 1.0 / tr.compressibility
            dV = np.array(s.V_V0.data) * V0 - Vmean
            dE = (np.array(s.etot_atom.data) - s.etot_atom.mean) * self.Nparticles
            s.volume_thermal_expansion = TimeSeries()
            s.volume_thermal_expansion.data = list(
                dV * (dE + self.pressure_barostat * dV) /
                (Vmean * Units.k_boltz * self.temperature_thermostat ** 2))
            s.volume_thermal_expansion.set_averages(cut, cut)
            tr.volume_thermal_expansion = s.volume_thermal_expansion.mean
        if d.ensamble in (MDDriver.NVT, MDDriver.NPT):
            self._elasticity(cut, kT, V0, npt)
        self._heat_capacity()
        tr.rg = self.rg

    def _elasticity(self, cut, kT, V0, npt):
        """Stress-fluctuation estimator (molecular_dynamics.py:825-1087)."""
        s, tr = self.series, self.series.thermal_results
        born_stress = s.born_stress_tensor.get_time_array()           # [6, n]
        fluc = born_stress.copy()
        if npt:
            # the ideal-gas pressure N kT / V fluctuates with V in NPT and stays in
            # the fluctuating stress (molecular_dynamics.py:829-836)
            fluc[:3, :] -= (np.asarray(s.particle_density.data) * kT)[None, :]
        fluc_mean = fluc[:, cut:].mean(axis=1)
        CB = s.born_elasticity_tensor.get_time_array()                 # [6, 6, n]
        s.born_elasticity_tensor.set_averages(cut, cut)
        for name, idx in (("11", ((0, 0), (1, 1), (2, 2))), ("12", ((0, 1), (0, 2), (1, 2))),
                          ("44", ((3, 3), (4, 4), (5, 5)))):
            ts = TimeSeries()
            ts.data = list(sum(CB[i, j, :] for i, j in idx) / 3.0)
            ts.set_averages(cut, cut)
            setattr(s, f"born_elasticity_tensor_{name}", ts)
        Vall = TimeSeries()
        Vall.data = list(np.array(s.V_V0.data) * V0)
        Vall.set_averages(cut, cut)
        rho_n = np.asarray(s.particle_density.data)
        s.stress_fluc = TimeSeriesMatrix(6)
        s.born_stress_fluc = TimeSeriesMatrix(6)
        s.elasticity_ideal_gas = TimeSeriesMatrix(6)
        s.elasticity_tensor = TimeSeriesMatrix(6)
        C = np.zeros((6, 6))
        for i in range(6):
            for j in range(6):
                sf = s.stress_fluc.data[i][j]
                sf.data = list(Vall.mean / kT * (fluc[i] * fluc[j] - fluc_mean[i] * fluc_mean[j]))
                sf.set_averages(cut, cut)
                bsf = s.born_stress_fluc.data[i][j]
                bsf.data = list(Vall.mean / kT * (
                    born_stress[i] * born_stress[j] -
                    s.born_stress_tensor.data[i].mean * s.born_stress_tensor.data[j].mean))
                bsf.set_averages(cut, cut)
                factor = (2.0 if i < 3 else 1.0) if i == j else 0.0
                kin = s.elasticity_ideal_gas.data[i][j]
                kin.data = list(rho_n * kT * factor)
                kin.set_averages(cut, cut)
                K = tr.bulk_modulus if (npt and i < 3 and j < 3) else 0.0
                el = s.elasticity_tensor.data[i][j]
                el.data = list(CB[i, j, :] - np.array(sf.data) + np.array(kin.data) + K)
                el.set_averages(cut, cut)
                C[i, j] = s.born_elasticity_tensor.data[i][j].mean - sf.mean + kin.mean + K
        tr.elasticity_tensor = C
        tr.elasticity_tensor_11 = (C[0, 0] + C[1, 1] + C[2, 2]) / 3.0
        tr.elasticity_tensor_12 = (C[0, 1] + C[0, 2] + C[1, 2]) / 3.0
        tr.elasticity_tensor_44 = (C[3, 3] + C[4, 4] + C[5, 5]) / 3.0
        for name, idx in (("11", ((0, 0), (1, 1), (2, 2))), ("12", ((0, 1), (0, 2), (1, 2))),
                          ("44", ((3, 3), (4, 4), (5, 5)))):
            ts = TimeSeries()
            ts.data = list(sum(np.array(s.elasticity_tensor.data[i][j].data) for i, j in idx) / 3.0)
            ts.set_averages(cut, cut)
            setattr(s, f"elasticity_tensor_{name}", ts)

    def _heat_capacity(self):
        """molecular_dynamics.py:1110-1168."""
        s, tr, N = self.series, self.series.thermal_results, self.Nparticles
        kB = Units.k_boltz
        tr.cv_id = 1.5 * kB
        T = s.temperature_kin.mean
        if self.driver.ensamble == MDDriver.NVE:
            tr.epot_ekin_std_err = abs(s.epot_atom.std / s.ekin_atom.std - 1.0)
            tr.epot_ekin_slope_fit = np.polyfit(s.ekin_atom.data, s.epot_atom.data, deg=1)[0]

            def cv_nve(sigma):
                return 1.5 * kB / (1.0 - sigma ** 2 / (1.5 * N * (kB * T) ** 2))

            tr.cv_ekin = cv_nve(s.ekin_atom.std * N)
            tr.cv_epot = cv_nve(s.epot_atom.std * N)
            tr.cv_epot_ekin_err = abs(tr.cv_epot / tr.cv_ekin - 1.0)
        elif self.driver.ensamble == MDDriver.NVT:
            tr.ekin_atom_std_exact = np.sqrt(1.5 * N * (kB * T) ** 2) / N
            tr.ekin_atom_std_err = abs(s.ekin_atom.std / tr.ekin_atom_std_exact - 1.0)
            tr.cv_etot = (s.etot_atom.std * N) ** 2 / (kB * T ** 2) / N
            tr.cv_epot = (s.epot_atom.std * N) ** 2 / (kB * T ** 2) / N + 1.5 * kB
            tr.cv_etot_epot_err = abs(tr.cv_epot / tr.cv_etot - 1.0)
