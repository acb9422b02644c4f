"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py use this module, as the checker; the product never imports it.
The library is built by oracle/Makefile into oracle/build/liborc.so.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liborc.so")

QUINTIC, BSPLINE, MORSE, LJ = 0, 1, 2, 3
NVE, NVT, NPT = 0, 1, 2
KINETIC, CONFIGURATIONAL, AVERAGE = 0, 1, 2

# orc_md_scalars layout (mdelast_oracle.h)
S = dict(epot_eam_atom=0, epot_atom=1, ekin_atom=2, etot_atom=3, ekin_tot=4,
         etot_cons_initial=5, etot_cons=6, etot_cons_err=7, virial=8, total_laplacian=9,
         total_force2=10, temperature_kin=11, temperature_conf_inverse=12, pressure=13,
         time=14, cycle=15, box0=16, box1=17, box2=18, bar_eps=19, bar_peps=20)
A = dict(positions=0, coordinates=1, velocities=2, forces=3, rho=4, epot=5, ekin=6, F=7,
         dF=8, d2F=9)

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)


class _Config(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ntypes", C.c_int),
        ("pos", _dp), ("vel", _dp), ("types", _up), ("mass", _dp),
        ("box", C.c_double * 3),
        ("ensemble", C.c_int), ("coupling", C.c_int),
        ("th_T", C.c_double), ("th_Nf", C.c_double), ("th_M", C.c_int),
        ("th_q", _dp), ("th_p_eta", _dp),
        ("bar_P", C.c_double), ("bar_W", C.c_double), ("bar_p_eps", C.c_double),
        ("bth_T", C.c_double), ("bth_Nf", C.c_double), ("bth_M", C.c_int),
        ("bth_q", _dp), ("bth_p_eta", _dp),
        ("Nf", C.c_double), ("dt", C.c_double),
        ("neighbors_list", C.c_int),
        ("r_skin", C.c_double), ("fatal_conservation", C.c_double),
        ("pair", C.POINTER(C.c_void_p)), ("pair_rcut", _dp), ("pair_rmin", _dp),
        ("has_eam", C.c_int),
        ("rho_tab", C.POINTER(C.c_void_p)), ("F_tab", C.POINTER(C.c_void_p)),
        ("eam_rcut", _dp), ("eam_rhomax", _dp),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_table_new.restype = C.c_void_p
        L.orc_table_new.argtypes = [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_table_analytic.restype = C.c_void_p
        L.orc_table_analytic.argtypes = [C.c_int, _dp]
        L.orc_table_free.argtypes = [C.c_void_p]
        for f in ("orc_table_value", "orc_table_d1", "orc_table_d2"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p, C.c_double]
        L.orc_table_coef.restype = _dp
        L.orc_table_coef.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_md_new.restype = C.c_void_p
        L.orc_md_new.argtypes = [C.POINTER(_Config), C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.orc_md_free.argtypes = [C.c_void_p]
        L.orc_md_error.restype = C.c_char_p
        L.orc_md_error.argtypes = [C.c_void_p]
        for f in ("orc_md_integrate", "orc_md_calculate_neighbors", "orc_md_calculate_forces"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_md_set_sample.argtypes = [C.c_void_p, C.c_int]
        L.orc_md_scalars.argtypes = [C.c_void_p, _dp]
        L.orc_md_scales.argtypes = [C.c_void_p, _dp]
        L.orc_md_atoms.restype = C.c_int
        L.orc_md_atoms.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_md_chain.restype = C.c_int
        L.orc_md_chain.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_md_nlist_size.restype = C.c_long
        L.orc_md_nlist_size.argtypes = [C.c_void_p]
        L.orc_md_warnings.restype = C.c_long
        L.orc_md_warnings.argtypes = [C.c_void_p, C.c_int]
        L.orc_md_pair_correlation.restype = C.c_int
        L.orc_md_pair_correlation.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


class Table:
    """A tabulated function f with node derivatives (QuinticTable / B-spline)."""

    def __init__(self, kind, h, f, f1, f2):
        self._keep = [np.ascontiguousarray(a, dtype=float) for a in (f, f1, f2)]
        self.h = float(h)
        self.kind = kind
        self.handle = lib().orc_table_new(kind, self.h, len(self._keep[0]),
                                          *(_ptr(a) for a in self._keep))
        if not self.handle:
            raise OracleError(1, "invalid table")

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.orc_table_free(self.handle)
            self.handle = None

    def value(self, x):
        return lib().orc_table_value(self.handle, float(x))

    def d1(self, x):
        return lib().orc_table_d1(self.handle, float(x))

    def d2(self, x):
        return lib().orc_table_d2(self.handle, float(x))

    def coef(self):
        n = C.c_int(0)
        p = lib().orc_table_coef(self.handle, C.byref(n))
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(n.value, 6)).copy()


class AnalyticPair(Table):
    """Analytic Morse {D, alpha, r0} or Lennard-Jones {epsilon, sigma, rcut} pair
    form (potential.hpp:29-75) behind the table evaluation calls."""

    def __init__(self, kind, param, rcut):
        self._keep = [np.ascontiguousarray(param, dtype=float)]
        self.h, self.kind, self.rcut = 0.0, kind, float(rcut)
        self.handle = lib().orc_table_analytic(kind, _ptr(self._keep[0]))
        if not self.handle:
            raise OracleError(1, "invalid analytic pair form")


def tables_from_setfl(data, interpolation="quintic"):
    """Oracle tables (pair slots, rho per type, F per type) from SETFL node data."""
    kind = QUINTIC if interpolation == "quintic" else BSPLINE
    pair = [Table(kind, data.dr, p.phi, p.dphi, p.d2phi) for p in data.pairs]
    rho = [Table(kind, data.dr, e.rho, e.drho, e.d2rho) for e in data.elements]
    F = [Table(kind, data.drho, e.F, e.dF, e.d2F) for e in data.elements]
    return pair, rho, F


class MD:
    """Oracle MD driver; the keyword arguments mirror the MDDriver ctor."""

    def __init__(self, *, positions, velocities, types, mass, box, data=None, tables=None,
                 interpolation="quintic", ensemble=NVE, coupling=KINETIC,
                 thermostat=(300.0, [1e-30], [0.0], None), barostat=(0.0, 1e-30, 0.0),
                 baro_thermostat=(300.0, [1e-30], [0.0], 1.0), Nf=None, dt=1e-15,
                 neighbors_list=False, r_skin=1.0, fatal_conservation=1e3, has_eam=True):
        L = lib()
        pos = np.ascontiguousarray(positions, dtype=float).reshape(-1, 3)
        vel = np.ascontiguousarray(velocities, dtype=float).reshape(-1, 3)
        N = len(pos)
        typ = np.ascontiguousarray(types, dtype=np.uint32)
        mass = np.ascontiguousarray(mass, dtype=float)
        if tables is None:
            tables = tables_from_setfl(data, interpolation)
        self.tables = tables
        pair, rho, F = tables
        ntypes = len(mass)
        if Nf is None:
            Nf = 3.0 * N
        thT, thq, thp, thNf = thermostat
        thNf = Nf if thNf is None else thNf
        bP, bW, bpe = barostat
        bT, bq, bp, bNf = baro_thermostat
        self._keep = keep = {}
        keep["pos"], keep["vel"], keep["typ"], keep["mass"] = pos, vel, typ, mass
        keep["thq"] = np.ascontiguousarray(thq, dtype=float)
        keep["thp"] = np.ascontiguousarray(thp, dtype=float)
        keep["bq"] = np.ascontiguousarray(bq, dtype=float)
        keep["bp"] = np.ascontiguousarray(bp, dtype=float)
        rcut = data.rcut if data is not None else None
        keep["prc"] = np.array([rcut] * len(pair) if rcut is not None else
                               [t.rcut for t in pair], dtype=float)
        keep["prm"] = np.array([data.rmin_fatal if data is not None else 0.0] * len(pair))
        keep["erc"] = np.array([rcut if rcut is not None else 0.0] * ntypes, dtype=float)
        keep["erm"] = np.array([data.rhomax if data is not None else 1e300] * ntypes, dtype=float)
        keep["pair"] = (C.c_void_p * len(pair))(*[t.handle for t in pair])
        keep["rho"] = (C.c_void_p * max(1, len(rho)))(*[t.handle for t in rho])
        keep["F"] = (C.c_void_p * max(1, len(F)))(*[t.handle for t in F])
        cfg = _Config()
        cfg.N, cfg.ntypes = N, ntypes
        cfg.pos, cfg.vel = _ptr(pos), _ptr(vel)
        cfg.types = typ.ctypes.data_as(_up)
        cfg.mass = _ptr(mass)
        cfg.box = (C.c_double * 3)(*[float(b) for b in box])
        cfg.ensemble, cfg.coupling = int(ensemble), int(coupling)
        cfg.th_T, cfg.th_Nf, cfg.th_M = float(thT), float(thNf), len(keep["thq"])
        cfg.th_q, cfg.th_p_eta = _ptr(keep["thq"]), _ptr(keep["thp"])
        cfg.bar_P, cfg.bar_W, cfg.bar_p_eps = float(bP), float(bW), float(bpe)
        cfg.bth_T, cfg.bth_Nf, cfg.bth_M = float(bT), float(bNf), len(keep["bq"])
        cfg.bth_q, cfg.bth_p_eta = _ptr(keep["bq"]), _ptr(keep["bp"])
        cfg.Nf, cfg.dt = float(Nf), float(dt)
        cfg.neighbors_list = int(bool(neighbors_list))
        cfg.r_skin, cfg.fatal_conservation = float(r_skin), float(fatal_conservation)
        cfg.pair = C.cast(keep["pair"], C.POINTER(C.c_void_p))
        cfg.pair_rcut, cfg.pair_rmin = _ptr(keep["prc"]), _ptr(keep["prm"])
        cfg.has_eam = int(bool(has_eam))
        cfg.rho_tab = C.cast(keep["rho"], C.POINTER(C.c_void_p))
        cfg.F_tab = C.cast(keep["F"], C.POINTER(C.c_void_p))
        cfg.eam_rcut, cfg.eam_rhomax = _ptr(keep["erc"]), _ptr(keep["erm"])
        self.N, self.ntypes = N, ntypes
        self.M_th, self.M_b = len(keep["thq"]), len(keep["bq"])
        err = C.c_int(0)
        msg = C.create_string_buffer(256)
        self.handle = L.orc_md_new(C.byref(cfg), C.byref(err), msg, 256)
        if not self.handle:
            raise OracleError(err.value, msg.value.decode())

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.orc_md_free(self.handle)
            self.handle = None

    def _check(self, rc):
        if rc:
            raise OracleError(rc, lib().orc_md_error(self.handle).decode())

    def integrate(self, n=1):
        for _ in range(n):
            self._check(lib().orc_md_integrate(self.handle))

    def calculate_neighbors(self):
        self._check(lib().orc_md_calculate_neighbors(self.handle))

    def calculate_forces(self):
        self._check(lib().orc_md_calculate_forces(self.handle))

    def set_sample(self, flag):
        lib().orc_md_set_sample(self.handle, int(bool(flag)))

    def scalars(self):
        out = np.zeros(72)
        lib().orc_md_scalars(self.handle, _ptr(out))
        d = {k: out[v] for k, v in S.items()}
        d["born_stress_tensor"] = out[24:30].copy()
        d["stress_tensor"] = out[30:36].copy()
        d["born_elasticity_tensor"] = out[36:72].reshape(6, 6).copy()
        d["cycle"] = int(d["cycle"])
        return d

    def scales(self):
        out = np.zeros(44)
        lib().orc_md_scales(self.handle, _ptr(out))
        return {"virial": out[0], "born_stress": out[1:7].copy(),
                "born_elasticity": out[7:43].reshape(6, 6).copy(), "laplacian": out[43]}

    def atoms(self, what):
        n = 3 * self.N if what in ("positions", "coordinates", "velocities", "forces") else self.N
        out = np.zeros(n)
        self._check(lib().orc_md_atoms(self.handle, A[what], _ptr(out)))
        return out.reshape(self.N, 3) if n == 3 * self.N else out

    def chain(self, which=0):
        M = self.M_th if which == 0 else self.M_b
        p, e = np.zeros(M), np.zeros(M)
        lib().orc_md_chain(self.handle, which, _ptr(p), _ptr(e))
        return p, e

    def nlist_size(self):
        return lib().orc_md_nlist_size(self.handle)

    def warnings(self, which):
        return lib().orc_md_warnings(self.handle, which)

    def pair_correlation(self, rg, mode=1):
        """(gr, gr_partial [S(S+1)/2, nbin]) of md_driver.cpp:732-863; mode 0 fills
        gr only, 2 gr_partial only (the other array is returned as zeros)."""
        rg = np.ascontiguousarray(rg, dtype=float)
        gr = np.zeros(len(rg))
        gp = np.zeros((self.ntypes * (self.ntypes + 1) // 2, len(rg)))
        self._check(lib().orc_md_pair_correlation(self.handle, _ptr(rg), len(rg), int(mode),
                                                   _ptr(gr), _ptr(gp)))
        return gr, gp
