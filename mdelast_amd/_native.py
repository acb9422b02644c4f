"""ctypes binding of libmdelast_hip.so (the C-ABI of include/mdelast.h).

There is no fallback: if the library is missing or cannot be loaded, every
use raises, so nothing silently runs anywhere but on the HIP path.
"""

from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmdelast_hip.so")

QUINTIC, BSPLINE, MORSE, LJ = 0, 1, 2, 3
NVE, NVT, NPT, NPT_FLEX = 0, 1, 2, 3
KINETIC, CONFIGURATIONAL, AVERAGE = 0, 1, 2
FP64, MIXED = 0, 1

# MDL_S_* record (include/mdelast.h)
S_INDEX = dict(epot_eam_atom=0, epot_atom=1, ekin_atom=2, etot_atom=3, ekin_tot=4,
               etot_cons_initial=5, etot_cons=6, etot_cons_err=7, virial=8, total_laplacian=9,
               total_force2=10, temperature_kin=11, temperature_conf_inverse=12, pressure=13,
               time=14, cycle=15, box0=16, box1=17, box2=18, bar_eps=19, bar_peps=20,
               mom0=21, mom1=22, mom2=23)
S_BORN_STRESS, S_STRESS, S_BORN_ELAST = 24, 30, 36
NRECORD, NSCALARS = 72, 82
S_BOX_INITIAL, S_NLIST_BUILDS, S_WARN_VIRIAL, S_WARN_RMIN, S_NLIST_ENTRIES = 72, 75, 76, 77, 78
S_NHC_GENERAL = 79
S_NMID_BUILDS = 80
S_LANES_PER_ATOM = 81

A_INDEX = dict(positions=0, coordinates=1, velocities=2, forces=3, rho=4, epot=5, ekin=6,
               pairs_rc=7, pairs_mid=8)

ERRORS = {1: "ARG", 2: "POSITION", 3: "TYPE", 4: "ENSEMBLE", 5: "RHOMAX", 6: "VOLUME",
          7: "MOMENTUM", 8: "PRESSURE", 9: "CONSERVATION", 10: "STATE", 11: "DEVICE",
          12: "UNSUPPORTED", 13: "NEIGHBORS", 14: "TABLE"}

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class MdlError(RuntimeError):
    """A failed C-ABI call; ``code`` is the MDL_ERR_* value, ``kind`` its name."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = ERRORS.get(code, str(code))
        super().__init__(f"[{self.kind}] {message}")


class Config(C.Structure):
    _fields_ = [
        ("N", C.c_int64), ("ntypes", C.c_int32),
        ("positions", _dp), ("velocities", _dp), ("types", C.POINTER(C.c_uint32)),
        ("mass", _dp), ("box", C.c_double * 3),
        ("ensemble", C.c_int32), ("coupling", C.c_int32),
        ("th_T", C.c_double), ("th_Nf", C.c_double), ("th_M", C.c_int32),
        ("th_q", _dp), ("th_p_eta", _dp),
        ("bar_P", C.c_double), ("bar_W", C.c_double), ("bar_p_eps", C.c_double),
        ("bth_T", C.c_double), ("bth_Nf", C.c_double), ("bth_M", C.c_int32),
        ("bth_q", _dp), ("bth_p_eta", _dp),
        ("Nf", C.c_double), ("dt", C.c_double),
        ("neighbors_list", C.c_int32), ("r_skin", C.c_double),
        ("fatal_conservation", C.c_double),
        ("pair", C.POINTER(_vp)), ("pair_rcut", _dp), ("pair_rmin", _dp),
        ("has_eam", C.c_int32),
        ("rho_tab", C.POINTER(_vp)), ("F_tab", C.POINTER(_vp)),
        ("eam_rcut", _dp), ("eam_rhomax", _dp),
        ("precision", C.c_int32), ("lanes_per_atom", C.c_int32),
    ]


# every symbol declared in include/mdelast.h, with (restype, argtypes)
SIGNATURES = {
    "mdl_last_error": (C.c_char_p, []),
    "mdl_device_info": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32)]),
    "mdl_table_create": (C.c_int32, [C.c_int32, C.c_double, C.c_int64, _dp, _dp, _dp,
                                     C.POINTER(_vp)]),
    "mdl_table_create_analytic": (C.c_int32, [C.c_int32, _dp, C.POINTER(_vp)]),
    "mdl_table_destroy": (C.c_int32, [_vp]),
    "mdl_setfl_read": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(_vp)]),
    "mdl_setfl_destroy": (C.c_int32, [_vp]),
    "mdl_setfl_grid": (C.c_int32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), _dp,
                                   C.POINTER(C.c_int64), _dp, _dp, _dp, _dp]),
    "mdl_setfl_element": (C.c_int32, [_vp, C.c_int32, C.c_char_p, C.c_int32, _dp, _dp, _dp,
                                      C.c_char_p, C.c_int32]),
    "mdl_setfl_nodes": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "mdl_spline5_node_derivatives": (C.c_int32, [C.c_int64, _dp, _dp, _dp, _dp]),
    "mdl_table_coefficients": (C.c_int32, [_vp, _dp, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int32)]),
    "mdl_table_eval": (C.c_int32, [_vp, C.c_int64, _dp, _dp, _dp, _dp]),
    "mdl_driver_create": (C.c_int32, [C.POINTER(Config), C.POINTER(_vp)]),
    "mdl_driver_destroy": (C.c_int32, [_vp]),
    "mdl_driver_integrate": (C.c_int32, [_vp, C.c_int64]),
    "mdl_driver_run": (C.c_int32, [_vp, C.c_int64, C.c_int64, C.c_int64, _dp,
                                   C.POINTER(C.c_int64)]),
    "mdl_drivers_run": (C.c_int32, [C.POINTER(_vp), C.c_int32, C.c_int64, C.c_int64, C.c_int64,
                                    C.POINTER(_dp), C.POINTER(C.c_int64)]),
    "mdl_replicas_run": (C.c_int32, [C.POINTER(_vp), C.c_int32, C.c_int64, C.c_int64, C.c_int64,
                                    C.POINTER(_dp), C.POINTER(C.c_int64)]),
    "mdl_driver_calculate_neighbors": (C.c_int32, [_vp]),
    "mdl_driver_pair_correlation": (C.c_int32, [_vp, _dp, C.c_int32, C.c_int32, _dp, _dp]),
    "mdl_driver_pair_correlation_noalloc": (C.c_int32, [_vp, _dp, C.c_int32, C.c_int32, _dp,
                                                        _dp]),
    "mdl_driver_set_sample_observables": (C.c_int32, [_vp, C.c_int32]),
    "mdl_driver_get_sample_observables": (C.c_int32, [_vp, C.POINTER(C.c_int32)]),
    "mdl_driver_get_scalars": (C.c_int32, [_vp, _dp]),
    "mdl_driver_get_atoms": (C.c_int32, [_vp, C.c_int32, _dp]),
    "mdl_driver_get_chain": (C.c_int32, [_vp, C.c_int32, _dp, _dp]),
    "mdl_driver_kernel_times": (C.c_int32, [_vp, _dp]),
    "mdl_driver_set_graph_steps": (C.c_int32, [_vp, C.c_int32]),
    "mdl_expx_eval": (C.c_int32, [C.c_int32, C.c_int64, _dp, _dp, _dp]),
    # learned control variate (include/mdelast_cv.h); device pointers as c_void_p
    "mdl_cv_descriptors": (C.c_int32, [_vp, _vp, C.c_int32, _dp, C.c_float, C.c_float,
                                       C.c_float, C.c_float, _vp, _vp, _vp, _vp]),
    "mdl_cv_stress_basis_sums": (C.c_int32, [_vp, _vp, C.c_int32, C.c_int32, _dp, C.c_double,
                                             C.c_double, C.c_double, C.c_double, _vp, _vp,
                                             C.c_int32, _vp, C.c_int32, _vp]),
    "mdl_cv_stress_proxy": (C.c_int32, [_vp, _vp, C.c_int32, C.c_int32, _dp, C.c_double,
                                        C.c_double, C.c_double, C.c_double, _vp, _vp, _vp,
                                        _vp]),
    "mdl_cv_stress_fields": (C.c_int32, [_vp, _vp, C.c_double, _vp, C.c_int32, C.c_int32, _dp,
                                         C.c_double, C.c_double, C.c_double, C.c_double, _vp,
                                         _vp, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    "mdl_driver_cv_descriptors": (C.c_int32, [_vp, C.c_float, C.c_float, C.c_float, C.c_float,
                                              _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mdl_driver_set_cv_capture": (C.c_int32, [_vp, C.c_float, C.c_float, C.c_float, C.c_float,
                                              _vp, _vp, _vp, _vp, _vp, C.c_int32]),
    "mdl_driver_set_cv_capture_pairs": (C.c_int32, [_vp, _vp, _vp, C.c_int32]),
    "mdl_driver_stream": (C.c_int32, [_vp, C.POINTER(_vp)]),
    "mdl_driver_copy_atoms": (C.c_int32, [_vp, C.c_int32, _vp]),
    "mdl_cv_gemm": (C.c_int32, [_vp, _vp, C.c_int32, _vp, _vp, C.c_int64, C.c_int32, C.c_int32,
                                _vp]),
    "mdl_cv_gemm_tn": (C.c_int32, [_vp, _vp, _vp, _vp, C.c_int64, C.c_int32, C.c_int32,
                                   C.c_int32, _vp]),
    "mdl_cv_colsum": (C.c_int32, [_vp, _vp, _vp, C.c_int64, C.c_int32, _vp]),
    "mdl_cv_act_forward": (C.c_int32, [_vp, _vp, _vp, _vp, C.c_int64, C.c_int32, _vp]),
    "mdl_cv_field_contract": (C.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_float, C.c_int64,
                                          C.c_int32, _vp, _vp, _vp, _vp]),
    "mdl_cv_field_contract_backward": (C.c_int32, [_vp, _vp, _vp, _vp, C.c_float, _vp, _vp, _vp,
                                                   C.c_int64, C.c_int32, _vp, _vp, _vp]),
    "mdl_cv_act_backward": (C.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int64, C.c_int32,
                                        _vp]),
}

_lib = None
# whether torch's HIP runtime was in the process when the library was loaded
BOUND_WITH_TORCH = None


def lib():
    """The loaded library; raises if it is not built or not loadable."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MdlError(11, f"{LIB_PATH} is not built (run `python -m mdelast_amd.build`)")
        # The library binds the HIP runtime already in the process (torch's,
        # when torch was imported first) or else the system one. torch asks for
        # its bundled runtime by another file name, so importing torch after
        # this point would map a second runtime that finds no device: a
        # process that uses the control variate imports torch (mdelast_amd.cv
        # does) before the library's first use.
        global BOUND_WITH_TORCH
        BOUND_WITH_TORCH = "torch" in sys.modules
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != 0:
        raise MdlError(status, lib().mdl_last_error().decode(errors="replace"))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int32(0)
    check(lib().mdl_device_info(name, 256, C.byref(cu)))
    return name.value.decode(), cu.value


def expx_eval(x, form=0):
    """exp(-x) and (1 - exp(-x))/x on the GPU as the integrator takes them: form 0
    the barostat's operator, 1 the chain sweeps' wide form, 2 their series form."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    e = np.empty_like(x)
    op = np.empty_like(x)
    check(lib().mdl_expx_eval(int(form), x.size, dptr(x), dptr(e), dptr(op)))
    return e, op
