"""Host-side checks that need no GPU: the C-ABI library, the Python mirror of
md_driver_module, the SETFL reader, and the bench entry's argument contract."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import cells
from mdelast_amd import _native as nat
from mdelast_amd import md_driver_module as mdm
from mdelast_amd import setfl, units as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc))
                   if f.endswith(".h"))
    return sorted(set(re.findall(r"\b(mdl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 15
    lib = ctypes.CDLL(nat.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(nat.SIGNATURES), set(syms) ^ set(nat.SIGNATURES)
    nat.lib()   # binds every signature


def test_units_are_the_reference_constants():
    # src/units.hpp:6-24
    assert mdm.Units.k_boltz == 1.38064852e-16
    assert mdm.Units.ev == 1.6021773e-12
    assert mdm.Units.Navogadro == 6.02214086e23
    assert mdm.Units.GPa == 1e9 * (1e-5 * 1e6)
    assert mdm.Units.fs == 1e-15 and mdm.Units.angstrom == 1e-8


def test_api_objects_on_host():
    v = mdm.Vector3d(1.0, 2.0, 2.0)
    assert v.length() == 3.0 and v.length2() == 9.0
    assert repr(v) == "(1.000000,2.000000,2.000000)"
    b = mdm.Box(np.diag([2.0, 3.0, 4.0]))
    assert b.volume() == 24.0 and b.length() == 3.0
    th = mdm.Thermostat(300.0, [1.0, 2.0], [0.1, 0.2], 30.0)
    assert th.nchain == 2 and th.eta == [0.0, 0.0]
    ba = mdm.Barostat(0.0, 1.0, 0.5, th)
    assert ba.epsilon == 0.0 and ba.p_epsilon == 0.5
    assert int(mdm.MDDriver.NPT) == 2 and int(mdm.MDDriver.THERMOSTAT_AVERAGE) == 2
    assert mdm.MDDriver.Ensamble.NVT == mdm.MDDriver.NVT
    data = cells.setfl_data("ZnCu.eam.alloy")
    p = data.pairs[0]
    t = mdm.TablePotential(data.rcut, data.rmin_fatal, data.dr, p.phi, p.dphi, p.d2phi)
    assert t.interpolation == "quintic"
    with pytest.raises(ValueError):
        mdm.TablePotential(data.rcut, 0.0, data.dr, p.phi, p.dphi, p.d2phi, "cubic")
    # analytic potentials (potential.hpp:29-75) evaluate on the host
    m = mdm.Morse(10.0, 1.0, 1.0, 2.0)
    assert abs(m.calculate(2.0).V + 1.0) < 1e-15 and m.calculate(11.0).V == 0.0
    lj = mdm.LennardJones(3.0, 1.0, 1.0)
    assert abs(lj.calculate(3.0).V) < 1e-15


def test_get_potential_interface():
    from mdelast_amd import get_potential
    el, pots, eams = get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")
    assert [e[0] for e in el] == ["Zn", "Cu"] and len(pots) == 3 and len(eams) == 2
    assert pots[0].interpolation == "quintic" and eams[1].interpolation == "quintic"
    assert el[1][2] == 63.55   # Cu mass as written in ZnCu.eam.alloy
    with pytest.raises(ValueError):
        get_potential.get_EAM_alloy_file("ZnCu.eam.alloy", interpolation="cubic")
    saved = get_potential.DEFAULT_INTERPOLATION
    try:
        get_potential.DEFAULT_INTERPOLATION = "bspline"
        assert get_potential.get_EAM_alloy_file("ZnCu.eam.alloy")[1][0].interpolation == "bspline"
    finally:
        get_potential.DEFAULT_INTERPOLATION = saved


def test_setfl_reader_matches_loadtxt_layout():
    # the reference reads each block with numpy.loadtxt (get_potential.py:110, :126)
    data = setfl.read_setfl("ZnCu.eam.alloy")
    path = data.path
    Nrho, Nr = data.Nrho, data.Nr
    blk = np.loadtxt(path, skiprows=6, max_rows=Nrho + Nr)
    assert np.array_equal(blk[:Nrho] * U.ev, data.elements[0].F)
    assert np.array_equal(blk[Nrho:] / U.angstrom ** 3, data.elements[0].rho)
    r = np.linspace(0.0, data.dr * (Nr - 1.0), Nr)
    idx = 5 + 2 * (1 + Nrho + Nr)
    phir = np.loadtxt(path, skiprows=idx, max_rows=Nr)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = phir / r * U.ev * U.angstrom
    phi[0] = phi[1]
    assert np.array_equal(phi, data.pairs[0].phi)
    assert data.rcut == r[-1] and data.rcut != 6.8034017008504248e-08


def test_no_silent_fallback_without_gpu():
    """Without a GPU the driver must fail loudly, never compute on the CPU."""
    try:
        nat.device_info()
        pytest.skip("a GPU is present")
    except nat.MdlError as e:
        assert e.kind == "DEVICE"
    cell = cells.smoke_cell()
    import harness as H
    with pytest.raises(nat.MdlError):
        H.gpu_md(cell, H.md_setup(cell, static=True))


def test_driver_rejects_analytic_potentials_and_bad_boxes():
    th = mdm.Thermostat(300.0, [1.0], [0.0], 3.0)
    ba = mdm.Barostat(0.0, 1.0, 0.0, th)
    pos = [mdm.Vector3d(0, 0, 0)]
    with pytest.raises(mdm.MDDriverError):
        mdm.MDDriver(pos, pos, mdm.PotentialVector([mdm.Morse(5.0, 1.0, 1.0, 1.0)]),
                     mdm.EAMVector([]), mdm.MDDriver.NVE, 0, th, ba,
                     mdm.Box(np.diag([10.0, 10.0, 10.0])), [0.1], 3.0, [1.0], [0], 1e-15,
                     False, 1.0, 1e3)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
