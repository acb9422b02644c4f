"""The NPT -> NVT handoff of the two-stage calculation (sample_case/run.py:251-283)
on a synthetic finished NPT stage: the last dump by number (not by name), its
species mapped to types, positions and box scaled by <V/V0>^(1/3)."""

import types

import numpy as np
import pytest

from mdelast_amd.md_driver_module import Units
from mdelast_amd.molecular_dynamics import Material, npt_to_nvt_material


def _write_dump(path, rows):
    with open(path, "w") as f:
        f.write(f"{len(rows)}\na= 3.6 Properties=species:S:1:pos:R:3:vel:R:3 Time=1.0\n")
        for name, p in rows:
            f.write(f"{name} {p[0]} {p[1]} {p[2]} 0.1 0.2 0.3\n")


def _npt(tmp_path, v_v0, last_rows, n=3):
    dump = tmp_path / "dump"
    dump.mkdir()
    # 9.xyz sorts after 10.xyz by name; the handoff must take 10.xyz
    _write_dump(dump / "9.xyz", [("Zn", (9.0, 9.0, 9.0))] * n)
    _write_dump(dump / "10.xyz", last_rows)
    mat = Material(potential=["p0", "p1"], atomic_weight=[65.38, 63.546], atom_name="Zn-Cu",
                   eam=["e0", "e1"], lattice=[None] * n, atom_type=[0] * n,
                   box_length=(1.0, 1.0, 1.0))
    series = types.SimpleNamespace(V_V0=types.SimpleNamespace(mean=v_v0))
    L = 2.0e-7
    return types.SimpleNamespace(dump_dir=str(dump), material=mat, series=series,
                                 Nparticles=n, Lx=L, Ly=1.5 * L, Lz=0.5 * L)


def test_handoff_scales_last_dump(tmp_path):
    rows = [("Cu", (1.0, -2.0, 3.0)), ("Zn", (0.5, 0.25, -0.125)), ("Cu", (0.0, 4.0, -1.0))]
    npt = _npt(tmp_path, 1.030301, rows)          # 1.01^3
    m = npt_to_nvt_material(npt)
    s = 1.01
    assert m.atom_type == [1, 0, 1]
    got = np.array([[p.x, p.y, p.z] for p in m.lattice])
    want = np.array([r[1] for r in rows]) * Units.angstrom * s
    np.testing.assert_allclose(got, want, rtol=1e-14)
    np.testing.assert_allclose(m.box_length, (2.0e-7 * s, 3.0e-7 * s, 1.0e-7 * s), rtol=1e-14)
    assert m.potential == ["p0", "p1"] and m.eam == ["e0", "e1"]
    assert m.atomic_weight == [65.38, 63.546] and m.atom_name == "Zn-Cu"


def test_handoff_rejects_short_dump(tmp_path):
    npt = _npt(tmp_path, 1.0, [("Cu", (0.0, 0.0, 0.0))], n=3)
    with pytest.raises(ValueError, match="holds 1 atoms"):
        npt_to_nvt_material(npt)


def test_handoff_needs_a_dump(tmp_path):
    npt = _npt(tmp_path, 1.0, [("Cu", (0.0, 0.0, 0.0))] * 3)
    for f in (tmp_path / "dump").iterdir():
        f.unlink()
    with pytest.raises(ValueError, match="no trajectory dump"):
        npt_to_nvt_material(npt)
