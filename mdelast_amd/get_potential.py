"""Potential loaders (mirror of src/python/get_potential.py of mdelast v1.2).

``get_EAM_alloy_file(file_name, interpolation=None)`` returns
``(elements, potentials, eams)`` exactly as the reference does
(get_potential.py:58-147): element records ``[symbol, Z, mass, a]``, the pair
potentials in the driver's triangular slot order, one EAM per species.
``get_EAM_file(*, file_name, a_factor=1.0, interpolation=None)`` returns the
one-element ``Material`` of get_potential.py:150-257. The node data come from
:mod:`mdelast_amd.setfl`.
"""

from __future__ import annotations

from . import md_driver_module as mdm
from . import setfl
from . import units as U

DEFAULT_INTERPOLATION = "quintic"
INTERPOLATIONS = ("quintic", "bspline")


def _interp(interpolation):
    choice = DEFAULT_INTERPOLATION if interpolation is None else interpolation
    if choice not in INTERPOLATIONS:
        raise ValueError(f"interpolation must be one of {INTERPOLATIONS}, got {choice!r}")
    return choice


def tables_from_setfl(data: setfl.SetflData, interpolation: str):
    """Driver objects (pair potentials by slot, EAM per species) for SETFL node data."""
    potentials = [mdm.TablePotential(data.rcut, data.rmin_fatal, data.dr, p.phi, p.dphi,
                                     p.d2phi, interpolation) for p in data.pairs]
    eams = [mdm.EAM(data.rcut, data.rhomax, data.dr, data.drho, e.rho, e.drho, e.d2rho, e.F,
                    e.dF, e.d2F, interpolation) for e in data.elements]
    return potentials, eams


def get_EAM_alloy_file(file_name, interpolation=None):
    """Load a multi-element LAMMPS SETFL file (get_potential.py:58-147)."""
    interpolation = _interp(interpolation)
    data = setfl.read_setfl(file_name)
    elements = [[e.name, e.atomic_number, e.atomic_weight, e.a_lattice_angstrom]
                for e in data.elements]
    potentials, eams = tables_from_setfl(data, interpolation)
    return elements, potentials, eams


def get_EAM_file(*, file_name, a_factor=1.0, interpolation=None):
    """One-element EAM file (SETFL ``*.eam.alloy.txt`` or FS ``*.eam.fs.txt``) as a
    ``Material`` (get_potential.py:150-257): one pair table and one EAM, the
    file's atomic weight, atom name "Cu" as in the reference, and the file's
    lattice constant times ``a_factor`` (CGS)."""
    from .molecular_dynamics import Material
    interpolation = _interp(interpolation)
    data = setfl.read_eam_file(file_name)
    potentials, eams = tables_from_setfl(data, interpolation)
    el = data.elements[0]
    return Material(atomic_weight=el.atomic_weight, atom_name="Cu", potential=potentials[0],
                    eam=eams[0], a_lattice=el.a_lattice_angstrom * U.angstrom * a_factor)


def get_Ar_LJ(*, epsilon=None, sigma=None, rcut_fac=3.5):
    """Argon Lennard-Jones ``Material`` (get_potential.py:260-273)."""
    from .molecular_dynamics import get_Ar_LJ as _get
    return _get(epsilon=epsilon, sigma=sigma, rcut_fac=rcut_fac)
