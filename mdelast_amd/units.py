"""CGS unit constants of the MD driver.

The values are the driver's own (src/units.hpp:6-24), not CODATA: every table,
force and energy crosses the C-ABI in these units, so they must be bit-identical
to the ones the reference uses (eV = 1.6021773e-12 erg, k_B = 1.38064852e-16 erg/K).
"""

k_boltz = 1.38064852e-16
Navogadro = 6.02214086e23

ev = 1.6021773e-12
joule = 1e7

angstrom = 1e-8
nm = 1e-7

bar = 1e6
kbar = 1e3 * bar
atm = 1.01325 * bar
Pa = 1e-5 * bar
GPa = 1e9 * Pa

mus = 1e-6
ns = 1e-9
ps = 1e-12
fs = 1e-15
