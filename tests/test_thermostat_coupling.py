"""Thermostat coupling of the oracle (CONFIGURATIONAL and AVERAGE,
md_driver.cpp:512-531 of the reference; thermostat.cpp:40-66), CPU only.

With a chain of one thermostat the U4 propagator reduces to
p_eta += dt G with G = 2 e_couple - N_f kT (thermostat.cpp:53-64), and the
Suzuki-Yoshida weights sum to one. That fixes one NVT step from outside the
oracle:

* CONFIGURATIONAL: e_couple = (k_B / T_conf^-1) N_f / 2 does not change while the
  velocities are scaled, so over a whole step
  p_eta(1) = p_eta(0) + dt/2 (k_B N_f T_conf(0) - N_f kT) + dt/2 (k_B N_f T_conf(1) - N_f kT)
  with T_conf(0) of the initial state and T_conf(1) of the step's forces;
* every coupling: the half-step propagation is a scalar recurrence in
  (p_eta, the velocity scale S, K) that the step's velocity-Verlet part joins:
  v(1) = S_2 (S_1 v(0) + dt/2 (F(0) + F(1)) / m). The recurrence below takes the
  oracle's forces before and after the step and returns v(1) and p_eta(1);
  a coupling that took the wrong energy (kinetic for configurational, one
  term of the average, a factor of two) differs from the oracle by far more
  than round-off.
"""

from __future__ import annotations

import numpy as np
import pytest

import cells
import harness as H
from mdelast_amd import units as U

# Suzuki-Yoshida seven-stage weights (Yoshida, Phys. Lett. A 150 (1990) 262;
# Tuckerman, Statistical Mechanics, eq. 4.11.12)
_W = [0.784513610477560, 0.235573213359357, -1.17767998417887]
WSY7 = _W + [1.0 - 2.0 * sum(_W)] + _W[::-1]


def _half_step(p, K, econf, coupling, q, Nf, kT, dt):
    """The NHC half step (dt) for one thermostat: returns p, the velocity scale, K."""
    S = 1.0

    def e(K):
        return K if coupling == 0 else econf if coupling == 1 else 0.5 * (K + econf)

    for w in WSY7:
        dtw = w * dt
        p += 0.5 * dtw * (2.0 * e(K) - Nf * kT)
        f = np.exp(-dtw * p / q)
        S *= f
        K *= f * f
        p += 0.5 * dtw * (2.0 * e(K) - Nf * kT)
    return p, S, K


def _setup():
    cell = cells.zn06_cell(5)
    setup = H.md_setup(cell, T=300.0, M=1, seed=5)
    return cell, setup


@pytest.mark.parametrize("coupling", [0, 1, 2])
def test_one_step_against_the_scalar_recurrence(coupling):
    cell, setup = _setup()
    dt = 2e-15
    o = H.oracle_md(cell, setup, ensemble="NVT", dt=dt, coupling=coupling)
    m = np.array(cells.masses(cell["data"]))[cell["types"]][:, None]
    T, q, p0, Nf = setup["thermostat"]
    q, p0, kT = float(q[0]), float(p0[0]), U.k_boltz * T
    s0 = o.scalars()
    v0, F0 = o.atoms("velocities"), o.atoms("forces")
    econf0 = U.k_boltz / s0["temperature_conf_inverse"] * Nf / 2.0
    K0 = 0.5 * np.sum(m * v0 ** 2)
    p, S1, _ = _half_step(p0, K0, econf0, coupling, q, Nf, kT, 0.5 * dt)
    o.integrate()
    s1 = o.scalars()
    F1 = o.atoms("forces")
    vmid = S1 * v0 + 0.5 * dt * (F0 + F1) / m
    econf1 = U.k_boltz / s1["temperature_conf_inverse"] * Nf / 2.0
    Kmid = 0.5 * np.sum(m * vmid ** 2)
    p, S2, _ = _half_step(p, Kmid, econf1, coupling, q, Nf, kT, 0.5 * dt)
    v1 = S2 * vmid
    po, _ = o.chain(0)
    assert abs(po[0] / p - 1) < 1e-11, (po[0], p)
    assert np.abs(o.atoms("velocities") - v1).max() < 1e-12 * np.abs(v1).max()
    if coupling == 1:
        closed = p0 + 0.5 * dt * (2 * econf0 - Nf * kT) + 0.5 * dt * (2 * econf1 - Nf * kT)
        assert abs(po[0] / closed - 1) < 1e-12
    # the other couplings' recurrences miss the oracle's p_eta by far more
    for other in {0, 1, 2} - {coupling}:
        pa, Sa, _ = _half_step(p0, K0, econf0, other, q, Nf, kT, 0.5 * dt)
        va = Sa * v0 + 0.5 * dt * (F0 + F1) / m
        pa, _, _ = _half_step(pa, 0.5 * np.sum(m * va ** 2), econf1, other, q, Nf, kT, 0.5 * dt)
        assert abs(po[0] / pa - 1) > 1e-6, other


def test_couplings_differ_and_share_the_first_forces():
    cell, setup = _setup()
    runs = []
    for c in (0, 1, 2):
        o = H.oracle_md(cell, setup, ensemble="NVT", dt=2e-15, coupling=c)
        o.integrate(5)
        runs.append((o.chain(0)[0][0], o.atoms("velocities")))
    assert len({round(r[0] / runs[0][0], 12) for r in runs}) == 3
