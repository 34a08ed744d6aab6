"""Pins of the oracle's MD integrators (orc_md_run: prop_velocity_verlet,
nhc2/chain_nhc2, nhcn/chain_nhcn, src/integration.f90:131-713) by what the
mathematics fixes, on the CPU:

* velocity Verlet is exact for a constant force (r0 + v0 t + F t^2 / 2m);
* velocity Verlet is time reversible;
* NVE energy drift, and the conserved quantity of the Nose-Hoover chains,
  fluctuate as dt^2 (a dropped half-step factor or a wrong exponent breaks
  the scaling or the conservation);
* chain_nhc2's extended energy, written with the chain velocities of its
  own equations of motion (Frenkel & Smit), is conserved;
* one chain_nhcn call from rest against the closed form of its first
  sweeps, which fixes the reference's stale middle-chain forces;
* the Yoshida weights sum to 1, are symmetric (order 5 keeps the
  reference's order, middle weight last) and cancel the third-order error
  (sum w^3 = 0).
"""
import math

import numpy as np
import pytest

from mdff_amd import configs as C
from mdff_amd import params as P


def _lj_small(jitter=0.05, seed=3):
    s = C.lj_fcc(ncell=3)  # 108 atoms
    # linear truncation: the energy is shifted to 0 at the cutoff, so it is
    # conserved up to the integrator's own error and the force jump at rc
    # (the reference never truncates the force; its quadratic truncation
    # adds uc1 r^2 to the energy without the matching force, so it is not
    # conserved)
    s.params.trunc = 1
    rng = np.random.default_rng(seed)
    s.r = s.r + jitter * rng.standard_normal(s.r.shape)
    return s


def _run(O, s, mdp, v, nsteps, f=None, xi=None, vxi=None, itime0=1):
    orc = O.Oracle(s.params, s.itype, s.A)
    if f is None:
        f = orc.engforce(s.r, itime=0)["f"]
    return orc.md_run(mdp, s.r, v, f, nsteps, itime0=itime0, xi=xi, vxi=vxi)


def test_velocity_verlet_exact_for_constant_force(oracle_lib):
    s = _lj_small()
    s.params.lnmlj = 0  # no potential: the forces stay what they are
    rng = np.random.default_rng(1)
    F = rng.standard_normal(s.r.shape)
    v0 = rng.standard_normal(s.r.shape)
    m = 2.5
    dt, n = 0.01, 25
    mdp = P.md_params("nve-vv", dt=dt, mass=(m,), lreduced=True)
    r, v, f, _, _, steps = _run(oracle_lib, s, mdp, v0, n, f=F)
    t = dt * n
    np.testing.assert_allclose(r, s.r + v0 * t + F * t * t / (2 * m), rtol=0, atol=1e-12)
    np.testing.assert_allclose(v, v0 + F * t / m, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(f, F)
    ek = 0.5 * m * (v * v).sum()
    assert steps[-1]["e_kin"] == pytest.approx(ek, rel=1e-13)
    assert steps[-1]["temp_r"] == pytest.approx(2 * ek / (3 * s.natm), rel=1e-13)


def test_velocity_verlet_time_reversible(oracle_lib):
    s = _lj_small()
    v0 = C.maxwell_velocities(s, 0.72, seed=2)
    mdp = P.md_params("nve-vv", dt=0.005, mass=(1.0,), lreduced=True)
    r1, v1, f1, _, _, _ = _run(oracle_lib, s, mdp, v0, 40)
    back = _lj_small()
    back.r = r1
    r2, v2, _, _, _, _ = _run(oracle_lib, back, mdp, -v1, 40, f=f1)
    np.testing.assert_allclose(r2, s.r, rtol=0, atol=1e-9)
    np.testing.assert_allclose(-v2, v0, rtol=0, atol=1e-9)


def _fluct(x):
    x = np.asarray(x)
    return float(np.max(np.abs(x - x[0])))


@pytest.mark.parametrize("integ,nhc_n", [("nve-vv", 0), ("nvt-nhcn", 2)])
def test_conserved_energy_scales_as_dt2(oracle_lib, integ, nhc_n):
    s = _lj_small()
    v0 = C.maxwell_velocities(s, 1.2, seed=4)
    fl = []
    for dt, nst in ((0.004, 100), (0.002, 200)):
        mdp = P.md_params(integ, dt=dt, temp=0.72, timesca_thermo=0.2, nhc_n=max(nhc_n, 2),
                          mass=(1.0,), lreduced=True)
        _, _, _, _, _, steps = _run(oracle_lib, s, mdp, v0, nst)
        h = [st["u_lj"] + st["e_kin"] + st["e_nvt"] for st in steps]
        fl.append(_fluct(h))
    # second order: halving dt divides the error by about four
    assert 2.5 < fl[0] / fl[1] < 6.0, fl
    assert fl[0] < 2e-3 * abs(steps[0]["u_lj"])


def test_nhcn_thermostat_moves_towards_target(oracle_lib):
    s = _lj_small()
    v0 = C.maxwell_velocities(s, 2.0, seed=5)
    mdp = P.md_params("nvt-nhcn", dt=0.004, temp=0.5, timesca_thermo=0.05, nhc_n=4,
                      mass=(1.0,), lreduced=True)
    _, _, _, xi, vxi, steps = _run(oracle_lib, s, mdp, v0, 300)
    t = np.array([st["temp_r"] for st in steps])
    # the system starts at 2.0 (and heats from the potential); the chain
    # drives it to 0.5 on average over the last third
    assert abs(t[200:].mean() - 0.5) < 0.2, t[200:].mean()
    assert np.all(np.isfinite(xi[:4])) and np.all(np.isfinite(vxi[:4]))


def test_nhc2_extended_energy_conserved(oracle_lib):
    """chain_nhc2 integrates Frenkel & Smit's two-link chain, whose chain
    'velocities' are d xi / dt: H = K + U + Q1 vxi1^2/2 + Q2 vxi2^2/2 +
    L kT xi1 + kT xi2 is conserved (the reference's e_nvt uses vxi^2 / 2Q
    instead, which is not)."""
    s = _lj_small()
    v0 = C.maxwell_velocities(s, 1.2, seed=6)
    fl = []
    for dt, nst in ((0.004, 100), (0.002, 200)):
        mdp = P.md_params("nvt-nhc2", dt=dt, temp=0.72, timesca_thermo=0.2, mass=(1.0,),
                          lreduced=True)
        orc = oracle_lib.Oracle(s.params, s.itype, s.A)
        f = orc.engforce(s.r, itime=0)["f"]
        r, v, xi, vxi = s.r, v0, np.zeros(16), np.zeros(16)
        L = 3.0 * (s.natm - 1)
        Q1 = mdp.timesca_thermo ** 2 * mdp.temp * L
        Q2 = mdp.timesca_thermo ** 2 * mdp.temp
        h = []
        for _ in range(nst):
            r, v, f, xi, vxi, st = orc.md_run(mdp, r, v, f, 1, xi=xi, vxi=vxi)
            st = st[0]
            h.append(st["u_lj"] + st["e_kin"] + 0.5 * Q1 * vxi[0] ** 2 + 0.5 * Q2 * vxi[1] ** 2
                     + L * mdp.temp * xi[0] + mdp.temp * xi[1])
            # the reference's own e_nvt formula (src/integration.f90:237-243)
            assert st["e_nvt"] == pytest.approx(L * mdp.temp * xi[0] + 0.5 * vxi[0] ** 2 / Q1
                                                + 0.5 * vxi[1] ** 2 / Q2 + mdp.temp * xi[1],
                                                rel=1e-12, abs=1e-15)
        fl.append(_fluct(h))
    assert 2.5 < fl[0] / fl[1] < 6.0, fl


def test_chain_nhcn_first_sweep_closed_form(oracle_lib):
    """One chain_nhcn call (Yoshida order 1, one step) of a three-link chain
    from rest.  G(2) is still zero when the descending sweep uses it (the
    reference zeroes G per call and sets only G(1) and G(nhc_n) first), so
    the middle link starts at 0; with the chain at rest every exponential
    of the descending sweep is 1."""
    K, L, T, dt = 3.7, 30.0, 0.9, 0.01
    Q = np.array([2.0 * L, 2.0, 2.0])
    vxi = np.zeros(3)
    xi = np.zeros(3)
    s = oracle_lib.chain_nhcn(K, vxi, xi, Q, L, T, dt, 3, yosh=1, mults=1)
    d2, d4, d8 = dt / 2, dt / 4, dt / 8
    g1 = 2 * K - L * T
    v3 = -T * d4
    v1 = g1 * d4
    s_ref = math.exp(-v1 * d2 / Q[0])
    xi_ref = np.array([v1 * d2 / Q[0], 0.0, v3 * d2 / Q[2]])
    v1 = v1 + (2 * K * s_ref ** 2 - L * T) * d4
    g2 = v1 * v1 / Q[0] - T
    v2 = g2 * d4 * math.exp(-v3 * d8 / Q[2])
    v3 = v3 + (v2 * v2 / Q[1] - T) * d4
    assert s == pytest.approx(s_ref, rel=1e-15)
    np.testing.assert_allclose(vxi, [v1, v2, v3], rtol=1e-14, atol=0)
    np.testing.assert_allclose(xi, xi_ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("order", [1, 3, 5, 7, 9])
def test_yoshida_weights(oracle_lib, order):
    w = oracle_lib.yoshida(order)
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    if order == 5:
        # the reference puts Suzuki's middle weight 1 - 4w last
        # (src/integration.f90:611-616), not in the middle
        np.testing.assert_array_equal(w[:4], np.full(4, w[0]))
        assert w[4] == pytest.approx(1.0 - 4.0 * w[0], abs=1e-15)
    else:
        np.testing.assert_array_equal(w, w[::-1])
    if order > 1:
        assert abs((w ** 3).sum()) < 2e-13
