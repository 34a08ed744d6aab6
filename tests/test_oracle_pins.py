"""Pins the oracle functions that the reference's example runs do not reach:
the dipole extrapolation of the first SCF iterate, the Lennard-Jones
truncation constants and the surface term.  Each check comes from outside
the oracle's own formulas: Kolafa's published closed form for the ASPC
coefficients, exactness of the predictors on a history linear in time, the
conditions that define the truncated potentials at the cutoff, and
finite-difference derivatives of the energy.  CPU only.
"""
import math

import numpy as np
import pytest

from mdff_amd import configs as C
from mdff_amd import params as P


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# --------------------------------------------------------------------------
# dipole extrapolation (src/field.f90:3807-3992)

def _ext_params(n, algo, order, ntype=2):
    p = P.field_defaults(n, ntype)
    p.algo_ext_dipole = algo
    p.extrapolate_order = order
    p.calc_mode = P.CALC_MD
    return p


def kolafa_b(m):
    """ASPC predictor coefficients of Kolafa, J. Comput. Chem. 25, 335
    (2004), for the reference's order m (m + 1 history points):
    B_j = (-1)^(j+1) j C(2m+2, m+1-j) / C(2m, m), j = 1..m+1."""
    return np.array([(-1) ** (j + 1) * j * math.comb(2 * m + 2, m + 1 - j) / math.comb(2 * m, m)
                     for j in range(1, m + 2)])


def kolafa_w(m):
    """Kolafa's corrector weight (m + 1) / (2m + 1)."""
    return (m + 1) / (2 * m + 1)


@pytest.mark.parametrize("m", range(1, 6))
def test_aspc_predictor_matches_kolafa_closed_form(O, m):
    n = 3
    p = _ext_params(n, P.EXT_ASPC, m)
    assert O.aspc_order(p, m + 1) == m
    got = []
    for j in range(m + 1):
        h = np.zeros((m + 1, n, 3), order="F")
        h[j] = 1.0  # unit history at lag j: the predictor returns B_{j+1}
        mu = O.extrapolate(p, m + 1, h)
        assert np.all(mu == mu[0, 0])
        got.append(mu[0, 0])
    assert np.allclose(got, kolafa_b(m), rtol=0, atol=1e-14), (got, kolafa_b(m))
    b, w = O.aspc_coeffs(m)
    assert abs(w - kolafa_w(m)) < 1e-15
    assert np.allclose(b[:m + 1], kolafa_b(m), rtol=0, atol=1e-14)
    assert np.all(b[m + 1:] == 0.0)


def test_aspc_order_zero_is_the_previous_step_without_corrector(O):
    """Order 0 (and the first two MD steps of any order): mu = h_0, and the
    reference's w_ASPC = 0 leaves the predictor unchanged by the corrector
    (Kolafa's formula would give w = 1 here)."""
    b, w = O.aspc_coeffs(0)
    assert b[0] == 1.0 and np.all(b[1:] == 0.0) and w == 0.0
    p = _ext_params(4, P.EXT_ASPC, 3)
    for itime in (0, 1):
        assert O.aspc_order(p, itime) == 0
    rng = np.random.default_rng(0)
    h = np.asfortranarray(rng.standard_normal((4, 4, 3)))
    mu = O.extrapolate(p, 1, h)
    assert np.array_equal(mu, h[0])


@pytest.mark.parametrize("algo,order", [(P.EXT_POLY, 1)] + [(P.EXT_ASPC, m) for m in range(1, 6)])
def test_predictor_exact_for_linear_history(O, algo, order):
    """A history linear in time, h_k = a + b (t - 1 - k), is extrapolated to
    a + b t exactly, at the full order and while the order ramps up."""
    n = 5
    rng = np.random.default_rng(order)
    a = rng.standard_normal((n, 3))
    b = rng.standard_normal((n, 3))
    t = 7.0
    p = _ext_params(n, algo, order)
    for itime in range(2, order + 4):
        h = np.asfortranarray(np.stack([a + b * (t - 1 - k) for k in range(order + 1)]))
        mu = O.extrapolate(p, itime, h)
        assert np.abs(mu - (a + b * t)).max() < 1e-12, (itime, np.abs(mu - (a + b * t)).max())


@pytest.mark.parametrize("algo,order", [(P.EXT_POLY, 0), (P.EXT_POLY, 1)] +
                         [(P.EXT_ASPC, m) for m in range(0, 6)])
def test_history_shift(O, algo, order):
    """After the predictor the history moves back by one step over the order
    in use (h_k <- h_{k-1}, src/field.f90:3836, 3977-3979), h_0 keeps the
    last converged moments until the SCF stores the new ones, and entries
    beyond the order in use are untouched; poly order 1 at the first step is
    zeroth order."""
    n = 3
    rng = np.random.default_rng(10 + order)
    p = _ext_params(n, algo, order)
    for itime in range(0, order + 3):
        h0 = np.asfortranarray(rng.standard_normal((order + 1, n, 3)))
        h = h0.copy(order="F")
        mu = O.extrapolate(p, itime, h)
        if algo == P.EXT_POLY:
            eo = order
            if order == 1 and itime <= 1:
                assert np.array_equal(mu, h0[0])
            elif order == 1:
                assert np.abs(mu - (2 * h0[0] - h0[1])).max() < 1e-15
            else:
                assert np.array_equal(mu, h0[0])
        else:
            eo = O.aspc_order(p, itime)
            assert eo == (order if itime >= order + 1 else (0 if itime <= 1 else itime - 1))
        assert np.array_equal(h[0], h0[0])
        for k in range(1, order + 1):
            if k <= eo:
                assert np.array_equal(h[k], h0[k - 1]), (itime, k)
            else:
                assert np.array_equal(h[k], h0[k]), (itime, k)


@pytest.mark.parametrize("m", range(1, 6))
def test_aspc_corrector_is_kolafa_mixing(O, m):
    """Corrector: mu = w alpha E + (1 - w) mu_p with Kolafa's weight, on
    polarizable atoms; u_pol = 1/2 sum alpha |E|^2 (that of alpha E)."""
    n = 6
    p = _ext_params(n, P.EXT_ASPC, m)
    alpha = 1.3
    p.ldip_polar[1] = 1
    P.field_check_tag(p, poldip_iso=[0.0, alpha])
    itype = np.array([1, 2, 2, 1, 2, 2], np.int32)
    rng = np.random.default_rng(m)
    E = rng.standard_normal((n, 3))
    mup = rng.standard_normal((n, 3))
    mu, upol = O.aspc_correct(p, itype, m + 1, E, mup)
    w = kolafa_w(m)
    pol = (itype == 2)[:, None]
    ref = w * np.where(pol, alpha * E, 0.0) + (1 - w) * mup
    assert np.abs(mu - ref).max() < 1e-14
    assert abs(upol - 0.5 * alpha * (E[itype == 2] ** 2).sum()) < 1e-13


# --------------------------------------------------------------------------
# Lennard-Jones truncations (src/field.f90:1043-1064, 1282-1290)

def _lj_dimer(O, d, trunc, p_exp=6.0, q_exp=12.0, rc=2.5):
    p = P.field_defaults(2, 1)
    P.reduced_units(p)
    p.lnmlj = 1
    p.cutshortrange = rc
    p.trunc = trunc
    k = P.pidx(1, 1)
    p.epslj[k], p.sigmalj[k], p.plj[k], p.qlj[k] = 1.3, 1.1, p_exp, q_exp
    P.field_check_tag(p)
    A = np.eye(3) * 20.0
    r = np.array([[1.0, 1.0, 1.0], [1.0 + d, 1.0, 1.0]])
    return O.Oracle(p, np.array([1, 1], np.int32), A).engforce(r)


@pytest.mark.parametrize("pq", [(6.0, 12.0), (8.0, 14.0)])
def test_lj_linear_truncation_vanishes_at_cutoff(O, pq):
    """trunc 1: U(r) = V(r) - V(rc), so U -> 0 at the cutoff and U - V is the
    same constant at every distance; forces are untruncated."""
    rc = 2.5
    for d in (rc * (1 - 1e-9), rc * (1 - 1e-7)):
        u = _lj_dimer(O, d, 1, *pq, rc=rc)["u_lj"]
        f = abs(_lj_dimer(O, d, 0, *pq, rc=rc)["f"][0, 0])
        assert abs(u) <= 2.0 * f * (rc - d) + 1e-15, (d, u)
    shifts = [_lj_dimer(O, d, 1, *pq)["u_lj"] - _lj_dimer(O, d, 0, *pq)["u_lj"] for d in (1.0, 1.7, 2.3)]
    assert np.ptp(shifts) < 1e-14 and shifts[0] != 0.0
    for d in (1.0, 2.3):
        assert np.array_equal(_lj_dimer(O, d, 1, *pq)["f"], _lj_dimer(O, d, 0, *pq)["f"])


@pytest.mark.parametrize("pq", [(6.0, 12.0), (8.0, 14.0)])
def test_lj_quadratic_truncation_smooth_at_cutoff(O, pq):
    """trunc 2: U(r) = V(r) + c1 r^2 - c2 with U(rc) = 0 and dU/dr(rc) = 0,
    so U vanishes quadratically at the cutoff; U - V is linear in r^2."""
    rc = 2.5
    h = 1e-3
    u1 = _lj_dimer(O, rc - h, 2, *pq, rc=rc)["u_lj"]
    u2 = _lj_dimer(O, rc - 2 * h, 2, *pq, rc=rc)["u_lj"]
    # U(rc - h) = U''(rc) h^2 / 2 + O(h^3): ratio 4 for a double root
    assert abs(u2 / u1 - 4.0) < 0.02, (u1, u2)
    # the linear truncation at the same point is first order
    v1 = _lj_dimer(O, rc - h, 1, *pq, rc=rc)["u_lj"]
    assert abs(u1) < 0.05 * abs(v1)
    ds = np.array([1.0, 1.6, 2.2])
    g = [_lj_dimer(O, d, 2, *pq)["u_lj"] - _lj_dimer(O, d, 0, *pq)["u_lj"] for d in ds]
    # linear in d^2: equal slopes between the three points
    s1 = (g[1] - g[0]) / (ds[1] ** 2 - ds[0] ** 2)
    s2 = (g[2] - g[1]) / (ds[2] ** 2 - ds[1] ** 2)
    assert abs(s1 - s2) < 1e-12 * abs(s1)
    assert np.array_equal(_lj_dimer(O, 1.6, 2, *pq)["f"], _lj_dimer(O, 1.6, 0, *pq)["f"])


# --------------------------------------------------------------------------
# surface term (src/field.f90:1928-1949, 1974-1990)

def _surf_params(n, lsurf, dip=True):
    p = P.field_defaults(n, 2)
    p.lcoulomb = 1
    p.cutlongrange = 3.99
    p.qch[0], p.qch[1] = 1.3, -0.65
    p.alphaES = 0.9
    p.kES[:] = [9, 9, 9]
    p.doefield = 1
    p.doefg = 1
    p.lsurf = lsurf
    P.field_check_tag(p)
    p.task_coul[:] = [1, int(dip), int(dip)]
    return p


def _surf_system(seed=8):
    rng = np.random.default_rng(seed)
    n = 18
    L = 8.0
    r = C.rsa_positions(n, L, 1.8, rng)
    itype = np.array([1] * 6 + [2] * 12, np.int32)
    q = np.where(itype == 1, 1.3, -0.65)
    mu = 0.2 * rng.standard_normal((n, 3))
    return r, itype, np.eye(3) * L, q, mu


def test_surface_energy_of_a_charge_pair(O):
    """u_surf = (2 pi / 3V) |sum q r + sum mu|^2 coul_unit: a +q/-q pair in a
    large box, with and without dipoles."""
    L = 12.0
    A = np.eye(3) * L
    r = np.array([[1.0, 2.0, 3.0], [3.5, 2.5, 1.0]])
    itype = np.array([1, 2], np.int32)
    q = np.array([1.3, -1.3])
    V = L ** 3
    for mu in (np.zeros((2, 3)), np.array([[0.1, -0.2, 0.3], [0.05, 0.4, -0.1]])):
        p0 = _surf_params(2, 0)
        p1 = _surf_params(2, 1)
        a = O.multipole_es(p0, itype, A, r, q, mu, [1, 1, 1])
        b = O.multipole_es(p1, itype, A, r, q, mu, [1, 1, 1])
        M = (q[:, None] * r).sum(axis=0) + mu.sum(axis=0)
        want = 2 * np.pi / (3 * V) * (M @ M) * p1.coul_unit
        assert abs((b["u_coul"] - a["u_coul"]) - want) < 1e-12 * abs(want)
        assert abs(b["u_surf"] * p1.coul_unit - want) < 1e-12 * abs(want)


def test_surface_forces_are_minus_energy_gradient(O):
    """F = -dU/dr with lsurf = 1 (charges and dipoles, moments fixed): the
    surface force -(4 pi / 3V) q_i (sum q r + sum mu) coul_unit is the
    gradient of u_surf."""
    r, itype, A, q, mu = _surf_system()
    p = _surf_params(len(q), 1)

    def U(rr):
        return O.multipole_es(p, itype, A, rr, q, mu, [1, 1, 1], do_efg=False)["u_coul"]

    out = O.multipole_es(p, itype, A, r, q, mu, [1, 1, 1])
    p0 = _surf_params(len(q), 0)
    out0 = O.multipole_es(p0, itype, A, r, q, mu, [1, 1, 1])
    h = 1e-5
    for i, a in [(0, 0), (4, 1), (9, 2), (17, 0)]:
        rp = r.copy()
        rp[i, a] += h
        rm = r.copy()
        rm[i, a] -= h
        fd = -(U(rp) - U(rm)) / (2 * h)
        assert abs(fd - out["f"][i, a]) < 2e-6 * np.abs(out["f"]).max(), (i, a, fd, out["f"][i, a])
    # the surface force differs from the plain one
    assert np.abs(out["f"] - out0["f"]).max() > 1e-3 * np.abs(out0["f"]).max()


def test_surface_field_is_unscaled_total_moment(O):
    """The reference adds sum q r + sum mu to the field as it is, without the
    4 pi / 3V factor of its energy and forces (src/field.f90:1940-1943, kept
    as a quirk): ef(lsurf) - ef(no lsurf) equals that vector at every atom,
    and the EFG and stress have no surface part."""
    r, itype, A, q, mu = _surf_system(seed=9)
    a = O.multipole_es(_surf_params(len(q), 0), itype, A, r, q, mu, [1, 1, 1])
    b = O.multipole_es(_surf_params(len(q), 1), itype, A, r, q, mu, [1, 1, 1])
    M = (q[:, None] * r).sum(axis=0) + mu.sum(axis=0)
    assert np.abs((b["ef"] - a["ef"]) - M[None, :]).max() < 1e-12 * np.abs(M).max()
    assert np.array_equal(b["efg"], a["efg"])
    assert np.array_equal(b["tau"], a["tau"])


# --------------------------------------------------------------------------
# pair potentials whose reference routines cannot run as written: the
# undamped BMHFT branch reads uninitialised damping factors
# (src/field.f90:2900, 2975) and the Morse routine is never dispatched and
# reads an uninitialised exponential in its force (:3124, 3132).  Both are
# pinned by their textbook closed forms on a dimer.

def _dimer(O, d, term, **kw):
    p = P.field_defaults(2, 1)
    p.cutshortrange = 9.0
    k = P.pidx(1, 1)
    if term == "morse":
        p.lmorse = 1
        p.epsmor[k], p.sigmamor[k], p.rhomor[k] = kw["eps"], kw["sig"], kw["rho"]
    else:
        setattr(p, term, 1)
        p.Abmhftd[k], p.Bbmhftd[k], p.Cbmhftd[k], p.Dbmhftd[k] = kw["A"], kw["B"], kw["C"], kw["D"]
        p.BDbmhftd[k] = kw.get("BD", 0.0)
    P.field_check_tag(p)
    r = np.array([[1.0, 1.0, 1.0], [1.0 + d, 1.0, 1.0]])
    return O.Oracle(p, np.array([1, 1], np.int32), np.eye(3) * 30.0).engforce(r)


def test_morse_dimer_closed_form(O):
    """V(r) = eps (exp(-2 rho (r - sigma)) - 2 exp(-rho (r - sigma))): depth
    eps at r = sigma with zero force there, a zero at sigma - ln 2 / rho, and
    the force -dV/dr on the second atom at every distance."""
    eps, sig, rho = 0.37, 2.1, 1.45
    kw = dict(eps=eps, sig=sig, rho=rho)
    out = _dimer(O, sig, "morse", **kw)
    assert abs(out["u_morse"] + eps) < 1e-14
    assert np.abs(out["f"]).max() < 1e-13
    assert abs(_dimer(O, sig - math.log(2.0) / rho, "morse", **kw)["u_morse"]) < 1e-14
    for d in (1.6, 2.5, 3.7, 6.0):
        x = math.exp(-rho * (d - sig))
        v = eps * (x * x - 2.0 * x)
        dv = eps * (-2.0 * rho * x * x + 2.0 * rho * x)
        out = _dimer(O, d, "morse", **kw)
        assert abs(out["u_morse"] - v) <= 1e-13 * max(1.0, abs(v))
        # atom 2 sits at +d along x: its force is -dV/dr
        assert abs(out["f"][1, 0] + dv) <= 1e-12 * max(1.0, abs(dv))
        assert abs(out["f"][0, 0] - dv) <= 1e-12 * max(1.0, abs(dv))


def test_bmhft_dimer_closed_form_and_damped_limit(O):
    """Undamped BMHFT: V(r) = A exp(-B r) - C / r^6 - D / r^8 and its
    derivative; the Tang-Toennies damped form tends to it when the damping
    constant grows (f_n(b r) -> 1)."""
    kw = dict(A=1500.0, B=3.3, C=21.0, D=4.5)
    for d in (1.4, 2.0, 3.1, 5.5):
        v = kw["A"] * math.exp(-kw["B"] * d) - kw["C"] / d ** 6 - kw["D"] / d ** 8
        dv = (-kw["B"] * kw["A"] * math.exp(-kw["B"] * d) + 6.0 * kw["C"] / d ** 7
              + 8.0 * kw["D"] / d ** 9)
        out = _dimer(O, d, "lbmhft", **kw)
        assert abs(out["u_bmhft"] - v) <= 1e-13 * max(1.0, abs(v)), (d, out["u_bmhft"], v)
        assert abs(out["f"][1, 0] + dv) <= 1e-12 * max(1.0, abs(dv)), (d, out["f"][1, 0], dv)
        damped = _dimer(O, d, "lbmhftd", BD=400.0, **kw)
        assert abs(damped["u_bmhft"] - v) <= 1e-13 * max(1.0, abs(v))
        assert abs(damped["f"][1, 0] + dv) <= 1e-12 * max(1.0, abs(dv))
        # a finite damping constant lowers the dispersion attraction
        assert _dimer(O, d, "lbmhftd", BD=1.0, **kw)["u_bmhft"] > v
