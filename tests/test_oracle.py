"""Pins the CPU oracle: the reference's own known answers (examples ex0,
ex2, ex4, ex9), closed forms, and invariants chosen so that a dropped term,
a wrong sign or index or a transposed operand fails one of them.  CPU only.
"""
import math

import numpy as np
import pytest

import golden_cases as G
from mdff_amd import configs as C
from mdff_amd import params as P


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# --------------------------------------------------------------------------
# helpers: erfc and Tang-Toennies

def test_errfc_is_numerical_recipes_erfcc(O):
    xs = np.linspace(0.0, 6.0, 301)
    for x in xs:
        e = O.errfc(x)
        ref = math.erfc(x)
        # NR erfcc: fractional error below 1.2e-7 everywhere
        assert abs(e - ref) <= 1.2e-7 * ref + 1e-300
        if x > 0.0:
            assert abs(O.errfc(-x) - (2.0 - e)) < 1e-15
    # the formula itself, not libm: differs from erfc at 1e-8 level somewhere
    assert max(abs(O.errfc(x) - math.erfc(x)) / math.erfc(x) for x in xs[1:]) > 1e-9


@pytest.mark.parametrize("order", range(1, 9))
def test_tang_toennies_limits_and_derivative(O, order):
    b, c = 2.3, 1.4
    f0, _ = O.tt_damp(b, c, 0.0, order)
    assert abs(f0 - (1.0 - c)) < 1e-15
    finf, fdinf = O.tt_damp(b, c, 40.0, order)
    assert abs(finf - 1.0) < 1e-12 and abs(fdinf) < 1e-12
    for r in (0.3, 1.1, 2.7):
        h = 1e-6
        fp, _ = O.tt_damp(b, c, r + h, order)
        fm, _ = O.tt_damp(b, c, r - h, order)
        _, fd = O.tt_damp(b, c, r, order)
        assert abs((fp - fm) / (2 * h) - fd) < 1e-8 * max(1.0, abs(fd))
        # closed form 1 - c e^{-br} sum_k (br)^k/k!
        br = b * r
        s = sum(br ** k / math.factorial(k) for k in range(order + 1))
        f, _ = O.tt_damp(b, c, r, order)
        assert abs(f - (1.0 - c * math.exp(-br) * s)) < 1e-14


def test_kmesh_count_and_half_space(O):
    import ctypes
    k = (ctypes.c_int32 * 3)(3, 2, 4)
    assert O.lib().orc_kmesh_count(k) == P.num_kvectors([3, 2, 4]) == 4 + 2 * 9 + 3 * 5 * 9


# --------------------------------------------------------------------------
# Lennard-Jones: ex0 (DL_POLY) and ex2 (Cambridge Cluster Database)

def test_ex0_lj_fcc_matches_dlpoly(O):
    p, itype, A, r, d = G.ex0()
    out = O.Oracle(p, itype, A).engforce(r)
    # the reference's own value for this input (Utot, E19.12)
    assert abs(out["u_lj"] - (-21.9404397839)) < 1e-9
    tail = G.lj_tail(p, [256], float(np.linalg.det(A)))
    assert abs(tail - float(d["dlpoly_lr_energy"])) < 1e-6
    # DL_POLY's total configurational energy includes the tail; the two codes
    # agree to 1.2e-6 relative (DL_POLY prints 7 digits)
    assert abs(out["u_lj"] + tail - float(d["dlpoly_eng_tot"])) < 3e-5
    # perfect lattice: no net force on any atom
    assert np.abs(out["f"]).max() < 1e-13


@pytest.mark.parametrize("name", ["13", "38", "55", "75", "98", "100", "147", "38i", "75i"])
def test_ex2_lj_clusters_match_ccd(O, name):
    p, itype, A, r, e_ccd = G.ex2(name)
    out = O.Oracle(p, itype, A).engforce(r)
    n = r.shape[0]
    assert abs(out["u_lj"] - e_ccd) < 2e-6 * max(1.0, n / 50)
    # the CCD structures are (near) minima: small residual forces
    assert np.abs(out["f"]).max() < 1e-2


# --------------------------------------------------------------------------
# Ewald: ex4 known answers (Nymand & Linse; GULP), Madelung constant

def test_ex4_two_charges_nymand_linse(O):
    p, itype, A, r, e = G.ex4_twocharges()
    out = O.Oracle(p, itype, A).engforce(r)
    assert abs(out["efg"][0, 0, 0] - e["V1xx"]) < 1e-7
    assert abs(out["efg"][0, 1, 1] - e["V1yy"]) < 1e-7
    # the reference's own digits for this input
    assert abs(out["efg"][0, 0, 0] - 2.000374925518976) < 1e-14
    assert abs(out["efg"][0, 1, 1] + 1.000187462857358) < 1e-14


def _nmr_principal(V):
    w = np.linalg.eigvalsh(V)
    iso = w.mean()
    w = sorted(w, key=lambda x: abs(x - iso))
    return np.array(w)  # xx, yy, zz with |zz| >= |yy| >= |xx|


def test_ex4_cluster_efg_matches_gulp(O):
    p, itype, A, r, gulp = G.ex4_cluster()
    out = O.Oracle(p, itype, A).engforce(r)
    for i in range(r.shape[0]):
        e = out["efg"][i]
        V = np.triu(e) + np.triu(e, 1).T  # the six components the reference reports
        pr = -_nmr_principal(V)  # lefg_vasp_sign
        pr = np.array(sorted(pr, key=abs))
        assert np.allclose(pr, gulp[i], atol=2e-3), (i, pr, gulp[i])


def test_madelung_nacl(O):
    """Rock salt, 8-ion cubic cell of side a: E = -N/2 M q^2 / (a/2)."""
    a = 2.0
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    r = np.vstack([base, base + [0.5, 0, 0]]) * a
    p = P.field_defaults(8, 2)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.cutlongrange = 1.0
    p.qch[0], p.qch[1] = 1.0, -1.0
    # rc = a/2 keeps one of the two images at exactly a/2: alpha large
    # enough that erfc(alpha a/2) is negligible
    p.alphaES = 5.0
    p.kES[:] = [26, 26, 26]
    P.field_check_tag(p)
    A = np.eye(3) * a
    P.ewald_param(p, A)
    out = O.Oracle(p, np.array([1] * 4 + [2] * 4, np.int32), A).engforce(r)
    M = 1.747564594633182
    assert abs(out["u_coul"] - (-8 / 2 * M / (a / 2))) < 1e-9
    assert np.abs(out["f"]).max() < 1e-10


# --------------------------------------------------------------------------
# ex9 test 11: BMHFTD + polarizable ions with Tang-Toennies damping

def test_ex9_pim_damped_matches_reference_run(O):
    p, itype, A, r, meta = G.ex9_test11()
    assert abs(p.alphaES - meta["reference_alpha"]) < 5e-6
    assert list(p.kES) == [meta["reference_kmax"]] * 3
    out = O.Oracle(p, itype, A).engforce(r, itime=0)
    assert out["n_scf"] == meta["reference_nscf"]
    # Utot = u_bmhft + u_coul, printed by the reference with 12 digits
    assert abs(out["u_bmhft"] + out["u_coul"] - meta["reference_utot"]) < 1e-7


# --------------------------------------------------------------------------
# invariants: forces = -grad U, field = -dU/dmu, virials, alpha independence

def _fd_forces(energy, r, idx, h=1e-5):
    out = []
    for i, a in idx:
        rp = r.copy()
        rp[i, a] += h
        rm = r.copy()
        rm[i, a] -= h
        out.append(-(energy(rp) - energy(rm)) / (2 * h))
    return np.array(out)


def _small_system(ntype=2, n=24, L=9.0, seed=5):
    rng = np.random.default_rng(seed)
    r = C.rsa_positions(n, L, 1.7, rng)
    itype = np.array([1 + (i % ntype) for i in range(n)], np.int32)
    return r, itype, np.eye(3) * L


@pytest.mark.parametrize("term", ["lj", "bmhftd", "bmhft", "morse"])
def test_short_range_forces_and_virial_by_fd(O, term):
    r, itype, A = _small_system()
    n = r.shape[0]
    p = P.field_defaults(n, 2)
    p.cutshortrange = 4.4
    for (i, j) in ((1, 1), (1, 2), (2, 2)):
        k = P.pidx(i, j)
        p.epslj[k], p.sigmalj[k] = 0.02 * (i + j), 1.6 + 0.1 * j
        p.Abmhftd[k], p.Bbmhftd[k], p.Cbmhftd[k], p.Dbmhftd[k], p.BDbmhftd[k] = \
            900.0 * i, 3.1 + 0.2 * j, 12.0, 3.0 * i, 2.4
        p.epsmor[k], p.sigmamor[k], p.rhomor[k] = 0.1 * i, 2.0, 1.3 + 0.1 * j
    setattr(p, {"lj": "lnmlj", "bmhftd": "lbmhftd", "bmhft": "lbmhft", "morse": "lmorse"}[term], 1)
    P.field_check_tag(p)
    key = {"lj": "u_lj", "bmhftd": "u_bmhft", "bmhft": "u_bmhft", "morse": "u_morse"}[term]
    vkey = "vir" + key[1:]
    orc = O.Oracle(p, itype, A)
    out = orc.engforce(r)
    if term == "lj":
        # truncation changes the energy smoothly only inside; keep notrunc
        pass
    fd = _fd_forces(lambda x: orc.engforce(x)[key], r, [(0, 0), (3, 1), (7, 2), (11, 0)])
    an = np.array([out["f"][0, 0], out["f"][3, 1], out["f"][7, 2], out["f"][11, 0]])
    assert np.allclose(fd, an, rtol=1e-6, atol=1e-8 * np.abs(out["f"]).max())
    # virial = -(1/3) dU/ds under uniform scaling r -> (1+s) r, A -> (1+s) A
    h = 1e-6
    up = O.Oracle(p, itype, A * (1 + h)).engforce(r * (1 + h))[key]
    um = O.Oracle(p, itype, A * (1 - h)).engforce(r * (1 - h))[key]
    assert abs(-(up - um) / (2 * h) / 3.0 - out[vkey]) < 1e-6 * max(1.0, abs(out[vkey]))
    # stress trace = 3 vir / Omega / press_unit; symmetric
    tau = out["tau_nonb"]
    assert abs(np.trace(tau) - 3 * out[vkey] / np.linalg.det(A) / p.press_unit) < 1e-9 * max(1, abs(np.trace(tau)))
    assert np.allclose(tau, tau.T)
    # Newton's third law
    assert np.abs(out["f"].sum(axis=0)).max() < 1e-10 * max(1.0, np.abs(out["f"]).max())


def _ewald_params(n, L, alpha, kmax, rc, dip=True):
    p = P.field_defaults(n, 2)
    p.lcoulomb = 1
    p.cutlongrange = rc
    p.qch[0], p.qch[1] = 1.3, -0.65
    p.alphaES = alpha
    p.kES[:] = [kmax] * 3
    p.doefield = 1
    p.doefg = 1
    P.field_check_tag(p)
    p.task_coul[:] = [1, int(dip), int(dip)]
    return p


def _ewald_system(seed=3):
    rng = np.random.default_rng(seed)
    n = 18
    L = 8.0
    r = C.rsa_positions(n, L, 1.8, rng)
    itype = np.array([1] * 6 + [2] * 12, np.int32)
    q = np.where(itype == 1, 1.3, -0.65)
    mu = 0.2 * rng.standard_normal((n, 3))
    return r, itype, np.eye(3) * L, q, mu


def test_ewald_alpha_independence_with_dipoles(O):
    r, itype, A, q, mu = _ewald_system()
    n = r.shape[0]
    outs = []
    for alpha in (1.3, 1.5):
        p = _ewald_params(n, 8.0, alpha, 24, 3.99)
        outs.append(O.multipole_es(p, itype, A, r, q, mu, [1, 1, 1]))
    a, b = outs
    assert abs(a["u_coul"] - b["u_coul"]) < 1e-7 * abs(a["u_coul"])
    assert np.abs(a["f"] - b["f"]).max() < 1e-6 * np.abs(a["f"]).max()
    assert np.abs(a["ef"] - b["ef"]).max() < 1e-6 * np.abs(a["ef"]).max()
    for i in range(n):
        ea = np.triu(a["efg"][i])
        eb = np.triu(b["efg"][i])
        assert np.abs(ea - eb).max() < 1e-5 * np.abs(ea).max()


def test_ewald_forces_field_by_fd(O):
    """F = -dU/dr at fixed moments; E = -dU/dmu (field without coul_unit,
    energy with it); q-q, q-d and d-d terms all present."""
    r, itype, A, q, mu = _ewald_system(seed=4)
    n = r.shape[0]
    p = _ewald_params(n, 8.0, 0.7, 9, 3.99)
    cu = p.coul_unit

    def U(rr=r, mm=mu):
        return O.multipole_es(p, itype, A, rr, q, mm, [1, 1, 1], do_efg=False)["u_coul"]

    out = O.multipole_es(p, itype, A, r, q, mu, [1, 1, 1])
    idx = [(0, 0), (5, 2), (9, 1), (17, 0)]
    fd = _fd_forces(lambda x: U(rr=x), r, idx, h=1e-5)
    an = np.array([out["f"][i, a] for i, a in idx])
    assert np.allclose(fd, an, rtol=2e-6, atol=1e-7 * np.abs(out["f"]).max())
    for i, a in [(2, 0), (11, 1), (16, 2)]:
        h = 1e-5
        mp = mu.copy()
        mp[i, a] += h
        mm = mu.copy()
        mm[i, a] -= h
        fdE = -(U(mm=mp) - U(mm=mm)) / (2 * h) / cu
        assert abs(fdE - out["ef"][i, a]) < 2e-6 * max(1.0, abs(out["ef"][i, a]))


def test_ewald_efg_traceless_and_symmetric_for_charges(O):
    r, itype, A, q, mu = _ewald_system(seed=6)
    n = r.shape[0]
    p = _ewald_params(n, 8.0, 1.4, 22, 3.99, dip=False)
    out = O.multipole_es(p, itype, A, r, q, np.zeros((n, 3)), [1, 0, 0])
    for i in range(n):
        e = out["efg"][i]
        V = np.triu(e) + np.triu(e, 1).T
        assert abs(np.trace(V)) < 1e-6 * np.abs(V).max()


def test_ewald_virial_theorem_for_charges(O):
    """Coulomb energy is homogeneous of degree -1: the virial trace times the
    volume equals U (converged Ewald sums)."""
    r, itype, A, q, mu = _ewald_system(seed=7)
    n = r.shape[0]
    p = _ewald_params(n, 8.0, 1.4, 22, 3.99, dip=False)
    out = O.multipole_es(p, itype, A, r, q, np.zeros((n, 3)), [1, 0, 0])
    V = np.linalg.det(A)
    # uniform scaling derivative of U
    h = 1e-6
    up = O.multipole_es(p, itype, A * (1 + h), r * (1 + h), q, np.zeros((n, 3)), [1, 0, 0])["u_coul"]
    um = O.multipole_es(p, itype, A * (1 - h), r * (1 - h), q, np.zeros((n, 3)), [1, 0, 0])["u_coul"]
    dUds = (up - um) / (2 * h)
    assert abs(dUds + out["u_coul"]) < 1e-6 * abs(out["u_coul"])
    # the reference's stress: tr(tau_coul) Omega press_unit = -dU/ds
    assert abs(np.trace(out["tau"]) * V * p.press_unit + dUds) < 1e-6 * abs(dUds)


def test_pim_scf_fixed_point_and_variational_forces(O):
    s = C.silica(20, 40, seed=21, rc=5.0, kmax=6, pim=True, efg=False)
    p = s.params
    p.doefield = 1
    p.conv_tol_ind = 1e-11
    p.min_scf_pol_iter = 3
    p.max_scf_pol_iter = 400
    p.calc_mode = P.CALC_OTHER
    orc = O.Oracle(p, s.itype, s.A)
    out = orc.engforce(s.r)
    alpha_o = p.poldip[1 + 16 * 0]
    o_atoms = s.itype == 2
    # mu = alpha E at convergence on polarizable atoms, zero elsewhere
    assert np.abs(out["mu"][o_atoms] - alpha_o * out["ef"][o_atoms]).max() < 1e-8
    assert np.abs(out["mu"][~o_atoms]).max() == 0.0
    # u_pol = 1/2 sum mu^2 / alpha
    assert abs(out["u_pol"] - 0.5 * (out["mu"][o_atoms] ** 2).sum() / alpha_o) < 1e-9 * out["u_pol"]
    # total energy stationary in mu: forces are -grad of u_coul + u_bmhft

    def Utot(x):
        o = O.Oracle(p, s.itype, s.A).engforce(x)
        return o["u_coul"] + o["u_bmhft"]

    idx = [(0, 0), (25, 1), (41, 2)]
    fd = _fd_forces(Utot, s.r, idx, h=1e-5)
    an = np.array([out["f"][i, a] for i, a in idx])
    assert np.allclose(fd, an, rtol=1e-5, atol=1e-6 * np.abs(out["f"]).max())


def test_translation_and_permutation_invariance(O):
    s = C.silica(10, 20, seed=22, rc=4.5, kmax=5, pim=True, efg=True)
    p = s.params
    p.calc_mode = P.CALC_OTHER
    a = O.Oracle(p, s.itype, s.A).engforce(s.r)
    b = O.Oracle(p, s.itype, s.A).engforce(s.r + np.array([3.1, -7.3, 12.4]))
    assert abs(a["u_coul"] - b["u_coul"]) < 1e-9 * abs(a["u_coul"])
    assert np.abs(a["f"] - b["f"]).max() < 1e-8 * np.abs(a["f"]).max()
    # swap two O atoms
    perm = np.arange(s.natm)
    perm[[12, 25]] = perm[[25, 12]]
    c = O.Oracle(p, s.itype[perm], s.A).engforce(s.r[perm])
    assert abs(a["u_coul"] - c["u_coul"]) < 1e-9 * abs(a["u_coul"])
    assert np.abs(a["f"][perm] - c["f"]).max() < 1e-8 * np.abs(a["f"]).max()
    assert np.abs(a["efg"][perm] - c["efg"]).max() < 1e-8 * np.abs(a["efg"]).max()
