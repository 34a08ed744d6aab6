"""The oracle's standalone EFG (efgcalc: multipole_efg_ES / multipole_efg_DS,
efg_stat, nmr_convention) pinned by the reference's ex4 known answers and by
what the mathematics fixes:

* ex4 two charges in a box of 10 (reduced units): V1,xx and V1,yy of the
  reference's own Ewald run to 16 digits (2.000374925518976,
  -1.000187462857358) and the Nymand-Linse values (run_ex4.sh:78-79) for
  both the Ewald and the direct sum;
* ex4 9-site cluster: principal components and asymmetry of every site
  against GULP (run_ex4.sh:161-169), Ewald and direct sums, through the
  EFGALL file and efg_stat;
* independence of the Ewald EFG from alpha (charges and point dipoles),
  agreement of the Ewald and direct sums for a neutral cell with dipoles,
  agreement with the MD-path EFG (multipole_ES) when the real-space cutoff
  holds every pair, tracelessness, the it_efg restriction equal to zeroing
  the other species' moments, the charged-cell correction.
"""
import copy
import os

import numpy as np
import pytest

import golden_cases as G
from mdff_amd import efg as E
from mdff_amd import io
from mdff_amd import params as P

FILES = os.path.join(os.path.dirname(__file__), "golden", "files")


@pytest.fixture(scope="module")
def O():
    from oracle import oracle
    return oracle


def _two_charges(longrange):
    fr = next(io.read_trajff(os.path.join(FILES, "ex4_TRAJFF.twocharges")))
    p = P.field_defaults(2, 2)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.qch[0], p.qch[1] = 1.0, -1.0
    p.cutlongrange = 1000.0
    if longrange == "ewald":
        p.kES[:] = [11, 11, 11]
        p.alphaES = 0.8
        P.field_check_tag(p)
        P.ewald_param(p, fr.A)
    else:
        p.kES[:] = [0, 0, 0]
        P.field_check_tag(p)
    return p, fr, P.efg_opts(longrange, ncelldirect=40, cutlongrange=1000.0)


def test_two_charges_ewald_reference_digits(O):
    p, fr, o = _two_charges("ewald")
    e = O.efgcalc(p, o, fr.itype(), fr.A, fr.r)["efg"]
    # the reference's own Ewald run of this input
    assert abs(e[0, 0, 0] - 2.000374925518976) <= 2e-15
    assert abs(e[0, 1, 1] - -1.000187462857358) <= 2e-15
    g = G.ex4()["twocharges_expected"]
    # the published values are truncated to seven decimals
    assert abs(e[0, 0, 0] - g["V1xx"]) < 1e-7 and abs(e[0, 1, 1] - g["V1yy"]) < 1e-7
    # traceless up to the erfc approximation's own error (1.2e-7 relative,
    # src/functions.f90:28-40)
    assert abs(np.trace(e[0])) < 1e-8 and np.array_equal(e[0], e[0].T)


def test_two_charges_direct_sum(O):
    p, fr, o = _two_charges("direct")
    e = O.efgcalc(p, o, fr.itype(), fr.A, fr.r)["efg"]
    g = G.ex4()["twocharges_expected"]
    assert abs(e[0, 0, 0] - g["V1xx"]) <= 5e-8 and abs(e[0, 1, 1] - g["V1yy"]) <= 5e-8


def _cluster(longrange):
    g = G.ex4()
    fr = next(io.read_trajff(os.path.join(FILES, "ex4_TRAJFF.cluster")))
    p = P.field_defaults(fr.natm, 2)
    p.lcoulomb = 1
    p.qch[0], p.qch[1] = g["cluster_charges"]
    if longrange == "ewald":
        p.cutlongrange = g["cluster_ewald"]["cutlongrange"]
        P.field_check_tag(p)
        P.ewald_param(p, fr.A, epsw=g["cluster_ewald"]["epsw"], lautoES=True)
    else:
        p.cutlongrange = 100.0
        p.kES[:] = [0, 0, 0]
        P.field_check_tag(p)
    o = P.efg_opts(longrange, ncelldirect=40, cutlongrange=100.0, lefg_vasp_sign=True)
    return p, fr, o, np.array(g["gulp_principal"]), np.array(g["gulp_eta"])


def oracle_nmr(efg6):
    """nmr_fn of efg_stat built on the oracle: LAPACK eigenvalues (numpy's
    eigvalsh, the dsyevd family of the reference's DSYEV) and the oracle's
    nmr_convention."""
    from oracle import oracle as O
    n = efg6.shape[0]
    nmr = np.zeros((n, 4))
    U = np.zeros((n, 5))
    S = np.zeros(n)
    bad = 0
    sq3 = 1.0 / np.sqrt(3.0)
    for i, (xx, yy, zz, xy, xz, yz) in enumerate(efg6):
        w = np.linalg.eigvalsh(np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]))
        v, ok = O.nmr_convention(w)
        nmr[i] = v
        bad += 0 if ok else 1
        U[i] = [zz * 0.5, xz * sq3, yz * sq3, xy * sq3, (xx - yy) * sq3 * 0.5]
        S[i] = np.sqrt(np.sum(U[i] ** 2))
    return nmr, U, S, bad


@pytest.mark.parametrize("longrange", ["ewald", "direct"])
def test_cluster_matches_gulp_through_efgall_and_efg_stat(O, tmp_path, longrange):
    p, fr, o, gulp, gulp_eta = _cluster(longrange)
    out = O.efgcalc(p, o, fr.itype(), fr.A, fr.r)
    path = str(tmp_path / "EFGALL")
    with open(path, "w") as fh:
        io.write_efgall(fh, fr, out["efg"], True, efgcalc=True)
    opt = E.EfgStatOptions(umin=-15.0, smax=25.0, vzzmin=-50.0)
    st = E.efg_stat(path, 1, fr.itype(), opt, nmrff=str(tmp_path / "NMRFF"), nmr_fn=oracle_nmr)
    nmr = st.nmr_last
    assert np.allclose(nmr[:, :3], gulp, atol=1e-3), np.abs(nmr[:, :3] - gulp).max()
    assert np.allclose(nmr[:, 3], gulp_eta, atol=2e-4), np.abs(nmr[:, 3] - gulp_eta).max()
    lines = open(tmp_path / "NMRFF").read().splitlines()
    assert len(lines) == 10 and lines[0].startswith("#     site")
    assert lines[1].split()[:2] == ["1", "A"]
    # every site is counted once in each distribution (it=0: all sites)
    for k, cnt in ((0, 9), (1, 8), (2, 1)):
        assert st.dibvzz[k].sum() == cnt and st.dibeta[k].sum() == cnt and st.dibS[k].sum() == cnt
        assert st.dibU[0, k].sum() == cnt and st.dibU[5, k].sum() == 4 * cnt
    E.efg_write_output(st, opt, str(tmp_path))
    eta = open(tmp_path / "DTETAFF").read().splitlines()
    assert len(eta) == 2 + opt.bins()[0] + 2
    # the eta density integrates to one per species
    dens = np.array([[float(x) for x in l.split()] for l in eta[2:2 + opt.bins()[0]]])
    assert np.allclose(dens[:, 1:].sum(axis=0) * opt.reseta, 1.0)


def _pim_cell(seed, n_si=12, n_o=24, kmax=8, dip=False):
    from mdff_amd import configs as C
    s = C.silica(n_si, n_o, seed=seed, rc=4.0, kmax=kmax, pim=True, efg=True)
    p = s.params
    p.calc_mode = P.CALC_OTHER
    if dip:
        p.dip[0], p.dip[0 + 16], p.dip[0 + 32] = 0.05, -0.02, 0.03
        P.field_check_tag(p)
    return s


@pytest.mark.parametrize("dip", [False, True])
def test_ewald_efg_independent_of_alpha(O, dip):
    """The total (real + reciprocal + self) is alpha-independent at
    converged sums: charges, induced and permanent dipoles."""
    s = _pim_cell(11, dip=dip)
    L = s.A[0, 0]
    o = P.efg_opts("ewald")
    res = []
    # erfc(alpha L / 2) and exp(-(pi kmax / (alpha L))^2) below 1e-11
    for aL, kmax in ((10.0, 18), (12.0, 22)):
        p = copy.deepcopy(s.params)
        p.alphaES = aL / L
        p.kES[:] = [kmax] * 3
        p.conv_tol_ind = 1e-12
        p.max_scf_pol_iter = 200
        res.append(O.efgcalc(p, o, s.itype, s.A, s.r))
    e0, e1 = res[0]["efg"], res[1]["efg"]
    assert np.abs(e0 - e1).max() <= 1e-7 * np.abs(e0).max()
    assert np.abs(res[0]["mu"] - res[1]["mu"]).max() <= 1e-7 * np.abs(res[0]["mu"]).max()
    tr = np.trace(e0, axis1=1, axis2=2)
    assert np.abs(tr).max() <= 1e-7 * np.abs(e0).max()


def test_ewald_and_direct_sums_agree_for_point_dipoles(O):
    """Neutral cell of charges and permanent point dipoles (no induction):
    the direct sum over a large cube of images converges to the Ewald
    value.  Pins the sign and index pattern of the rank-3 terms in both."""
    rng = np.random.default_rng(4)
    n = 6
    A = np.diag([7.0, 8.0, 7.5])
    r = (rng.uniform(-0.5, 0.5, (n, 3))) @ A.T
    p = P.field_defaults(n, 2)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.qch[0], p.qch[1] = 0.6, -0.6
    p.dip[0], p.dip[16], p.dip[32] = 0.3, -0.2, 0.1
    p.dip[1], p.dip[17], p.dip[33] = -0.1, 0.25, 0.2
    p.kES[:] = [16, 16, 16]
    p.alphaES = 0.9
    p.cutlongrange = 3.0
    P.field_check_tag(p)
    itype = np.array([1, 1, 1, 2, 2, 2], np.int32)
    # net dipole of the cell is nonzero: the direct cube sum converges to
    # the Ewald (tin-foil) value only with the surface term, which the EFG
    # does not see (second derivative), so they must agree
    es = O.efgcalc(p, P.efg_opts("ewald"), itype, A, r)["efg"]
    ds = O.efgcalc(p, P.efg_opts("direct", ncelldirect=30, cutlongrange=1e6), itype, A, r)["efg"]
    assert np.abs(es - ds).max() <= 2e-5 * np.abs(es).max(), np.abs(es - ds).max()


def test_matches_md_path_efg_when_the_cutoff_holds_every_pair(O):
    """multipole_efg_ES sums every minimum-image pair; multipole_ES only
    pairs within cutlongrange.  With alpha so large that erfc(alpha rc) is
    below 1e-16, the two agree (upper six components: the MD path's
    reciprocal EFG fills those only)."""
    s = _pim_cell(12, kmax=22)
    p = copy.deepcopy(s.params)
    L = s.A[0, 0]
    p.cutlongrange = L / 2
    p.alphaES = 6.3 / p.cutlongrange
    p.kES[:] = [22, 22, 22]
    p.conv_tol_ind = 1e-13
    p.max_scf_pol_iter = 300
    md = O.Oracle(p, s.itype, s.A).engforce(s.r)
    st = O.efgcalc(p, P.efg_opts("ewald"), s.itype, s.A, s.r)
    iu = np.triu_indices(3)
    a = md["efg"][:, iu[0], iu[1]]
    b = st["efg"][:, iu[0], iu[1]]
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), np.abs(a - b).max()


def test_it_efg_restriction_and_charged_correction(O):
    s = _pim_cell(13, dip=True)
    p = copy.deepcopy(s.params)
    p.conv_tol_ind = 1e-12
    base = O.efgcalc(p, P.efg_opts("ewald"), s.itype, s.A, s.r)
    only1 = O.efgcalc(p, P.efg_opts("ewald", lefg_it_contrib=True, it_efg=1), s.itype, s.A, s.r)
    only2 = O.efgcalc(p, P.efg_opts("ewald", lefg_it_contrib=True, it_efg=2), s.itype, s.A, s.r)
    # the sources are split by species (the self term of a site goes with
    # its own species): the two restricted sums add up to the full one
    assert np.abs(only1["efg"] + only2["efg"] - base["efg"]).max() <= 1e-10 * np.abs(base["efg"]).max()
    assert np.allclose(only1["mu"], base["mu"], rtol=0, atol=0)
    # species 1 alone: a site of species 2 sees no self term
    assert not np.allclose(only1["efg"], 0.0) and not np.allclose(only2["efg"], 0.0)
    # charged-cell correction: + 4 pi Q / (3 V) on the diagonal
    corr = O.efgcalc(p, P.efg_opts("ewald", lmp_correction=True), s.itype, s.A, s.r)["efg"]
    Q = sum(p.qch[t - 1] for t in s.itype)
    V = abs(np.linalg.det(s.A))
    d = 4.0 * np.pi * Q / (3.0 * V) * p.coul_unit
    assert np.allclose(corr - base["efg"], d * np.eye(3)[None], rtol=0, atol=1e-12 * max(1, abs(d)))
    sign = O.efgcalc(p, P.efg_opts("ewald", lefg_vasp_sign=True), s.itype, s.A, s.r)["efg"]
    assert np.array_equal(sign, -base["efg"])


def test_nmr_convention_closed_forms(O):
    # axial tensor: eta 0, Vzz the distinct eigenvalue
    v, ok = O.nmr_convention([-1.0, -1.0, 2.0])
    assert ok and v == (-1.0, -1.0, 2.0, 0.0)
    v, ok = O.nmr_convention([-2.0, 1.0, 1.0])
    assert ok and v[2] == -2.0 and v[3] == 0.0
    # eta = (Vxx - Vyy) / Vzz with |Vzz| >= |Vyy| >= |Vxx|
    v, ok = O.nmr_convention([-3.0, 1.0, 2.0])
    assert ok and v[:3] == (1.0, 2.0, -3.0) and abs(v[3] - (1.0 - 2.0) / -3.0) < 1e-16
    # zero tensor
    v, ok = O.nmr_convention([0.0, 0.0, 0.0])
    assert ok and v[3] == 0.0
