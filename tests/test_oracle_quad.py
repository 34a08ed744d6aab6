"""Pins of the oracle's quadrupole terms (multipole_ES_dir :2542-2601,
multipole_ES_rec :2723-2740, induced quadrupoles :1716-1803) and of its
Wannier-centre dipoles (moment_from_WFc :3492-3672), by what the
mathematics fixes, on the CPU:

* the pair interaction energies q-Q, mu-Q and Q-Q, E(A+B) - E(A) - E(B), do
  not depend on the Ewald alpha (the real-space tensors and the reciprocal
  structure-factor term must agree; to 1e-6, the accuracy of the
  reference's erfc approximation), and match the bare multipole energies
  (q/3) T2:Q, (mu/3).T3:Q and Q:T4:Q/9 of a pair in a large cell;
* the reference's real-space quadrupole field is three times the one its
  energy implies: E_dir / 3 + E_rec = -dU/dmu at a probe site (finite
  differences of the energy);
* its quadrupole EFG is the gradient of that real-space field, with the
  sign of the pair's order: EFG_dir = -s dE_dir/dr, s = +1 at the smaller
  index;
* at SCF convergence the induced quadrupoles are polquad . E;
* Wannier dipoles -2 sum (r_c - r_ion), the centre test on r^2 < rcut_wfc,
  and the error of a missing centre.
"""
import numpy as np
import pytest

from mdff_amd import params as P

L_BIG = 40.0


def _params(n, ntype=1, alpha=0.35, kmax=18, L=L_BIG, rc=None):
    p = P.field_defaults(n, ntype)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.alphaES = alpha
    p.kES[:] = [kmax, kmax, kmax]
    p.cutlongrange = 0.5 * L if rc is None else rc
    return p


def _sym(a):
    a = np.asarray(a, dtype=np.float64)
    return 0.5 * (a + a.T)


Q_A = _sym([[0.7, 0.2, -0.1], [0.0, -0.3, 0.25], [0.0, 0.0, -0.4]])
Q_B = _sym([[-0.2, 0.15, 0.3], [0.0, 0.5, -0.1], [0.0, 0.0, -0.3]])


def _es(O, p, r, q, mu, th, task=(1, 1, 1, 1, 1, 1), **kw):
    n = len(q)
    return O.multipole_es(p, np.ones(n, np.int32), np.eye(3) * kw.pop("L", L_BIG), r, q, mu,
                          task, damp_ind=False, theta=th, **kw)


def _pair_energy(O, p, r, site_a, site_b, L=L_BIG):
    """E(A+B) - E(A) - E(B) with (q, mu, theta) of each site."""
    def moms(sa, sb):
        q = np.array([sa[0], sb[0]], dtype=np.float64)
        mu = np.array([sa[1], sb[1]], dtype=np.float64)
        th = np.array([sa[2], sb[2]], dtype=np.float64)
        return q, mu, th
    zero = (0.0, np.zeros(3), np.zeros((3, 3)))
    e = []
    for sa, sb in ((site_a, site_b), (site_a, zero), (zero, site_b)):
        q, mu, th = moms(sa, sb)
        e.append(_es(O, p, r, q, mu, th, L=L, do_forces=False, do_stress=False)["u_coul"])
    return e[0] - e[1] - e[2]


def _bare_tensors(r):
    """Cartesian derivatives of 1/|r|: T2, T3, T4."""
    d2 = r @ r
    d = np.sqrt(d2)
    I = np.eye(3)
    T2 = (3.0 * np.einsum("a,b->ab", r, r) - d2 * I) / d ** 5
    T3 = -15.0 * np.einsum("a,b,c->abc", r, r, r) / d ** 7 + 3.0 * (
        np.einsum("ab,c->abc", I, r) + np.einsum("ac,b->abc", I, r) +
        np.einsum("bc,a->abc", I, r)) / d ** 5
    T4 = 105.0 * np.einsum("a,b,c,e->abce", r, r, r, r) / d ** 9
    rr = np.einsum("a,b->ab", r, r)
    T4 -= 15.0 * (np.einsum("ce,ab->abce", I, rr) + np.einsum("be,ac->abce", I, rr) +
                  np.einsum("bc,ae->abce", I, rr) + np.einsum("ae,bc->abce", I, rr) +
                  np.einsum("ac,be->abce", I, rr) + np.einsum("ab,ce->abce", I, rr)) / d ** 7
    T4 += 3.0 * (np.einsum("ab,ce->abce", I, I) + np.einsum("ac,be->abce", I, I) +
                 np.einsum("ae,bc->abce", I, I)) / d ** 5
    return T2, T3, T4


R_PAIR = np.array([[20.0, 20.0, 20.0], [21.3, 19.1, 21.7]])
SITES = {
    "q-Q": ((1.3, np.zeros(3), np.zeros((3, 3))), (0.0, np.zeros(3), Q_B)),
    "mu-Q": ((0.0, np.array([0.4, -0.7, 0.3]), np.zeros((3, 3))), (0.0, np.zeros(3), Q_B)),
    "Q-Q": ((0.0, np.zeros(3), Q_A), (0.0, np.zeros(3), Q_B)),
}


@pytest.mark.parametrize("kind", sorted(SITES))
def test_pair_energy_alpha_independent(oracle_lib, kind):
    a, b = SITES[kind]
    # k meshes long enough for the k^4 growth of the Q-Q reciprocal terms
    e1 = _pair_energy(oracle_lib, _params(2, alpha=0.30, kmax=22), R_PAIR, a, b)
    e2 = _pair_energy(oracle_lib, _params(2, alpha=0.42, kmax=30), R_PAIR, a, b)
    # to the accuracy of the reference's erfc (Numerical Recipes erfcc,
    # fractional error below 1.2e-7)
    assert abs(e1 - e2) <= 1e-6 * abs(e1), (e1, e2)


@pytest.mark.parametrize("kind", sorted(SITES))
def test_pair_energy_matches_bare_multipoles(oracle_lib, kind):
    a, b = SITES[kind]
    Lh = 80.0
    r = R_PAIR + 20.0
    e = _pair_energy(oracle_lib, _params(2, alpha=0.2, kmax=24, L=Lh), r, a, b, L=Lh)
    T2, T3, T4 = _bare_tensors(r[0] - r[1])
    if kind == "q-Q":
        ref = a[0] * np.sum(T2 * b[2]) / 3.0
    elif kind == "mu-Q":
        ref = a[1] @ np.einsum("abc,bc->a", T3, b[2]) / 3.0
    else:
        ref = np.einsum("ab,abce,ce->", a[2], T4, b[2]) / 9.0
    # the periodic images add O((d / L)^3) of the pair energy
    assert e == pytest.approx(ref, rel=2e-3), (e, ref)


def _probe_fields(O, p, r, th_b, probe_first=True):
    """Field and EFG at a probe (no moments) next to a quadrupole, with the
    full sum and with the reciprocal part alone (a real-space cutoff below
    the pair distance)."""
    q = np.zeros(2)
    mu = np.zeros((2, 3))
    th = np.zeros((2, 3, 3))
    ip, ib = (0, 1) if probe_first else (1, 0)
    th[ib] = th_b
    full = _es(O, p, r, q, mu, th, do_forces=False, do_stress=False)
    prec = _params(2, alpha=p.alphaES, kmax=p.kES[0], rc=0.01)
    rec = _es(O, prec, r, q, mu, th, do_forces=False, do_stress=False)
    return full["ef"][ip], rec["ef"][ip], full["efg"][ip], rec["efg"][ip]


def test_quadrupole_field_is_three_times_the_energy_gradient_in_real_space(oracle_lib):
    p = _params(2)
    r = R_PAIR
    ef, ef_rec, _, _ = _probe_fields(oracle_lib, p, r, Q_B)
    ef_dir = ef - ef_rec
    h = 1e-4
    grad = np.zeros(3)
    for a in range(3):
        e = []
        for sgn in (1.0, -1.0):
            mu = np.zeros((2, 3))
            mu[0, a] = sgn * h
            th = np.zeros((2, 3, 3))
            th[1] = Q_B
            e.append(_es(oracle_lib, p, r, np.zeros(2), mu, th, do_forces=False,
                         do_stress=False)["u_coul"])
        grad[a] = (e[0] - e[1]) / (2 * h)
    np.testing.assert_allclose(ef_dir / 3.0 + ef_rec, -grad, rtol=1e-7, atol=1e-12)
    assert np.abs(ef_dir).max() > 0.1 * np.abs(ef).max()  # the real-space part matters


@pytest.mark.parametrize("probe_first", [True, False])
def test_quadrupole_efg_is_signed_gradient_of_real_space_field(oracle_lib, probe_first):
    p = _params(2)
    ip = 0 if probe_first else 1
    r0 = R_PAIR.copy() if probe_first else R_PAIR[::-1].copy()
    _, _, efg, efg_rec = _probe_fields(oracle_lib, p, r0, Q_B, probe_first)
    efg_dir = efg - efg_rec
    h = 1e-4
    grad = np.zeros((3, 3))
    for b in range(3):
        f = []
        for sgn in (1.0, -1.0):
            r = r0.copy()
            r[ip, b] += sgn * h
            ef, ef_rec, _, _ = _probe_fields(oracle_lib, p, r, Q_B, probe_first)
            f.append(ef - ef_rec)
        grad[:, b] = (f[0] - f[1]) / (2 * h)
    s = 1.0 if probe_first else -1.0
    np.testing.assert_allclose(efg_dir, -s * grad, rtol=1e-6, atol=1e-9)


def _quad_melt(O, n_si=8, n_o=16, seed=3):
    """A small charged melt with static quadrupoles on Si and quadrupole
    polarizability on O (FP64 engforce)."""
    from mdff_amd import configs as C
    s = C.silica(n_si, n_o, seed=seed, rc=4.0, kmax=6, pim=True, efg=True, scf_tol=1e-10)
    p = s.params
    for a in range(1, 4):
        for b in range(1, 4):
            p.quad[P.qidx(1, a, b)] = 0.05 * Q_A[a - 1, b - 1]
    p.lquad_polar[1] = 1
    P.field_check_tag(p, polquad_iso=[0.0, 0.3])
    return s


def test_scf_induced_quadrupoles_are_polquad_times_field(oracle_lib):
    s = _quad_melt(oracle_lib)
    p = s.params
    assert list(p.task_coul_quad) == [1, 1, 1]
    orc = oracle_lib.Oracle(p, s.itype, s.A)
    o = orc.engforce(s.r, itime=0)
    assert o["rmsd"] <= 1e-10
    th_ind = o["theta"] - np.array([[[p.quad[P.qidx(t, a, b)] for b in (1, 2, 3)]
                                     for a in (1, 2, 3)] for t in s.itype])
    ox = s.itype == 2
    # diagonal polquad(g,g,g) = 0.3: theta_aa = 0.3 E_a, off-diagonal zero
    diag = np.einsum("iaa->ia", th_ind[ox])
    # the last iterate against the final field: within the SCF tolerance
    np.testing.assert_allclose(diag / 0.3, o["ef"][ox], rtol=0, atol=1e-9)
    off = th_ind[ox] - np.einsum("ia,ab->iab", diag, np.eye(3))
    assert np.abs(off).max() == 0.0
    assert np.abs(th_ind[~ox]).max() == 0.0


def test_wannier_dipoles_closed_form_and_r2_test(oracle_lib):
    # ions of species 1 with two centres (species 3) each; species 2 plain
    p = P.field_defaults(8, 3)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.qch[0], p.qch[1], p.qch[2] = 2.0, 1.0, -1.5
    p.lwfc[0], p.lwfc[1], p.lwfc[2] = 2, 0, -1
    p.rcut_wfc = 0.5  # compared with r^2: centres within 0.707
    p.alphaES, p.kES[:], p.cutlongrange = 0.6, [4, 4, 4], 5.0
    P.field_check_tag(p)
    L = 12.0
    ions = np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]])
    off = np.array([[[0.3, 0.2, -0.1], [-0.4, 0.35, 0.25]],
                    [[0.1, -0.6, 0.2], [0.5, 0.1, 0.1]]])  # |.| 0.374 0.592 0.648 0.520
    r = np.vstack([ions, [[3.0, 9.0, 4.0], [9.0, 3.0, 8.0]], (ions[:, None, :] + off).reshape(4, 3)])
    r[2:4] -= 0.0
    itype = np.array([1, 1, 2, 2, 3, 3, 3, 3], np.int32)
    # the first ion's second centre wraps through the cell boundary
    r[5] = np.mod(r[5], L)
    orc = oracle_lib.Oracle(p, itype, np.eye(3) * L)
    o = orc.engforce(r, itime=0)
    mu = o["mu"]
    np.testing.assert_allclose(mu[:2], -2.0 * off.sum(axis=1), rtol=1e-13, atol=1e-13)
    assert np.abs(mu[2:]).max() == 0.0
    # a centre at 0.75 (r^2 = 0.5625 > 0.5) is not found: the ion misses one
    r2 = r.copy()
    r2[7] = ions[1] + np.array([0.75, 0.0, 0.0])
    with pytest.raises(RuntimeError, match="status 8"):
        oracle_lib.Oracle(p, itype, np.eye(3) * L).engforce(r2, itime=0)
