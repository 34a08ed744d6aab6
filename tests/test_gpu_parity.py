"""FP64 GPU path (C ABI) against the CPU oracle on identical inputs, to
1e-10 relative (relative to the largest magnitude of each quantity)."""
import numpy as np
import pytest

from helpers import assert_close, compare_engforce

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def _run_both(sys_, O, steps=1, perturb=0.0, seed=0, r0=None):
    from mdff_amd.field import ForceField
    ff = ForceField(sys_.params, sys_.itype, sys_.A)
    orc = O.Oracle(sys_.params, sys_.itype, sys_.A)
    rng = np.random.default_rng(seed)
    r = (sys_.r if r0 is None else r0).copy()
    outs = []
    for it in range(1, steps + 1):
        g = ff.engforce(r, itime=it)
        o = orc.engforce(r, itime=it)
        outs.append((g, o))
        r = r + perturb * rng.standard_normal(r.shape)
    ff.close()
    return outs


def _jitter(s, amp, seed=1):
    rng = np.random.default_rng(seed)
    return s.r + amp * rng.standard_normal(s.r.shape)


def test_lj_fcc_500(oracle_lib):
    s = C.config(1)
    (g, o), = _run_both(s, oracle_lib, r0=_jitter(s, 0.05))
    compare_engforce(g, o)


@pytest.mark.parametrize("trunc", [1, 2])
def test_lj_truncations(oracle_lib, trunc):
    s = C.config(1)
    s.params.trunc = trunc
    (g, o), = _run_both(s, oracle_lib, r0=_jitter(s, 0.05))
    compare_engforce(g, o)


def test_silica_rigid_ion_648(oracle_lib):
    s = C.config(2)
    (g, o), = _run_both(s, oracle_lib)
    compare_engforce(g, o)


def test_silica_pim_efg_small(oracle_lib):
    s = C.silica(60, 120, seed=11, rc=6.0, kmax=5, pim=True, efg=True)
    for g, o in _run_both(s, oracle_lib, steps=3, perturb=0.01):
        compare_engforce(g, o)


@pytest.mark.parametrize("iters", [(3, 100), (1, 100), (1, 1)])
@pytest.mark.parametrize("algo,order", [(P.EXT_POLY, 1), (P.EXT_ASPC, 1), (P.EXT_ASPC, 2),
                                        (P.EXT_ASPC, 4)])
def test_pim_extrapolation_history(oracle_lib, algo, order, iters):
    """Extrapolated first iterates over 6 MD steps.  With min_scf_pol_iter = 1
    the first iteration may be the last, so the ASPC corrector's effect on
    the residual and on the stopping test is compared too (the reference
    tests convergence after the corrector, src/field.f90:3368-3400)."""
    s = C.silica(40, 80, seed=12, rc=5.5, kmax=4, pim=True, efg=False)
    s.params.algo_ext_dipole = algo
    s.params.extrapolate_order = order
    s.params.min_scf_pol_iter, s.params.max_scf_pol_iter = iters
    perturb = 0.005
    if iters[0] == 1:
        s.params.conv_tol_ind = 3e-3
        perturb = 0.001
    nscf = []
    for g, o in _run_both(s, oracle_lib, steps=6, perturb=perturb):
        compare_engforce(g, o)
        assert abs(g["rmsd"] - o["rmsd"]) <= 1e-6 * o["rmsd"], (g["rmsd"], o["rmsd"])
        nscf.append(o["n_scf"])
    if iters == (1, 100) and algo == P.EXT_ASPC:
        # the extrapolated iterate converges in one step on some MD step
        assert min(nscf) == 1, nscf


def test_pim_static_calc_mode(oracle_lib):
    s = C.silica(40, 80, seed=13, rc=5.5, kmax=4, pim=True, efg=True)
    s.params.calc_mode = P.CALC_OTHER
    (g, o), = _run_both(s, oracle_lib)
    compare_engforce(g, o)


def test_surface_term_and_efield(oracle_lib):
    """lsurf on a rigid-ion melt.  (With polarizable ions the reference's
    unscaled surface field, src/field.f90:1942, makes the SCF diverge.)"""
    s = C.silica(30, 60, seed=14, rc=5.0, kmax=4, pim=False, efg=True)
    s.params.lsurf = 1
    (g, o), = _run_both(s, oracle_lib)
    compare_engforce(g, o)


def test_two_charges_efg_known_answer(oracle_lib):
    """Nymand & Linse two charges in a 10 box (reference examples/ex4)."""
    from mdff_amd.field import ForceField
    p = P.field_defaults(2, 2)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.cutlongrange = 1000.0
    p.qch[0], p.qch[1] = 1.0, -1.0
    p.kES[:] = [11, 11, 11]
    p.alphaES = 0.8
    p.doefg = 1
    P.field_check_tag(p)
    A = np.eye(3) * 10.0
    P.ewald_param(p, A)
    ff = ForceField(p, np.array([1, 2], np.int32), A)
    g = ff.engforce(np.array([[0.5, 0, 0], [-0.5, 0, 0]]))
    assert abs(g["efg"][0, 0, 0] - 2.0003749) < 1e-7
    assert abs(g["efg"][0, 1, 1] + 1.0001874) < 1e-7
    o = oracle_lib.Oracle(p, np.array([1, 2], np.int32), A).engforce(np.array([[0.5, 0, 0], [-0.5, 0, 0]]))
    compare_engforce(g, o)


def test_lj_morse_bmhftd_mix_triclinic(oracle_lib):
    """All short-range terms at once in a skewed cell (the BMHFTD routine
    overwrites the LJ forces, Morse adds)."""
    rng = np.random.default_rng(3)
    n = 150
    A = np.array([[14.0, 0.0, 0.0], [2.0, 13.0, 0.0], [1.5, -1.0, 12.5]]).T
    s = C.rsa_positions(n, 1.0, 0.1, rng)
    r = s @ A.T
    p = P.field_defaults(n, 2)
    p.lnmlj = p.lbmhftd = p.lmorse = 1
    p.cutshortrange = 5.5
    for (i, j) in ((1, 1), (1, 2), (2, 2)):
        k = P.pidx(i, j)
        p.epslj[k], p.sigmalj[k] = 0.01 * (i + j), 2.0 + 0.1 * i
        p.Abmhftd[k], p.Bbmhftd[k], p.Cbmhftd[k], p.Dbmhftd[k], p.BDbmhftd[k] = 2e3, 3.5, 20.0, 2.0, 3.0
        p.epsmor[k], p.sigmamor[k], p.rhomor[k] = 0.05, 2.2, 1.7
    P.field_check_tag(p)
    itype = np.array([1 + (i % 2) for i in range(n)], np.int32)
    sysx = C.System(p, itype, A, r, "mix", {})
    (g, o), = _run_both(sysx, oracle_lib)
    compare_engforce(g, o)
    assert g["u_lj"] != 0.0 and g["u_morse"] != 0.0


def test_coulomb_only_accumulates_forces(oracle_lib):
    """Without a short-range term the Coulomb forces are added to the
    incoming ones (src/field.f90:1966-1968)."""
    from mdff_amd.field import ForceField
    s = C.silica(20, 40, seed=15, rc=5.0, kmax=4)
    s.params.lbmhftd = 0
    f_in = np.random.default_rng(0).standard_normal((s.natm, 3))
    ff = ForceField(s.params, s.itype, s.A)
    g = ff.engforce(s.r, f_in=f_in)
    o = oracle_lib.Oracle(s.params, s.itype, s.A).engforce(s.r, f_in=f_in)
    assert_close(g["f"], o["f"], what="forces")


def test_device_entry_matches_host_entry(oracle_lib):
    torch = pytest.importorskip("torch")
    from mdff_amd.field import ForceField
    s = C.silica(40, 80, seed=16, rc=5.5, kmax=4, pim=True, efg=True)
    ff1 = ForceField(s.params, s.itype, s.A)
    g = ff1.engforce(s.r, itime=1)
    ff2 = ForceField(s.params, s.itype, s.A)
    pos = torch.tensor(s.r.T.copy(), dtype=torch.float64, device="cuda")
    f = torch.zeros_like(pos)
    mu = torch.zeros_like(pos)
    efg = torch.zeros((3, 3, s.natm), dtype=torch.float64, device="cuda")
    en = ff2.engforce_device(pos, f, mu=mu, efg=efg, itime=1)
    torch.cuda.synchronize()
    assert_close(f.cpu().numpy().T, g["f"], 1e-13, "forces")
    assert_close(mu.cpu().numpy().T, g["mu"], 1e-13, "mu")
    assert_close(efg.cpu().numpy().transpose(2, 1, 0), g["efg"], 1e-13, "efg")
    assert abs(en.u_coul - g["u_coul"]) <= 1e-12 * abs(g["u_coul"])


def test_ex9_test11_reference_energy_on_gpu(oracle_lib):
    """examples/ex9 test 11 (BMHFTD + damped PIM + Ewald, 200 atoms):
    the GPU reproduces the reference run's Utot and SCF iteration count."""
    import golden_cases as G
    from mdff_amd.field import ForceField
    p, itype, A, r, meta = G.ex9_test11()
    g = ForceField(p, itype, A).engforce(r, itime=0)
    o = oracle_lib.Oracle(p, itype, A).engforce(r, itime=0)
    compare_engforce(g, o)
    assert g["n_scf"] == meta["reference_nscf"]
    assert abs(g["u_bmhft"] + g["u_coul"] - meta["reference_utot"]) < 1e-7


def test_ex4_cluster_efg_on_gpu(oracle_lib):
    import golden_cases as G
    from mdff_amd.field import ForceField
    p, itype, A, r, gulp = G.ex4_cluster()
    g = ForceField(p, itype, A).engforce(r)
    o = oracle_lib.Oracle(p, itype, A).engforce(r)
    compare_engforce(g, o)
