"""Mixed-precision path (FP32 real space, fp16-split MFMA k-space, FP64
reductions) against the FP64 GPU path on the same inputs.  Bar: relative
force RMS error <= 1e-5 (the stated target), dipoles / EFG / energies to
1e-5."""
import copy

import numpy as np
import pytest

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _pair(s, steps=1, perturb=0.0):
    from mdff_amd.field import ForceField
    p64 = copy.deepcopy(s.params)
    p64.precision = P.PREC_FP64
    pmx = copy.deepcopy(s.params)
    pmx.precision = P.PREC_MIXED
    f64 = ForceField(p64, s.itype, s.A)
    fmx = ForceField(pmx, s.itype, s.A)
    rng = np.random.default_rng(3)
    r = s.r.copy()
    out = []
    for it in range(1, steps + 1):
        out.append((fmx.engforce(r, itime=it), f64.engforce(r, itime=it)))
        r = r + perturb * rng.standard_normal(r.shape)
    f64.close()
    fmx.close()
    return out


@pytest.mark.parametrize("pim", [False, True])
def test_mixed_vs_fp64_silica(pim):
    s = C.silica(200, 400, seed=31, rc=9.0, kmax=8, pim=pim, efg=True)
    for g, o in _pair(s, steps=2, perturb=0.01):
        assert rel_rms(g["f"], o["f"]) <= 1e-5, rel_rms(g["f"], o["f"])
        assert abs(g["u_coul"] - o["u_coul"]) <= 1e-6 * abs(o["u_coul"])
        assert abs(g["u_bmhft"] - o["u_bmhft"]) <= 1e-6 * abs(o["u_bmhft"])
        assert rel_rms(g["efg"], o["efg"]) <= 1e-5
        if pim:
            assert rel_rms(g["mu"], o["mu"]) <= 1e-5
            # the mixed SCF ends at tolerance or when the FP32 noise floor stalls
            # it; either way well before the iteration cap
            assert g["n_scf"] < s.params.max_scf_pol_iter
        assert np.abs(g["tau_coul"] - o["tau_coul"]).max() <= 1e-5 * np.abs(o["tau_coul"]).max()


def test_mixed_ex9_reference_energy():
    import golden_cases as G
    from mdff_amd.field import ForceField
    p, itype, A, r, meta = G.ex9_test11(precision=P.PREC_MIXED)
    g = ForceField(p, itype, A).engforce(r, itime=0)
    assert abs(g["u_bmhft"] + g["u_coul"] - meta["reference_utot"]) <= 1e-6 * abs(meta["reference_utot"])


@pytest.mark.parametrize("precision", [P.PREC_FP64, P.PREC_MIXED])
def test_pcg_scf_matches_jacobi(precision):
    """The conjugate-gradient SCF reaches the Jacobi iteration's fixed point:
    forces, dipoles, u_pol and EFGs against the FP64 Jacobi solve (the
    reference's algorithm) to the 1e-5 bar, over two perturbed steps (the
    dipole history feeds the initial guess), in no more Ewald calls."""
    from mdff_amd.field import ForceField
    s = C.silica(200, 400, seed=31, rc=9.0, kmax=8, pim=True, efg=True)
    pj = copy.deepcopy(s.params)
    pj.precision = P.PREC_FP64
    pc = copy.deepcopy(s.params)
    pc.precision = precision
    pc.scf_solver = P.SCF_PCG
    fj = ForceField(pj, s.itype, s.A)
    fc = ForceField(pc, s.itype, s.A)
    rng = np.random.default_rng(5)
    r = s.r.copy()
    try:
        for it in range(1, 3):
            o = fj.engforce(r, itime=it)
            g = fc.engforce(r, itime=it)
            assert rel_rms(g["f"], o["f"]) <= 1e-5, rel_rms(g["f"], o["f"])
            assert rel_rms(g["mu"], o["mu"]) <= 1e-5, rel_rms(g["mu"], o["mu"])
            assert abs(g["u_pol"] - o["u_pol"]) <= 1e-5 * abs(o["u_pol"])
            assert rel_rms(g["efg"], o["efg"]) <= 1e-5
            assert g["n_scf"] <= o["n_scf"], (g["n_scf"], o["n_scf"])
            r = r + 0.01 * rng.standard_normal(r.shape)
    finally:
        fj.close()
        fc.close()


@pytest.mark.parametrize("solver", [P.SCF_JACOBI, P.SCF_PCG])
def test_mixed_vs_fp64_pim_3000(solver):
    """The 3000-atom polarizable silica workload (1000 Si + 2000 O, rc 10 A,
    kmax 10, SCF to 1e-8, EFG at every site): force, dipole and EFG error of
    the mixed path against the FP64 path (Jacobi, the reference's SCF), with
    either SCF solver in the mixed path."""
    from mdff_amd.field import ForceField
    s = C.config(3)
    p64 = copy.deepcopy(s.params)
    p64.precision = P.PREC_FP64
    pmx = copy.deepcopy(s.params)
    pmx.precision = P.PREC_MIXED
    pmx.scf_solver = solver
    f64 = ForceField(p64, s.itype, s.A)
    fmx = ForceField(pmx, s.itype, s.A)
    try:
        o = f64.engforce(s.r, itime=1)
        g = fmx.engforce(s.r, itime=1)
    finally:
        f64.close()
        fmx.close()
    ef, em, eg = rel_rms(g["f"], o["f"]), rel_rms(g["mu"], o["mu"]), rel_rms(g["efg"], o["efg"])
    print(f"pim 3000 solver {solver}: force {ef:.2e} dipole {em:.2e} efg {eg:.2e} "
          f"n_scf {g['n_scf']} (fp64 {o['n_scf']}) rmsd {g['rmsd']:.2e} "
          f"(fp64 {o['rmsd']:.2e}) scf_tol {g['scf_tol']:.2e}")
    assert o["rmsd"] <= s.params.conv_tol_ind
    assert g["rmsd"] <= g["scf_tol"] or g["n_scf"] < s.params.max_scf_pol_iter
    assert ef <= 1e-5 and em <= 1e-5 and eg <= 1e-5, (ef, em, eg)
    assert abs(g["u_pol"] - o["u_pol"]) <= 1e-5 * abs(o["u_pol"])


@pytest.mark.parametrize("part", [P.MIXED_FP64_SHORTRANGE, P.MIXED_FP64_DIRECT,
                                  P.MIXED_FP64_SCF_FIELD, P.MIXED_FP64_KSPACE])
def test_mixed_with_fp64_part(part):
    """Each part of the mixed path switched back to FP64 on its own
    (mixed_fp64_parts): the combinations route the SCF pair table through
    the separate table pass (FP64 static real space), the FP64 table, or the
    FP64 reciprocal space beside the FP32 SCF field.  Same bar as the full
    mixed path."""
    s = C.silica(200, 400, seed=31, rc=9.0, kmax=8, pim=True, efg=True)
    s.params.mixed_fp64_parts = part
    for g, o in _pair(s, steps=2, perturb=0.01):
        assert rel_rms(g["f"], o["f"]) <= 1e-5, (part, rel_rms(g["f"], o["f"]))
        assert rel_rms(g["mu"], o["mu"]) <= 1e-5, (part, rel_rms(g["mu"], o["mu"]))
        assert rel_rms(g["efg"], o["efg"]) <= 1e-5
        assert abs(g["u_pol"] - o["u_pol"]) <= 1e-5 * abs(o["u_pol"])


def _min_image_dist(r, i, L):
    d = r - r[i]
    d -= L * np.round(d / L)
    dist = np.sqrt((d ** 2).sum(axis=1))
    dist[i] = np.inf
    return float(dist.min())


def test_mixed_isolated_polarizable_atom():
    """A polarizable O atom with no neighbour inside the cutoff: its row of
    the compacted SCF pair table is empty (the table keeps one zero entry for
    it), and its dipole comes from the reciprocal-space field alone.  The
    600-atom melt sits in one octant of a box of twice its edge; the atom is
    moved to the opposite corner."""
    s = C.silica(200, 400, seed=31, rc=9.0, kmax=8, pim=True)
    L = float(s.A[0, 0])
    s.A = 2.0 * s.A
    s.r = s.r.copy()
    s.r[-1] = 1.5 * L
    assert _min_image_dist(s.r, s.natm - 1, 2.0 * L) > s.params.cutlongrange + s.params.skindiff
    (g, o), = _pair(s)
    assert rel_rms(g["f"], o["f"]) <= 1e-5, rel_rms(g["f"], o["f"])
    assert rel_rms(g["mu"], o["mu"]) <= 1e-5, rel_rms(g["mu"], o["mu"])
    assert np.abs(o["mu"][-1]).max() > 0.0
    assert np.abs(g["mu"][-1] - o["mu"][-1]).max() <= 1e-5 * np.abs(o["mu"]).max()


def test_mixed_pair_table_int32_slots():
    """More than 65,536 atoms: the compacted SCF pair table holds int32
    neighbour slots instead of 16-bit ones (21,846 Si + 43,692 O, rc 6 A,
    kmax 4)."""
    s = C.silica(21846, 43692, seed=5, rc=6.0, kmax=4, pim=True)
    assert s.natm > 65536
    (g, o), = _pair(s)
    assert rel_rms(g["f"], o["f"]) <= 1e-5, rel_rms(g["f"], o["f"])
    assert rel_rms(g["mu"], o["mu"]) <= 1e-5, rel_rms(g["mu"], o["mu"])
    assert abs(g["u_pol"] - o["u_pol"]) <= 1e-5 * abs(o["u_pol"])


def test_mixed_floor_not_raised_by_iteration_limit():
    """The mixed SCF's noise-floor estimate (the tolerance it may apply above
    conv_tol_ind) is raised only by a stalled solve: evaluations that stop at
    max_scf_pol_iter above the tolerance leave the next evaluation's
    tolerance at conv_tol_ind.  scf_tol reports the tolerance applied."""
    from mdff_amd.field import ForceField
    s = C.silica(200, 400, seed=31, rc=9.0, kmax=8, pim=True)
    p = copy.deepcopy(s.params)
    p.precision = P.PREC_MIXED
    p.conv_tol_ind = 1e-7
    p.min_scf_pol_iter = 1
    p.max_scf_pol_iter = 2
    ff = ForceField(p, s.itype, s.A)
    try:
        for it in range(1, 4):
            g = ff.engforce(s.r, itime=it)
            assert g["n_scf"] == 2 and g["rmsd"] > p.conv_tol_ind, (g["n_scf"], g["rmsd"])
            assert g["scf_tol"] == p.conv_tol_ind, (it, g["scf_tol"])
    finally:
        ff.close()
