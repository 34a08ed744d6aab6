"""The reciprocal sums and the SCF pair table cover only the atoms that need
them: the SCF calls' moments and fields sit on the polarizable atoms, which
the context's species-major numbering keeps contiguous.  The silica
workloads have the polarizable species last; here it is the first (Si
polarizable, O not), so that the covered range starts at atom 0 and the
other species follows it."""
import copy

import numpy as np
import pytest

from helpers import compare_engforce
from test_gpu_mixed import rel_rms

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def _si_polar(n_si, n_o, seed, rc, kmax, alpha_si=0.8):
    s = C.silica(n_si, n_o, seed=seed, rc=rc, kmax=kmax, pim=True, efg=True)
    p = s.params
    p.ldip_polar[0] = 1
    p.ldip_polar[1] = 0
    for k in range(len(p.ldip_damping)):
        p.ldip_damping[k] = 0
    for k in range(len(p.poldip)):
        p.poldip[k] = 0.0
    P.field_check_tag(p, poldip_iso=[alpha_si, 0.0])
    # species interleaved in the caller's order (the context renumbers them)
    order = np.random.default_rng(seed).permutation(s.natm)
    return C.System(p, s.itype[order], s.A, s.r[order], s.name, dict(s.info))


def test_first_species_polarizable_fp64_vs_oracle(oracle_lib):
    from mdff_amd.field import ForceField
    s = _si_polar(40, 80, seed=95, rc=5.5, kmax=5)
    ff = ForceField(s.params, s.itype, s.A)
    orc = oracle_lib.Oracle(s.params, s.itype, s.A)
    rng = np.random.default_rng(3)
    for it in range(1, 3):
        r = s.r + 0.01 * (it - 1) * rng.standard_normal(s.r.shape)
        print(f"Si-polarizable silica itime {it}", flush=True)
        compare_engforce(ff.engforce(r, itime=it), orc.engforce(r, itime=it))
    ff.close()


@pytest.mark.parametrize("solver", [P.SCF_JACOBI, P.SCF_PCG])
def test_first_species_polarizable_mixed_vs_fp64(solver):
    from mdff_amd.field import ForceField
    s = _si_polar(500, 1000, seed=96, rc=8.0, kmax=8)
    s.params.conv_tol_ind = 1e-8
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
    print(f"Si-polarizable 1500 solver {solver}: force {ef:.2e} dipole {em:.2e} efg {eg:.2e} "
          f"n_scf {g['n_scf']} (fp64 {o['n_scf']})", flush=True)
    assert np.abs(o["mu"][s.itype == 2]).max() == 0.0  # O carries no dipole
    assert ef <= 1e-5 and em <= 1e-5 and eg <= 1e-5, (ef, em, eg)


@pytest.mark.parametrize("slices", [1, 4])
def test_apply_slice_counts_vs_fp64(slices):
    """The other apply K-slice counts on the polarizable-first melt (their
    row blocks cover the polarizable rows only), against FP64."""
    from mdff_amd.field import ForceField
    s = _si_polar(500, 1000, seed=97, rc=8.0, kmax=8)
    s.params.conv_tol_ind = 1e-8
    p64 = copy.deepcopy(s.params)
    p64.precision = P.PREC_FP64
    pmx = copy.deepcopy(s.params)
    pmx.precision = P.PREC_MIXED
    f64 = ForceField(p64, s.itype, s.A)
    fmx = ForceField(pmx, s.itype, s.A)
    try:
        assert fmx.set_mixed_tuning(slices, 512)
        o = f64.engforce(s.r, itime=1)
        g = fmx.engforce(s.r, itime=1)
    finally:
        f64.close()
        fmx.close()
    ef, eg = rel_rms(g["f"], o["f"]), rel_rms(g["efg"], o["efg"])
    print(f"apply slices {slices}: force {ef:.2e} efg {eg:.2e}", flush=True)
    assert ef <= 1e-5 and eg <= 1e-5, (ef, eg)
