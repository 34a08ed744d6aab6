"""Callers whose atoms come in any order.  A context numbers its atoms
internally (species by species, by cell), so these cases check that every
output comes back in the caller's order, that the pair-order conventions
follow the caller's numbering (the Tang-Toennies damping slot of a pair, the
sign of the quadrupole-quadrupole EFG, the atom a Wannier error names) and
that the dipole history crosses the numbering unchanged."""
import numpy as np
import pytest

from helpers import assert_close, compare_engforce
from test_gpu_quad import _melt, _wfc_system

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def _shuffled(s, seed):
    order = np.random.default_rng(seed).permutation(s.natm)
    return C.System(s.params, s.itype[order], s.A, s.r[order], s.name, dict(s.info))


def test_shuffled_damped_pim_trajectory(oracle_lib):
    """Silica with O damped by Si and O (the damping slot of a Si-O pair
    depends on which atom comes first) and the species interleaved: three
    steps with ASPC extrapolation against the oracle."""
    from mdff_amd.field import ForceField
    s = _shuffled(C.silica(40, 80, seed=91, rc=5.5, kmax=5, pim=True, efg=True), 1)
    p = s.params
    p.algo_ext_dipole = P.EXT_ASPC
    p.extrapolate_order = 2
    ff = ForceField(p, s.itype, s.A)
    orc = oracle_lib.Oracle(p, s.itype, s.A)
    rng = np.random.default_rng(2)
    for it in range(1, 4):
        r = s.r + 0.01 * (it - 1) * rng.standard_normal(s.r.shape)
        print(f"shuffled silica itime {it}", flush=True)
        compare_engforce(ff.engforce(r, itime=it), orc.engforce(r, itime=it))
    ff.close()


def test_shuffled_quadrupoles(oracle_lib):
    from mdff_amd.field import ForceField
    s = _shuffled(_melt(True, True), 3)
    g = ForceField(s.params, s.itype, s.A).engforce(s.r)
    o = oracle_lib.Oracle(s.params, s.itype, s.A).engforce(s.r)
    compare_engforce(g, o)
    assert_close(g["theta"], o["theta"], what="theta")


def test_device_outputs_in_caller_order():
    """engforce_device and engforce on a shuffled melt give the same
    forces, dipoles, fields and EFGs."""
    import torch
    from mdff_amd.field import ForceField
    s = _shuffled(C.silica(40, 80, seed=92, rc=5.5, kmax=5, pim=True, efg=True), 4)
    n = s.natm
    host = ForceField(s.params, s.itype, s.A).engforce(s.r)
    ff = ForceField(s.params, s.itype, s.A)
    dev = lambda *shape: torch.zeros(shape, dtype=torch.float64, device="cuda")
    pos = torch.tensor(s.r.T.copy(), dtype=torch.float64, device="cuda").contiguous()
    f, mu, ef, efg = dev(3, n), dev(3, n), dev(3, n), dev(3, 3, n)
    ff.engforce_device(pos, f, mu=mu, ef=ef, efg=efg)
    torch.cuda.synchronize()
    assert_close(f.cpu().numpy().T, host["f"], 1e-12, "forces")
    assert_close(mu.cpu().numpy().T, host["mu"], 1e-12, "mu")
    assert_close(ef.cpu().numpy().T, host["ef"], 1e-12, "ef")
    assert_close(efg.cpu().numpy().transpose(2, 1, 0), host["efg"], 1e-12, "efg")
    ff.close()


def test_dipole_history_round_trip():
    """The history set before the first evaluation (caller's order) and
    after it (stored in the internal numbering) reads back unchanged."""
    from mdff_amd.field import ForceField
    s = _shuffled(C.silica(40, 80, seed=93, rc=5.5, kmax=5, pim=True), 5)
    s.params.algo_ext_dipole = P.EXT_ASPC
    s.params.extrapolate_order = 3
    ff = ForceField(s.params, s.itype, s.A)
    rng = np.random.default_rng(6)
    size = (s.params.extrapolate_order + 1) * 3 * s.natm
    h0 = rng.standard_normal(size)
    ff.set_dipole_history(h0)
    np.testing.assert_array_equal(ff.dipole_history(), h0)
    ff.engforce(s.r, itime=1)
    h1 = rng.standard_normal(size)
    ff.set_dipole_history(h1)
    np.testing.assert_array_equal(ff.dipole_history(), h1)
    ff.close()


def test_wannier_error_names_the_callers_atom():
    """Ion 2 of the caller (1-based) loses a centre: the error names it
    whatever the internal numbering."""
    from mdff_amd._lib import MdffError
    from mdff_amd.field import ForceField
    p, itype, A, r, off = _wfc_system()
    order = np.array([6, 1, 3, 0, 7, 2, 5, 4])
    inv = np.argsort(order)
    ff = ForceField(p, itype[order], A)
    ff.engforce(r[order], itime=0)
    r2 = r.copy()
    r2[7] = r[1] + np.array([0.75, 0.0, 0.0])
    with pytest.raises(MdffError) as e:
        ff.engforce(r2[order], itime=1)
    assert e.value.status == 8
    assert f"atom {inv[1] + 1} " in str(e.value), str(e.value)
    ff.close()


def test_shuffled_md_pim_aspc(oracle_lib):
    """md_run on interleaved species: positions, velocities and forces go in
    and come out in the caller's order across the ASPC-driven trajectory."""
    from test_gpu_md import _check, _md_both
    s = _shuffled(C.silica(100, 200, seed=11, rc=6.0, kmax=5, pim=True, efg=False,
                           scf_tol=1e-9), 7)
    p = s.params
    p.calc_mode = P.CALC_MD
    p.algo_ext_dipole = P.EXT_ASPC
    p.extrapolate_order = 2
    p.min_scf_pol_iter = 1
    v0 = C.maxwell_velocities(s, 3000.0, seed=3)
    mdp = P.md_params("nvt-nhc2", dt=0.001, temp=3000.0, timesca_thermo=0.05,
                      mass=C.species_masses(s))
    gs, os_, g, o, _, _ = _md_both(oracle_lib, s, mdp, v0, 6, chunks=2)
    _check(gs, os_, g, o, ("u_bmhft", "u_coul", "u_pol", "e_kin", "e_nvt"))


def test_shuffled_standalone_efg(oracle_lib):
    """The standalone EFG run (mdffgpu_efg) on interleaved species: the EFG
    and dipoles of every site in the caller's order."""
    import copy
    from mdff_amd import efg as E
    from mdff_amd.field import ForceField
    from test_oracle_efg import _pim_cell
    s = _shuffled(_pim_cell(21, n_si=30, n_o=60, kmax=7), 8)
    p = copy.deepcopy(s.params)
    o = P.efg_opts("ewald")
    ff = ForceField(E.context_params(p, o, s.A), s.itype, s.A)
    g = ff.efg(s.r, o)
    ff.close()
    ref = oracle_lib.efgcalc(p, o, s.itype, s.A, s.r)
    assert g["n_scf"] == ref["n_scf"]
    assert_close(g["efg"], ref["efg"], what="efg")
    assert_close(g["mu"], ref["mu"], what="mu")
