"""Device-resident MD (mdffgpu_md_run) against the oracle's integrators on
the same inputs, and the cell-grid neighbour list (nlist.hip) in cubic and
skewed cells.  FP64 parity bar 1e-10 relative, per step for the energies,
at the end for positions, velocities and forces (every step adds the
rounding of one force evaluation, far below the bar over these runs)."""
import numpy as np
import pytest

from helpers import assert_close, compare_engforce
from test_gpu_geometry import skewed_silica

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T),
                        device="cuda:0").contiguous()


def _host(t):
    return t.cpu().numpy().T.copy()


def _md_both(O, s, mdp, v0, nsteps, chunks=1):
    """The same trajectory on the GPU (in `chunks` md_run calls) and in the
    oracle; returns (gpu steps, oracle steps, gpu r v f, oracle r v f)."""
    from mdff_amd.field import ForceField
    ff = ForceField(s.params, s.itype, s.A)
    orc = O.Oracle(s.params, s.itype, s.A)
    g0 = ff.engforce(s.r, itime=0)
    o0 = orc.engforce(s.r, itime=0)
    compare_engforce(g0, o0, efg=False)
    pos, vel, frc = _dev(s.r), _dev(v0), _dev(g0["f"])
    xi_g, vxi_g = np.zeros(16), np.zeros(16)
    gsteps = []
    per = nsteps // chunks
    for c in range(chunks):
        print(f"md_run {s.name} integrator {mdp.integrator} chunk {c}", flush=True)
        gsteps += ff.md_run(mdp, pos, vel, frc, per, itime0=1 + c * per, xi=xi_g, vxi=vxi_g)
    info = ff.list_info()
    ff.close()
    r, v, f, xi_o, vxi_o, osteps = orc.md_run(mdp, s.r, v0, o0["f"], per * chunks, itime0=1)
    return gsteps, osteps, (_host(pos), _host(vel), _host(frc)), (r, v, f), \
        (xi_g, vxi_g, xi_o, vxi_o), info


def _check(gs, os_, g, o, keys):
    for k, (a, b) in enumerate(zip(gs, os_)):
        for key in keys:
            assert_close(a[key], b[key], what=f"step {k + 1} {key}")
        assert a["n_scf"] == b["n_scf"], (k, a["n_scf"], b["n_scf"])
    for a, b, what in zip(g, o, ("positions", "velocities", "forces")):
        assert_close(a, b, what=what)


def test_nve_vv_lj_fcc_500(oracle_lib):
    s = C.config(1)
    rng = np.random.default_rng(7)
    s.r = s.r + 0.03 * rng.standard_normal(s.r.shape)
    v0 = C.maxwell_velocities(s, 0.72, seed=1)
    mdp = P.md_params("nve-vv", dt=0.005, mass=(1.0,), lreduced=True)
    gs, os_, g, o, _, info = _md_both(oracle_lib, s, mdp, v0, 60, chunks=2)
    _check(gs, os_, g, o, ("u_lj", "vir_lj", "e_kin", "temp_r"))
    assert info["builds"] >= 2  # the skin was crossed during the run


def test_nvt_nhcn_silica_648(oracle_lib):
    s = C.config(2)
    v0 = C.maxwell_velocities(s, 3000.0, seed=2)
    mdp = P.md_params("nvt-nhcn", dt=0.001, temp=3000.0, timesca_thermo=0.05, nhc_n=4,
                      nhc_yosh_order=3, nhc_mults=2, mass=C.species_masses(s))
    gs, os_, g, o, th, _ = _md_both(oracle_lib, s, mdp, v0, 20)
    _check(gs, os_, g, o, ("u_bmhft", "u_coul", "e_kin", "e_nvt"))
    xi_g, vxi_g, xi_o, vxi_o = th
    assert_close(xi_g[:4], xi_o[:4], what="xi")
    assert_close(vxi_g[:4], vxi_o[:4], what="vxi")


def test_nvt_nhc2_pim_aspc(oracle_lib):
    """Polarizable ions with the ASPC predictor-corrector: the step index
    md_run passes to each evaluation drives the dipole history."""
    s = C.silica(100, 200, seed=11, rc=6.0, kmax=5, pim=True, efg=False, scf_tol=1e-9)
    p = s.params
    p.calc_mode = P.CALC_MD
    p.algo_ext_dipole = P.EXT_ASPC
    p.extrapolate_order = 2
    p.min_scf_pol_iter = 1
    v0 = C.maxwell_velocities(s, 3000.0, seed=3)
    mdp = P.md_params("nvt-nhc2", dt=0.001, temp=3000.0, timesca_thermo=0.05,
                      mass=C.species_masses(s))
    gs, os_, g, o, th, _ = _md_both(oracle_lib, s, mdp, v0, 8, chunks=2)
    _check(gs, os_, g, o, ("u_bmhft", "u_coul", "u_pol", "e_kin", "e_nvt"))
    assert_close(th[0][:2], th[2][:2], what="xi")


def test_cell_list_skewed_cell(oracle_lib):
    """A skewed cell whose widths admit the cell-grid stencil: the list of
    the grid search gives the oracle's all-pairs energies and forces."""
    from mdff_amd.field import ForceField
    s = skewed_silica(500, 1000, seed=21, kmax=(5, 4, 6), rc=8.0, pim=False, efg=False)
    ff = ForceField(s.params, s.itype, s.A)
    orc = oracle_lib.Oracle(s.params, s.itype, s.A)
    rng = np.random.default_rng(5)
    r = s.r
    for it in range(1, 4):
        g = ff.engforce(r, itime=it)
        o = orc.engforce(r, itime=it)
        compare_engforce(g, o, efg=False)
        r = r + 0.2 * rng.standard_normal(r.shape)  # past the skin: rebuilt each call
    info = ff.list_info()
    ff.close()
    assert info["cell_search"], "the stencil search did not apply"
    assert info["builds"] == 3


def test_cell_list_small_box_falls_back(oracle_lib):
    """500 LJ atoms in an 8.4 sigma box: cells too few for the stencil, the
    all-pairs build runs."""
    from mdff_amd.field import ForceField
    s = C.config(1)
    ff = ForceField(s.params, s.itype, s.A)
    ff.engforce(s.r, itime=1)
    info = ff.list_info()
    ff.close()
    assert not info["cell_search"]
