"""Standalone EFG (calc = 'efg') on the GPU against the oracle's efgcalc at
1e-10, and the ex4 known answers through the whole file pipeline (TRAJFF ->
EFGALL -> efg_stat -> NMRFF) with the GPU's NMR kernel."""
import copy
import os

import numpy as np
import pytest

from mdff_amd import configs as C
from mdff_amd import efg as E
from mdff_amd import io
from mdff_amd import params as P

from test_oracle_efg import _cluster, _pim_cell, _two_charges, oracle_nmr

pytestmark = pytest.mark.gpu

FILES = os.path.join(os.path.dirname(__file__), "golden", "files")


def close(g, o, tol=1e-10):
    err = np.abs(g - o).max()
    scale = np.abs(o).max()
    assert err <= tol * scale, (err, scale)


def _gpu_efg(p, itype, A, r, opts):
    from mdff_amd.field import ForceField
    ff = ForceField(E.context_params(p, opts, A), itype, A)
    try:
        return ff.efg(r, opts)
    finally:
        ff.close()


@pytest.mark.parametrize("longrange", ["ewald", "direct"])
def test_ex4_two_charges(oracle_lib, longrange):
    p, fr, o = _two_charges(longrange)
    print(f"[efg two charges] {longrange}", flush=True)
    g = _gpu_efg(p, fr.itype(), fr.A, fr.r, o)["efg"]
    ref = oracle_lib.efgcalc(p, o, fr.itype(), fr.A, fr.r)["efg"]
    close(g, ref)
    if longrange == "ewald":
        assert abs(g[0, 0, 0] - 2.000374925518976) <= 1e-13
        assert abs(g[0, 1, 1] - -1.000187462857358) <= 1e-13


@pytest.mark.parametrize("longrange", ["ewald", "direct"])
def test_ex4_cluster_file_pipeline_matches_gulp(oracle_lib, tmp_path, longrange):
    """efgcalc over examples/ex4/TRAJFF.cluster: EFGALL, NMRFF and the
    distributions written, principal values and eta against GULP."""
    p, fr, o, gulp, gulp_eta = _cluster(longrange)
    stat = E.EfgStatOptions(umin=-15.0, smax=25.0, vzzmin=-50.0)
    print(f"[efg cluster] {longrange}", flush=True)
    run = E.efgcalc(p, o, os.path.join(FILES, "ex4_TRAJFF.cluster"), ncefg=1, stat=stat,
                    workdir=str(tmp_path))
    ref = oracle_lib.efgcalc(p, o, fr.itype(), fr.A, fr.r)["efg"]
    close(run.efg[0], ref)
    nmr = run.stats.nmr_last
    assert np.allclose(nmr[:, :3], gulp, atol=1e-3)
    assert np.allclose(nmr[:, 3], gulp_eta, atol=2e-4)
    # the GPU's eigenvalues and ordering against LAPACK's on the same file
    back = next(io.read_efgall(str(tmp_path / "EFGALL")))
    nref = oracle_nmr(back.efg6)
    close(nmr, nref[0], 1e-12)
    for name in ("EFGALL", "NMRFF", "DTETAFF", "DTVZZFF", "DTIBUFF", "DTIBSFF"):
        assert (tmp_path / name).exists()


@pytest.mark.parametrize("variant", ["plain", "dip+restricted"])
def test_polarizable_skewed_cell(oracle_lib, variant):
    """Induced dipoles from the SCF (and permanent dipoles), a skewed cell
    with a different kmax per axis; the restricted variant adds it_efg,
    the charged-cell correction and the DFT sign."""
    from test_gpu_geometry import skewed_silica
    s = skewed_silica(40, 80, seed=81, kmax=(7, 5, 6))
    p = copy.deepcopy(s.params)
    p.calc_mode = P.CALC_MD  # the EFG run takes the moments as calc = 'efg'
    p.conv_tol_ind = 1e-12
    p.max_scf_pol_iter = 200
    o = P.efg_opts("ewald")
    if variant != "plain":
        p.dip[1], p.dip[17], p.dip[33] = 0.04, -0.03, 0.02
        P.field_check_tag(p)
        o = P.efg_opts("ewald", lefg_it_contrib=True, it_efg=2, lmp_correction=True,
                       lefg_vasp_sign=True)
    print(f"[efg pim skewed] {variant}", flush=True)
    g = _gpu_efg(p, s.itype, s.A, s.r, o)
    ref = oracle_lib.efgcalc(p, o, s.itype, s.A, s.r)
    assert g["n_scf"] == ref["n_scf"]
    close(g["mu"], ref["mu"])
    close(g["efg"], ref["efg"])


def test_trajectory_with_changing_cell(oracle_lib, tmp_path):
    """Three TRAJFF frames (unformatted) whose cell grows and shears: one
    context, its cell updated per frame, against the oracle frame by
    frame."""
    s = _pim_cell(21, n_si=30, n_o=60, kmax=7)
    p = copy.deepcopy(s.params)
    p.conv_tol_ind = 1e-12
    p.max_scf_pol_iter = 200
    rng = np.random.default_rng(6)
    frac = np.linalg.solve(s.A, s.r.T).T
    cells = [s.A, 1.01 * s.A, 1.01 * s.A @ np.array([[1, 0.04, 0], [0, 1, 0], [0, 0, 1.0]])]
    path = str(tmp_path / "TRAJFF")
    with open(path, "wb") as fh:
        for k, A in enumerate(cells):
            # the reference's TRAJFF writer stores the rows of A, its reader
            # takes them as columns: write A^T so that A comes back
            r = (frac + 0.003 * k * rng.standard_normal(frac.shape)) @ A.T
            fr = io.Frame("traj", A.T.copy(), ["Si", "O"], [30, 60], r)
            io.write_trajff(fh, fr, "rnn", formatted=False)
    frames = [(f.A, f.r) for f in io.read_trajff(path, "rnn", formatted=False)]
    for (A_read, _), A in zip(frames, cells):
        assert np.array_equal(A_read, A)
    o = P.efg_opts("ewald")
    print("[efg trajectory] 3 frames", flush=True)
    run = E.efgcalc(p, o, path, ncefg=3, traj_formatted=False, efgall_formatted=False,
                    workdir=str(tmp_path))
    back = list(io.read_efgall(str(tmp_path / "EFGALL"), formatted=False))
    assert len(back) == 3
    for k, (A, r) in enumerate(frames):
        ref = oracle_lib.efgcalc(p, o, s.itype, A, r)["efg"]
        close(run.efg[k], ref)
        assert np.array_equal(back[k].efg6, io.efg_six(run.efg[k]))


def test_nmr_kernel_against_lapack():
    rng = np.random.default_rng(3)
    n = 5000
    v = rng.normal(size=(n, 6)) * rng.uniform(0.01, 100.0, (n, 1))
    v[:, 2] = -(v[:, 0] + v[:, 1])  # traceless like an EFG
    # exact special cases: axial, zero, diagonal, tiny
    v[0] = [-1.0, -1.0, 2.0, 0, 0, 0]
    v[1] = 0.0
    v[2] = [3.0, -1.0, -2.0, 0, 0, 0]
    v[3] = [1e-7, -2e-7, 1e-7, 0, 0, 0]
    nmr, U, S, bad = E.efg_nmr(v)
    ref = oracle_nmr(v)
    assert bad == ref[3] == 0
    scale = np.abs(v).max(axis=1, keepdims=True)
    assert np.all(np.abs(nmr[:, :3] - ref[0][:, :3]) <= 1e-13 * np.maximum(scale, 1e-300))
    assert np.abs(nmr[4:, 3] - ref[0][4:, 3]).max() <= 1e-10
    assert np.array_equal(nmr[:4, 3], ref[0][:4, 3])
    assert np.allclose(U, ref[1], rtol=1e-15, atol=0) and np.allclose(S, ref[2], rtol=1e-14, atol=0)


def test_pim_3000_efg_matches_oracle(oracle_lib):
    """The 3000-atom polarizable silica workload (kmax 10, SCF to 1e-8)."""
    s = C.config(3)
    p = copy.deepcopy(s.params)
    p.calc_mode = P.CALC_OTHER
    o = P.efg_opts("ewald")
    print(f"[efg pim 3000] {s.natm} atoms", flush=True)
    g = _gpu_efg(p, s.itype, s.A, s.r, o)
    ref = oracle_lib.efgcalc(p, o, s.itype, s.A, s.r)
    assert g["n_scf"] == ref["n_scf"]
    close(g["efg"], ref["efg"])


def test_mixed_precision_context_efg():
    """A mixed-precision context takes the SCF in its own precision; the EFG
    sums are FP64: the result is within the mixed bar of the FP64 one."""
    s = _pim_cell(23, n_si=150, n_o=300, kmax=9)
    out = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        p = copy.deepcopy(s.params)
        p.precision = prec
        p.conv_tol_ind = 1e-7
        out[prec] = _gpu_efg(p, s.itype, s.A, s.r, P.efg_opts("ewald"))["efg"]
    err = np.sqrt(np.mean((out[1] - out[0]) ** 2)) / np.sqrt(np.mean(out[0] ** 2))
    assert err <= 1e-5, err


def test_refused_options():
    from mdff_amd._lib import MdffError
    s = _pim_cell(24)
    with pytest.raises(MdffError):
        o = P.efg_opts("direct", cutlongrange=10.0)  # induced dipoles need Ewald
        _gpu_efg(s.params, s.itype, s.A, s.r, o)
    o = P.efg_opts("ewald")
    o.lefg_old = 1
    with pytest.raises(MdffError):
        _gpu_efg(s.params, s.itype, s.A, s.r, o)
