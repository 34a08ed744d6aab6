"""Quadrupoles (task_coul(4:6), static and induced) and Wannier-centre
dipoles on the GPU against the oracle, FP64 at 1e-10, and the options the
library refuses around them."""
import numpy as np
import pytest

from helpers import assert_close, compare_engforce
from test_oracle_quad import Q_A, Q_B, _quad_melt

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu


def _both(O, s, frames, itimes=None):
    from mdff_amd.field import ForceField
    ff = ForceField(s.params, s.itype, s.A)
    orc = O.Oracle(s.params, s.itype, s.A)
    out = []
    for k, r in enumerate(frames):
        it = k + 1 if itimes is None else itimes[k]
        print(f"{s.name} frame {k} itime {it}", flush=True)
        out.append((ff.engforce(r, itime=it), orc.engforce(r, itime=it)))
    ff.close()
    return out


def _check(g, o):
    compare_engforce(g, o)
    assert_close(g["theta"], o["theta"], what="theta")


def _melt(static=True, polar=True, n_si=60, n_o=120, seed=5):
    s = C.silica(n_si, n_o, seed=seed, rc=6.0, kmax=6, pim=True, efg=True, scf_tol=1e-9)
    p = s.params
    if static:
        for a in range(1, 4):
            for b in range(1, 4):
                p.quad[P.qidx(1, a, b)] = 0.08 * Q_A[a - 1, b - 1]
                p.quad[P.qidx(2, a, b)] = 0.03 * Q_B[a - 1, b - 1]
    if polar:
        p.lquad_polar[1] = 1
        P.field_check_tag(p, polquad_iso=[0.0, 0.25])
    else:
        P.field_check_tag(p)
    p.calc_mode = P.CALC_OTHER
    return s


@pytest.mark.parametrize("static,polar", [(True, False), (False, True), (True, True)])
def test_quadrupoles_match_oracle(oracle_lib, static, polar):
    s = _melt(static, polar)
    assert list(s.params.task_coul_quad) == [1, 1, 1]
    (g, o), = _both(oracle_lib, s, [s.r])
    _check(g, o)
    assert np.abs(g["theta"]).max() > 0.0


def test_quadrupoles_aspc_trajectory(oracle_lib):
    """Induced quadrupoles with the ASPC predictor-corrector over a short
    trajectory (theta_ind zeroed by the predictor, polquad E from the
    corrector)."""
    s = _melt(True, True)
    p = s.params
    p.calc_mode = P.CALC_MD
    p.algo_ext_dipole = P.EXT_ASPC
    p.extrapolate_order = 2
    p.min_scf_pol_iter = 1
    rng = np.random.default_rng(3)
    frames = [s.r + 0.02 * k * rng.standard_normal(s.r.shape) for k in range(4)]
    for g, o in _both(oracle_lib, s, frames):
        _check(g, o)


def test_small_quad_melt_all_outputs(oracle_lib):
    s = _quad_melt(oracle_lib)
    (g, o), = _both(oracle_lib, s, [s.r])
    _check(g, o)


def _wfc_system():
    p = P.field_defaults(8, 3)
    P.reduced_units(p)
    p.lcoulomb = 1
    p.qch[0], p.qch[1], p.qch[2] = 2.0, 1.0, -1.5
    p.lwfc[0], p.lwfc[1], p.lwfc[2] = 2, 0, -1
    p.rcut_wfc = 0.5
    p.alphaES, p.kES[:], p.cutlongrange = 0.6, [4, 4, 4], 5.0
    P.field_check_tag(p)
    L = 12.0
    ions = np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]])
    off = np.array([[[0.3, 0.2, -0.1], [-0.4, 0.35, 0.25]],
                    [[0.1, -0.6, 0.2], [0.5, 0.1, 0.1]]])
    r = np.vstack([ions, [[3.0, 9.0, 4.0], [9.0, 3.0, 8.0]],
                   (ions[:, None, :] + off).reshape(4, 3)])
    r[5] = np.mod(r[5], L)
    itype = np.array([1, 1, 2, 2, 3, 3, 3, 3], np.int32)
    return p, itype, np.eye(3) * L, r, off


def test_wannier_dipoles_match_oracle(oracle_lib):
    from mdff_amd.field import ForceField
    p, itype, A, r, off = _wfc_system()
    ff = ForceField(p, itype, A)
    g = ff.engforce(r, itime=0)
    o = oracle_lib.Oracle(p, itype, A).engforce(r, itime=0)
    compare_engforce(g, o)
    np.testing.assert_allclose(g["mu"][:2], -2.0 * off.sum(axis=1), rtol=1e-13, atol=1e-13)
    # a centre beyond sqrt(rcut_wfc): the reference's STOP, status 8
    r2 = r.copy()
    r2[7] = r[1] + np.array([0.75, 0.0, 0.0])
    from mdff_amd._lib import MdffError
    with pytest.raises(MdffError) as e:
        ff.engforce(r2, itime=1)
    assert e.value.status == 8
    ff.close()


def test_quadrupole_refusals():
    from mdff_amd._lib import MdffError
    from mdff_amd.field import ForceField
    s = _melt(True, True)
    p = s.params
    p.precision = P.PREC_MIXED
    with pytest.raises(MdffError) as e:
        ForceField(p, s.itype, s.A)
    assert e.value.status == 6
    p.precision = P.PREC_FP64
    p.scf_solver = P.SCF_PCG
    with pytest.raises(MdffError) as e:
        ForceField(p, s.itype, s.A)
    assert e.value.status == 6
    p.scf_solver = P.SCF_JACOBI
    p.quad[P.qidx(1, 1, 2)] += 0.1  # no longer symmetric
    with pytest.raises(MdffError) as e:
        ForceField(p, s.itype, s.A)
    assert e.value.status == 1
