"""The allocation-free entry point mdffgpu_engforce_noalloc against the
existing device entry point mdffgpu_engforce_device, through the C ABI:
same outputs (FP64, 1e-10 relative, in practice bitwise) on two calls in a
row, the second with a smaller input; a trajectory with the dipole history
carried by the caller; refusals of what it does not run."""
import ctypes

import numpy as np
import pytest

from helpers import assert_close

from mdff_amd import configs as C
from mdff_amd import params as P

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _noalloc(s, pos, force, mu, ef, efg, itime=1, hist=None):
    from mdff_amd import _lib
    L = _lib.lib()
    en = P.MdffEnergies()
    Af = np.ascontiguousarray(np.asarray(s.A, dtype=np.float64).reshape(-1, order="F"))
    it = np.ascontiguousarray(s.itype, dtype=np.int32)
    ptr = (lambda t: None if t is None else ctypes.c_void_p(t.data_ptr()))
    st = L.mdffgpu_engforce_noalloc(ctypes.byref(s.params), it.ctypes.data_as(_ip),
                                    Af.ctypes.data_as(_dp), ptr(pos), int(itime), ptr(hist),
                                    ptr(force), ptr(mu), ptr(ef), ptr(efg), ctypes.byref(en),
                                    None)
    return st, en, L.mdffgpu_last_error(None).decode()


def _outputs(n):
    import torch
    return [torch.zeros((3, n), dtype=torch.float64, device="cuda") for _ in range(3)] + \
           [torch.zeros((3, 3, n), dtype=torch.float64, device="cuda")]


def _compare(en_a, outs_a, en_b, outs_b):
    for name, a, b in zip(("forces", "mu", "ef", "efg"), outs_a, outs_b):
        assert_close(a.cpu().numpy(), b.cpu().numpy(), 1e-10, name)
    da, db = en_a.as_dict(), en_b.as_dict()
    for k in ("u_bmhft", "vir_bmhft", "u_coul", "u_pol", "pvirial_coul"):
        assert abs(da[k] - db[k]) <= 1e-10 * max(abs(db[k]), 1e-300), k
    # the SCF residual is a small difference of O(1) fields: its last bits
    # follow the summation order of the block partials
    assert abs(da["rmsd"] - db["rmsd"]) <= 1e-6 * db["rmsd"], (da["rmsd"], db["rmsd"])
    assert_close(da["tau_coul"], db["tau_coul"], 1e-10, "tau_coul")
    assert_close(da["tau_nonb"], db["tau_nonb"], 1e-10, "tau_nonb")
    assert da["n_scf"] == db["n_scf"] and da["n_es_calls"] == db["n_es_calls"]


def test_noalloc_matches_device_entry_twice_smaller_second():
    import torch
    from mdff_amd.field import ForceField
    systems = [C.silica(60, 120, seed=81, rc=6.0, kmax=6, pim=True, efg=True),
               C.silica(30, 60, seed=82, rc=5.0, kmax=4, pim=True, efg=True)]
    for s in systems:
        print(f"[noalloc] {s.name}: {s.natm} atoms", flush=True)
        pos = _dev(s.r.T)
        f0 = np.random.default_rng(1).standard_normal((3, s.natm))
        ref_outs = _outputs(s.natm)
        ref_outs[0].copy_(_dev(f0))
        ff = ForceField(s.params, s.itype, s.A)
        en_ref = ff.engforce_device(pos, ref_outs[0], mu=ref_outs[1], ef=ref_outs[2],
                                    efg=ref_outs[3], itime=1)
        torch.cuda.synchronize()
        ff.close()
        outs = _outputs(s.natm)
        outs[0].copy_(_dev(f0))
        st, en, msg = _noalloc(s, pos, *outs, itime=1)
        assert st == 0, msg
        torch.cuda.synchronize()
        _compare(en, outs, en_ref, ref_outs)


def test_noalloc_trajectory_with_caller_history():
    """Four MD steps with ASPC extrapolation: the history moves through the
    caller's buffer and every step equals a persistent context's."""
    import torch
    from mdff_amd.field import ForceField
    s = C.silica(40, 80, seed=83, rc=5.5, kmax=5, pim=True, efg=True)
    s.params.algo_ext_dipole = P.EXT_ASPC
    s.params.extrapolate_order = 2
    ff = ForceField(s.params, s.itype, s.A)
    hist = torch.zeros((s.params.extrapolate_order + 1) * 3 * s.natm, dtype=torch.float64,
                       device="cuda")
    rng = np.random.default_rng(5)
    r = s.r.copy()
    for it in range(1, 5):
        pos = _dev(r.T)
        ref_outs = _outputs(s.natm)
        en_ref = ff.engforce_device(pos, ref_outs[0], mu=ref_outs[1], ef=ref_outs[2],
                                    efg=ref_outs[3], itime=it)
        outs = _outputs(s.natm)
        st, en, msg = _noalloc(s, pos, *outs, itime=it, hist=hist)
        assert st == 0, msg
        torch.cuda.synchronize()
        _compare(en, outs, en_ref, ref_outs)
        assert_close(hist.cpu().numpy(), ff.dipole_history(), 1e-10, "history")
        r = r + 0.003 * rng.standard_normal(r.shape)
    ff.close()


def test_noalloc_refusals():
    """Mixed precision and the conjugate-gradient SCF are refused; so is a
    workload whose buffers exceed the static scratch (a k mesh of kmax 100),
    with a message naming the limit; the next call still works."""
    import torch
    s = C.silica(10, 20, seed=84, rc=4.0, kmax=3, pim=True)
    pos = _dev(s.r.T)
    outs = _outputs(s.natm)
    s.params.precision = P.PREC_MIXED
    assert _noalloc(s, pos, *outs)[0] == 6
    s.params.precision = P.PREC_FP64
    s.params.scf_solver = P.SCF_PCG
    assert _noalloc(s, pos, *outs)[0] == 6
    s.params.scf_solver = P.SCF_JACOBI
    s.params.kES[:] = [100, 100, 100]
    st, _, msg = _noalloc(s, pos, *outs)
    assert st == 3 and "MiB of static scratch" in msg, (st, msg)
    s.params.kES[:] = [3, 3, 3]
    st, en, msg = _noalloc(s, pos, *outs)
    assert st == 0, msg
    torch.cuda.synchronize()
    assert en.n_scf > 0
