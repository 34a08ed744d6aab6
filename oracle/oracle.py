"""ctypes front-end of the CPU FP64 oracle (oracle/mdff_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The library is built by
oracle/Makefile into oracle/_build/liboracle.so.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from mdff_amd.params import MdffEfgOpts, MdffEnergies, MdffMdParams, MdffMdStep, MdffParams

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class _State(ctypes.Structure):
    _fields_ = [("hist", _dp), ("u_pol", ctypes.c_double)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_errfc.restype = ctypes.c_double
        L.orc_errfc.argtypes = [ctypes.c_double]
        L.orc_tt_damp.restype = None
        L.orc_tt_damp.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, _dp, _dp]
        L.orc_kmesh_count.restype = ctypes.c_int
        L.orc_kmesh_count.argtypes = [_ip]
        L.orc_engforce.restype = ctypes.c_int
        L.orc_engforce.argtypes = [ctypes.POINTER(MdffParams), _ip, _dp, _dp, _dp, _dp,
                                   ctypes.c_int, ctypes.POINTER(_State), _dp, _dp, _dp,
                                   _dp, _dp, _dp, _dp, ctypes.POINTER(MdffEnergies)]
        L.orc_multipole_es.restype = ctypes.c_int
        L.orc_multipole_es.argtypes = [ctypes.POINTER(MdffParams), _ip, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, _ip] + [ctypes.c_int] * 5 + [
                                           ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp,
                                           _dp, _dp, _dp]
        L.orc_aspc_order.restype = ctypes.c_int
        L.orc_aspc_order.argtypes = [ctypes.POINTER(MdffParams), ctypes.c_int]
        L.orc_aspc_coeffs.restype = None
        L.orc_aspc_coeffs.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_extrapolate.restype = ctypes.c_int
        L.orc_extrapolate.argtypes = [ctypes.POINTER(MdffParams), ctypes.c_int, _dp, _dp]
        L.orc_aspc_correct.restype = None
        L.orc_aspc_correct.argtypes = [ctypes.POINTER(MdffParams), _ip, ctypes.c_int, _dp, _dp,
                                       _dp]
        L.orc_efgcalc.restype = ctypes.c_int
        L.orc_efgcalc.argtypes = [ctypes.POINTER(MdffParams), ctypes.POINTER(MdffEfgOpts), _ip,
                                  _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp]
        L.orc_md_run.restype = ctypes.c_int
        L.orc_md_run.argtypes = [ctypes.POINTER(MdffParams), _ip, _dp, ctypes.POINTER(MdffMdParams),
                                 ctypes.c_int, ctypes.c_int] + [_dp] * 11 + [
                                     ctypes.POINTER(_State), ctypes.POINTER(MdffMdStep)]
        L.orc_yoshida.restype = ctypes.c_int
        L.orc_yoshida.argtypes = [ctypes.c_int, _dp]
        L.orc_chain_nhcn.restype = ctypes.c_double
        L.orc_chain_nhcn.argtypes = [ctypes.c_double, _dp, _dp, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
        L.orc_chain_nhc2.restype = ctypes.c_double
        L.orc_chain_nhc2.argtypes = [ctypes.c_double, _dp, _dp, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def errfc(x: float) -> float:
    """errfc of src/functions.f90:28-40."""
    return lib().orc_errfc(float(x))


def tt_damp(b: float, c: float, r: float, order: int):
    """(f, f') of TT_damping_functions, src/field.f90:3775-3805."""
    f = ctypes.c_double()
    fd = ctypes.c_double()
    lib().orc_tt_damp(b, c, r, order, ctypes.byref(f), ctypes.byref(fd))
    return f.value, fd.value


def aspc_order(params: MdffParams, itime: int) -> int:
    """Order switch of extrapolate_dipole_aspc, src/field.f90:3918-3927."""
    return lib().orc_aspc_order(ctypes.byref(params), int(itime))


def aspc_coeffs(order: int):
    """(B_ASPC(1..6), w_ASPC) of src/field.f90:3929-3967."""
    b = np.zeros(6)
    w = ctypes.c_double()
    lib().orc_aspc_coeffs(int(order), _p(b), ctypes.byref(w))
    return b, w.value


def extrapolate(params: MdffParams, itime: int, hist):
    """First SCF iterate in MD (extrapolate_dipole_poly / the ASPC
    predictor).  hist: dipia_ind_t as an (order+1, natm, 3) array, updated
    in place; returns mu (natm, 3)."""
    n = params.natm
    h = np.asfortranarray(hist, dtype=np.float64)
    if h.shape != (params.extrapolate_order + 1, n, 3):
        raise ValueError("hist must be (extrapolate_order+1, natm, 3)")
    mu = np.zeros(3 * n)
    rc = lib().orc_extrapolate(ctypes.byref(params), int(itime), _p(h), _p(mu))
    if rc != 0:
        raise RuntimeError(f"oracle extrapolation failed with status {rc}")
    hist[...] = h
    return mu.reshape(3, n).T.copy()


def aspc_correct(params: MdffParams, itype, itime: int, E, mu_pred):
    """ASPC corrector: returns (mu, u_pol) from the field E (natm,3) and the
    predictor mu_pred (natm,3)."""
    n = params.natm
    Ef = np.ascontiguousarray(np.asarray(E, dtype=np.float64).T.reshape(-1))
    mu = np.ascontiguousarray(np.asarray(mu_pred, dtype=np.float64).T.reshape(-1))
    it = np.ascontiguousarray(itype, dtype=np.int32)
    u = ctypes.c_double()
    lib().orc_aspc_correct(ctypes.byref(params), it.ctypes.data_as(_ip), int(itime), _p(Ef),
                           _p(mu), ctypes.byref(u))
    return mu.reshape(3, n).T.copy(), u.value


class Oracle:
    """Stateful oracle mirroring one MDFF process: keeps the induced-dipole
    history dipia_ind_t and the stale u_pol between calls."""

    def __init__(self, params: MdffParams, itype, A):
        self.p = params
        self.itype = np.ascontiguousarray(itype, dtype=np.int32)
        self.A = np.asfortranarray(np.asarray(A, dtype=np.float64).reshape(3, 3))
        n = params.natm
        self.hist = np.zeros((params.extrapolate_order + 1) * n * 3, dtype=np.float64)
        self.state = _State(_p(self.hist), 0.0)

    def engforce(self, r, itime: int = 0, f_in=None):
        """One engforce_driver call.  r is (natm, 3) Cartesian; returns a dict
        with f (natm,3), mu, ef (natm,3), efg (natm,3,3), theta and the
        energies."""
        n = self.p.natm
        r = np.asarray(r, dtype=np.float64)
        rx, ry, rz = (np.ascontiguousarray(r[:, k]) for k in range(3))
        f = np.zeros((3, n)) if f_in is None else np.ascontiguousarray(np.asarray(f_in, dtype=np.float64).T)
        mu = np.zeros(3 * n)
        theta = np.zeros(9 * n)
        ef = np.zeros(3 * n)
        efg = np.zeros(9 * n)
        en = MdffEnergies()
        Aflat = np.ascontiguousarray(self.A.reshape(-1, order="F"))
        rc = lib().orc_engforce(ctypes.byref(self.p), self.itype.ctypes.data_as(_ip), _p(Aflat),
                                _p(rx), _p(ry), _p(rz), int(itime), ctypes.byref(self.state),
                                _p(f[0]), _p(f[1]), _p(f[2]), _p(mu), _p(theta), _p(ef),
                                _p(efg), ctypes.byref(en))
        if rc != 0:
            raise RuntimeError(f"oracle engforce failed with status {rc}")
        out = en.as_dict()
        out["f"] = f.T.copy()
        out["mu"] = mu.reshape(3, n).T.copy()
        out["ef"] = ef.reshape(3, n).T.copy()
        out["efg"] = efg.reshape(3, 3, n).transpose(2, 1, 0).copy()
        out["theta"] = theta.reshape(3, 3, n).transpose(2, 1, 0).copy()
        return out


    def md_run(self, mdp: MdffMdParams, r, v, f, nsteps: int, itime0: int = 1, xi=None,
               vxi=None):
        """nsteps steps of the reference integrator (orc_md_run); r, v, f are
        (natm, 3) with f the forces at r.  Returns (r, v, f, xi, vxi, steps)
        with steps a list of dicts as ForceField.md_run's."""
        n = self.p.natm
        cols = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:, k])
                for a in (r, v, f) for k in range(3)]
        xi = np.zeros(16) if xi is None else np.array(xi, dtype=np.float64)
        vxi = np.zeros(16) if vxi is None else np.array(vxi, dtype=np.float64)
        steps = (MdffMdStep * max(1, int(nsteps)))()
        Aflat = np.ascontiguousarray(self.A.reshape(-1, order="F"))
        rc = lib().orc_md_run(ctypes.byref(self.p), self.itype.ctypes.data_as(_ip), _p(Aflat),
                              ctypes.byref(mdp), int(itime0), int(nsteps),
                              *[_p(c) for c in cols], _p(xi), _p(vxi),
                              ctypes.byref(self.state), steps)
        if rc != 0:
            raise RuntimeError(f"oracle md_run failed with status {rc}")
        out = [np.stack(cols[3 * k:3 * k + 3], axis=1) for k in range(3)]
        return (out[0], out[1], out[2], xi, vxi, [steps[k].as_dict() for k in range(int(nsteps))])


def yoshida(order: int) -> np.ndarray:
    w = np.zeros(9)
    if lib().orc_yoshida(int(order), _p(w)) != 0:
        raise ValueError("Yoshida order must be 1, 3, 5, 7 or 9")
    return w[:order].copy()


def chain_nhcn(kin, vxi, xi, Q, L, temp, dt, nhc_n, yosh=3, mults=2):
    """One chain_nhcn call; vxi and xi (float64 arrays) are updated in place;
    returns the velocity scale factor."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    return lib().orc_chain_nhcn(float(kin), _p(vxi), _p(xi), _p(Q), float(L), float(temp),
                                float(dt), int(nhc_n), int(yosh), int(mults))


def chain_nhc2(kin, vxi, xi, Q, L, temp, dt):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    return lib().orc_chain_nhc2(float(kin), _p(vxi), _p(xi), _p(Q), float(L), float(temp),
                                float(dt))


def multipole_es(params: MdffParams, itype, A, r, q, mu, task, damp_ind=True,
                 do_efield=True, do_efg=True, do_forces=True, do_stress=True, u_pol=0.0,
                 theta=None):
    """One multipole_ES call (src/field.f90:1817-2054) with explicit charges
    q (natm), dipoles mu (natm,3) and quadrupoles theta (natm,3,3) or None;
    task holds task_coul(1:3) or (1:6)."""
    n = params.natm
    r = np.asarray(r, dtype=np.float64)
    rx, ry, rz = (np.ascontiguousarray(r[:, k]) for k in range(3))
    q = np.ascontiguousarray(q, dtype=np.float64)
    muf = np.ascontiguousarray(np.asarray(mu, dtype=np.float64).T.reshape(-1))
    it = np.ascontiguousarray(itype, dtype=np.int32)
    task = list(task) + [0] * (6 - len(task))
    t = (ctypes.c_int32 * 6)(*[int(x) for x in task])
    thf = None if theta is None else np.ascontiguousarray(
        np.asarray(theta, dtype=np.float64).transpose(2, 1, 0).reshape(-1))
    ef = np.zeros(3 * n)
    efg = np.zeros(9 * n)
    f = np.zeros((3, n))
    u = ctypes.c_double()
    pv = ctypes.c_double()
    tau = np.zeros(9)
    parts = np.zeros(4)
    Aflat = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3).reshape(-1, order="F"))
    rc = lib().orc_multipole_es(ctypes.byref(params), it.ctypes.data_as(_ip), _p(Aflat), _p(rx),
                                _p(ry), _p(rz), _p(q), _p(muf),
                                None if thf is None else _p(thf), t, int(damp_ind),
                                int(do_efield), int(do_efg), int(do_forces), int(do_stress),
                                float(u_pol), _p(ef), _p(efg), _p(f[0]), _p(f[1]), _p(f[2]),
                                ctypes.byref(u), _p(tau), ctypes.byref(pv), _p(parts))
    if rc != 0:
        raise RuntimeError(f"oracle multipole_ES failed with status {rc}")
    return {"u_coul": u.value, "pvirial": pv.value, "tau": tau.reshape(3, 3, order="F"),
            "ef": ef.reshape(3, n).T.copy(), "efg": efg.reshape(3, 3, n).transpose(2, 1, 0).copy(),
            "f": f.T.copy(), "u_dir": parts[0], "u_rec": parts[1], "u_self": parts[2],
            "u_surf": parts[3]}


def efgcalc(params: MdffParams, opts: MdffEfgOpts, itype, A, r):
    """One configuration of the standalone EFG run (efgcalc,
    src/efg.f90:325-616).  Returns dict efg (natm,3,3), mu (natm,3),
    n_scf, rmsd."""
    n = params.natm
    r = np.asarray(r, dtype=np.float64)
    rx, ry, rz = (np.ascontiguousarray(r[:, k]) for k in range(3))
    it = np.ascontiguousarray(itype, dtype=np.int32)
    Af = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3).reshape(-1, order="F"))
    efg = np.zeros(9 * n)
    mu = np.zeros(3 * n)
    ns = ctypes.c_int32()
    rm = ctypes.c_double()
    rc = lib().orc_efgcalc(ctypes.byref(params), ctypes.byref(opts), it.ctypes.data_as(_ip),
                           _p(Af), _p(rx), _p(ry), _p(rz), _p(mu), _p(efg), ctypes.byref(ns),
                           ctypes.byref(rm))
    if rc != 0:
        raise RuntimeError(f"oracle efgcalc failed with status {rc}")
    return {"efg": efg.reshape(3, 3, n).transpose(2, 1, 0).copy(),
            "mu": mu.reshape(3, n).T.copy(), "n_scf": ns.value, "rmsd": rm.value}


def nmr_convention(w):
    """nmr_convention (src/efg.f90:2546-2648) on eigenvalues w in the
    ascending order DSYEV returns them: (Vxx, Vyy, Vzz, eta) and whether the
    ordering checks that stop the reference (:2618-2629) pass."""
    w = [float(x) for x in w]
    diso = (w[0] + w[1] + w[2]) / 3.0
    if abs(diso) != 0.0 and abs(diso) > 1e-3:
        diso = 0.0
    a = [abs(x - diso) for x in w]
    if a[0] >= a[1] and a[0] >= a[2]:
        d11 = w[0]
        d22, d33 = (w[1], w[2]) if a[1] >= a[2] else (w[2], w[1])
    elif a[1] >= a[0] and a[1] >= a[2]:
        d11 = w[1]
        d22, d33 = (w[0], w[2]) if a[0] >= a[2] else (w[2], w[0])
    else:
        d11 = w[2]
        d22, d33 = (w[0], w[1]) if a[0] >= a[1] else (w[1], w[0])
    xx, yy, zz = d33, d22, d11
    eta = (xx - yy) / zz if zz != 0.0 else 0.0
    if xx == yy:
        eta = 0.0
    if zz == 0.0:
        eta = 0.0
    if abs(xx) < 1.0e-5 and abs(yy) < 1.0e-5 and abs(zz) < 1.0e-5:
        eta = 0.0
    ok = not (abs(zz - diso) < abs(xx - diso) or abs(zz - diso) < abs(yy - diso)
              or abs(yy - diso) < abs(xx - diso))
    return (xx, yy, zz, eta), ok
