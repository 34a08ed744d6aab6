"""Host front-end of the force/field evaluation: the Python mirror of
MDFF's engforce_driver (src/field.f90:1111-1166) over the C ABI.

``ForceField`` owns one library context (one configuration on one GPU), the
equivalent of the state the reference keeps in the field and config modules
between calls (k mesh, dipole history dipia_ind_t, stale u_pol).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .params import MdffEnergies, MdffMdParams, MdffMdStep, MdffParams

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


_KERNEL_NAMES: list | None = None


def _kernel_index(name: str) -> int:
    """Index of a kernel name in the library's instrumentation table."""
    global _KERNEL_NAMES
    if _KERNEL_NAMES is None:
        L = _lib.lib()
        _KERNEL_NAMES = [L.mdffgpu_kernel_name(i).decode() for i in range(L.mdffgpu_kernel_count())]
    return _KERNEL_NAMES.index(name)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class ForceField:
    """engforce_driver on one MI355X.

    params: MdffParams after field_check_tag / ewald_param
    itype:  (natm,) 1-based species indices
    A:      (3,3) cell, lattice vectors as columns
    """

    def __init__(self, params: MdffParams, itype, A):
        L = _lib.lib()
        self.params = params
        self.n = params.natm
        self.itype = np.ascontiguousarray(itype, dtype=np.int32)
        if self.itype.shape != (self.n,):
            raise ValueError("itype must have natm entries")
        Af = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3).reshape(-1, order="F"))
        h = ctypes.c_void_p()
        st = L.mdffgpu_create(ctypes.byref(params), self.itype.ctypes.data_as(_ip), _ptr(Af),
                              ctypes.byref(h))
        _lib.check(st, None)
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().mdffgpu_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_cell(self, A) -> None:
        Af = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3).reshape(-1, order="F"))
        _lib.check(_lib.lib().mdffgpu_set_cell(self._h, _ptr(Af)), self._h)

    def num_kvectors(self) -> int:
        return int(_lib.lib().mdffgpu_num_kvectors(self._h))

    def engforce(self, r, itime: int = 0, f_in=None) -> dict:
        """One engforce_driver call on host arrays; r is (natm,3) Cartesian.
        Returns f, mu, ef (natm,3), efg, theta (natm,3,3) and the energies."""
        n = self.n
        r = np.asarray(r, dtype=np.float64)
        rx, ry, rz = (np.ascontiguousarray(r[:, k]) for k in range(3))
        f = np.zeros((3, n)) if f_in is None else np.ascontiguousarray(np.asarray(f_in, dtype=np.float64).T)
        mu = np.zeros(3 * n)
        theta = np.zeros(9 * n)
        ef = np.zeros(3 * n)
        efg = np.zeros(9 * n)
        en = MdffEnergies()
        st = _lib.lib().mdffgpu_engforce(self._h, _ptr(rx), _ptr(ry), _ptr(rz), int(itime),
                                         _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), _ptr(mu),
                                         _ptr(theta), _ptr(ef), _ptr(efg), ctypes.byref(en))
        _lib.check(st, self._h)
        out = en.as_dict()
        out["f"] = f.T.copy()
        out["mu"] = mu.reshape(3, n).T.copy()
        out["ef"] = ef.reshape(3, n).T.copy()
        out["efg"] = efg.reshape(3, 3, n).transpose(2, 1, 0).copy()
        out["theta"] = theta.reshape(3, 3, n).transpose(2, 1, 0).copy()
        return out

    def efg(self, r, opts) -> dict:
        """Standalone EFG of one configuration (efgcalc's per-frame body,
        src/efg.f90:404-488) on host positions r (natm,3); opts is an
        MdffEfgOpts (params.efg_opts).  Returns efg (natm,3,3), mu (natm,3)
        and the SCF statistics."""
        n = self.n
        r = np.asarray(r, dtype=np.float64)
        rx, ry, rz = (np.ascontiguousarray(r[:, k]) for k in range(3))
        mu = np.zeros(3 * n)
        efg = np.zeros(9 * n)
        en = MdffEnergies()
        st = _lib.lib().mdffgpu_efg(self._h, ctypes.byref(opts), _ptr(rx), _ptr(ry), _ptr(rz),
                                    _ptr(mu), _ptr(efg), ctypes.byref(en))
        _lib.check(st, self._h)
        return {"efg": efg.reshape(3, 3, n).transpose(2, 1, 0).copy(),
                "mu": mu.reshape(3, n).T.copy(), "n_scf": en.n_scf, "rmsd": en.rmsd,
                "scf_tol": en.scf_tol}

    def engforce_device(self, pos, force, mu=None, ef=None, efg=None, itime: int = 0,
                        stream=None) -> MdffEnergies:
        """Evaluation on device-resident torch tensors (float64, contiguous):
        pos and force are (3, natm) SoA, mu / ef (3, natm), efg (3, 3, natm)
        in Fortran (natm,3,3) order, i.e. efg[b, a, i].  force is in/out."""
        for t in (pos, force):
            if t.dtype.itemsize != 8 or not t.is_contiguous() or t.numel() != 3 * self.n:
                raise ValueError("pos and force must be contiguous float64 (3, natm) tensors")
        en = MdffEnergies()
        sp = None if stream is None else ctypes.c_void_p(stream)
        st = _lib.lib().mdffgpu_engforce_device(
            self._h, ctypes.c_void_p(pos.data_ptr()), int(itime), ctypes.c_void_p(force.data_ptr()),
            None if mu is None else ctypes.c_void_p(mu.data_ptr()),
            None if ef is None else ctypes.c_void_p(ef.data_ptr()),
            None if efg is None else ctypes.c_void_p(efg.data_ptr()), ctypes.byref(en), sp)
        _lib.check(st, self._h)
        return en

    def md_run(self, mdp: MdffMdParams, pos, vel, force, nsteps: int, itime0: int = 1,
               xi=None, vxi=None, stream=None) -> list:
        """nsteps MD steps on device-resident torch tensors (float64,
        contiguous (3, natm) SoA): pos, vel and force are in/out, force
        holding the forces at pos on entry.  xi / vxi: numpy arrays of the
        thermostat coordinates (nhc_n entries), updated in place.  Returns
        one dict per step (energies, e_kin, temp_r, e_nvt)."""
        for t in (pos, vel, force):
            if t.dtype.itemsize != 8 or not t.is_contiguous() or t.numel() != 3 * self.n:
                raise ValueError("pos, vel and force must be contiguous float64 (3, natm) tensors")
        steps = (MdffMdStep * max(1, int(nsteps)))()
        nvt = mdp.integrator != 0
        if nvt:
            for a in (xi, vxi):
                if a is None or a.dtype != np.float64 or a.size < mdp.nhc_n:
                    raise ValueError("a thermostat needs float64 xi and vxi of nhc_n entries")
        sp = None if stream is None else ctypes.c_void_p(stream)
        st = _lib.lib().mdffgpu_md_run(
            self._h, ctypes.byref(mdp), int(itime0), int(nsteps), ctypes.c_void_p(pos.data_ptr()),
            ctypes.c_void_p(vel.data_ptr()), ctypes.c_void_p(force.data_ptr()),
            _ptr(xi) if nvt else None, _ptr(vxi) if nvt else None, steps, sp)
        _lib.check(st, self._h)
        return [steps[k].as_dict() for k in range(int(nsteps))]

    def list_info(self) -> dict:
        """Neighbour-list builds so far and whether the last one searched
        the cell grid."""
        b = ctypes.c_int32()
        cs = ctypes.c_int32()
        _lib.check(_lib.lib().mdffgpu_list_info(self._h, ctypes.byref(b), ctypes.byref(cs)), self._h)
        return {"builds": b.value, "cell_search": bool(cs.value)}

    def profile(self, on: bool = True, only: str = None) -> None:
        """Enable (and reset) the per-kernel HIP-event instrumentation; with
        ``only``, bracket just the launches of that kernel."""
        L = _lib.lib()
        _lib.check(L.mdffgpu_profile_enable(self._h, int(on)), self._h)
        if on and only is not None:
            _lib.check(L.mdffgpu_profile_only(self._h, _kernel_index(only)), self._h)

    def profile_select(self, only: str = None) -> None:
        """While profiling is enabled, bracket only the launches of kernel
        ``only`` from now on, or none (None), keeping the totals."""
        k = -2 if only is None else _kernel_index(only)
        _lib.check(_lib.lib().mdffgpu_profile_only(self._h, k), self._h)

    def set_mixed_tuning(self, apply_slices: int, split_atoms: int) -> bool:
        """Launch shape of the mixed path's matrix-core reciprocal sums;
        False on an FP64 context."""
        st = _lib.lib().mdffgpu_set_mixed_tuning(self._h, int(apply_slices), int(split_atoms))
        if st == 6:  # MDFFGPU_ERR_UNSUPPORTED
            return False
        _lib.check(st, self._h)
        return True

    def autotune(self, pos, force, itime: int = 1, reps: int = 2,
                 candidates=None) -> dict:
        """Times engforce_device on one frame for every mixed-path launch
        shape (apply K slices x structure-factor split size) and keeps the
        fastest.  Each timed call is an evaluation of that frame (dipole
        history included), so tune a context that is not mid-trajectory."""
        import torch
        if candidates is None:
            candidates = [(a, b) for a in (1, 2, 4) for b in (512, 1024, 2048)]
        if not self.set_mixed_tuning(*candidates[0]):
            return {}
        times = []
        for cand in candidates:
            self.set_mixed_tuning(*cand)
            self.engforce_device(pos, force, itime=itime)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self.engforce_device(pos, force, itime=itime)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) / reps)
        best = candidates[int(np.argmin(times))]
        self.set_mixed_tuning(*best)
        return {"apply_slices": best[0], "split_atoms": best[1],
                "ms": [[list(c), t * 1e3] for c, t in zip(candidates, times)]}

    def profile_stats(self) -> dict:
        """{kernel name: (total ms, launches)} of the instrumented launches."""
        L = _lib.lib()
        out = {}
        for k in range(L.mdffgpu_kernel_count()):
            ms = ctypes.c_double()
            cnt = ctypes.c_int64()
            _lib.check(L.mdffgpu_profile_read(self._h, k, ctypes.byref(ms), ctypes.byref(cnt)), self._h)
            if cnt.value:
                out[L.mdffgpu_kernel_name(k).decode()] = (ms.value, cnt.value)
        return out

    def dipole_history(self) -> np.ndarray:
        p = self.params
        h = np.zeros((p.extrapolate_order + 1) * self.n * 3)
        _lib.check(_lib.lib().mdffgpu_get_dipole_history(self._h, _ptr(h)), self._h)
        return h

    def set_dipole_history(self, h) -> None:
        h = np.ascontiguousarray(h, dtype=np.float64)
        _lib.check(_lib.lib().mdffgpu_set_dipole_history(self._h, _ptr(h)), self._h)


def device_count() -> int:
    return int(_lib.lib().mdffgpu_device_count())
