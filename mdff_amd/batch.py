"""Batched evaluation of many configurations and the force-matching adjoint
over the C ABI (mdffgpu_batch_*), with a torch.autograd.Function for the
loss and a small autotuner for the launch shape.

The reference evaluates one configuration per process (engforce_driver,
src/field.f90:1111-1166); the batch shares one model (BMHFTD pair
potential, point charges, isotropic induced dipoles with Tang-Toennies
damping) over configurations of uneven atom counts, each with its own cell,
Ewald parameter, k mesh and cutoffs.

Parameter vector theta (mdffgpu_batch_num_params entries):
    [A, B, C6, C8, BD] for each unordered species pair (1,1), (1,2), (2,2)
    [eps, sigma, rho] (Morse) for each unordered pair, when lmorse
    qch for each species
    alpha for each polarizable species (isotropic poldip)
    b (Tang-Toennies dipole damping, dip_pol_damp_b(t,t,:)) for each species
      whose own field terms are damped, when some species is polarizable
"""
from __future__ import annotations

import copy
import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from .params import MdffBatchConfig, MdffParams

_ERR_UNSUPPORTED = 6  # MDFFGPU_ERR_UNSUPPORTED of include/mdffgpu.h

_SR_NAMES = ("A", "B", "C6", "C8", "BD")
_MO_NAMES = ("eps_mo", "sigma_mo", "rho_mo")


def _pairs(ntype: int):
    return [(ti, tj) for ti in range(ntype) for tj in range(ti, ntype)]


def _damp_entries(params: MdffParams) -> dict:
    """Species whose damping b is a parameter, with the damping-table
    entries their centres use: (t, t, u) when the centre has the smaller
    index and (t, u, t) otherwise (src/field.f90:2207-2211), on damped pairs
    with c != 0; only when some species is polarizable."""
    nt = params.ntype
    if not any(params.ldip_polar[t] for t in range(nt)):
        return {}
    L = params.ldip_damping

    def idx(x, y, z):
        return x + 16 * (y + 16 * z)

    out = {}
    for t in range(nt):
        ent = []
        for u in range(nt):
            for e, act in ((idx(t, t, u), L[idx(t, t, u)] or L[idx(u, t, u)]),
                           (idx(t, u, t), L[idx(u, u, t)] or L[idx(t, u, t)])):
                if act and params.dip_pol_damp_c[e] != 0.0 and e not in ent:
                    ent.append(e)
        if ent:
            out[t] = ent
    return out


def _damp_types(params: MdffParams) -> list:
    return sorted(_damp_entries(params))


def theta_names(params: MdffParams) -> list:
    names = []
    for ti, tj in _pairs(params.ntype):
        names += [f"{n}[{ti + 1}{tj + 1}]" for n in _SR_NAMES]
    if params.lmorse:
        for ti, tj in _pairs(params.ntype):
            names += [f"{n}[{ti + 1}{tj + 1}]" for n in _MO_NAMES]
    names += [f"q[{t + 1}]" for t in range(params.ntype)]
    names += [f"alpha[{t + 1}]" for t in range(params.ntype) if params.ldip_polar[t]]
    names += [f"b_TT[{t + 1}]" for t in _damp_types(params)]
    return names


def get_theta(params: MdffParams) -> np.ndarray:
    """The parameter vector of a model."""
    out = []
    for ti, tj in _pairs(params.ntype):
        k = ti + 16 * tj
        out += [params.Abmhftd[k], params.Bbmhftd[k], params.Cbmhftd[k], params.Dbmhftd[k],
                params.BDbmhftd[k]]
    if params.lmorse:
        for ti, tj in _pairs(params.ntype):
            k = ti + 16 * tj
            out += [params.epsmor[k], params.sigmamor[k], params.rhomor[k]]
    out += [params.qch[t] for t in range(params.ntype)]
    out += [params.poldip[t] for t in range(params.ntype) if params.ldip_polar[t]]
    ent = _damp_entries(params)
    out += [params.dip_pol_damp_b[ent[t][0]] for t in sorted(ent)]
    return np.array(out, dtype=np.float64)


def set_theta(params: MdffParams, theta) -> MdffParams:
    """A copy of params with the parameter vector theta (symmetric pair
    entries and the isotropic polarizability diagonal kept consistent)."""
    p = copy.deepcopy(params)
    th = np.asarray(theta, dtype=np.float64)
    x = 0
    for ti, tj in _pairs(p.ntype):
        for k in (ti + 16 * tj, tj + 16 * ti):
            p.Abmhftd[k], p.Bbmhftd[k], p.Cbmhftd[k], p.Dbmhftd[k], p.BDbmhftd[k] = th[x:x + 5]
        x += 5
    if p.lmorse:
        for ti, tj in _pairs(p.ntype):
            for k in (ti + 16 * tj, tj + 16 * ti):
                p.epsmor[k], p.sigmamor[k], p.rhomor[k] = th[x:x + 3]
            x += 3
    for t in range(p.ntype):
        p.qch[t] = th[x]
        x += 1
    for t in range(p.ntype):
        if p.ldip_polar[t]:
            for a in range(3):
                p.poldip[t + 16 * (a + 3 * a)] = th[x]
            x += 1
    ent = _damp_entries(p)
    for t in sorted(ent):
        for e in ent[t]:
            p.dip_pol_damp_b[e] = th[x]
        x += 1
    return p


class Batch:
    """Configurations evaluated together on one GPU.

    model:   MdffParams of the shared model (precision selects FP64 or the
             FP32 path with FP64 loss/gradient reductions)
    configs: list of dicts with natm, A (3x3, columns = lattice vectors),
             kmax (3), alpha, rc (real-space Ewald cutoff), rsr (short-range
             cutoff)
    itype:   concatenated 1-based species of all atoms
    """

    def __init__(self, model: MdffParams, configs, itype):
        L = _lib.lib()
        n = len(configs)
        arr = (MdffBatchConfig * n)()
        for c, g in enumerate(configs):
            arr[c].natm = int(g["natm"])
            for a in range(3):
                arr[c].kmax[a] = int(g["kmax"][a])
            Af = np.asarray(g["A"], dtype=np.float64).reshape(3, 3).reshape(-1, order="F")
            for k in range(9):
                arr[c].A[k] = float(Af[k])
            arr[c].alphaES = float(g["alpha"])
            arr[c].cutlongrange = float(g["rc"])
            arr[c].cutshortrange = float(g["rsr"])
        self.itype = np.ascontiguousarray(itype, dtype=np.int32)
        self.model = copy.deepcopy(model)
        self.configs = [dict(g) for g in configs]
        self.offsets = np.concatenate([[0], np.cumsum([g["natm"] for g in configs])]).astype(int)
        h = ctypes.c_void_p()
        st = L.mdffgpu_batch_create(ctypes.byref(self.model), n, arr,
                                    self.itype.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    ctypes.byref(h))
        _lib.check_batch(st, None)
        self._h = h
        self.natm = int(L.mdffgpu_batch_num_atoms(h))
        self.nparam = int(L.mdffgpu_batch_num_params(h))
        self.apply_lanes = 4

    @classmethod
    def from_systems(cls, systems, precision=None) -> "Batch":
        """Batch of configs.System sharing the first system's model."""
        model = copy.deepcopy(systems[0].params)
        if precision is not None:
            model.precision = precision
        cfgs = []
        for s in systems:
            p = s.params
            cfgs.append({"natm": s.natm, "A": s.A, "kmax": list(p.kES), "alpha": p.alphaES,
                         "rc": p.cutlongrange, "rsr": p.cutshortrange})
        b = cls(model, cfgs, np.concatenate([s.itype for s in systems]))
        return b

    # -- lifecycle ----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().mdffgpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        _lib.check_batch(st, self._h)

    # -- inputs -------------------------------------------------------------
    def set_positions(self, pos) -> None:
        """pos: torch float64 tensor (3, natm) on the GPU (SoA)."""
        import torch
        if pos.shape != (3, self.natm) or pos.dtype != torch.float64 or not pos.is_cuda:
            raise ValueError("positions must be a (3, natm) float64 CUDA tensor")
        pos = pos.contiguous()
        torch.cuda.synchronize()
        self._check(_lib.lib().mdffgpu_batch_set_positions(self._h, pos.data_ptr(), None))

    def set_model(self, model: MdffParams) -> None:
        self.model = copy.deepcopy(model)
        self._check(_lib.lib().mdffgpu_batch_set_model(self._h, ctypes.byref(self.model)))

    def set_theta(self, theta) -> None:
        self.set_model(set_theta(self.model, theta))

    def theta(self) -> np.ndarray:
        return get_theta(self.model)

    def graph(self, on: bool = True) -> None:
        self._check(_lib.lib().mdffgpu_batch_graph(self._h, 1 if on else 0))

    def set_tuning(self, apply_lanes: int) -> None:
        self._check(_lib.lib().mdffgpu_batch_set_tuning(self._h, int(apply_lanes)))
        self.apply_lanes = int(apply_lanes)

    def set_mixed_tuning(self, apply_slices: int, split_atoms: int) -> bool:
        """Launch shape of the FP32 matrix-core reciprocal sums; False when
        the batch has no such path (FP64, or configurations whose kmax
        differ)."""
        st = _lib.lib().mdffgpu_batch_set_mixed_tuning(self._h, int(apply_slices),
                                                       int(split_atoms))
        if st == _ERR_UNSUPPORTED:
            return False
        self._check(st)
        return True

    # -- evaluation ---------------------------------------------------------
    def forces(self, scf_iters: int = 0):
        """Forces (3, natm) float64 CUDA tensor, and (n_scf, rmsd)."""
        import torch
        f = torch.empty((3, self.natm), dtype=torch.float64, device="cuda")
        ns = ctypes.c_int32(0)
        rm = ctypes.c_double(0.0)
        torch.cuda.synchronize()
        self._check(_lib.lib().mdffgpu_batch_forces(self._h, f.data_ptr(), int(scf_iters),
                                                    ctypes.byref(ns), ctypes.byref(rm), None))
        return f, (ns.value, rm.value)

    def loss_grad(self, target, scf_iters: int = 0, want_forces: bool = False):
        """Loss, gradient (numpy, nparam) and optionally the forces."""
        import torch
        if target.shape != (3, self.natm) or target.dtype != torch.float64:
            raise ValueError("target must be a (3, natm) float64 CUDA tensor")
        target = target.contiguous()
        f = torch.empty((3, self.natm), dtype=torch.float64, device="cuda") if want_forces else None
        loss = ctypes.c_double(0.0)
        grad = np.zeros(self.nparam, dtype=np.float64)
        ns = ctypes.c_int32(0)
        rm = ctypes.c_double(0.0)
        torch.cuda.synchronize()
        self._check(_lib.lib().mdffgpu_batch_loss_grad(
            self._h, target.data_ptr(), f.data_ptr() if f is not None else None,
            ctypes.byref(loss), grad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            int(scf_iters), ctypes.byref(ns), ctypes.byref(rm), None))
        self.last_n_scf = ns.value
        self.last_rmsd = rm.value
        if want_forces:
            return loss.value, grad, f
        return loss.value, grad


def _loss_function():
    import torch

    class ForceMatchLoss(torch.autograd.Function):
        """L(theta) = sum |F(theta) - F_target|^2 / (3 N) over a Batch; the
        gradient comes from the library's adjoint (BMHFTD and Morse pair
        terms recomputed, exact tangents for charges, polarizabilities and
        the dipole part of the damping b)."""

        @staticmethod
        def forward(ctx, theta, batch, target, scf_iters):
            th = theta.detach().to("cpu", torch.float64).numpy()
            batch.set_theta(th)
            loss, grad = batch.loss_grad(target, scf_iters)
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
            return torch.as_tensor(loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None, None, None

    return ForceMatchLoss


_FML = None


def force_match_loss(theta, batch: Batch, target, scf_iters: int = 0):
    """Differentiable force-matching loss of theta (a torch tensor)."""
    global _FML
    if _FML is None:
        _FML = _loss_function()
    return _FML.apply(theta, batch, target, scf_iters)


# ---------------------------------------------------------------------------
# autotuner

def _cache_path() -> str:
    return os.environ.get("MDFF_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mdff_amd", "tune.json"))


def _key(batch: Batch, scf_iters: int) -> str:
    import torch
    kmax = max(max(g["kmax"]) for g in batch.configs)
    return (f"{torch.cuda.get_device_name(0)}|nconf={len(batch.configs)}|natm={batch.natm}|"
            f"kmax={kmax}|prec={batch.model.precision}|scf={scf_iters}")


MIXED_CANDIDATES = tuple({"apply_slices": a, "split_atoms": b}
                         for a in (1, 2, 4) for b in (512, 1024, 2048))
FP64_CANDIDATES = tuple({"apply_lanes": k} for k in (2, 4, 8, 16))


def _apply(batch: Batch, cand: dict) -> None:
    if "apply_lanes" in cand:
        batch.set_tuning(cand["apply_lanes"])
    else:
        batch.set_mixed_tuning(cand["apply_slices"], cand["split_atoms"])


def autotune(batch: Batch, target, scf_iters: int, candidates=None, reps: int = 3,
             use_cache: bool = True, verbose: bool = False) -> dict:
    """Times one loss+gradient evaluation for every launch shape in
    candidates and keeps the fastest: for the FP32 batch the K slices of the
    matrix-core field apply and the atoms per structure-factor split, for
    the FP64 batch the lanes per atom of the reciprocal-space kernel.
    Results are cached per device and batch shape in a JSON file
    (MDFF_TUNE_CACHE)."""
    import torch
    if candidates is None:
        mixed = batch.set_mixed_tuning(2, 2048)
        candidates = MIXED_CANDIDATES if mixed else FP64_CANDIDATES
    key = _key(batch, scf_iters)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache and "best" in cache[key]:
        _apply(batch, cache[key]["best"])
        return cache[key]
    times = []
    for cand in candidates:
        _apply(batch, cand)
        batch.loss_grad(target, scf_iters)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.loss_grad(target, scf_iters)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) / reps)
        if verbose:
            print(f"[autotune] {cand}: {times[-1] * 1e3:.3f} ms")
    best = candidates[int(np.argmin(times))]
    _apply(batch, best)
    rec = {"best": dict(best), "ms": [[dict(c), t * 1e3] for c, t in zip(candidates, times)]}
    if use_cache:
        cache[key] = rec
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            json.dump(cache, open(path, "w"), indent=1)
        except OSError:
            pass
    return rec
