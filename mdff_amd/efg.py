"""Standalone electric-field-gradient run (calc = 'efg'): the host side of
efgcalc (src/efg.f90:325-616), efg_stat (:2057-2525) and efg_write_output
(:1632-1741) over the library's mdffgpu_efg / mdffgpu_efg_nmr.

    opts = params.efg_opts("ewald")
    stat = EfgStatOptions(umin=-15, smax=25, vzzmin=-50)
    run = efgcalc(field_params, opts, "TRAJFF", ncefg=10, stat=stat, workdir=".")

reads ncefg frames from TRAJFF, evaluates every site's EFG on the GPU (one
library context, its cell updated per frame), writes EFGALL, and with
stat.lefg_stat reads EFGALL back to write NMRFF and the distribution files
DTETAFF, DTVZZFF, DTIBUFF and DTIBSFF, as the reference does.
"""
from __future__ import annotations

import copy
import ctypes
import dataclasses
import os
from typing import Callable, Optional

import numpy as np

from . import _lib, io
from .params import MdffEfgOpts, MdffParams

# cq_unit = e / h / 1000 with eh = evtoj / 1e14 / hplanck
# (src/constants.f90:78-81)
_EVTOJ = 1.602176487e-19
_HPLANCK = 6.62617636e-34
CQ_UNIT = _EVTOJ / 1e14 / _HPLANCK / 1000.0

_dp = ctypes.POINTER(ctypes.c_double)


def efg_nmr(efg6: np.ndarray):
    """Per-site NMR parameters on the GPU (mdffgpu_efg_nmr): efg6 (n,6) as
    vxx vyy vzz vxy vxz vyz.  Returns nmr (n,4) = Vxx, Vyy, Vzz, eta in the
    NMR order, U (n,5), S (n) and the number of sites failing the
    reference's ordering checks."""
    e = np.asfortranarray(np.asarray(efg6, dtype=np.float64).reshape(-1, 6))
    n = e.shape[0]
    nmr = np.zeros((n, 4), order="F")
    U = np.zeros((n, 5), order="F")
    S = np.zeros(n)
    bad = ctypes.c_int32()
    st = _lib.lib().mdffgpu_efg_nmr(n, e.ctypes.data_as(_dp), nmr.ctypes.data_as(_dp),
                                    U.ctypes.data_as(_dp), S.ctypes.data_as(_dp),
                                    ctypes.byref(bad))
    _lib.check(st, None)
    return np.ascontiguousarray(nmr), np.ascontiguousarray(U), S, int(bad.value)


def context_params(params: MdffParams, opts: MdffEfgOpts, A) -> MdffParams:
    """Parameters of the library context for a standalone EFG run.  The
    direct sum (longrange 'direct') uses none of the context's Ewald
    settings, and the reference leaves cutlongrange unclamped there
    (ewald_param runs for 'ewald' only, src/field.f90:532): the context gets
    placeholder Ewald values and a cutoff within half the cell, while the
    direct sum takes its cutoff from opts.cutdirect."""
    if opts.longrange == 0:
        return params
    p = copy.deepcopy(params)
    A = np.asarray(A, dtype=np.float64).reshape(3, 3)
    V = abs(np.linalg.det(A))
    widths = [V / np.linalg.norm(np.cross(A[:, (k + 1) % 3], A[:, (k + 2) % 3])) for k in range(3)]
    p.cutlongrange = min(p.cutlongrange, 0.5 * min(widths)) if p.cutlongrange > 0 else 0.5 * min(widths)
    p.kES[:] = [max(1, k) for k in p.kES]
    if not p.alphaES > 0.0:
        p.alphaES = 1.0
    return p


@dataclasses.dataclass
class EfgStatOptions:
    """&efgtag statistics options (efg_default_tag, src/efg.f90:171-195)."""
    lefg_stat: bool = True
    reseta: float = 0.1
    resvzz: float = 0.1
    resu: float = 0.1
    umin: float = 0.0
    smax: float = 0.0
    vzzmin: float = 0.0

    def bins(self):
        """PANeta, PANvzz, PANU, PANS of efg_check_tag (src/efg.f90:223-226)."""
        return (int(1.0 / self.reseta), int(2.0 * abs(self.vzzmin) / self.resvzz),
                int(2.0 * abs(self.umin) / self.resu), int(self.smax / self.resu))


class EfgStatError(RuntimeError):
    """A value outside its distribution, or a site whose principal values
    fail the NMR ordering checks: the reference stops there
    (src/efg.f90:2359-2362, 2392-2396, 2402-2420, 2618-2629)."""


@dataclasses.dataclass
class EfgStats:
    nconf: int
    natmi: list                 # per species, index 0 = all
    atypei: list                # index 0 = 'ALL'
    per_config: list            # per frame: per species dict of min/max/means
    m_U: np.ndarray             # (ntype+1, 5)
    m2_U: np.ndarray            # (ntype+1, 5, 5), upper triangle
    dibU: np.ndarray            # (6, ntype+1, PANU+1)
    dibS: np.ndarray            # (ntype+1, PANS+1)
    dibvzz: np.ndarray          # (ntype+1, PANvzz+1)
    dibeta: np.ndarray          # (ntype+1, PANeta+1)
    nmr_last: np.ndarray        # (natm, 4) of the last frame


def _ftrunc(x: float) -> int:
    """Fortran INT: truncation toward zero."""
    return int(np.trunc(x))


def efg_stat(efgall: str, nconf: int, itype: np.ndarray, opts: EfgStatOptions,
             formatted: bool = True, quad_efg=None, nmrff: Optional[str] = None,
             lwfc=None, nmr_fn: Optional[Callable] = None) -> EfgStats:
    """efg_stat (src/efg.f90:2057-2525) over the first nconf frames of an
    EFGALL file.  itype: 1-based species of every site; quad_efg: nuclear
    quadrupole moment per species (Cq column of NMRFF); nmr_fn (efg6 ->
    nmr, U, S, bad) defaults to the GPU's efg_nmr."""
    nmr_fn = nmr_fn or efg_nmr
    itype = np.asarray(itype, dtype=np.int64)
    PANeta, PANvzz, PANU, PANS = opts.bins()
    frames = io.read_efgall(efgall, formatted)
    out_lines = []
    first = None
    per_config = []
    dibU = dibS = dibvzz = dibeta = m_U = m2_U = None
    for ic in range(nconf):
        try:
            fr = next(frames)
        except StopIteration:
            raise EfgStatError(f"EFGALL holds {ic} frames, {nconf} requested") from None
        natm, ntype = fr.natm, len(fr.atypei)
        if first is None:
            first = fr
            if itype.shape != (natm,):
                raise ValueError("itype must have one entry per EFGALL site")
            nt1 = ntype + 1
            dibU = np.zeros((6, nt1, PANU + 1), dtype=np.int64)
            dibS = np.zeros((nt1, PANS + 1), dtype=np.int64)
            dibvzz = np.zeros((nt1, PANvzz + 1), dtype=np.int64)
            dibeta = np.zeros((nt1, PANeta + 1), dtype=np.int64)
            m_U = np.zeros((nt1, 5))
            m2_U = np.zeros((nt1, 5, 5))
            # the per-species sums are zeroed once, before the frame loop, and
            # each frame divides them in place (src/efg.f90:2152-2163,
            # 2300-2318): a later frame's figures carry the earlier ones
            acc = {k: np.zeros(nt1) for k in ("vzzm", "vzzma", "etam", "vzzsq", "pvzz",
                                              "vzzmini", "vzzmaxi")}
        nmr, U, S, bad = nmr_fn(fr.efg6)
        if bad:
            raise EfgStatError(f"frame {ic + 1}: {bad} sites fail the NMR ordering checks")
        out_lines.append("#     site           xx          yy          zz           Cq(MHz)       eta")
        natmi = np.array([natm] + list(fr.natmi), dtype=np.float64)
        stats = acc
        for ia in range(natm):
            it = int(itype[ia])
            for t in (0, it):
                m_U[t] += U[ia]
                for u1 in range(5):
                    for u2 in range(u1, 5):
                        m2_U[t, u1, u2] += U[ia, u1] * U[ia, u2]
            vzz, eta = nmr[ia, 2], nmr[ia, 3]
            for t in (0, it):
                stats["vzzm"][t] += vzz
                stats["etam"][t] += eta
                stats["vzzma"][t] += abs(vzz)
                if vzz <= stats["vzzmini"][t]:
                    stats["vzzmini"][t] = vzz
                if vzz >= stats["vzzmaxi"][t]:
                    stats["vzzmaxi"][t] = vzz
                if vzz >= 0.0:
                    stats["pvzz"][t] += 1.0
                stats["vzzsq"][t] += vzz * vzz
            if lwfc is None or lwfc[it - 1] >= 0:
                q = 0.0 if quad_efg is None else float(quad_efg[it - 1])
                out_lines.append(f"{ia + 1:7d} {fr.atypei[it - 1][:3]:<3s}"
                                 + "".join(io.fmt_f(x, 12, 4) for x in nmr[ia, :3])
                                 + io.fmt_f(vzz * CQ_UNIT * q, 18, 4) + io.fmt_f(eta, 18, 4))
        for k in ("vzzm", "vzzsq", "vzzma", "etam", "pvzz"):
            stats[k] /= natmi
        with np.errstate(divide="ignore", invalid="ignore"):
            sigma = np.sqrt(stats["vzzsq"] - stats["vzzma"] * stats["vzzma"])
            rho_z = sigma / stats["vzzma"]
        snap = {k: v.copy() for k, v in stats.items()}
        snap["rho_z"] = rho_z
        per_config.append(snap)
        # distributions (src/efg.f90:2345-2433)
        for ia in range(natm):
            it = int(itype[ia])
            if lwfc is not None and lwfc[it - 1] == -1:
                continue
            for ui in range(5):
                ku = _ftrunc((U[ia, ui] - opts.umin) / opts.resu) + 1
                if ku < 0 or ku > PANU:
                    raise EfgStatError(f"site {ia + 1}: U{ui + 1} = {U[ia, ui]} outside the U "
                                       "distribution (umin)")
                dibU[ui, 0, ku] += 1
                dibU[ui, it, ku] += 1
                if ui != 0:
                    dibU[5, 0, ku] += 1
                    dibU[5, it, ku] += 1
            ks = _ftrunc(S[ia] / opts.resu) + 1
            if ks < 0 or ks > PANS:
                raise EfgStatError(f"site {ia + 1}: S = {S[ia]} outside the S distribution (smax)")
            dibS[0, ks] += 1
            dibS[it, ks] += 1
            kvzz = _ftrunc((nmr[ia, 2] - opts.vzzmin) / opts.resvzz) + 1
            keta = _ftrunc(nmr[ia, 3] / opts.reseta)
            if kvzz < 0 or kvzz > PANvzz:
                raise EfgStatError(f"site {ia + 1}: Vzz = {nmr[ia, 2]} outside the Vzz "
                                   "distribution (vzzmin)")
            if keta < 0 or keta > PANeta:
                raise EfgStatError(f"site {ia + 1}: eta = {nmr[ia, 3]} outside [0, 1]")
            dibvzz[0, kvzz] += 1
            dibeta[0, keta] += 1
            dibvzz[it, kvzz] += 1
            dibeta[it, keta] += 1
        nmr_last = nmr
    natm, ntype = first.natm, len(first.atypei)
    m_U[0] /= float(nconf) * float(natm)
    for it in range(1, ntype + 1):
        m_U[it] /= float(nconf) * float(first.natmi[it - 1])
        m2_U[it] /= float(nconf) * float(first.natmi[it - 1])
    if nmrff:
        with open(nmrff, "w") as fh:
            fh.write("\n".join(out_lines) + "\n")
    return EfgStats(nconf, [natm] + list(first.natmi), ["ALL"] + list(first.atypei), per_config,
                    m_U, m2_U, dibU, dibS, dibvzz, dibeta, nmr_last)


def efg_write_output(st: EfgStats, opts: EfgStatOptions, workdir: str = ".",
                     lwfc=None) -> None:
    """DTETAFF, DTVZZFF, DTIBUFF, DTIBSFF (efg_write_output,
    src/efg.f90:1632-1741): normalised distributions per species (column 0
    all non-Wannier sites)."""
    PANeta, PANvzz, PANU, PANS = opts.bins()
    ntype = len(st.natmi) - 1
    natmi = np.array(st.natmi, dtype=np.float64)
    natmi[0] = sum(st.natmi[it] for it in range(1, ntype + 1)
                   if lwfc is None or lwfc[it - 1] != -1)
    nc = float(st.nconf)

    def row(vals):
        return "".join(io.fmt_f(v, 15, 8) for v in vals)

    with np.errstate(divide="ignore", invalid="ignore"):
        eta = ["#" + row([opts.reseta] + list(natmi)), row([0.0] * (ntype + 2))]
        for i in range(PANeta):
            eta.append(row([(i + 1 - 0.5) * opts.reseta] +
                           [st.dibeta[it, i] / (opts.reseta * natmi[it] * nc) for it in range(ntype + 1)]))
        vzz = [row([opts.vzzmin + i * opts.resvzz] +
                   [st.dibvzz[it, i] / (opts.resvzz * natmi[it] * nc) for it in range(ntype + 1)])
               for i in range(PANvzz + 1)]
        uu = []
        for i in range(PANU + 1):
            vals = [opts.umin + i * opts.resu]
            for it in range(ntype + 1):
                den = opts.resu * natmi[it] * nc
                vals += [st.dibU[k, it, i] / den for k in range(5)]
                vals.append(st.dibU[5, it, i] / (opts.resu * 4.0 * natmi[it] * nc))
            uu.append(row(vals))
        ss = [row([i * opts.resu] + [st.dibS[it, i] / (opts.resu * natmi[it] * nc)
                                     for it in range(ntype + 1)]) for i in range(PANS + 1)]
    for name, lines in (("DTETAFF", eta), ("DTVZZFF", vzz), ("DTIBUFF", uu), ("DTIBSFF", ss)):
        with open(os.path.join(workdir, name), "w") as fh:
            fh.write("\n".join(lines) + "\n\n\n")


@dataclasses.dataclass
class EfgRun:
    efg: list           # per frame (natm,3,3)
    mu: list            # per frame (natm,3)
    n_scf: list
    stats: Optional[EfgStats]


def efgcalc(params: MdffParams, opts: MdffEfgOpts, trajff: str, ncefg: int,
            data: str = "rnn", traj_formatted: bool = True, efgall_formatted: bool = True,
            lefgprintall: bool = True, stat: Optional[EfgStatOptions] = None,
            workdir: str = ".", quad_efg=None, forcefield=None) -> EfgRun:
    """efgcalc (src/efg.f90:325-616) over the first ncefg frames of trajff.
    params describes the field (charges, polarizabilities, Ewald alpha and
    k mesh, SCF controls); its natm is taken from the trajectory.  The EFG
    of each frame is written to workdir/EFGALL when lefgprintall; with
    stat.lefg_stat the statistics are taken from that file and NMRFF and the
    distribution files are written (the reference's order)."""
    from .field import ForceField
    if ncefg <= 0:
        raise ValueError("ncefg is zero but calc = 'efg' requested (src/efg.f90:228-231)")
    frames = io.read_trajff(trajff, data, traj_formatted)
    ff = forcefield
    efgs, mus, nscf = [], [], []
    itype = None
    A_cur = None  # cell the context holds (unknown for a context passed in)
    path = os.path.join(workdir, "EFGALL")
    mode = "w" if efgall_formatted else "wb"
    with open(path, mode) as fh:
        for k in range(ncefg):
            try:
                fr = next(frames)
            except StopIteration:
                raise ValueError(f"TRAJFF holds {k} frames, ncefg = {ncefg}") from None
            if itype is None:
                itype = fr.itype()
            if ff is None:
                p = copy.deepcopy(context_params(params, opts, fr.A))
                p.natm = fr.natm
                ff = ForceField(p, itype, fr.A)
            elif A_cur is None or not np.array_equal(fr.A, A_cur):
                ff.set_cell(fr.A)
            A_cur = fr.A.copy()
            out = ff.efg(fr.r, opts)
            efgs.append(out["efg"])
            mus.append(out["mu"])
            nscf.append(out["n_scf"])
            if lefgprintall:
                io.write_efgall(fh, fr, out["efg"], efgall_formatted, efgcalc=True)
    if forcefield is None and ff is not None:
        ff.close()
    stats = None
    if stat is not None and stat.lefg_stat:
        stats = efg_stat(path, ncefg, itype, stat, efgall_formatted, quad_efg,
                         nmrff=os.path.join(workdir, "NMRFF"))
        efg_write_output(stats, stat, workdir)
    return EfgRun(efgs, mus, nscf, stats)
