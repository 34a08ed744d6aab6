"""Force-field parameters on the C-ABI boundary, and the host-side set-up the
reference performs before ``engforce_driver`` runs.

``MdffParams`` mirrors ``struct mdffgpu_params`` of ``include/mdffgpu.h``
field by field.  The helpers restate the reference's parameter processing:

* ``field_defaults``  -- field_default_tag / control_default_tag
  (src/field.f90:190-273, src/control.f90:199-241)
* ``field_check_tag`` -- symmetrisation of the pair tables, isotropic
  polarizabilities, Tang-Toennies order check and task_coul
  (src/field.f90:282-433)
* ``lattice``         -- reciprocal basis, volume, widths
  (src/lattice.f90:60-168)
* ``ewald_param``     -- clamp of cutlongrange and the dl_poly-style alpha /
  kmax used with lautoES (src/field.f90:860-908)

Index conventions are Fortran's, 0-based: a pair table entry (it, jt) sits at
``it + 16*jt``; damping tables (a, b, c) at ``a + 16*(b + 16*c)``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

NTYPEMAX = 16
TT_MAX_ORDER = 8

PREC_FP64 = 0
PREC_MIXED = 1
# mixed_fp64_parts bits (parts of the mixed path evaluated in FP64)
MIXED_FP64_SHORTRANGE = 1
MIXED_FP64_DIRECT = 2
MIXED_FP64_SCF_FIELD = 4
MIXED_FP64_KSPACE = 8
EXT_POLY = 0
EXT_ASPC = 1
# SCF solver of the induced dipoles (mdffgpu_scf_solver): the reference's
# Jacobi iteration, or the preconditioned conjugate gradient of this library
SCF_JACOBI = 0
SCF_PCG = 1
CALC_OTHER = 0
CALC_MD = 1
CALC_OPT = 2

COUL_UNIT = 14.3996441494161  # src/constants.f90:89
PRESS_UNIT = 0.00624150964712042  # src/constants.f90:85
BOLTZ_UNIT = 8.61734229648141e-05  # src/constants.f90:86
TIME_UNIT = 98.2269514139276  # src/constants.f90:87

_D256 = ctypes.c_double * 256
_D4096 = ctypes.c_double * 4096
_I4096 = ctypes.c_int32 * 4096


class MdffParams(ctypes.Structure):
    """``struct mdffgpu_params`` (include/mdffgpu.h)."""

    _fields_ = [
        ("natm", ctypes.c_int32),
        ("ntype", ctypes.c_int32),
        ("lnmlj", ctypes.c_int32),
        ("lmorse", ctypes.c_int32),
        ("lbmhft", ctypes.c_int32),
        ("lbmhftd", ctypes.c_int32),
        ("lcoulomb", ctypes.c_int32),
        ("lsurf", ctypes.c_int32),
        ("calc_mode", ctypes.c_int32),
        ("trunc", ctypes.c_int32),
        ("cutshortrange", ctypes.c_double),
        ("cutlongrange", ctypes.c_double),
        ("skindiff", ctypes.c_double),
        ("coul_unit", ctypes.c_double),
        ("press_unit", ctypes.c_double),
        ("plj", _D256),
        ("qlj", _D256),
        ("epslj", _D256),
        ("sigmalj", _D256),
        ("epsmor", _D256),
        ("sigmamor", _D256),
        ("rhomor", _D256),
        ("Abmhftd", _D256),
        ("Bbmhftd", _D256),
        ("Cbmhftd", _D256),
        ("Dbmhftd", _D256),
        ("BDbmhftd", _D256),
        ("qch", ctypes.c_double * 16),
        ("dip", ctypes.c_double * 48),
        ("poldip", ctypes.c_double * 144),
        ("ldip_polar", ctypes.c_int32 * 16),
        ("ldip_damping", _I4096),
        ("dip_pol_damp_b", _D4096),
        ("dip_pol_damp_c", _D4096),
        ("dip_pol_damp_k", _I4096),
        ("alphaES", ctypes.c_double),
        ("kES", ctypes.c_int32 * 3),
        ("task_coul", ctypes.c_int32 * 3),
        ("conv_tol_ind", ctypes.c_double),
        ("min_scf_pol_iter", ctypes.c_int32),
        ("max_scf_pol_iter", ctypes.c_int32),
        ("extrapolate_order", ctypes.c_int32),
        ("algo_ext_dipole", ctypes.c_int32),
        ("doefield", ctypes.c_int32),
        ("doefg", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("mixed_fp64_parts", ctypes.c_int32),
        ("scf_solver", ctypes.c_int32),
        ("task_coul_quad", ctypes.c_int32 * 3),
        ("reserved", ctypes.c_int32 * 2),
        ("quad", ctypes.c_double * 144),
        ("polquad", ctypes.c_double * 432),
        ("lquad_polar", ctypes.c_int32 * 16),
        ("lwfc", ctypes.c_int32 * 16),
        ("ldip_wfc", ctypes.c_int32),
        ("reserved2", ctypes.c_int32),
        ("rcut_wfc", ctypes.c_double),
    ]


class MdffEnergies(ctypes.Structure):
    """``struct mdffgpu_energies`` (include/mdffgpu.h)."""

    _fields_ = [
        ("u_lj", ctypes.c_double),
        ("vir_lj", ctypes.c_double),
        ("u_morse", ctypes.c_double),
        ("vir_morse", ctypes.c_double),
        ("u_bmhft", ctypes.c_double),
        ("vir_bmhft", ctypes.c_double),
        ("u_coul", ctypes.c_double),
        ("u_pol", ctypes.c_double),
        ("pvirial_coul", ctypes.c_double),
        ("tau_nonb", ctypes.c_double * 9),
        ("tau_coul", ctypes.c_double * 9),
        ("n_scf", ctypes.c_int32),
        ("n_es_calls", ctypes.c_int32),
        ("rmsd", ctypes.c_double),
        ("scf_tol", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(v[:]).reshape(3, 3, order="F") if name.startswith("tau") else v
        return out


def pidx(it: int, jt: int) -> int:
    """Flat index of pair-table entry (it, jt), 1-based Fortran types."""
    return (it - 1) + NTYPEMAX * (jt - 1)


def didx(a: int, b: int, c: int) -> int:
    """Flat index of damping-table entry (a, b, c), 1-based Fortran types."""
    return (a - 1) + NTYPEMAX * ((b - 1) + NTYPEMAX * (c - 1))


def field_defaults(natm: int, ntype: int) -> MdffParams:
    """Defaults of field_default_tag and control_default_tag
    (src/field.f90:190-273, src/control.f90:199-241)."""
    if not 1 <= ntype <= NTYPEMAX:
        raise ValueError(f"ntype must be in 1..{NTYPEMAX}")
    p = MdffParams()
    p.natm = natm
    p.ntype = ntype
    p.skindiff = 0.15
    p.coul_unit = COUL_UNIT
    p.press_unit = PRESS_UNIT
    for k in range(256):
        p.qlj[k] = 12.0
        p.plj[k] = 6.0
    p.alphaES = 1.0
    p.kES[:] = [10, 10, 10]
    p.conv_tol_ind = 1e-6
    p.ldip_wfc = 1  # src/field.f90:266-267
    p.rcut_wfc = 0.5
    p.min_scf_pol_iter = 3
    p.max_scf_pol_iter = 100
    p.extrapolate_order = 0
    p.algo_ext_dipole = EXT_POLY
    p.calc_mode = CALC_MD
    return p


def reduced_units(p: MdffParams) -> None:
    """lreduced (src/constants.f90:135-146)."""
    p.coul_unit = 1.0
    p.press_unit = 1.0


_PAIR_TABLES = ("plj", "qlj", "epslj", "sigmalj", "epsmor", "sigmamor", "rhomor",
                "Abmhftd", "Bbmhftd", "Cbmhftd", "Dbmhftd", "BDbmhftd")


def qidx(it: int, a: int, b: int) -> int:
    """Flat index of quad(it, a, b), all 1-based."""
    return (it - 1) + NTYPEMAX * ((a - 1) + 3 * (b - 1))


def pqidx(it: int, g: int, a: int, b: int) -> int:
    """Flat index of polquad(it, g, a, b), all 1-based."""
    return (it - 1) + NTYPEMAX * ((g - 1) + 3 * ((a - 1) + 3 * (b - 1)))


def field_check_tag(p: MdffParams, poldip_iso=None, symmetric_pot: bool = True,
                    polquad_iso=None) -> None:
    """Parameter checks and derived flags of field_check_tag
    (src/field.f90:282-433).

    * pair tables: for it = 1..ntype, column it takes row it, i.e. the lower
      triangle receives the upper one (:332-344);
    * damping tables: for every first index, the same on the (b, c) slice
      (:345-351);
    * Tang-Toennies orders above 8 stop the reference (:355-358): ValueError;
    * task_coul from the non-zero charges, static dipoles and dipole
      polarizabilities (:367-399);
    * isotropic polarizabilities expand to a diagonal tensor (:418-431),
      the quadrupole ones to polquad(g,g,g) = polquad_iso (:422-426);
    * task_coul(4:6) from the quadrupoles and quadrupole polarizabilities
      (:376, 383-395).
    """
    nt = p.ntype
    if symmetric_pot:
        for name in _PAIR_TABLES:
            t = getattr(p, name)
            for it in range(1, nt + 1):
                row = [t[pidx(it, jt)] for jt in range(1, nt + 1)]
                for jt in range(1, nt + 1):
                    t[pidx(jt, it)] = row[jt - 1]
        for name in ("ldip_damping", "dip_pol_damp_b", "dip_pol_damp_c", "dip_pol_damp_k"):
            t = getattr(p, name)
            for it2 in range(1, nt + 1):
                for it in range(1, nt + 1):
                    row = [t[didx(it2, it, c)] for c in range(1, nt + 1)]
                    for c in range(1, nt + 1):
                        t[didx(it2, c, it)] = row[c - 1]
    if any(p.dip_pol_damp_k[k] > TT_MAX_ORDER for k in range(4096)):
        raise ValueError("Tang-Toennies expansion order too large (dip_pol_damp_k > 8)")
    if poldip_iso is not None:
        for it, a in enumerate(poldip_iso):
            if a != 0.0:
                for x in range(3):
                    for y in range(3):
                        p.poldip[it + NTYPEMAX * (x + 3 * y)] = a if x == y else 0.0
    if polquad_iso is not None:
        for it, a in enumerate(polquad_iso):
            if a != 0.0:
                for k in range(27):
                    p.polquad[it + NTYPEMAX * k] = 0.0
                for g in range(1, 4):
                    p.polquad[pqidx(it + 1, g, g, g)] = a
    if p.lcoulomb:
        lqch = any(p.qch[k] != 0.0 for k in range(16))
        ldip = any(p.dip[k] != 0.0 for k in range(48))
        ldip = ldip or any(p.ldip_polar[it] for it in range(nt))
        lqua = any(p.quad[k] != 0.0 for k in range(144))
        lqua = lqua or any(p.lquad_polar[it] for it in range(nt))
        p.task_coul[0] = int(lqch)
        p.task_coul[1] = int(ldip and lqch)
        p.task_coul[2] = int(ldip)
        p.task_coul_quad[0] = int(lqua and lqch)
        p.task_coul_quad[1] = int(lqua and ldip)
        p.task_coul_quad[2] = int(lqua)


def lattice(A) -> dict:
    """Reciprocal basis B (columns, divided by the volume), volume, widths and
    lattice-vector norms as src/lattice.f90:60-168 computes them.  A is 3x3
    with the lattice vectors as columns (Fortran simu_cell%A)."""
    A = np.asarray(A, dtype=np.float64).reshape(3, 3)
    a1, a2, a3 = A[:, 0], A[:, 1], A[:, 2]
    B = np.stack([np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)], axis=1)
    omega = float(B[:, 0] @ a1)
    W = [omega / math.sqrt(float(B[:, i] @ B[:, i])) for i in range(3)]
    B = B / omega
    anorm = [math.sqrt(float(A[:, i] @ A[:, i])) for i in range(3)]
    return {"A": A, "B": B, "omega": omega, "W": W, "anorm": anorm}


def ewald_param(p: MdffParams, A, epsw: float = 1e-6, lautoES: bool = False) -> None:
    """src/field.f90:860-908: cutlongrange is always clamped to half the
    smallest perpendicular width; with lautoES, alphaES and kES come from
    the dl_poly-style estimate, which divides by that half width rather than
    by the cutoff (:875,884)."""
    lat = lattice(A)
    rcut = 0.5 * min(lat["W"])
    p.cutlongrange = min(p.cutlongrange, rcut)
    eps = min(abs(epsw), 0.5)
    tol = math.sqrt(abs(math.log(eps * p.cutlongrange)))
    alpha = math.sqrt(abs(math.log(eps * p.cutlongrange * tol))) / rcut
    tol1 = math.sqrt(-math.log(eps * p.cutlongrange * (2.0 * tol * alpha) ** 2))
    kmax = [_nint(0.25 + lat["anorm"][i] * alpha * tol1 / math.pi) for i in range(3)]
    if lautoES:
        p.alphaES = alpha
        p.kES[:] = kmax


def _nint(x: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def num_kvectors(kmax) -> int:
    """src/field.f90:1648."""
    kx, ky, kz = kmax
    return kz + ky * (2 * kz + 1) + kx * (2 * ky + 1) * (2 * kz + 1)


class MdffBatchConfig(ctypes.Structure):
    """Mirror of mdffgpu_batch_config (include/mdffgpu.h)."""

    _fields_ = [
        ("natm", ctypes.c_int32),
        ("kmax", ctypes.c_int32 * 3),
        ("A", ctypes.c_double * 9),
        ("alphaES", ctypes.c_double),
        ("cutlongrange", ctypes.c_double),
        ("cutshortrange", ctypes.c_double),
    ]


class MdffEfgOpts(ctypes.Structure):
    """Mirror of mdffgpu_efg_opts (include/mdffgpu.h): the &efgtag options of
    the standalone EFG run (src/efg.f90:102-164) that reach the per-frame
    evaluation, and longrange / ncelldirect / cutlongrange of the
    control and field tags."""

    _fields_ = [
        ("longrange", ctypes.c_int32),
        ("ncelldirect", ctypes.c_int32),
        ("cutdirect", ctypes.c_double),
        ("lefg_it_contrib", ctypes.c_int32),
        ("it_efg", ctypes.c_int32),
        ("lmp_correction", ctypes.c_int32),
        ("lefg_vasp_sign", ctypes.c_int32),
        ("lefg_old", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


EFG_EWALD = 0
EFG_DIRECT = 1


def efg_opts(longrange: str = "ewald", ncelldirect: int = 2, cutlongrange: float = 0.0,
             lefg_it_contrib: bool = False, it_efg: int = 0, lmp_correction: bool = False,
             lefg_vasp_sign: bool = False) -> MdffEfgOpts:
    """Options with the reference defaults (efg_default_tag,
    src/efg.f90:171-195; ncelldirect 2, src/field.f90:225)."""
    if longrange not in ("ewald", "direct"):
        raise ValueError("longrange must be 'ewald' or 'direct'")
    o = MdffEfgOpts()
    o.longrange = EFG_EWALD if longrange == "ewald" else EFG_DIRECT
    o.ncelldirect = int(ncelldirect)
    o.cutdirect = float(cutlongrange)
    o.lefg_it_contrib = int(bool(lefg_it_contrib))
    o.it_efg = int(it_efg)
    o.lmp_correction = int(bool(lmp_correction))
    o.lefg_vasp_sign = int(bool(lefg_vasp_sign))
    return o


# integrators of the device-resident MD loop (mdffgpu_integrator)
NVE_VV = 0
NVT_NHC2 = 1
NVT_NHCN = 2
INTEGRATORS = {"nve-vv": NVE_VV, "nvt-nhc2": NVT_NHC2, "nvt-nhcn": NVT_NHCN}
NHC_MAX = 16


class MdffMdParams(ctypes.Structure):
    """Mirror of mdffgpu_md_params (include/mdffgpu.h): &mdtag after the
    unit conversions of md_check_tag (src/md.f90:316-321)."""

    _fields_ = [
        ("integrator", ctypes.c_int32),
        ("nhc_n", ctypes.c_int32),
        ("nhc_yosh_order", ctypes.c_int32),
        ("nhc_mults", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("temp", ctypes.c_double),
        ("timesca_thermo", ctypes.c_double),
        ("boltz_unit", ctypes.c_double),
        ("mass", ctypes.c_double * NTYPEMAX),
        ("reserved", ctypes.c_int32 * 4),
    ]


class MdffMdStep(ctypes.Structure):
    """Mirror of mdffgpu_md_step (include/mdffgpu.h)."""

    _fields_ = [
        ("en", MdffEnergies),
        ("e_kin", ctypes.c_double),
        ("temp_r", ctypes.c_double),
        ("e_nvt", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        out = self.en.as_dict()
        out.update(e_kin=self.e_kin, temp_r=self.temp_r, e_nvt=self.e_nvt)
        return out


def md_params(integrator: str = "nve-vv", dt: float = 0.0, temp: float = 1.0,
              timesca_thermo: float = 1.0, nhc_n: int = 4, nhc_yosh_order: int = 3,
              nhc_mults: int = 2, mass=(1.0,), lreduced: bool = False) -> MdffMdParams:
    """&mdtag values in the reference's input units (dt and timesca_thermo
    in ps, temp in K; reduced units with lreduced), converted as
    md_check_tag does (src/md.f90:316-321, units src/constants.f90:85-87,
    139-142).  Defaults as md_default_tag (src/md.f90:191-203)."""
    if integrator not in INTEGRATORS:
        raise ValueError(f"integrator must be one of {sorted(INTEGRATORS)}")
    boltz = 1.0 if lreduced else BOLTZ_UNIT
    tu = 1.0 if lreduced else TIME_UNIT
    m = MdffMdParams()
    m.integrator = INTEGRATORS[integrator]
    m.nhc_n = 2 if integrator == "nvt-nhc2" else int(nhc_n)
    m.nhc_yosh_order = int(nhc_yosh_order)
    m.nhc_mults = int(nhc_mults)
    m.dt = float(dt) * tu
    m.temp = float(temp) * boltz
    m.timesca_thermo = float(timesca_thermo) * tu
    m.boltz_unit = boltz
    for k, v in enumerate(mass):
        m.mass[k] = float(v)
    return m
