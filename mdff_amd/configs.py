"""Seeded synthetic configurations of the benchmark and parity workloads.

Both the GPU library and the oracle read the same arrays produced here, so
the reference's compiler-dependent random numbers never matter.

* ``lj_fcc``         -- n-m LJ fluid on an fcc lattice, reduced units
                        (500 atoms = 5x5x5 cells x 4, rho* = 0.8442, rc 2.5,
                        skin 0.3)
* ``silica``         -- rigid-ion or polarizable silica-like melt: Si(+2.4)
                        and O(-1.2) placed by random sequential addition with
                        a 1.2 A minimum image separation at 2.2 g/cm3,
                        BMHFTD pair parameters drawn from the seed, Ewald
                        alpha balanced between real and reciprocal space
* ``batch``          -- 64 configurations with atom counts drawn uniformly
                        from [96, 1536], Si:O = 1:2
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import params as P

AMU_G = 1.66053906660e-24  # g per amu
MASS_SI = 28.0855
MASS_O = 15.9994
Q_SI = 2.4
Q_O = -1.2
DENSITY = 2.2  # g/cm3


@dataclasses.dataclass
class System:
    params: P.MdffParams
    itype: np.ndarray  # (n,) 1-based
    A: np.ndarray  # (3,3) lattice vectors as columns
    r: np.ndarray  # (n,3) Cartesian
    name: str
    info: dict

    @property
    def natm(self) -> int:
        return int(self.params.natm)


def lj_fcc(ncell: int = 5, rho: float = 0.8442, rc: float = 2.5, skin: float = 0.3) -> System:
    """fcc lattice of 4*ncell^3 LJ atoms (sigma = eps = 1, reduced units)."""
    n = 4 * ncell ** 3
    L = (n / rho) ** (1.0 / 3.0)
    a = L / ncell
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)])
    r = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    p = P.field_defaults(n, 1)
    P.reduced_units(p)
    p.lnmlj = 1
    p.cutshortrange = rc
    p.skindiff = skin
    p.epslj[0] = 1.0
    p.sigmalj[0] = 1.0
    P.field_check_tag(p)
    return System(p, np.ones(n, np.int32), np.eye(3) * L, r, f"lj_fcc_{n}",
                  {"L": L, "rho": rho, "rc": rc})


def box_length(n_si: int, n_o: int, density: float = DENSITY) -> float:
    mass_g = (n_si * MASS_SI + n_o * MASS_O) * AMU_G
    vol_A3 = mass_g / density * 1e24
    return vol_A3 ** (1.0 / 3.0)


def rsa_positions(n: int, L: float, dmin, rng: np.random.Generator, itype=None) -> np.ndarray:
    """Random sequential addition in a cubic box with a minimum-image
    separation dmin (a number, or a dict {(ti,tj): d} per species pair with
    itype giving the species of each atom in placement order), using a cell
    grid."""
    if isinstance(dmin, dict):
        dmax = max(dmin.values())
        dm2 = {k: v * v for k, v in dmin.items()}
        dm2.update({(b, a): v for (a, b), v in dm2.items()})
    else:
        dmax = dmin
        dm2 = None
    nc = max(1, int(L // dmax))
    h = L / nc
    grid = {}
    out = np.empty((n, 3))
    d2 = dmax * dmax
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n + 1000:
            raise RuntimeError("random sequential addition did not converge")
        x = rng.random(3) * L
        ci = tuple((x // h).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cell = ((ci[0] + dx) % nc, (ci[1] + dy) % nc, (ci[2] + dz) % nc)
                    for j in grid.get(cell, ()):
                        d = out[j] - x
                        d -= L * np.round(d / L)
                        lim = d2 if dm2 is None else dm2[(int(itype[k]), int(itype[j]))]
                        if d @ d < lim:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out[k] = x
            grid.setdefault(ci, []).append(k)
            k += 1
    return out


def balanced_alpha(L: float, rc: float, kmax: int) -> float:
    """Ewald alpha with exp(-(alpha rc)^2) = exp(-(pi kmax / (alpha L))^2),
    i.e. equal real- and reciprocal-space truncation."""
    return math.sqrt(math.pi * kmax / (L * rc))


def bmhftd_params(rng: np.random.Generator) -> dict:
    """BMHFTD parameters for the (Si,Si), (Si,O), (O,O) pairs: A in
    [1e3, 2e4] eV, B in [2.5, 5] 1/A, C6 in [0, 100] eV A^6, C8 = 0.1 C6
    eV A^8, damping BD = B."""
    out = {}
    for pair in ((1, 1), (1, 2), (2, 2)):
        A = rng.uniform(1e3, 2e4)
        B = rng.uniform(2.5, 5.0)
        C6 = rng.uniform(0.0, 100.0)
        out[pair] = (A, B, C6, 0.1 * C6, B)
    return out


# Minimum separations of the polarizable configurations.  A common 1.2 A
# floor lets O-O pairs come so close that 2 alpha_O / r^3 > 1 and the
# Jacobi dipole iteration of the reference diverges (polarization
# catastrophe); these species-pair floors keep the SCF contractive.
PIM_DMIN = {(1, 1): 2.6, (1, 2): 1.5, (2, 2): 2.2}


def silica(n_si: int, n_o: int, seed: int, rc: float = 10.0, kmax: int = 10,
           pim: bool = False, efg: bool = False, scf_tol: float = 1e-8,
           damping=(2.0, 1.0, 4), alpha_o: float = 1.0, pot_seed=None,
           precision: int = P.PREC_FP64, dmin=None) -> System:
    """Silica-like melt: Si block then O block (itype 1 = Si, 2 = O).
    Rigid-ion melts use a 1.2 A minimum separation; polarizable ones use
    PIM_DMIN unless dmin is given."""
    rng = np.random.default_rng(seed)
    n = n_si + n_o
    L = box_length(n_si, n_o)
    itype = np.array([1] * n_si + [2] * n_o, np.int32)
    if dmin is None:
        dmin = PIM_DMIN if pim else 1.2
    r = rsa_positions(n, L, dmin, rng, itype)
    prng = rng if pot_seed is None else np.random.default_rng(pot_seed)
    pot = bmhftd_params(prng)
    p = P.field_defaults(n, 2)
    p.lbmhftd = 1
    p.lcoulomb = 1
    p.cutshortrange = min(rc, 0.5 * L)
    p.cutlongrange = rc
    p.skindiff = 0.5
    for (i, j), (A, B, C6, C8, BD) in pot.items():
        k = P.pidx(i, j)
        p.Abmhftd[k], p.Bbmhftd[k], p.Cbmhftd[k], p.Dbmhftd[k], p.BDbmhftd[k] = A, B, C6, C8, BD
    p.qch[0] = Q_SI
    p.qch[1] = Q_O
    if pim:
        p.ldip_polar[1] = 1
        b, c, k = damping
        for jt in (1, 2):
            idx = P.didx(2, 2, jt)
            p.ldip_damping[idx] = 1
            p.dip_pol_damp_b[idx] = b
            p.dip_pol_damp_c[idx] = c
            p.dip_pol_damp_k[idx] = k
        p.conv_tol_ind = scf_tol
        p.min_scf_pol_iter = 3
        p.max_scf_pol_iter = 100
    p.doefg = int(efg)
    p.doefield = int(efg)
    p.precision = precision
    P.field_check_tag(p, poldip_iso=[0.0, alpha_o] if pim else None)
    A = np.eye(3) * L
    P.ewald_param(p, A)
    p.alphaES = balanced_alpha(L, p.cutlongrange, kmax)
    p.kES[:] = [kmax, kmax, kmax]
    return System(p, itype, A, r, f"silica_{n}{'_pim' if pim else ''}",
                  {"L": L, "seed": seed, "pot": pot, "alpha": p.alphaES, "kmax": kmax})


def species_masses(sys_: System) -> tuple:
    """Masses per species: 1 for the reduced-unit LJ fluid, Si and O (amu)
    for the silica-like melts."""
    return (1.0,) if sys_.params.ntype == 1 else (MASS_SI, MASS_O)


def maxwell_velocities(sys_: System, temp: float, seed: int, lreduced=None) -> np.ndarray:
    """Seeded Maxwell-Boltzmann velocities (natm, 3) at temp (K, or reduced
    with lreduced) with zero net momentum, rescaled to exactly temp as
    calc_temp measures it (src/kinetic.f90:531-556)."""
    if lreduced is None:
        lreduced = sys_.params.ntype == 1
    kb = 1.0 if lreduced else P.BOLTZ_UNIT
    m = np.asarray(species_masses(sys_))[sys_.itype - 1]
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((sys_.natm, 3)) * np.sqrt(kb * temp / m)[:, None]
    v -= (m[:, None] * v).sum(axis=0) / m.sum()
    t_now = (m[:, None] * v * v).sum() / (3.0 * sys_.natm * kb)
    return v * math.sqrt(temp / t_now)


def config(k: int, precision: int = P.PREC_FP64) -> System:
    """The five workloads (1-based)."""
    if k == 1:
        return lj_fcc()
    if k == 2:
        return silica(216, 432, seed=2, rc=10.0, kmax=10, precision=precision)
    if k == 3:
        return silica(1000, 2000, seed=3, rc=10.0, kmax=10, pim=True, efg=True, scf_tol=1e-8,
                      precision=precision)
    if k == 5:
        return silica(8192, 16384, seed=5, rc=12.0, kmax=24, pim=True, efg=True, scf_tol=1e-7,
                      precision=precision)
    raise ValueError("config 4 is the batch, see batch()")


def batch(nconf: int = 64, seed: int = 4, pot_seed: int = 40, nmin: int = 96,
          nmax: int = 1536, kmax: int = 8, precision: int = P.PREC_MIXED):
    """The batched force-matching set: atom counts uniform in [nmin, nmax]
    rounded to Si:O = 1:2, each with its own seed; all share the potential
    drawn from pot_seed.  Cutoffs are min(10 A, L/2)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(nconf):
        n = int(rng.integers(nmin, nmax + 1))
        n_si = max(1, n // 3)
        n_o = 2 * n_si
        L = box_length(n_si, n_o)
        rc = min(10.0, 0.5 * L - 1e-9)
        out.append(silica(n_si, n_o, seed=1000 + c, rc=rc, kmax=kmax, pim=True, efg=False,
                          scf_tol=1e-8, pot_seed=pot_seed, precision=precision))
    return out
