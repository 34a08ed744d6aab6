"""Algorithmic work per unit of the hot kernels, fixed once here and used by
bench.py's roofline and by the documentation.

Convention: add/sub/mul = 1 flop, FMA = 2 flops.  The k-space kernels are
written with separable phases e^{i k.r} = e1^n1 e2^n2 e3^n3, so the
per-(atom, k) work is a complex multiply by the step phase plus the
accumulations; no transcendental is evaluated per (atom, k).

unit of the k-space kernels: one (atom, k) pair, k over the half-space mesh
(src/field.f90:1648: K = kx + ky(2kz+1) + kx(2ky+1)(2kz+1)).
unit of the real-space kernels: one directed neighbour pair inside the
cutoff (each pair is evaluated from both atoms).
"""

# structure factor rho(k) = sum_j w_j e^{ik.r_j}
RHO_CHG = 10      # phase step 6 + 2 FMA (q weight)
RHO_DIP = 11      # phase step 6 + t update 1 + 2 FMA (i k.mu weight)
RHO_BOTH = 15     # phase step 6 + t update 1 + 4 FMA
# apply: field sums U_c (13) and EFG / dipole-force sums V_cd (22)
APPLY_FIELD = 13  # phase step 6, recarg 3, R0 1, R3 2, n3 1
APPLY_EFG = 22    # + recarg2 3, Q0 1, Q3 2, Q33 3

# bytes per unit moved by the hot k-space kernels from HBM/L2 are O(1/K) and
# O(1/N): both are compute bound (arithmetic intensity > 1e3 flop/B).

KERNEL_FLOPS_PER_UNIT = {
    "k_rho64<chg>": RHO_CHG,
    "k_rho64<dip>": RHO_DIP,
    "k_rho64<chg+dip>": RHO_BOTH,
    "k_apply64<field>": APPLY_FIELD,
    "k_apply64<efg>": APPLY_EFG,
    "mx_rho": None,     # MFMA path, flops from the GEMM shape
    "mx_apply": None,
}

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6     # FP64 vector and FP64 matrix
PEAK_FP32_TFLOPS = 157.3    # FP32 vector (packed)
PEAK_FP16_MFMA_TFLOPS = 2500.0  # dense FP16/BF16 matrix
PEAK_HBM_GBS = 8000.0


def kspace_algorithmic_flops(kernel: str, natm: int, nk: int, steps: int, n_ind_total: int,
                             efg: bool, n_dip: int = None, n_fld: int = None) -> float:
    """Algorithmic flops of the reciprocal-space sums over `steps`
    evaluations with `n_ind_total` induced-dipole Ewald calls in all, from
    the per-(atom, k) weights above: per evaluation one static call
    (charges), the induced-dipole calls and one final call (charges and
    dipoles; field, and EFG / dipole forces when requested).
      rho:   RHO_CHG, RHO_DIP per SCF call, RHO_BOTH
      apply: APPLY_FIELD, APPLY_FIELD per SCF call, APPLY_EFG (or FIELD)
    With n_dip and n_fld (the mixed path), the work that path does: its
    dipole-only structure factors (the SCF calls and the final call, whose
    charge sums are the static call's) run over the n_dip atoms that can
    carry a dipole, and the static and SCF field calls over the n_fld
    polarizable atoms."""
    if n_dip is None or n_fld is None:
        nK = float(natm) * nk
        if kernel in ("mx_rho", "k_rho64"):
            w = steps * (RHO_CHG + RHO_BOTH) + n_ind_total * RHO_DIP
        else:
            w = steps * (APPLY_FIELD + (APPLY_EFG if efg else APPLY_FIELD)) + n_ind_total * APPLY_FIELD
        return w * nK
    if kernel in ("mx_rho", "k_rho64"):
        return float(nk) * (steps * RHO_CHG * natm + (n_ind_total + steps) * RHO_DIP * n_dip)
    final = APPLY_EFG if efg else APPLY_FIELD
    return float(nk) * ((steps + n_ind_total) * APPLY_FIELD * n_fld + steps * final * natm)


def mixed_kspace_flops(kernel: str, natm: int, kmax, steps: int, n_scf_total: int,
                       efg: bool, n_dip: int = None, n_fld: int = None) -> float:
    """Executed matrix-core flops of the mixed k-space GEMMs over a run of
    `steps` evaluations with `n_scf_total` SCF iterations in all: per call
    3 fp16 products of a real GEMM over the conjugate-paired half of the
    (n2, n3) columns, CH = ((2ky+1)(2kz+1) + 1) / 2.
      mx_rho:   rows 2*NW*(kx+1) (weight, n1, re|im), cols 2*CH, K atoms
      mx_apply: rows atoms, K 2*CH, cols 2*W*(kx+1)
    NW = 1 (static call, all atoms), 3 (SCF calls and the dipole-only final
    call, over the n_dip atoms); W = 3 (static and SCF calls, over the n_fld
    atoms), or 6 in the final call (all atoms) when EFGs or dipole forces
    are needed."""
    kx, ky, kz = kmax
    n1, c23 = kx + 1, (2 * ky + 1) * (2 * kz + 1)
    ch = (c23 + 1) // 2
    nd = natm if n_dip is None else n_dip
    nf = natm if n_fld is None else n_fld
    if kernel == "mx_rho":
        calls = [(2 * 1 * n1, natm)] * steps + [(2 * 3 * n1, nd)] * (n_scf_total + steps)
        return sum(3 * 2.0 * r * 2 * ch * k for r, k in calls)
    calls = [(2 * 3 * n1, nf)] * (steps + n_scf_total) + [(2 * (6 if efg else 3) * n1, natm)] * steps
    return sum(3 * 2.0 * rows * 2 * ch * c for c, rows in calls)
