// tables.h -- host-side derivation of the per-type-pair and per-type tables
// and of the cell (shared by the single-configuration context and the
// batched path).
#pragma once

#include "../../include/mdffgpu.h"
#include "common.h"

namespace mdff {

// Pair, type and Tang-Toennies tables of the field parameters
// (initialize_param_non_bonded, src/field.f90:1021-1064; typeinfo_init,
// src/read_pos.f90:175-210).  Returns -1, or the first Tang-Toennies order
// outside 0..MDFFGPU_TT_MAX_ORDER.
int make_tables(const mdffgpu_params& P, PairTables& T, TypeTables& Y, TTTable& tt);

// Cell from the lattice vectors (columns of A): B = A^-T, volume and the
// three widths (src/lattice.f90:60-120).  False when the volume is not
// positive.
bool make_cell(const double* A, Cell& cl, double W[3]);

void inv3(const double* m, double* out);

}  // namespace mdff
