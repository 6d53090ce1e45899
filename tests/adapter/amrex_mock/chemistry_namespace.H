// Compile-only stand-in for FHDeX's chemistry namespace (src_chemistry/chemistry_namespace.H):
// stoich_coeffs_R/P are two-index arrays read as (reaction, species).
#pragma once

#include "AMReX_MultiFab.H"

namespace chemistry {
struct IntTable {
    int operator()(int, int) const { return 0; }
};
extern int nreaction, reaction_type;
extern IntTable stoich_coeffs_R, stoich_coeffs_P;
extern amrex::GpuArray<amrex::Real, MAX_REACTION> rate_const, alpha_param, beta_param;
extern amrex::Real T0_chem;
}  // namespace chemistry
