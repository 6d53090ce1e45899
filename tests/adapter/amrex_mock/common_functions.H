// Compile-only stand-in for FHDeX's common namespace: the variables the adapters read, with the
// types src_common/common_namespace.H gives them.
#pragma once

#include <string>

#include "AMReX_MultiFab.H"

namespace common {
extern int nspecies;
extern amrex::GpuArray<int, 3> n_cells;
extern amrex::IntVect ngc;
extern int cross_cell;
extern std::string plot_base_name;
extern int restart, reset_stats, plot_means, plot_vars, seed;
extern amrex::GpuArray<amrex::Real, 3> grav, p_lo, p_hi, t_lo, t_hi, rho_lo, rho_hi;
extern amrex::GpuArray<amrex::Real, MAX_SPECIES> molmass, diameter, dof, e0, hcv, hcp;
extern amrex::GpuArray<amrex::Real, MAX_SPECIES> bc_Yk_x_lo, bc_Yk_x_hi, bc_Yk_y_lo, bc_Yk_y_hi, bc_Yk_z_lo,
    bc_Yk_z_hi;
extern amrex::GpuArray<amrex::Real, MAX_SPECIES> bc_Xk_x_lo, bc_Xk_x_hi, bc_Xk_y_lo, bc_Xk_y_hi, bc_Xk_z_lo,
    bc_Xk_z_hi;
extern amrex::Real variance_coef_mom, variance_coef_ener, variance_coef_mass, k_B, Runiv;
extern int algorithm_type, advection_type, visc_type, stoch_stress_form;
extern amrex::GpuArray<int, 3> bc_vel_lo, bc_vel_hi, bc_mass_lo, bc_mass_hi, bc_therm_lo, bc_therm_hi;
}  // namespace common
