// fhdx_amrex.H -- header-only AMReX adapter: the reference's RK3step / calculateFlux /
// conservedToPrimitive / calculateTransportCoeffs signatures (src_compressible/
// compressible_functions.H:23-55) implemented over the C ABI of fhdx.h.
//
// Include it from src_compressible instead of linking timeStep.cpp, flux.cpp, transCoeffs.cpp
// and conservedPrimitiveConversions.cpp, build AMReX with USE_HIP=TRUE for gfx950, and link
// fhdex_amd/libfhdx.so.  Requirements: a fully periodic domain and one box per rank
// (max_grid_size >= n_cells), which is what the library's index checks accept.  Compiled only
// where AMREX_HOME exists; it is not part of the library build.
#pragma once

#include <AMReX_MultiFab.H>
#include <AMReX_Geometry.H>

#include <array>

#include "common_functions.H"
#include "compressible_namespace.H"
#include "fhdx.h"

namespace fhdx_amrex {

inline fhdx_fab fab(amrex::Array4<amrex::Real> const& a)
{
    fhdx_fab f;
    f.p = a.p;
    f.lo[0] = a.begin.x; f.lo[1] = a.begin.y; f.lo[2] = a.begin.z;
    f.hi[0] = a.end.x - 1; f.hi[1] = a.end.y - 1; f.hi[2] = a.end.z - 1;
    f.ncomp = a.ncomp;
    f.pad = 0;
    return f;
}

inline fhdx_fab fab(amrex::MultiFab const& mf, amrex::MFIter const& mfi)
{
    return fab(const_cast<amrex::MultiFab&>(mf).array(mfi));
}

// Parameters from the reference's namespaces (common_namespace.H, compressible_namespace.H).
inline fhdx_params params_from_namespaces(amrex::Geometry const& geom)
{
    using namespace common;
    fhdx_params p;
    fhdx_params_default(&p);
    p.nspecies = nspecies;
    for (int d = 0; d < 3; ++d) {
        p.n_cells[d] = n_cells[d];
        p.dx[d] = geom.CellSize(d);
        p.bc_vel_lo[d] = bc_vel_lo[d];
        p.bc_vel_hi[d] = bc_vel_hi[d];
        p.bc_mass_lo[d] = bc_mass_lo[d];
        p.bc_mass_hi[d] = bc_mass_hi[d];
        p.bc_therm_lo[d] = bc_therm_lo[d];
        p.bc_therm_hi[d] = bc_therm_hi[d];
        p.grav[d] = grav[d];
    }
    p.ngc = ngc[0];
    p.algorithm_type = algorithm_type;
    p.advection_type = advection_type;
    p.visc_type = visc_type;
    p.transport_type = compressible::transport_type;
    p.stoch_stress_form = stoch_stress_form;
    p.nreaction = 0;
    for (int n = 0; n < nspecies; ++n) {
        p.molmass[n] = molmass[n];
        p.diameter[n] = diameter[n];
        p.dof[n] = dof[n];
        p.hcv[n] = hcv[n];
        p.hcp[n] = hcp[n];
        p.e0[n] = e0[n];
    }
    p.k_B = k_B;
    p.Runiv = Runiv;
    p.variance_coef_mom = variance_coef_mom;
    p.variance_coef_ener = variance_coef_ener;
    p.variance_coef_mass = variance_coef_mass;
    p.zeta_ratio = compressible::zeta_ratio;
    fhdx_params_finalize(&p);
    return p;
}

inline fhdx_ctx* context(amrex::Geometry const& geom)
{
    static fhdx_ctx* ctx = nullptr;
    if (!ctx) {
        fhdx_params p = params_from_namespaces(geom);
        if (fhdx_ctx_create(&p, &ctx) != FHDX_OK) amrex::Abort(fhdx_last_error());
    }
    return ctx;
}

inline void check(int status)
{
    if (status != FHDX_OK) amrex::Abort(fhdx_last_error());
}

inline void* stream() { return (void*)amrex::Gpu::gpuStream(); }

}  // namespace fhdx_amrex

// ---- the reference's entry points ----------------------------------------------------------

inline void conservedToPrimitive(amrex::MultiFab& prim, const amrex::MultiFab& cons)
{
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        fhdx_amrex::check(fhdx_cons_to_prim(fhdx_amrex::context(amrex::DefaultGeometry()),
                                            fhdx_amrex::fab(prim, mfi), fhdx_amrex::fab(cons, mfi),
                                            fhdx_amrex::stream()));
}

inline void calculateTransportCoeffs(const amrex::MultiFab& prim, amrex::MultiFab& eta, amrex::MultiFab& zeta,
                                     amrex::MultiFab& kappa, amrex::MultiFab& chi, amrex::MultiFab& D)
{
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        fhdx_amrex::check(fhdx_transport(fhdx_amrex::context(amrex::DefaultGeometry()), fhdx_amrex::fab(prim, mfi),
                                         fhdx_amrex::fab(eta, mfi), fhdx_amrex::fab(zeta, mfi),
                                         fhdx_amrex::fab(kappa, mfi), fhdx_amrex::fab(chi, mfi),
                                         fhdx_amrex::fab(D, mfi), fhdx_amrex::stream()));
}

inline void calculateFlux(const amrex::MultiFab& cons, const amrex::MultiFab& prim, const amrex::MultiFab& eta,
                          const amrex::MultiFab& zeta, const amrex::MultiFab& kappa, const amrex::MultiFab& chi,
                          const amrex::MultiFab& D, std::array<amrex::MultiFab, 3>& flux,
                          std::array<amrex::MultiFab, 3>& stochFlux, std::array<amrex::MultiFab, 3>& cornx,
                          std::array<amrex::MultiFab, 3>& corny, std::array<amrex::MultiFab, 3>& cornz,
                          amrex::MultiFab& visccorn, amrex::MultiFab& rancorn, const amrex::Geometry& geom,
                          const amrex::Vector<amrex::Real>& /*stoch_weights*/, const amrex::Real dt)
{
    using namespace fhdx_amrex;
    for (amrex::MFIter mfi(cons); mfi.isValid(); ++mfi) {
        fhdx_fab fl[3], sf[3], cx[3], cy[3], cz[3];
        for (int d = 0; d < 3; ++d) {
            fl[d] = fab(flux[d], mfi);
            sf[d] = fab(stochFlux[d], mfi);
            cx[d] = fab(cornx[d], mfi);
            cy[d] = fab(corny[d], mfi);
            cz[d] = fab(cornz[d], mfi);
        }
        check(fhdx_calculate_flux(context(geom), fab(cons, mfi), fab(prim, mfi), fab(eta, mfi), fab(zeta, mfi),
                                  fab(kappa, mfi), fab(chi, mfi), fab(D, mfi), fl, sf, cx, cy, cz,
                                  fab(visccorn, mfi), fab(rancorn, mfi), dt, stream()));
    }
}

// RK3step with Philox noise keyed by (common::seed, a step counter kept here).
inline void RK3step(amrex::MultiFab& cu, amrex::MultiFab& cup, amrex::MultiFab& cup2, amrex::MultiFab& /*cup3*/,
                    amrex::MultiFab& prim, amrex::MultiFab& /*source*/, amrex::MultiFab& eta, amrex::MultiFab& zeta,
                    amrex::MultiFab& kappa, amrex::MultiFab& chi, amrex::MultiFab& D,
                    std::array<amrex::MultiFab, 3>& flux, std::array<amrex::MultiFab, 3>& stochFlux,
                    std::array<amrex::MultiFab, 3>& cornx, std::array<amrex::MultiFab, 3>& corny,
                    std::array<amrex::MultiFab, 3>& cornz, amrex::MultiFab& visccorn, amrex::MultiFab& rancorn,
                    amrex::MultiFab& /*ranchem*/, const amrex::Geometry& geom, const amrex::Real dt)
{
    using namespace fhdx_amrex;
    static uint64_t step = 0;
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        fhdx_fab fl[3], sf[3], cx[3], cy[3], cz[3];
        for (int d = 0; d < 3; ++d) {
            fl[d] = fab(flux[d], mfi);
            sf[d] = fab(stochFlux[d], mfi);
            cx[d] = fab(cornx[d], mfi);
            cy[d] = fab(corny[d], mfi);
            cz[d] = fab(cornz[d], mfi);
        }
        check(fhdx_rk3_step_ref(context(geom), fab(cu, mfi), fab(cup, mfi), fab(cup2, mfi), fab(prim, mfi),
                                fab(eta, mfi), fab(zeta, mfi), fab(kappa, mfi), fab(chi, mfi), fab(D, mfi), fl, sf,
                                cx, cy, cz, fab(visccorn, mfi), fab(rancorn, mfi), nullptr, nullptr, nullptr,
                                nullptr, dt, (uint64_t)common::seed, step, stream()));
    }
    ++step;
}
