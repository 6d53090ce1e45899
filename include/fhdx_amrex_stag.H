// fhdx_amrex_stag.H -- header-only AMReX adapter for the staggered scheme: the reference's
// RK3stepStag / calculateFluxStag / conservedToPrimitiveStag signatures
// (src_compressible_stag/compressible_functions_stag.H:113-143 and
// conservedPrimitiveConversionsStag.cpp:5-6) over the C ABI of fhdx.h.
//
// Include it from src_compressible_stag instead of linking timeStepStag.cpp, fluxStag.cpp and
// conservedPrimitiveConversionsStag.cpp (with fhdx_amrex.H for the collocated helpers it
// shares: calculateTransportCoeffs, the context cache, the Philox seed), build AMReX with
// USE_HIP=TRUE for gfx950 and link fhdex_amd/libfhdx.so.  Scope of the library's staggered
// path: one box per rank covering the whole periodic domain (max_grid_size >= n_cells), the
// 3-D branch with chemistry, no reservoir, membrane or turbulence forcing; anything else ends
// in amrex::Abort with the library's message.  Compiled only where AMREX_HOME exists.
#pragma once

#include <array>

#include "TurbForcingComp.H"
#include "fhdx_amrex.H"

namespace fhdx_amrex {

// The staggered entry points take the whole domain as one box.
inline void single_box(const amrex::MultiFab& mf, const char* who)
{
    if (mf.boxArray().size() != 1)
        amrex::Abort(std::string(who) + ": the staggered path takes one box covering the domain "
                     "(set max_grid_size >= n_cells)");
}

inline void fabs3(const std::array<amrex::MultiFab, 3>& mf, const amrex::MFIter& mfi, fhdx_fab out[3])
{
    for (int d = 0; d < 3; ++d) out[d] = fab(mf[d], mfi);
}

// edgeflux_x[0], edgeflux_x[1], edgeflux_y[0], edgeflux_y[1], edgeflux_z[0], edgeflux_z[1]
inline void edge6(const std::array<amrex::MultiFab, 2>& ex, const std::array<amrex::MultiFab, 2>& ey,
                  const std::array<amrex::MultiFab, 2>& ez, const amrex::MFIter& mfi, fhdx_fab out[6])
{
    out[0] = fab(ex[0], mfi);
    out[1] = fab(ex[1], mfi);
    out[2] = fab(ey[0], mfi);
    out[3] = fab(ey[1], mfi);
    out[4] = fab(ez[0], mfi);
    out[5] = fab(ez[1], mfi);
}

}  // namespace fhdx_amrex

// conservedToPrimitiveStag (conservedPrimitiveConversionsStag.cpp:5-144); the library also
// fills the periodic ghosts, which the reference's callers do next with FillBoundary.
inline void conservedToPrimitiveStag(amrex::MultiFab& prim, std::array<amrex::MultiFab, AMREX_SPACEDIM>& velStag,
                                     amrex::MultiFab& cons,
                                     const std::array<amrex::MultiFab, AMREX_SPACEDIM>& momStag)
{
    using namespace fhdx_amrex;
    single_box(cons, "conservedToPrimitiveStag");
    for (amrex::MFIter mfi(cons); mfi.isValid(); ++mfi) {
        fhdx_fab v[3], m[3];
        fabs3(velStag, mfi, v);
        fabs3(momStag, mfi, m);
        check(fhdx_cons_to_prim_stag(context(amrex::DefaultGeometry()), fab(prim, mfi), v, fab(cons, mfi), m,
                                     stream()));
    }
}

// calculateFluxStag (fluxStag.cpp:5-1868).  The reference hands in the weighted noise MultiFabs;
// here the noise of RK3stepStag's current step (philox_seed(), philox_step()) is evaluated in
// place with the stage weight stoch_weights[1], so the stochface/stochedge/stochcen arguments
// are not read.
inline void calculateFluxStag(const amrex::MultiFab& cons_in, const std::array<amrex::MultiFab, AMREX_SPACEDIM>& momStag_in,
                              const amrex::MultiFab& prim_in, const std::array<amrex::MultiFab, AMREX_SPACEDIM>& velStag_in,
                              const amrex::MultiFab& eta_in, const amrex::MultiFab& zeta_in,
                              const amrex::MultiFab& kappa_in, const amrex::MultiFab& chi_in,
                              const amrex::MultiFab& D_in, std::array<amrex::MultiFab, AMREX_SPACEDIM>& faceflux_in,
                              std::array<amrex::MultiFab, 2>& edgeflux_x_in, std::array<amrex::MultiFab, 2>& edgeflux_y_in,
                              std::array<amrex::MultiFab, 2>& edgeflux_z_in,
                              std::array<amrex::MultiFab, AMREX_SPACEDIM>& cenflux_in,
                              std::array<amrex::MultiFab, AMREX_SPACEDIM>& /*stochface_in*/,
                              std::array<amrex::MultiFab, 2>& /*stochedge_x_in*/,
                              std::array<amrex::MultiFab, 2>& /*stochedge_y_in*/,
                              std::array<amrex::MultiFab, 2>& /*stochedge_z_in*/,
                              std::array<amrex::MultiFab, AMREX_SPACEDIM>& /*stochcen_in*/,
                              const amrex::Geometry& geom, const amrex::Vector<amrex::Real>& stoch_weights,
                              const amrex::Real dt)
{
    using namespace fhdx_amrex;
    single_box(cons_in, "calculateFluxStag");
    for (amrex::MFIter mfi(cons_in); mfi.isValid(); ++mfi) {
        fhdx_fab m[3], v[3], ff[3], cf[3], ef[6];
        fabs3(momStag_in, mfi, m);
        fabs3(velStag_in, mfi, v);
        fabs3(faceflux_in, mfi, ff);
        fabs3(cenflux_in, mfi, cf);
        edge6(edgeflux_x_in, edgeflux_y_in, edgeflux_z_in, mfi, ef);
        check(fhdx_calculate_flux_stag(context(geom), fab(cons_in, mfi), m, fab(prim_in, mfi), v, fab(eta_in, mfi),
                                       fab(zeta_in, mfi), fab(kappa_in, mfi), fab(chi_in, mfi), fab(D_in, mfi), ff,
                                       ef, cf, stoch_weights[1], dt, philox_seed(), philox_step(), stream()));
    }
}

// RK3stepStag (timeStepStag.cpp:13-1245) with Philox noise keyed by (philox_seed(),
// philox_step()); a device error flag (a mass fraction outside (0,1) in a stage's transport
// step, where the reference aborts at transCoeffs.cpp:45-46) ends in amrex::Abort.
inline void RK3stepStag(amrex::MultiFab& cu, std::array<amrex::MultiFab, AMREX_SPACEDIM>& cumom,
                        amrex::MultiFab& prim, std::array<amrex::MultiFab, AMREX_SPACEDIM>& facevel,
                        amrex::MultiFab& source, amrex::MultiFab& eta, amrex::MultiFab& zeta,
                        amrex::MultiFab& kappa, amrex::MultiFab& chi, amrex::MultiFab& D,
                        std::array<amrex::MultiFab, AMREX_SPACEDIM>& faceflux,
                        std::array<amrex::MultiFab, 2>& edgeflux_x, std::array<amrex::MultiFab, 2>& edgeflux_y,
                        std::array<amrex::MultiFab, 2>& edgeflux_z,
                        std::array<amrex::MultiFab, AMREX_SPACEDIM>& cenflux, amrex::MultiFab& ranchem,
                        const amrex::Geometry& geom, const amrex::Real dt, const int /*step*/,
                        TurbForcingComp& /*turbforce*/)
{
    using namespace fhdx_amrex;
    single_box(cu, "RK3stepStag");
    fhdx_ctx* ctx = context(geom);
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        fhdx_fab m[3], v[3], ff[3], cf[3], ef[6];
        fabs3(cumom, mfi, m);
        fabs3(facevel, mfi, v);
        fabs3(faceflux, mfi, ff);
        fabs3(cenflux, mfi, cf);
        edge6(edgeflux_x, edgeflux_y, edgeflux_z, mfi, ef);
        const bool chem = chemistry::nreaction > 0;
        check(fhdx_rk3_step_stag(ctx, fab(cu, mfi), m, fab(prim, mfi), v, fab(eta, mfi), fab(zeta, mfi),
                                 fab(kappa, mfi), fab(chi, mfi), fab(D, mfi), ff, ef, cf,
                                 chem ? fab(source, mfi) : fhdx_fab{}, chem ? fab(ranchem, mfi) : fhdx_fab{}, dt,
                                 philox_seed(), philox_step(), stream()));
    }
    ++philox_step();
    check(fhdx_ctx_check(ctx, stream()));
}
