// fhdx_amrex.H -- header-only AMReX adapter: the reference's RK3step / calculateFlux /
// conservedToPrimitive / calculateTransportCoeffs / evaluateStats / WritePlotFile /
// WriteCheckPoint / ReadCheckPoint / compute_compressible_chemistry_source_CLE signatures (src_compressible/compressible_functions.H:23-135)
// implemented over the C ABI of fhdx.h.
//
// Include it from src_compressible instead of linking timeStep.cpp, flux.cpp, transCoeffs.cpp,
// conservedPrimitiveConversions.cpp, stats.cpp, writePlotFile.cpp and checkpoint.cpp, build
// AMReX with USE_HIP=TRUE for gfx950, and link
// fhdex_amd/libfhdx.so.  Any BoxArray works (max_grid_size may split the domain, main_driver.cpp:
// 90-117): RK3step runs each stage per FAB (fhdx_rk3_stage_ref) and exchanges ghosts with
// MultiFab::FillBoundary between stages, as timeStep.cpp does; periodic directions and walls
// (setBC per FAB) as fhdx_params_check accepts them.  A device error flag (a mass fraction
// outside (0,1), where the reference aborts at transCoeffs.cpp:45-46) ends in amrex::Abort after
// RK3step and calculateTransportCoeffs.  WritePlotFile / WriteCheckPoint / ReadCheckPoint and
// evaluateStats take a single-box domain.  Compiled only where AMREX_HOME exists; it is not part
// of the library build.
#pragma once

#include <AMReX_MultiFab.H>
#include <AMReX_Geometry.H>

#include <array>
#include <chrono>
#include <cstring>
#include <vector>

#include "chemistry_namespace.H"
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
        p.t_lo[d] = t_lo[d];
        p.t_hi[d] = t_hi[d];
        // reservoir gas, as SetupCWall (boundary.cpp:14-170, run by main_driver) left it
        p.p_lo[d] = p_lo[d];
        p.p_hi[d] = p_hi[d];
        p.rho_lo[d] = rho_lo[d];
        p.rho_hi[d] = rho_hi[d];
    }
    // wall compositions (common_namespace.H:173-184), face 2*direction + side
    const amrex::GpuArray<amrex::Real, MAX_SPECIES>* Yw[6] = {&bc_Yk_x_lo, &bc_Yk_x_hi, &bc_Yk_y_lo,
                                                              &bc_Yk_y_hi, &bc_Yk_z_lo, &bc_Yk_z_hi};
    const amrex::GpuArray<amrex::Real, MAX_SPECIES>* Xw[6] = {&bc_Xk_x_lo, &bc_Xk_x_hi, &bc_Xk_y_lo,
                                                              &bc_Xk_y_hi, &bc_Xk_z_lo, &bc_Xk_z_hi};
    for (int f = 0; f < 6; ++f)
        for (int n = 0; n < nspecies; ++n) {
            p.bc_Yk[f][n] = (*Yw[f])[n];
            p.bc_Xk[f][n] = (*Xw[f])[n];
        }
    p.ngc = ngc[0];
    p.algorithm_type = algorithm_type;
    p.advection_type = advection_type;
    p.visc_type = visc_type;
    p.transport_type = compressible::transport_type;
    p.stoch_stress_form = stoch_stress_form;
    p.nreaction = chemistry::nreaction;
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

inline void check(int status);

// One context per parameter set (the namespaces can change between calls, e.g. a test driver that
// runs several decks in one process): contexts are kept for the process, keyed by the finalized
// parameters.
inline fhdx_ctx* context(amrex::Geometry const& geom)
{
    struct Entry {
        fhdx_params p;
        fhdx_ctx* ctx;
    };
    static std::vector<Entry> cache;
    fhdx_params p = params_from_namespaces(geom);
    for (const Entry& e : cache)
        if (std::memcmp(&e.p, &p, sizeof(fhdx_params)) == 0) return e.ctx;
    fhdx_ctx* ctx = nullptr;
    check(fhdx_ctx_create(&p, &ctx));
    if (p.nreaction > 0) {  // reactions of chemistry_namespace.H
        fhdx_chemistry c{};
        c.nreaction = chemistry::nreaction;
        c.reaction_type = chemistry::reaction_type;
        for (int m = 0; m < c.nreaction; ++m) {
            for (int n = 0; n < p.nspecies; ++n) {
                c.stoich_R[m][n] = chemistry::stoich_coeffs_R(m, n);
                c.stoich_P[m][n] = chemistry::stoich_coeffs_P(m, n);
            }
            c.rate_const[m] = chemistry::rate_const[m];
            c.alpha_param[m] = chemistry::alpha_param[m];
            c.beta_param[m] = chemistry::beta_param[m];
        }
        c.T0_chem = chemistry::T0_chem;
        check(fhdx_ctx_set_chemistry(ctx, &c));
    }
    cache.push_back({p, ctx});
    return ctx;
}

inline void check(int status)
{
    if (status != FHDX_OK) amrex::Abort(fhdx_last_error());
}

inline void* stream() { return (void*)amrex::Gpu::gpuStream(); }

// Philox key of the next RK3step: the seed and the number of steps taken.  ReadCheckPoint sets
// both from the checkpoint, so a restarted run continues the noise sequence exactly.  As
// main_driver.cpp:123-140: seed > 0 is the key, seed = 0 takes the clock (the same value on every
// rank: the noise is keyed by the global face index, not per rank), seed < 0 aborts.
inline uint64_t& philox_seed()
{
    static uint64_t s = [] {
        if (common::seed > 0) return (uint64_t)common::seed;
        if (common::seed < 0) amrex::Abort("Must supply non-negative seed");
        long long now = (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
                            std::chrono::system_clock::now().time_since_epoch())
                            .count();
        amrex::ParallelDescriptor::Bcast(&now, 1, amrex::ParallelDescriptor::IOProcessorNumber());
        return (uint64_t)now;
    }();
    return s;
}
inline uint64_t& philox_step()
{
    static uint64_t s = 0;
    return s;
}

// The MultiFabs of WritePlotFile / WriteCheckPoint at iterator mfi (miscStats may be absent).
inline fhdx_fields fields(amrex::MFIter const& mfi, const amrex::MultiFab& cu, const amrex::MultiFab& cuMeans,
                          const amrex::MultiFab& cuVars, const amrex::MultiFab& prim,
                          const amrex::MultiFab& primMeans, const amrex::MultiFab& primVars,
                          const amrex::MultiFab& spatialCross, const amrex::MultiFab* miscStats,
                          const amrex::MultiFab& eta, const amrex::MultiFab& kappa)
{
    fhdx_fields f;
    f.cu = fab(cu, mfi);
    f.cuMeans = fab(cuMeans, mfi);
    f.cuVars = fab(cuVars, mfi);
    f.prim = fab(prim, mfi);
    f.primMeans = fab(primMeans, mfi);
    f.primVars = fab(primVars, mfi);
    f.spatialCross = fab(spatialCross, mfi);
    f.miscStats = miscStats ? fab(*miscStats, mfi) : fhdx_fab{};
    f.eta = fab(eta, mfi);
    f.kappa = fab(kappa, mfi);
    return f;
}

}  // namespace fhdx_amrex

// ---- the reference's entry points ----------------------------------------------------------

inline void conservedToPrimitive(amrex::MultiFab& prim, const amrex::MultiFab& cons)
{
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        fhdx_amrex::check(fhdx_cons_to_prim(fhdx_amrex::context(amrex::DefaultGeometry()),
                                            fhdx_amrex::fab(prim, mfi), fhdx_amrex::fab(cons, mfi),
                                            fhdx_amrex::stream()));
}

// setBC (boundary.cpp:172-1209), after prim.FillBoundary / cons.FillBoundary.
inline void setBC(amrex::MultiFab& prim, amrex::MultiFab& cons)
{
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        fhdx_amrex::check(fhdx_set_bc(fhdx_amrex::context(amrex::DefaultGeometry()), fhdx_amrex::fab(prim, mfi),
                                      fhdx_amrex::fab(cons, mfi), fhdx_amrex::stream()));
}

inline void calculateTransportCoeffs(const amrex::MultiFab& prim, amrex::MultiFab& eta, amrex::MultiFab& zeta,
                                     amrex::MultiFab& kappa, amrex::MultiFab& chi, amrex::MultiFab& D)
{
    fhdx_ctx* ctx = fhdx_amrex::context(amrex::DefaultGeometry());
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        fhdx_amrex::check(fhdx_transport(ctx, fhdx_amrex::fab(prim, mfi), fhdx_amrex::fab(eta, mfi),
                                         fhdx_amrex::fab(zeta, mfi), fhdx_amrex::fab(kappa, mfi),
                                         fhdx_amrex::fab(chi, mfi), fhdx_amrex::fab(D, mfi), fhdx_amrex::stream()));
    // transCoeffs.cpp:45-46 aborts on a mass fraction outside (0,1)
    fhdx_amrex::check(fhdx_ctx_check(ctx, fhdx_amrex::stream()));
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

// RK3step (timeStep.cpp:11-427) with Philox noise keyed by (philox_seed(), philox_step()): the
// step's noise sets defined like the reference's temporaries (timeStep.cpp:38-72) and filled per
// FAB, then per stage every FAB (fhdx_rk3_stage_ref: transport, weighted noise, calculateFlux,
// chemistry source, stage combination, conservedToPrimitive) followed by FillBoundary of the
// stage output and prim and setBC (timeStep.cpp:204-209, 309-314, 417-425).
inline void RK3step(amrex::MultiFab& cu, amrex::MultiFab& cup, amrex::MultiFab& cup2, amrex::MultiFab& /*cup3*/,
                    amrex::MultiFab& prim, amrex::MultiFab& source, amrex::MultiFab& eta, amrex::MultiFab& zeta,
                    amrex::MultiFab& kappa, amrex::MultiFab& chi, amrex::MultiFab& D,
                    std::array<amrex::MultiFab, 3>& flux, std::array<amrex::MultiFab, 3>& stochFlux,
                    std::array<amrex::MultiFab, 3>& cornx, std::array<amrex::MultiFab, 3>& corny,
                    std::array<amrex::MultiFab, 3>& cornz, amrex::MultiFab& visccorn, amrex::MultiFab& rancorn,
                    amrex::MultiFab& ranchem, const amrex::Geometry& geom, const amrex::Real dt)
{
    using namespace fhdx_amrex;
    fhdx_ctx* ctx = context(geom);
    const int nv = cu.nComp(), nr = chemistry::nreaction, g = cu.nGrow();
    std::array<amrex::MultiFab, 3> sA, sB;
    for (int d = 0; d < 3; ++d) {
        sA[d].define(stochFlux[d].boxArray(), stochFlux[d].DistributionMap(), nv, 0);
        sB[d].define(stochFlux[d].boxArray(), stochFlux[d].DistributionMap(), nv, 0);
    }
    amrex::MultiFab rA(rancorn.boxArray(), rancorn.DistributionMap(), 1, 0);
    amrex::MultiFab rB(rancorn.boxArray(), rancorn.DistributionMap(), 1, 0);
    amrex::MultiFab chA, chB;  // the reaction noise sets, with the ghost width of ranchem
    if (nr > 0) {
        chA.define(ranchem.boxArray(), ranchem.DistributionMap(), nr, g);
        chB.define(ranchem.boxArray(), ranchem.DistributionMap(), nr, g);
    }
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        fhdx_fab a[3], b[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = fab(sA[d], mfi);
            b[d] = fab(sB[d], mfi);
        }
        check(fhdx_fill_noise(ctx, philox_seed(), philox_step(), 0, a, b, fab(rA, mfi), fab(rB, mfi), stream()));
        if (nr > 0)
            check(fhdx_fill_chem_noise(ctx, philox_seed(), philox_step(), 0, fab(chA, mfi), fab(chB, mfi), stream()));
    }
    amrex::MultiFab* out[3] = {&cup, &cup2, &cu};
    for (int st = 0; st < 3; ++st) {
        for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
            fhdx_fab fl[3], sf[3], cx[3], cy[3], cz[3], a[3], b[3];
            for (int d = 0; d < 3; ++d) {
                fl[d] = fab(flux[d], mfi);
                sf[d] = fab(stochFlux[d], mfi);
                cx[d] = fab(cornx[d], mfi);
                cy[d] = fab(corny[d], mfi);
                cz[d] = fab(cornz[d], mfi);
                a[d] = fab(sA[d], mfi);
                b[d] = fab(sB[d], mfi);
            }
            const fhdx_fab none{};
            check(fhdx_rk3_stage_ref(ctx, st, fab(cu, mfi), fab(cup, mfi), fab(cup2, mfi), fab(prim, mfi),
                                     fab(eta, mfi), fab(zeta, mfi), fab(kappa, mfi), fab(chi, mfi), fab(D, mfi), fl,
                                     sf, cx, cy, cz, fab(visccorn, mfi), fab(rancorn, mfi), a, b, fab(rA, mfi),
                                     fab(rB, mfi), nr > 0 ? fab(source, mfi) : none,
                                     nr > 0 ? fab(ranchem, mfi) : none, nr > 0 ? fab(chA, mfi) : none,
                                     nr > 0 ? fab(chB, mfi) : none, dt, stream()));
        }
        out[st]->FillBoundary(geom.periodicity());
        prim.FillBoundary(geom.periodicity());
        setBC(prim, *out[st]);
    }
    // the reference aborts inside the step on a mass fraction outside (0,1) (transCoeffs.cpp:45-46)
    check(fhdx_ctx_check(ctx, stream()));
    ++philox_step();
}

// compute_compressible_chemistry_source_CLE (chemistry_functions.cpp:128-193).
inline void compute_compressible_chemistry_source_CLE(amrex::Real dt, amrex::Real /*dV*/, amrex::MultiFab& prim,
                                                      amrex::MultiFab& source, amrex::MultiFab& ranchem)
{
    using namespace fhdx_amrex;
    for (amrex::MFIter mfi(prim); mfi.isValid(); ++mfi)
        check(fhdx_chemistry_source(context(amrex::DefaultGeometry()), fab(prim, mfi), fab(source, mfi),
                                    fab(ranchem, mfi), dt, stream()));
}

// evaluateStats (stats.cpp:5-363); cross_cell from common_namespace.H:35.
inline void evaluateStats(const amrex::MultiFab& cons, amrex::MultiFab& consMean, amrex::MultiFab& consVar,
                          const amrex::MultiFab& prim, amrex::MultiFab& primMean, amrex::MultiFab& primVar,
                          amrex::MultiFab& spatialCross, amrex::MultiFab& miscStats, amrex::Real* miscVals,
                          const int steps, const amrex::Real* /*dx*/)
{
    using namespace fhdx_amrex;
    for (amrex::MFIter mfi(cons); mfi.isValid(); ++mfi)
        check(fhdx_evaluate_stats(context(amrex::DefaultGeometry()), fab(cons, mfi), fab(consMean, mfi),
                                  fab(consVar, mfi), fab(prim, mfi), fab(primMean, mfi), fab(primVar, mfi),
                                  fab(spatialCross, mfi), fab(miscStats, mfi), miscVals, steps, common::cross_cell,
                                  stream()));
}

// WritePlotFile (writePlotFile.cpp:7-190): <plot_base_name><step, 9 digits>.
inline void WritePlotFile(int step, const amrex::Real time, const amrex::Geometry& geom, const amrex::MultiFab& cu,
                          const amrex::MultiFab& cuMeans, const amrex::MultiFab& cuVars, const amrex::MultiFab& prim,
                          const amrex::MultiFab& primMeans, const amrex::MultiFab& primVars,
                          const amrex::MultiFab& spatialCross, const amrex::MultiFab& eta,
                          const amrex::MultiFab& kappa)
{
    using namespace fhdx_amrex;
    const fhdx_params p = params_from_namespaces(geom);
    const double lo[3] = {geom.ProbLo(0), geom.ProbLo(1), geom.ProbLo(2)};
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        const fhdx_fields f = fields(mfi, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, nullptr,
                                     eta, kappa);
        check(fhdx_write_plotfile(&p, common::plot_base_name.c_str(), step, time, lo, &f, common::plot_means,
                                  common::plot_vars, 0, stream()));
    }
}

// WriteCheckPoint (checkpoint.cpp:22-133): chk<step, 9 digits>, plus the Philox seed.
inline void WriteCheckPoint(int step, const amrex::Real time, int statsCount, const amrex::Geometry& geom,
                            const amrex::MultiFab& cu, const amrex::MultiFab& cuMeans, const amrex::MultiFab& cuVars,
                            const amrex::MultiFab& prim, const amrex::MultiFab& primMeans,
                            const amrex::MultiFab& primVars, const amrex::MultiFab& spatialCross,
                            const amrex::MultiFab& miscStats, const amrex::MultiFab& eta,
                            const amrex::MultiFab& kappa)
{
    using namespace fhdx_amrex;
    const fhdx_params p = params_from_namespaces(geom);
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        const fhdx_fields f = fields(mfi, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, &miscStats,
                                     eta, kappa);
        check(fhdx_write_checkpoint(&p, "chk", step, time, statsCount, philox_seed(), &f, 0, stream()));
    }
}

// ReadCheckPoint (checkpoint.cpp:135-255): defines the MultiFabs on the one-box domain, reads
// checkpoint chk<restart>, returns step + 1, and resumes the Philox key at that step.
inline void ReadCheckPoint(int& step, amrex::Real& time, int& statsCount, amrex::Geometry& geom,
                           amrex::MultiFab& cu, amrex::MultiFab& cuMeans, amrex::MultiFab& cuVars,
                           amrex::MultiFab& prim, amrex::MultiFab& primMeans, amrex::MultiFab& primVars,
                           amrex::MultiFab& spatialCross, amrex::MultiFab& miscStats, amrex::MultiFab& eta,
                           amrex::MultiFab& kappa)
{
    using namespace fhdx_amrex;
    const fhdx_params p = params_from_namespaces(geom);
    const int nv = 5 + p.nspecies, np = 6 + 2 * p.nspecies, g = p.ngc;
    amrex::BoxArray ba(geom.Domain());
    amrex::DistributionMapping dm(ba, amrex::ParallelDescriptor::NProcs());
    cu.define(ba, dm, nv, g);
    cuMeans.define(ba, dm, nv, g);
    cuVars.define(ba, dm, nv, g);
    prim.define(ba, dm, np, g);
    primMeans.define(ba, dm, np, g);
    primVars.define(ba, dm, np + 5, g);
    spatialCross.define(ba, dm, 8, g);
    miscStats.define(ba, dm, 10, g);
    eta.define(ba, dm, 1, g);
    kappa.define(ba, dm, 1, g);
    for (amrex::MFIter mfi(cu); mfi.isValid(); ++mfi) {
        const fhdx_fields f = fields(mfi, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, &miscStats,
                                     eta, kappa);
        uint64_t seed = philox_seed();
        int has_seed = 0;
        check(fhdx_read_checkpoint(&p, "chk", common::restart, common::reset_stats, &step, &time, &statsCount, &seed,
                                   &has_seed, &f, 0, stream()));
        if (has_seed) philox_seed() = seed;
        philox_step() = (uint64_t)(step - 1);
    }
}
