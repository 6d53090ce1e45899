// Compile-only translation unit: includes both AMReX adapters against the stand-in headers in
// amrex_mock/ and calls every entry point they define with the reference's argument lists, so
// that a signature mismatch with fhdx.h or with the reference's call sites fails the build.
#include "fhdx_amrex.H"
#include "fhdx_amrex_stag.H"

void call_every_entry_point()
{
    using amrex::MultiFab;
    const amrex::Geometry& geom = amrex::DefaultGeometry();
    MultiFab cu, cup, cup2, cup3, prim, source, eta, zeta, kappa, chi, D, visccorn, rancorn, ranchem;
    MultiFab cuMeans, cuVars, primMeans, primVars, spatialCross, miscStats;
    std::array<MultiFab, 3> flux, stochFlux, cornx, corny, cornz;
    amrex::Vector<amrex::Real> weights(3, 0.);
    const amrex::Real dt = 1e-12, time = 0.;
    amrex::Real miscVals[20] = {};
    const amrex::Real dx[3] = {1., 1., 1.};

    conservedToPrimitive(prim, cu);
    setBC(prim, cu);
    calculateTransportCoeffs(prim, eta, zeta, kappa, chi, D);
    calculateFlux(cu, prim, eta, zeta, kappa, chi, D, flux, stochFlux, cornx, corny, cornz, visccorn, rancorn,
                  geom, weights, dt);
    RK3step(cu, cup, cup2, cup3, prim, source, eta, zeta, kappa, chi, D, flux, stochFlux, cornx, corny, cornz,
            visccorn, rancorn, ranchem, geom, dt);
    compute_compressible_chemistry_source_CLE(dt, 1., prim, source, ranchem);
    evaluateStats(cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, miscStats, miscVals, 1, dx);
    WritePlotFile(1, time, geom, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, eta, kappa);
    WriteCheckPoint(1, time, 1, geom, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, miscStats, eta,
                    kappa);
    int step = 0, statsCount = 0;
    amrex::Real t = 0.;
    amrex::Geometry g2;
    ReadCheckPoint(step, t, statsCount, g2, cu, cuMeans, cuVars, prim, primMeans, primVars, spatialCross, miscStats,
                   eta, kappa);

    // staggered scheme
    std::array<MultiFab, 3> cumom, vel, faceflux, cenflux, stochface, stochcen;
    std::array<MultiFab, 2> ex, ey, ez, sx, sy, sz;
    TurbForcingComp turb;
    conservedToPrimitiveStag(prim, vel, cu, cumom);
    calculateFluxStag(cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, ex, ey, ez, cenflux, stochface, sx,
                      sy, sz, stochcen, geom, weights, dt);
    RK3stepStag(cu, cumom, prim, vel, source, eta, zeta, kappa, chi, D, faceflux, ex, ey, ez, cenflux, ranchem, geom,
                dt, 1, turb);
}

// Each entry point has the function type the reference declares for it
// (src_compressible/compressible_functions.H:23-135, src_compressible_stag/
// compressible_functions_stag.H:113-143, conservedPrimitiveConversionsStag.cpp:5-6): taking its
// address as a pointer of that type fails to compile if a parameter differs.
namespace ref_types {
using amrex::Geometry;
using amrex::MultiFab;
using amrex::Real;
using amrex::Vector;
using A3 = std::array<MultiFab, AMREX_SPACEDIM>;
using A2 = std::array<MultiFab, 2>;

void (*const calculateFlux_p)(const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&,
                              const MultiFab&, const MultiFab&, A3&, A3&, A3&, A3&, A3&, MultiFab&, MultiFab&,
                              const Geometry&, const Vector<Real>&, const Real) = &calculateFlux;
void (*const calculateTransportCoeffs_p)(const MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&,
                                         MultiFab&) = &calculateTransportCoeffs;
void (*const RK3step_p)(MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&,
                        MultiFab&, MultiFab&, MultiFab&, A3&, A3&, A3&, A3&, A3&, MultiFab&, MultiFab&, MultiFab&,
                        const Geometry&, const Real) = &RK3step;
void (*const conservedToPrimitive_p)(MultiFab&, const MultiFab&) = &conservedToPrimitive;
void (*const setBC_p)(MultiFab&, MultiFab&) = &setBC;
void (*const evaluateStats_p)(const MultiFab&, MultiFab&, MultiFab&, const MultiFab&, MultiFab&, MultiFab&,
                              MultiFab&, MultiFab&, Real*, const int, const Real*) = &evaluateStats;
void (*const WritePlotFile_p)(int, const Real, const Geometry&, const MultiFab&, const MultiFab&, const MultiFab&,
                              const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&,
                              const MultiFab&) = &WritePlotFile;
void (*const WriteCheckPoint_p)(int, const Real, int, const Geometry&, const MultiFab&, const MultiFab&,
                                const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&,
                                const MultiFab&, const MultiFab&, const MultiFab&) = &WriteCheckPoint;
void (*const ReadCheckPoint_p)(int&, Real&, int&, Geometry&, MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&,
                               MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&) = &ReadCheckPoint;
void (*const RK3stepStag_p)(MultiFab&, A3&, MultiFab&, A3&, MultiFab&, MultiFab&, MultiFab&, MultiFab&, MultiFab&,
                            MultiFab&, A3&, A2&, A2&, A2&, A3&, MultiFab&, const Geometry&, const Real, const int,
                            TurbForcingComp&) = &RK3stepStag;
void (*const calculateFluxStag_p)(const MultiFab&, const A3&, const MultiFab&, const A3&, const MultiFab&,
                                  const MultiFab&, const MultiFab&, const MultiFab&, const MultiFab&, A3&, A2&, A2&,
                                  A2&, A3&, A3&, A2&, A2&, A2&, A3&, const Geometry&, const Vector<Real>&,
                                  const Real) = &calculateFluxStag;
void (*const conservedToPrimitiveStag_p)(MultiFab&, A3&, MultiFab&, const A3&) = &conservedToPrimitiveStag;
// src_chemistry/chemistry_functions.H:20-21
void (*const chemistry_source_p)(Real, Real, MultiFab&, MultiFab&, MultiFab&) =
    &compute_compressible_chemistry_source_CLE;
#ifdef PLANT_MISMATCH
// a parameter changed on purpose: this must not compile
void (*const planted_p)(MultiFab&, const MultiFab&, const MultiFab&) = &conservedToPrimitive;
#endif
}  // namespace ref_types
