// kernels.hpp -- host launchers of the device kernels (ref_kernels.hip, fused_kernels.hip).
#pragma once

#include "common.hpp"

namespace fhdx {

// Views of every array calculateFlux touches (flux.cpp:4-16).
// With weigh set, ran[] and rancorn are outputs: the stochastic pass reads the step's noise sets
// A, B and stores ran = 1*A + w*B (component 0 zero) for each face it computes, and the node pass
// rancorn likewise (the reference's MultiFab::LinComb, timeStep.cpp:99-107), so no separate
// weighting launch runs; otherwise ran[] and rancorn hold the weighted noise already.
struct FluxArgs {
    Arr cons, prim, eta, zeta, kappa, chi, D;
    Arr flux[3], ran[3];
    Arr corn[3][3];  // [derivative direction][velocity component]: cornx/corny/cornz
    Arr visccorn, rancorn;
    double dt;
    bool weigh = false;
    Arr nA[3], nB[3], rA, rB;
    double w = 0.;
};

int ref_c2p(const DevParams& P, Arr prim, Arr cons, hipStream_t s);
int ref_transport(const DevParams& P, Arr prim, Arr eta, Arr zeta, Arr kappa, Arr chi, Arr Dm, int* err,
                  hipStream_t s);
int ref_fill_boundary(const DevParams& P, Arr a, int nc, hipStream_t s);
// The same for two arrays in one launch (a stage's state and its primitive variables).
int ref_fill_boundary2(const DevParams& P, Arr a, int nca, Arr b, int ncb, hipStream_t s);
// setBC: the ghost cells outside walls, faces lo x .. hi z in turn (no-op when periodic).
int ref_set_bc(const DevParams& P, Arr prim, Arr cons, hipStream_t s);
int fill_value(double* p, long n, double v, hipStream_t s);
// flux_ncomp: the components of each flux FAB (>= nvars + 4)
int ref_calculate_flux(const DevParams& P, const FluxArgs& a, int flux_ncomp[3], hipStream_t s);
// source: the source MultiFab (p == nullptr: zero)
int ref_stage(const DevParams& P, Arr out, Arr u0, Arr us, const Arr flux[3], Arr source, double dt, int mode,
              hipStream_t s);
// compute_compressible_chemistry_source_CLE over the valid cells; reaction noise sets A, B.
int ref_chem_source(const DevParams& P, Arr prim, Arr source, Arr ranchem, double dt, hipStream_t s);
int ref_fill_chem_noise(const DevParams& P, const NoiseKey& K, Arr A, Arr B, hipStream_t s);
int ref_lincomb(double* dst, const double* a, const double* b, long npts, int nc, int c0, double w,
                hipStream_t s);
// The same for four arrays at once (a stage's three face noise arrays and the node array).
int ref_lincomb4(double* const dst[4], const double* const a[4], const double* const b[4], const long npts[4],
                 const int nc[4], const int c0[4], double w, hipStream_t s);
int ref_fill_noise(const DevParams& P, const NoiseKey& K, const Arr A[3], const Arr B[3], Arr rA, Arr rB,
                   hipStream_t s);

// evaluateStats passes (stats.hip).  yz: n0*19 doubles, part: nparts*n0*19, misc: 20 (device).
struct StatsArgs {
    Arr cu, cm, cv, prim, pm, pv, sc, ms;
    double *yz, *part, *misc;
    int nparts, steps, cross_cell;
};
int stats_launch(const DevParams& P, const StatsArgs& a, hipStream_t s);

}  // namespace fhdx
