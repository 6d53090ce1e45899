// stag.hpp -- host launchers of the staggered-scheme kernels (stag_kernels.hip).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace fhdx {

// Every array one calculateFluxStag pass touches (fluxStag.cpp:5-21) plus its stress scratch:
// td (cells grown by 1): tau_stoch xx yy zz, tau xx yy zz; to[e] (edges xy, xz, yz): tau_stoch,
// tau.  ef in the reference's order edgeflux_x[0..1], _y[0..1], _z[0..1].
struct StagArrays {
    Arr cu, prim, eta, zeta, kappa, chi, D;
    Arr mom[3], vel[3];
    Arr ff[3], ef[6], cf[3];
    Arr td, to[3];
};

// The arrays of one stage combination: out/outm = f(u0/m0, us/ms, fluxes).  With chemistry
// (chem != 0) the cell update adds dt times the reaction source of the stage's prim, with the
// reaction noise A + w B of key K; source (nvars comps) and ranchem (nreaction) receive what
// the reference leaves in them when their pointers are set.
struct StagUpdate {
    Arr out, outm[3], u0, m0[3], us, ms[3];
    Arr ff[3], ef[6], cf[3];
    int chem;
    Arr prim, source, ranchem;
    NoiseKey K;
    double w;
};

// Periodic ghost fill: array a (cell: nodal -1, face: nodal d), components c0 .. c0+nc-1.
struct FillItem {
    Arr a;
    int nodal, c0, nc;
};
struct FillList {
    FillItem item[13];
    int g;  // ghost layers
};

// The transport arrays conservedToPrimitiveStag may also fill (valid cells) from the prim it
// computes: eta, zeta, kappa, chi, D and the device error flags.
struct StagTransport {
    Arr eta, zeta, kappa, chi, D;
    int* err;
};

// calculateFluxStag with the noise A + w B of the step's Philox sets (key K).
int stag_flux(const DevParams& P, const StagArrays& A, const NoiseKey& K, double w, double dt, hipStream_t s);
// Stage st (0, 1, 2) of RK3stepStag's combination.
int stag_update(const DevParams& P, const StagUpdate& U, int st, double dt, hipStream_t s);
// conservedToPrimitiveStag on the valid box (cons' density ghosts must be current); with tr set,
// also calculateTransportCoeffs of each valid cell's new prim (their ghosts are periodic images).
int stag_c2p(const DevParams& P, Arr prim, const Arr vel[3], Arr cons, const Arr mom[3], const StagTransport* tr,
             hipStream_t s);
// FillBoundary(periodicity) of the first `count` items.
int stag_fill(const DevParams& P, const FillList& L, int count, hipStream_t s);

}  // namespace fhdx
