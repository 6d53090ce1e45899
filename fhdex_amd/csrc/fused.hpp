// fused.hpp -- the fast RK3 step (fhdx_rk3_step / fhdx_rk3_step_injected).
#pragma once

#include <vector>

#include "common.hpp"

namespace fhdx {

// Per-cell record written once per stage (after conservedToPrimitive and the transport
// coefficients) and read by the node and face kernels; SoA components of a ghost-padded FAB.
struct RecLayout {
    int ns;
    static constexpr int U = 0, V = 1, W = 2, T = 3, P = 4, ETA = 5, KAP = 6, ZETA = 7, Y = 8;
    __host__ __device__ int X() const { return 8 + ns; }
    __host__ __device__ int YN() const { return 8 + 2 * ns; }    // clamped, renormalised Y
    __host__ __device__ int CHI() const { return 8 + 3 * ns; }
    __host__ __device__ int D() const { return 8 + 4 * ns; }
    __host__ __device__ int ncomp() const { return 8 + 4 * ns + ns * ns; }
};

// Node values: corner velocity gradients [derivative dir][component] (flux.cpp:1684-1703),
// visccorn (:1705) and the stage-weighted node noise (rancorn).
constexpr int NODE_VC = 9, NODE_W = 10, NODE_NCOMP = 11;

// Device workspace owned by a context for the fast path (allocated on first use).
struct FusedWorkspace {
    bool ready = false;
    std::vector<void*> blocks;
    fhdx_fab cup, cup2, rec, node;
    fhdx_fab F[3];
    fhdx_fab sA[3], sB[3], rA, rB;  // Philox noise of the step
};

void fused_free(FusedWorkspace& w);

// One RK3 step: cu advanced in place, prim recomputed, ghosts of both filled.  Noise from
// Philox (key K) unless sA is non-null (caller's A/B sets).
int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s);

}  // namespace fhdx
