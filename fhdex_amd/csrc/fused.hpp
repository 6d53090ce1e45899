// fused.hpp -- the fast RK3 step (fhdx_rk3_step / fhdx_rk3_step_injected).
#pragma once

#include <vector>

#include "common.hpp"

namespace fhdx {

// Device workspace owned by a context for the fast path (allocated on first use).
struct FusedWorkspace {
    bool ready = false;
    std::vector<void*> blocks;
    fhdx_fab cup, cup2, eta, zeta, kappa, chi, D;
    fhdx_fab flux[3], stoch[3], sA[3], sB[3], corn[3][3];
    fhdx_fab visccorn, rancorn, rA, rB;
};

void fused_free(FusedWorkspace& w);

// One RK3 step: cu advanced in place, prim recomputed, ghosts filled.  Noise from Philox
// (key K) unless sA is non-null (injected sets A/B).
int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s);

}  // namespace fhdx
