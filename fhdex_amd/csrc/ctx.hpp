// ctx.hpp -- the context behind the C ABI's fhdx_ctx handle (capi.hip, capi_stag.hip).
#pragma once

#include <vector>

#include "fused.hpp"
#include "kernels.hpp"

// The staggered scheme's stage arrays (capi_stag.hip).
struct StagWorkspace;

namespace fhdx {
void stag_workspace_free(StagWorkspace* w);
}

// Device buffers a context owns for the reference-shaped RK3 step: the step's two noise sets
// when the caller supplies none, and the chemistry arrays.  Every other array of that step is
// the caller's (the reference's RK3step argument list).
struct RefWorkspace {
    bool ready = false;
    std::vector<double*> blocks;
    fhdx_fab sA[3], sB[3], rA, rB;
    fhdx_fab source, ranA, ranB, ran;  // chemistry (nreaction > 0): source, reaction noise A, B, A + w B
};

struct fhdx_ctx {
    fhdx_params p;
    fhdx::DevParams dp;
    int device = 0;
    int* d_err = nullptr;  // device-side error flags (bit 0: mass fraction outside (0,1))
    RefWorkspace ref;
    fhdx::FusedWorkspace fused;
    double *st_yz = nullptr, *st_part = nullptr, *st_misc = nullptr;  // evaluateStats scratch
    int st_nparts = 0;
    bool chem_set = false;  // fhdx_ctx_set_chemistry called (required when p.nreaction > 0)
    StagWorkspace* stag = nullptr;  // staggered-scheme stage arrays (capi_stag.hip), on first use
};

