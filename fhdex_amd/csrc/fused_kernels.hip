// fused_kernels.hip -- the fast RK3 step.
#include "fused.hpp"
#include "kernels.hpp"

namespace fhdx {

void fused_free(FusedWorkspace& w)
{
    for (void* p : w.blocks) (void)hipFree(p);
    w.blocks.clear();
    w.ready = false;
}

static int alloc(FusedWorkspace& w, fhdx_fab& f, int l0, int l1, int l2, int h0, int h1, int h2, int nc)
{
    f.lo[0] = l0; f.lo[1] = l1; f.lo[2] = l2;
    f.hi[0] = h0; f.hi[1] = h1; f.hi[2] = h2;
    f.ncomp = nc;
    f.pad = 0;
    size_t n = (size_t)(h0 - l0 + 1) * (h1 - l1 + 1) * (h2 - l2 + 1) * nc;
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of fast-path workspace failed");
        return FHDX_ERR_NOMEM;
    }
    (void)hipMemset(p, 0, n * sizeof(double));
    f.p = (double*)p;
    w.blocks.push_back(p);
    return FHDX_OK;
}

static int ensure(const DevParams& P, FusedWorkspace& w)
{
    if (w.ready) return FHDX_OK;
    const int g = P.ngc, NS = P.ns, nv = P.nvars;
    const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    int r = 0;
    auto cell = [&](fhdx_fab& f, int nc) { if (!r) r = alloc(w, f, -g, -g, -g, n0 - 1 + g, n1 - 1 + g, n2 - 1 + g, nc); };
    auto face = [&](fhdx_fab& f, int d, int nc) {
        if (!r) r = alloc(w, f, 0, 0, 0, n0 - 1 + (d == 0), n1 - 1 + (d == 1), n2 - 1 + (d == 2), nc);
    };
    auto node = [&](fhdx_fab& f) { if (!r) r = alloc(w, f, 0, 0, 0, n0, n1, n2, 1); };
    cell(w.cup, nv); cell(w.cup2, nv); cell(w.eta, 1); cell(w.zeta, 1); cell(w.kappa, 1);
    cell(w.chi, NS); cell(w.D, NS * NS);
    for (int d = 0; d < 3; ++d) {
        face(w.flux[d], d, nv + 4); face(w.stoch[d], d, nv); face(w.sA[d], d, nv); face(w.sB[d], d, nv);
        for (int a = 0; a < 3; ++a) node(w.corn[a][d]);
    }
    node(w.visccorn); node(w.rancorn); node(w.rA); node(w.rB);
    if (r) return r;
    FHDX_HIP_TRY(hipDeviceSynchronize());
    w.ready = true;
    return FHDX_OK;
}

static long npts(const fhdx_fab& f)
{
    return (long)(f.hi[0] - f.lo[0] + 1) * (f.hi[1] - f.lo[1] + 1) * (f.hi[2] - f.lo[2] + 1);
}

int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s)
{
    (void)p;
    int r = ensure(P, w);
    if (r) return r;
    const int nv = P.nvars;
    const fhdx_fab *A = sA, *B = sB, *RA = rA, *RB = rB;
    if (!A) {
        Arr a3[3], b3[3];
        for (int d = 0; d < 3; ++d) { a3[d] = make_arr(w.sA[d]); b3[d] = make_arr(w.sB[d]); }
        if ((r = ref_fill_noise(P, K, a3, b3, make_arr(w.rA), make_arr(w.rB), s))) return r;
        A = w.sA; B = w.sB; RA = &w.rA; RB = &w.rB;
    }
    // state prim of cu (the caller's prim may be stale), then the three stages
    if ((r = ref_fill_boundary(P, make_arr(cu), nv, s))) return r;
    if ((r = ref_c2p(P, make_arr(prim), make_arr(cu), s))) return r;
    if ((r = ref_fill_boundary(P, make_arr(prim), P.nprim, s))) return r;
    double ws[3];
    stage_weights(ws);
    fhdx_fab in[3] = {cu, w.cup, w.cup2};
    fhdx_fab out[3] = {w.cup, w.cup2, cu};
    FluxArgs a;
    a.prim = make_arr(prim); a.eta = make_arr(w.eta); a.zeta = make_arr(w.zeta); a.kappa = make_arr(w.kappa);
    a.chi = make_arr(w.chi); a.D = make_arr(w.D);
    Arr fl[3];
    for (int d = 0; d < 3; ++d) {
        a.flux[d] = fl[d] = make_arr(w.flux[d]);
        a.ran[d] = make_arr(w.stoch[d]);
        for (int q = 0; q < 3; ++q) a.corn[q][d] = make_arr(w.corn[q][d]);
    }
    a.visccorn = make_arr(w.visccorn);
    a.rancorn = make_arr(w.rancorn);
    a.dt = dt;
    for (int st = 0; st < 3; ++st) {
        if ((r = ref_transport(P, a.prim, a.eta, a.zeta, a.kappa, a.chi, a.D, d_err, s))) return r;
        for (int d = 0; d < 3; ++d)
            if ((r = ref_lincomb(w.stoch[d].p, A[d].p, B[d].p, npts(w.stoch[d]), nv, 1, ws[st], s))) return r;
        if ((r = ref_lincomb(w.rancorn.p, RA->p, RB->p, npts(w.rancorn), 1, 0, ws[st], s))) return r;
        a.cons = make_arr(in[st]);
        if ((r = ref_calculate_flux(P, a, s))) return r;
        if ((r = ref_stage(P, make_arr(out[st]), make_arr(cu), make_arr(in[st]), fl, dt, st, s))) return r;
        if ((r = ref_c2p(P, a.prim, make_arr(out[st]), s))) return r;
        if ((r = ref_fill_boundary(P, make_arr(out[st]), nv, s))) return r;
        if ((r = ref_fill_boundary(P, a.prim, P.nprim, s))) return r;
    }
    return FHDX_OK;
}

}  // namespace fhdx
