// capi_stag.hip -- the C ABI of the staggered scheme (include/fhdx.h, "staggered scheme").
//
// Host C++ over stag_kernels.hip: shape checks of the reference's staggered arrays
// (main_driver.cpp:616-631, 1023-1050 of src_compressible_stag), the stage driver of
// RK3stepStag (timeStepStag.cpp:13-1245) and the context-owned stage arrays (cup, cup2 and
// their face momenta, allocated once on first use instead of per call as the reference does).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "stag.hpp"

using namespace fhdx;

struct StagWorkspace {
    std::vector<double*> blocks;
    fhdx_fab cup, cup2, cupm[3], cup2m[3];
    fhdx_fab td, to[3];
};

namespace fhdx {

void stag_workspace_free(StagWorkspace* w)
{
    if (!w) return;
    for (double* p : w->blocks) (void)hipFree(p);
    delete w;
}

}  // namespace fhdx

namespace {

hipStream_t SS(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Cell FAB of the domain with g ghost layers and at least nc components.
bool stag_cell(const fhdx_ctx* c, const fhdx_fab& f, int nc, int g, const char* what)
{
    bool ok = f.p != nullptr && f.ncomp >= nc;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == -g && f.hi[d] == c->p.n_cells[d] - 1 + g;
    if (!ok) set_error(std::string("bad staggered cell array shape: ") + what);
    return ok;
}

// Face FAB nodal in dir with g ghost layers.
bool stag_face(const fhdx_ctx* c, const fhdx_fab& f, int dir, int nc, int g, const char* what)
{
    bool ok = f.p != nullptr && f.ncomp >= nc;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == -g && f.hi[d] == c->p.n_cells[d] - 1 + g + (d == dir);
    if (!ok) set_error(std::string("bad staggered face array shape: ") + what);
    return ok;
}

// Edge FAB of type e (xy, xz, yz): nodal along its two axes, no ghosts.
bool stag_edge(const fhdx_ctx* c, const fhdx_fab& f, int e, const char* what)
{
    const int a = (e == 2) ? 1 : 0, b = (e == 0) ? 1 : 2;
    bool ok = f.p != nullptr && f.ncomp >= 1;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == 0 && f.hi[d] == c->p.n_cells[d] - 1 + (d == a || d == b);
    if (!ok) set_error(std::string("bad staggered edge array shape: ") + what);
    return ok;
}

// edgeflux_x[0] xy, [1] xz; edgeflux_y[0] xy, [1] yz; edgeflux_z[0] xz, [1] yz
constexpr int kEdgeFluxType[6] = {0, 1, 0, 2, 1, 2};

int stag_scope(const fhdx_ctx* c, const char* who)
{
    for (int d = 0; d < 3; ++d)
        if (!c->dp.periodic[d]) {
            set_error(std::string(who) + ": the staggered scheme is implemented for periodic boxes only");
            return FHDX_ERR_UNSUPPORTED;
        }
    return FHDX_OK;
}

int alloc(StagWorkspace* w, fhdx_fab& f, int l0, int l1, int l2, int h0, int h1, int h2, int nc)
{
    f.lo[0] = l0; f.lo[1] = l1; f.lo[2] = l2;
    f.hi[0] = h0; f.hi[1] = h1; f.hi[2] = h2;
    f.ncomp = nc;
    f.pad = 0;
    const size_t n = (size_t)(h0 - l0 + 1) * (h1 - l1 + 1) * (h2 - l2 + 1) * nc;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(p, 0, n * sizeof(double));
    if (e != hipSuccess) {
        if (p) (void)hipFree(p);
        set_error(std::string("staggered workspace allocation failed: ") + hipGetErrorString(e));
        return FHDX_ERR_NOMEM;
    }
    f.p = (double*)p;
    w->blocks.push_back(f.p);
    return FHDX_OK;
}

int ensure_stag(fhdx_ctx* c)
{
    if (c->stag) return FHDX_OK;
    StagWorkspace* w = new StagWorkspace();
    const int g = c->p.ngc, nv = c->dp.nvars;
    const int n0 = c->p.n_cells[0], n1 = c->p.n_cells[1], n2 = c->p.n_cells[2];
    int r = alloc(w, w->cup, -g, -g, -g, n0 - 1 + g, n1 - 1 + g, n2 - 1 + g, nv);
    if (!r) r = alloc(w, w->cup2, -g, -g, -g, n0 - 1 + g, n1 - 1 + g, n2 - 1 + g, nv);
    for (int d = 0; d < 3 && !r; ++d) {
        r = alloc(w, w->cupm[d], -g, -g, -g, n0 - 1 + g + (d == 0), n1 - 1 + g + (d == 1), n2 - 1 + g + (d == 2), 1);
        if (!r) r = alloc(w, w->cup2m[d], -g, -g, -g, n0 - 1 + g + (d == 0), n1 - 1 + g + (d == 1), n2 - 1 + g + (d == 2), 1);
        if (!r) {
            const int a = (d == 2) ? 1 : 0, b = (d == 0) ? 1 : 2;
            r = alloc(w, w->to[d], 0, 0, 0, n0 - 1 + (a == 0 || b == 0), n1 - 1 + (a == 1 || b == 1),
                      n2 - 1 + (a == 2 || b == 2), 2);
        }
    }
    if (!r) r = alloc(w, w->td, -1, -1, -1, n0, n1, n2, 6);
    if (r) {
        stag_workspace_free(w);
        return r;
    }
    c->stag = w;
    return FHDX_OK;
}

// FillBoundary(periodicity) of a stage's output (timeStepStag.cpp:526-535): state, momenta,
// velocities, prim; with tr, also the transport arrays the conversion filled (a ghost cell's
// coefficients are those of its periodic image, whose prim it holds).
int fill_stage_end(const fhdx_ctx* c, Arr cons, const Arr mom[3], const Arr vel[3], Arr prim, const StagTransport* tr,
                   hipStream_t s)
{
    FillList L{};
    L.g = c->p.ngc;
    L.item[0] = FillItem{cons, -1, 0, c->dp.nvars};
    for (int d = 0; d < 3; ++d) {
        L.item[1 + d] = FillItem{mom[d], d, 0, 1};
        L.item[4 + d] = FillItem{vel[d], d, 0, 1};
    }
    L.item[7] = FillItem{prim, -1, 0, c->dp.nprim};
    int count = 8;
    if (tr) {
        const int NS = c->p.nspecies;
        L.item[8] = FillItem{tr->eta, -1, 0, 1};
        L.item[9] = FillItem{tr->zeta, -1, 0, 1};
        L.item[10] = FillItem{tr->kappa, -1, 0, 1};
        L.item[11] = FillItem{tr->chi, -1, 0, NS};
        L.item[12] = FillItem{tr->D, -1, 0, NS * NS};
        count = 13;
    }
    return stag_fill(c->dp, L, count, s);
}

// The density ghosts conservedToPrimitiveStag reads (face velocities at index 0).
int fill_density(const fhdx_ctx* c, Arr cons, hipStream_t s)
{
    FillList L{};
    L.g = c->p.ngc;
    L.item[0] = FillItem{cons, -1, 0, 1};
    return stag_fill(c->dp, L, 1, s);
}

int c2p_with_ghosts(const fhdx_ctx* c, Arr prim, const Arr vel[3], Arr cons, const Arr mom[3], const StagTransport* tr,
                    hipStream_t s)
{
    int r = fill_density(c, cons, s);
    if (!r) r = stag_c2p(c->dp, prim, vel, cons, mom, tr, s);
    if (!r) r = fill_stage_end(c, cons, mom, vel, prim, tr, s);
    return r;
}

}  // namespace

extern "C" {

int fhdx_cons_to_prim_stag(fhdx_ctx* c, fhdx_fab prim, const fhdx_fab vel[3], fhdx_fab cu, const fhdx_fab cumom[3],
                           void* stream)
{
    if (!c || !vel || !cumom) return FHDX_ERR_ARG;
    if (int r = stag_scope(c, "fhdx_cons_to_prim_stag")) return r;
    const int g = c->p.ngc;
    bool ok = stag_cell(c, prim, c->dp.nprim, g, "prim") && stag_cell(c, cu, c->dp.nvars, g, "cu");
    for (int d = 0; d < 3 && ok; ++d)
        ok = stag_face(c, vel[d], d, 1, g, "vel") && stag_face(c, cumom[d], d, 1, g, "cumom");
    if (!ok) return FHDX_ERR_ARG;
    Arr v[3], m[3];
    for (int d = 0; d < 3; ++d) {
        v[d] = make_arr(vel[d]);
        m[d] = make_arr(cumom[d]);
    }
    return c2p_with_ghosts(c, make_arr(prim), v, make_arr(cu), m, nullptr, SS(stream));
}

static int check_stag_step_args(fhdx_ctx* c, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3],
                                fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                                const fhdx_fab faceflux[3], const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3],
                                double dt)
{
    if (!cumom || !vel || !faceflux || !edgeflux || !cenflux || !(dt > 0.)) return FHDX_ERR_ARG;
    const int g = c->p.ngc, NS = c->p.nspecies, nv = c->dp.nvars;
    bool ok = stag_cell(c, cu, nv, g, "cu") && stag_cell(c, prim, c->dp.nprim, g, "prim") &&
              stag_cell(c, eta, 1, g, "eta") && stag_cell(c, zeta, 1, g, "zeta") &&
              stag_cell(c, kappa, 1, g, "kappa") && stag_cell(c, chi, NS, g, "chi") && stag_cell(c, D, NS * NS, g, "D");
    for (int d = 0; d < 3 && ok; ++d)
        ok = stag_face(c, cumom[d], d, 1, g, "cumom") && stag_face(c, vel[d], d, 1, g, "vel") &&
             stag_face(c, faceflux[d], d, nv + 4, 0, "faceflux") && stag_cell(c, cenflux[d], 1, 1, "cenflux");
    for (int e = 0; e < 6 && ok; ++e) ok = stag_edge(c, edgeflux[e], kEdgeFluxType[e], "edgeflux");
    return ok ? FHDX_OK : FHDX_ERR_ARG;
}

static StagArrays stag_arrays(fhdx_ctx* c, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3],
                              fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                              const fhdx_fab faceflux[3], const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3])
{
    StagArrays A;
    A.cu = make_arr(cu);
    A.prim = make_arr(prim);
    A.eta = make_arr(eta);
    A.zeta = make_arr(zeta);
    A.kappa = make_arr(kappa);
    A.chi = make_arr(chi);
    A.D = make_arr(D);
    for (int d = 0; d < 3; ++d) {
        A.mom[d] = make_arr(cumom[d]);
        A.vel[d] = make_arr(vel[d]);
        A.ff[d] = make_arr(faceflux[d]);
        A.cf[d] = make_arr(cenflux[d]);
        A.to[d] = make_arr(c->stag->to[d]);
    }
    for (int e = 0; e < 6; ++e) A.ef[e] = make_arr(edgeflux[e]);
    A.td = make_arr(c->stag->td);
    return A;
}

int fhdx_calculate_flux_stag(fhdx_ctx* c, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3],
                             fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                             const fhdx_fab faceflux[3], const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3],
                             double weight, double dt, uint64_t seed, uint64_t step, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (int r = stag_scope(c, "fhdx_calculate_flux_stag")) return r;
    if (int r = check_stag_step_args(c, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, edgeflux, cenflux, dt))
        return r;
    if (int r = ensure_stag(c)) return r;
    const StagArrays A = stag_arrays(c, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, edgeflux, cenflux);
    return stag_flux(c->dp, A, make_noise_key(seed, step, 0), weight, dt, SS(stream));
}

int fhdx_rk3_step_stag(fhdx_ctx* c, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3],
                       fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                       const fhdx_fab faceflux[3], const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3],
                       fhdx_fab source, fhdx_fab ranchem, double dt, uint64_t seed, uint64_t step, void* stream)
{
    if (!c) return FHDX_ERR_ARG;
    if (int r = stag_scope(c, "fhdx_rk3_step_stag")) return r;
    if (int r = check_stag_step_args(c, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, edgeflux, cenflux, dt))
        return r;
    const bool chem = c->p.nreaction > 0;
    if (chem) {
        if (!c->chem_set) {
            set_error("nreaction > 0: call fhdx_ctx_set_chemistry before stepping");
            return FHDX_ERR_ARG;
        }
        const int g = c->p.ngc;
        if ((source.p && !stag_cell(c, source, c->dp.nvars, g, "source")) ||
            (ranchem.p && !stag_cell(c, ranchem, c->p.nreaction, g, "ranchem")))
            return FHDX_ERR_ARG;
    }
    if (int r = ensure_stag(c)) return r;
    hipStream_t s = SS(stream);
    const NoiseKey K = make_noise_key(seed, step, 0);
    double wts[3];
    stage_weights(wts);
    StagArrays A = stag_arrays(c, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, edgeflux, cenflux);
    const StagWorkspace* w = c->stag;
    const Arr u0 = make_arr(cu), c1 = make_arr(w->cup), c2 = make_arr(w->cup2);
    Arr m0[3], m1[3], m2[3];
    for (int d = 0; d < 3; ++d) {
        m0[d] = make_arr(cumom[d]);
        m1[d] = make_arr(w->cupm[d]);
        m2[d] = make_arr(w->cup2m[d]);
    }
    // calculateTransportCoeffs of each stage's prim (timeStepStag.cpp): on the caller's prim before
    // stage 0, then inside the conversion that ends stages 0 and 1.  The last conversion leaves
    // them alone, so the step returns the stage-2 coefficients, as the reference does.
    const StagTransport tc{A.eta, A.zeta, A.kappa, A.chi, A.D, c->d_err};
    for (int st = 0; st < 3; ++st) {
        int r = 0;
        if (st == 0 && (r = ref_transport(c->dp, A.prim, A.eta, A.zeta, A.kappa, A.chi, A.D, c->d_err, s))) return r;
        const Arr us = st == 0 ? u0 : st == 1 ? c1 : c2;
        const Arr* ms = st == 0 ? m0 : st == 1 ? m1 : m2;
        const Arr out = st == 0 ? c1 : st == 1 ? c2 : u0;
        const Arr* outm = st == 0 ? m1 : st == 1 ? m2 : m0;
        A.cu = us;
        for (int d = 0; d < 3; ++d) A.mom[d] = ms[d];
        if ((r = stag_flux(c->dp, A, K, wts[st], dt, s))) return r;
        StagUpdate U;
        U.chem = chem;
        U.prim = A.prim;
        U.source = source.p && chem ? make_arr(source) : Arr{};
        U.ranchem = ranchem.p && chem ? make_arr(ranchem) : Arr{};
        U.K = K;
        U.w = wts[st];
        U.out = out;
        U.u0 = u0;
        U.us = us;
        for (int d = 0; d < 3; ++d) {
            U.outm[d] = outm[d];
            U.m0[d] = m0[d];
            U.ms[d] = ms[d];
            U.ff[d] = A.ff[d];
            U.cf[d] = A.cf[d];
        }
        for (int e = 0; e < 6; ++e) U.ef[e] = A.ef[e];
        if ((r = stag_update(c->dp, U, st, dt, s))) return r;
        if ((r = c2p_with_ghosts(c, A.prim, A.vel, out, outm, st < 2 ? &tc : nullptr, s))) return r;
    }
    return FHDX_OK;
}

}  // extern "C"
