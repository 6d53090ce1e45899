// fused_kernels.hip -- the fast RK3 step on gfx950.
//
// Noise: each face and node evaluates its Philox A/B values in place, per stage, or (a tunable)
// k_noise writes them to arrays once per step; caller-supplied noise sets are copied into the
// grid layout and read.  Per stage:
//   k_cells    conservedToPrimitive + transport coefficients, once per valid cell, into a
//              ghost-padded record (u, v, w, T, p, eta, kappa, zeta, Y, X, clamped Y, chi, D);
//   k_nodes    corner velocity gradients, visccorn and the stage-weighted node noise;
//   k_faces    all four flux families (stochastic, diffusive, corner, hyperbolic) of the three
//              lower faces of each cell, accumulated in registers in the reference's
//              per-component order with the energy fold;
//   k_update   divergence + RK3 stage combination (+ the next stage's cell records after stages 1
//              and 2, primitive variables after stage 3).
// Every array of the box shares one layout (Grid), so a kernel computes each cell offset once
// and addresses all arrays as scalar base + 32-bit offset.  Records, stage states and prim are
// written with their periodic ghost images by the producing kernel; only the caller's cu gets
// a ghost fill, once per step.  Every value is computed with the reference's operation order, so the step is
// bit-identical to the reference-shaped path.
#include "fused.hpp"
#include "kernels.hpp"
#include "stage_bodies.hpp"

namespace fhdx {

constexpr int FB = 256;

__device__ inline void decode3(long t, int nx, int ny, int& i, int& j, int& k)
{
    i = (int)(t % nx);
    long r = t / nx;
    j = (int)(r % ny);
    k = (int)(r / ny);
}

// ---- periodic ghost fill: one thread per ghost cell --------------------------------------

__global__ void __launch_bounds__(FB) k_fill_ghosts(Grid G, double* a, int nc, int g, int n0, int n1, int n2)
{
    ghost_body(G.view(a), nc, g, n0, n1, n2, (long)blockIdx.x * FB + threadIdx.x);
}

static int fill_ghosts(const DevParams& P, const Grid& G, double* a, int nc, hipStream_t s)
{
    const long G0 = P.n[0] + 2 * P.ngc, G1 = P.n[1] + 2 * P.ngc, g2 = 2L * P.ngc;
    const long n = g2 * G0 * G1 + g2 * G0 * P.n[2] + g2 * P.n[1] * P.n[2];
    k_fill_ghosts<<<grid_for(n, FB), FB, 0, s>>>(G, a, nc, P.ngc, P.n[0], P.n[1], P.n[2]);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

// ---- valid-region copy of caller noise (face / node FABs) into the grid layout -------------

__global__ void __launch_bounds__(FB) k_copy_valid(Arr src, Grid G, double* dst, int nc, int n0, int n1, int n2)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)n0 * n1 * n2) return;
    int i, j, k;
    decode3(t, n0, n1, i, j, k);
    const Arr d = G.view(dst);
    for (int c = 0; c < nc; ++c) d(i, j, k, c) = src(i, j, k, c);
}

// The z levels of the block that owns tile cell (i, j): k runs over [k0, min(k0 + kz, nz)).  k is
// wave-uniform; taking it through readfirstlane keeps every address scalar.
// ZM false: one level (kz == 1), no loop, so the compiler keeps nothing live across levels.
#define TILE_LEVELS(TM, ...)                                             \
    int i, j, k0;                                                        \
    if (!tile_cell(TM, P.n[0], P.n[1], i, j, k0)) return;                \
    if (!ZM) {                                                           \
        const int k = k0;                                                \
        __VA_ARGS__                                                      \
    } else {                                                             \
        const int k1 = min(k0 + (TM).kz, P.n[2]);                        \
        for (int kl = k0; kl < k1; ++kl) {                               \
            const int k = __builtin_amdgcn_readfirstlane(kl);            \
            __VA_ARGS__                                                  \
        }                                                                \
    }

// ---- K0: the step's noise into arrays (when not evaluated in place): one thread per cell, its
// three lower faces and its node; the same values the faces and nodes evaluate in place.

template <int NV>
__global__ void __launch_bounds__(FB) k_noise(DevParams P, NoiseKey K, Grid G, double* A0, double* A1, double* A2,
                                              double* B0, double* B1, double* B2, double* rA, double* rB)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    const unsigned cb = G.off(i, j, k);
    const uint32_t lin = (uint32_t)t;
    double* As[3] = {A0, A1, A2};
    double* Bs[3] = {B0, B1, B2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const Fld A = G.fld(As[d]), B = G.fld(Bs[d]);
#pragma unroll
        for (int c = 1; c < NV; ++c) {
            const double sc = (c <= 3) ? P.var_mom : (c == 4) ? P.var_ener : P.var_mass;
            double za, zb;
            noise_ab(K, (uint32_t)(16 * d + c), lin, za, zb);
            A(cb, c) = za * sc;
            B(cb, c) = zb * sc;
        }
    }
    double za, zb;
    noise_ab(K, NODE_FIELD, lin, za, zb);
    G.fld(rA)(cb, 0) = za * P.var_mom;
    G.fld(rB)(cb, 0) = zb * P.var_mom;
}

// ---- K1: cell records ------------------------------------------------------------------

template <int NS, bool VW>
__global__ void __launch_bounds__(FB) k_cells(DevParams P, Grid G, double* U, double* R, int* err)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    cell_body<NS, VW>(P, G.fld(U), G.fld(R), err, i, j, k, G.off(i, j, k));
}

// ---- tiles: 2-D (x, y) blocks of 256 threads, one z level per block, region-major order ----
//
// A tile is tw x th cells (tw = 1 << tx_shift = min(64, next power of two >= nx), th = 256/tw).
// A region is a column of rh tiles stacked in y.  Block numbers run over the rh tiles of one z
// level of a region, then up z, then to the next region.  Workgroup b runs on XCD b % 8, and XCD
// x receives the contiguous number range x*nb/8 .. (x+1)*nb/8: the blocks an XCD holds at once
// sit on neighbouring z levels of one region, so the z-1 / z+1 and y-1 neighbours a face or node
// reads were fetched into that XCD's L2 a few blocks earlier.  With kz > 1 a block marches
// through kz consecutive z levels (TILE_LEVELS), reusing the level below from cache.

struct TileMap {
    int tx_shift;  // tile width 1 << tx_shift (x), height 256 >> tx_shift (y)
    int rh;        // tiles per region (y)
    int nrx, nry;  // regions in x and y
    int n2;        // z chunks of kz levels
    int kz;        // z levels a block marches through
    unsigned blocks() const { return (unsigned)((long)nrx * nry * n2 * rh); }
};

static TileMap make_tiles(int n0, int n1, int n2, int rh, int tw_max, int kz)
{
    TileMap m;
    m.kz = kz < 1 ? 1 : kz;
    n2 = (n2 + m.kz - 1) / m.kz;
    int smax = 6;
    while (smax > 4 && (1 << smax) > tw_max) --smax;
    int s = 0;
    while ((1 << s) < n0 && s < smax) ++s;
    m.tx_shift = s;
    const int tx = 1 << s, ty = FB >> s;
    const int nty = (n1 + ty - 1) / ty;
    m.rh = rh < 1 ? 1 : (rh > nty ? nty : rh);
    m.nrx = (n0 + tx - 1) / tx;
    m.nry = (nty + m.rh - 1) / m.rh;
    m.n2 = n2;
    return m;
}

__device__ inline bool tile_cell(const TileMap& m, int n0, int n1, int& i, int& j, int& k)
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int xcd = b & 7, r = b >> 3, per = nb >> 3, rem = nb & 7;
    const int t = xcd * per + (xcd < rem ? xcd : rem) + r;
    const int col = m.n2 * m.rh;
    const int region = t / col, q = t - region * col;
    k = q / m.rh;
    const int sub = q - k * m.rh;
    k *= m.kz;  // first z level of the block
    const int rx = region % m.nrx, ry = region / m.nrx;
    i = (rx << m.tx_shift) + (threadIdx.x & ((1 << m.tx_shift) - 1));
    j = (ry * m.rh + sub) * (FB >> m.tx_shift) + (threadIdx.x >> m.tx_shift);
    return i < n0 && j < n1;
}

// ---- K2: nodes (flux.cpp:1627-1713) and the stochastic edge terms of each node -------------

template <bool INL, bool ZM>
__global__ void __launch_bounds__(FB) k_nodes(DevParams P, TileMap tm, Grid G, double* R, double* N, double* rA,
                                              double* rB, NoiseKey K, double w)
{
    TILE_LEVELS(tm, node_body<INL>(P, G.fld(R), G.fld(N), G.fld(rA), G.fld(rB), K, w, i, j, k, G.off(i, j, k));)
}

// ---- K3: the three lower faces of each cell (flux.cpp:71-2372) -------------------------------

struct FaceArgs {
    double *Us, *R, *N;
    double *A[3], *B[3], *F[3];
    NoiseKey K;
    double w, volinv, dtinv, dt;
};

template <int NS, bool INL, bool ZM>
__global__ void __launch_bounds__(FB) k_faces(DevParams P, TileMap tm, Grid G, FaceArgs a)
{
    constexpr int nv = NS + 5;
    const Fld Us = G.fld(a.Us), R = G.fld(a.R), N = G.fld(a.N);
    TILE_LEVELS(tm, {
        const unsigned cb = G.off(i, j, k);
        double f[nv];
        face_flux<NS, 0, INL>(P, Us, R, N, G.fld(a.A[0]), G.fld(a.B[0]), a.K, a.w, a.volinv, a.dtinv, a.dt, i, j, k,
                              cb, f);
        store_face<NS, 0>(P, G.fld(a.F[0]), f, i, j, k, cb);
        face_flux<NS, 1, INL>(P, Us, R, N, G.fld(a.A[1]), G.fld(a.B[1]), a.K, a.w, a.volinv, a.dtinv, a.dt, i, j, k,
                              cb, f);
        store_face<NS, 1>(P, G.fld(a.F[1]), f, i, j, k, cb);
        face_flux<NS, 2, INL>(P, Us, R, N, G.fld(a.A[2]), G.fld(a.B[2]), a.K, a.w, a.volinv, a.dtinv, a.dt, i, j, k,
                              cb, f);
        store_face<NS, 2>(P, G.fld(a.F[2]), f, i, j, k, cb);
    })
}

// ---- K4: stage combination (+ primitive variables after the last stage) --------------------

// REC: also write the next stage's cell records (k_cells of that stage is then skipped).
// With chemistry (P.chem.nreaction > 0) the species source of the stage is evaluated from the
// stage's records (read before REC overwrites a cell's record) and in-place noise of key K.
template <int NS, bool ZM, bool REC, bool CHEM>
__global__ void __launch_bounds__(FB) k_update(DevParams P, TileMap tm, Grid G, double* out, double* u0,
                                               double* us, double* fx, double* fy, double* fz, double dt, int mode,
                                               double* prim, int write_prim, double* rec, int* err, NoiseKey K,
                                               double w, double dV)
{
    const Fld rn = G.fld(rec);
    const ChemStage ch{rn, K, w, dV};
    TILE_LEVELS(tm, update_body<NS>(P, G.fld(out), G.fld(u0), G.fld(us), G.fld(fx), G.fld(fy), G.fld(fz), dt, mode,
                                    G.fld(prim), write_prim, i, j, k, G.off(i, j, k), REC ? &rn : nullptr, err,
                                    CHEM ? &ch : nullptr);)
}

// ---- host side ---------------------------------------------------------------------------

void fused_free(FusedWorkspace& w)
{
    w.timing.destroy();
    for (void* p : w.blocks) (void)hipFree(p);
    w.blocks.clear();
    w.ready = false;
}

static int alloc_grid(FusedWorkspace& w, const Grid& G, double*& p, int nc)
{
    void* q = nullptr;
    const size_t n = (size_t)G.sn * nc;
    if (hipMalloc(&q, n * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc of fast-path workspace failed");
        return FHDX_ERR_NOMEM;
    }
    (void)hipMemset(q, 0, n * sizeof(double));
    p = (double*)q;
    w.blocks.push_back(q);
    return FHDX_OK;
}

static int ensure(const DevParams& P, FusedWorkspace& w)
{
    if (w.ready) return FHDX_OK;
    const Grid G = grid_of(P);
    const int nv = P.nvars;
    const RecLayout L{P.ns};
    int r = 0;
    auto a = [&](double*& p, int nc) { if (!r) r = alloc_grid(w, G, p, nc); };
    a(w.cup, nv);
    a(w.cup2, nv);
    a(w.rec, L.ncomp());
    a(w.node, NODE_NCOMP);
    for (int d = 0; d < 3; ++d) a(w.F[d], nv);
    if (r) return r;
    FHDX_HIP_TRY(hipDeviceSynchronize());
    w.ready = true;
    return FHDX_OK;
}

// Arrays for caller-supplied noise (the Philox path evaluates its noise in place and needs none).
static int ensure_noise_arrays(const DevParams& P, FusedWorkspace& w)
{
    if (w.sA[0]) return FHDX_OK;
    const Grid G = grid_of(P);
    int r = 0;
    for (int d = 0; d < 3 && !r; ++d) {
        r = alloc_grid(w, G, w.sA[d], P.nvars);
        if (!r) r = alloc_grid(w, G, w.sB[d], P.nvars);
    }
    if (!r) r = alloc_grid(w, G, w.rA, 1);
    if (!r) r = alloc_grid(w, G, w.rB, 1);
    return r;
}

template <int NS, bool INL>
static void launch_stage(const DevParams& P, FusedWorkspace& w, const Grid& G, double* cu, double* us, double* out,
                         double* prim, const NoiseKey& K, double wst, double dt, int st, int* d_err, hipStream_t s)
{
    const long ncell = (long)P.n[0] * P.n[1] * P.n[2];
    const RecLayout L{NS};
    const TileMap tm = make_tiles(P.n[0], P.n[1], P.n[2], w.rh, w.tw_max, w.kz);
    const unsigned ntile = tm.blocks();
    Timing& T = w.timing;
    // transport_type 2 takes its own cell kernel; the update kernels carry the default transport
    // only, so its records are never written there
    const bool vw = P.transport_type == 2;
    const bool rec_in_update = w.rec_in_update && !vw;
    if (st == 0 || !rec_in_update) {  // else the previous stage's k_update wrote them
        T.begin(KT_CELLS, s);
        if (vw) k_cells<NS, true><<<grid_for(ncell, FB), FB, 0, s>>>(P, G, us, w.rec, d_err);
        else k_cells<NS, false><<<grid_for(ncell, FB), FB, 0, s>>>(P, G, us, w.rec, d_err);  // records + ghost images
        T.end(s);
    }
    const bool zm = tm.kz > 1;
    T.begin(KT_NODES, s);
    if (zm) k_nodes<INL, true><<<ntile, FB, 0, s>>>(P, tm, G, w.rec, w.node, w.rA, w.rB, K, wst);
    else k_nodes<INL, false><<<ntile, FB, 0, s>>>(P, tm, G, w.rec, w.node, w.rA, w.rB, K, wst);
    T.end(s);
    FaceArgs fa;
    fa.Us = us;
    fa.R = w.rec;
    fa.N = w.node;
    for (int d = 0; d < 3; ++d) {
        fa.A[d] = w.sA[d];
        fa.B[d] = w.sB[d];
        fa.F[d] = w.F[d];
    }
    fa.K = K;
    fa.w = wst;
    fa.volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
    fa.dtinv = 1. / dt;
    fa.dt = dt;
    T.begin(KT_FACES, s);
    // a block cap per CU is imposed by reserving LDS the kernel does not use (160 KiB per CU)
    const unsigned lds = w.face_bpc == 2 ? 64u << 10 : (w.face_bpc == 3 ? 48u << 10 : 0u);
    if (zm) k_faces<NS, INL, true><<<ntile, FB, lds, s>>>(P, tm, G, fa);
    else k_faces<NS, INL, false><<<ntile, FB, lds, s>>>(P, tm, G, fa);
    T.end(s);
    T.begin(KT_UPDATE, s);
#define FHDX_UPD(ZMV, RECV, CHEMV)                                                                     \
    k_update<NS, ZMV, RECV, CHEMV><<<ntile, FB, 0, s>>>(P, tm, G, out, cu, us, w.F[0], w.F[1], w.F[2], dt, st, \
                                                        prim, st == 2, w.rec, d_err, K, wst,                   \
                                                        P.dx[0] * P.dx[1] * P.dx[2])
    const bool rec_next = rec_in_update && st < 2;
    const int variant = (zm ? 4 : 0) + (rec_next ? 2 : 0) + (P.chem.nreaction > 0 ? 1 : 0);
    switch (variant) {
        case 0: FHDX_UPD(false, false, false); break;
        case 1: FHDX_UPD(false, false, true); break;
        case 2: FHDX_UPD(false, true, false); break;
        case 3: FHDX_UPD(false, true, true); break;
        case 4: FHDX_UPD(true, false, false); break;
        case 5: FHDX_UPD(true, false, true); break;
        case 6: FHDX_UPD(true, true, false); break;
        default: FHDX_UPD(true, true, true); break;
    }
#undef FHDX_UPD
    T.end(s);
}

int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s)
{
    (void)p;
    int r = ensure(P, w);
    if (r) return r;
    const Grid G = grid_of(P);
    const long ncell = (long)P.n[0] * P.n[1] * P.n[2];
    const bool philox = (sA == nullptr);
    const bool inl = philox && w.noise_inline && P.ns <= kInlineNoiseMaxNS;
    if (philox && !inl) {
        if ((r = ensure_noise_arrays(P, w))) return r;
        w.timing.begin(KT_NOISE, s);
        switch (P.ns) {
            case 2: k_noise<7><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 3: k_noise<8><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 4: k_noise<9><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 5: k_noise<10><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 6: k_noise<11><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 7: k_noise<12><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            case 8: k_noise<13><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, G, w.sA[0], w.sA[1], w.sA[2], w.sB[0], w.sB[1], w.sB[2], w.rA, w.rB); break;
            default: return FHDX_ERR_UNSUPPORTED;
        }
        w.timing.end(s);
        FHDX_HIP_TRY(hipGetLastError());
    }
    if (!philox) {
        // caller noise: face / node FABs, valid entries copied into the grid layout
        if ((r = ensure_noise_arrays(P, w))) return r;
        w.timing.begin(KT_NOISE, s);
        for (int d = 0; d < 3; ++d) {
            k_copy_valid<<<grid_for(ncell, FB), FB, 0, s>>>(make_arr(sA[d]), G, w.sA[d], P.nvars, P.n[0], P.n[1],
                                                            P.n[2]);
            k_copy_valid<<<grid_for(ncell, FB), FB, 0, s>>>(make_arr(sB[d]), G, w.sB[d], P.nvars, P.n[0], P.n[1],
                                                            P.n[2]);
        }
        k_copy_valid<<<grid_for(ncell, FB), FB, 0, s>>>(make_arr(*rA), G, w.rA, 1, P.n[0], P.n[1], P.n[2]);
        k_copy_valid<<<grid_for(ncell, FB), FB, 0, s>>>(make_arr(*rB), G, w.rB, 1, P.n[0], P.n[1], P.n[2]);
        w.timing.end(s);
        FHDX_HIP_TRY(hipGetLastError());
    }
    double ws[3];
    stage_weights(ws);
    double* U = cu.p;
    w.timing.begin(KT_GHOSTS, s);
    r = fill_ghosts(P, G, U, P.nvars, s);
    w.timing.end(s);
    if (r) return r;
    double* in[3] = {U, w.cup, w.cup2};
    double* out[3] = {w.cup, w.cup2, U};
#define FHDX_STAGE(NSV, INLV) \
    launch_stage<NSV, INLV>(P, w, G, U, in[st], out[st], prim.p, K, ws[st], dt, st, d_err, s)
    for (int st = 0; st < 3; ++st) {
        switch (P.ns * 2 + (inl ? 1 : 0)) {
            case 4: FHDX_STAGE(2, false); break;
            case 5: FHDX_STAGE(2, true); break;
            case 6: FHDX_STAGE(3, false); break;
            case 7: FHDX_STAGE(3, true); break;
            case 8: FHDX_STAGE(4, false); break;
            case 9: FHDX_STAGE(4, true); break;
            case 10: FHDX_STAGE(5, false); break;
            case 11: FHDX_STAGE(5, true); break;
            case 12: FHDX_STAGE(6, false); break;
            case 14: FHDX_STAGE(7, false); break;
            case 16: FHDX_STAGE(8, false); break;
            default: return FHDX_ERR_UNSUPPORTED;
        }
        FHDX_HIP_TRY(hipGetLastError());
    }
#undef FHDX_STAGE
    return FHDX_OK;
}

size_t fused_workspace_bytes(const DevParams& P)
{
    const Grid G = grid_of(P);
    const RecLayout L{P.ns};
    size_t comps = 2 * (size_t)P.nvars + L.ncomp() + NODE_NCOMP + 3 * (size_t)P.nvars;
    if (P.ns > kInlineNoiseMaxNS) comps += 6 * (size_t)P.nvars + 2;  // the step's noise arrays
    return comps * (size_t)G.sn * sizeof(double);
}

int fused_rk3_step_in(const fhdx_params& p, const DevParams& P, const FusedWorkspace& shape, double* base,
                      size_t bytes, fhdx_fab cu, fhdx_fab prim, double dt, const NoiseKey& K, int* d_err,
                      hipStream_t s)
{
    const size_t need = fused_workspace_bytes(P);
    if (need > bytes) {
        set_error("the box needs " + std::to_string(need >> 20) + " MiB of workspace; the no-allocation scratch holds " +
                  std::to_string(bytes >> 20) + " MiB");
        return FHDX_ERR_UNSUPPORTED;
    }
    FusedWorkspace w;
    w.tw_max = shape.tw_max;
    w.rh = shape.rh;
    w.kz = shape.kz;
    w.face_bpc = shape.face_bpc;
    w.rec_in_update = shape.rec_in_update;
    w.noise_inline = true;
    const Grid G = grid_of(P);
    const RecLayout L{P.ns};
    double* q = base;
    auto take = [&](double*& a, int nc) {
        a = q;
        q += (size_t)G.sn * nc;
    };
    take(w.cup, P.nvars);
    take(w.cup2, P.nvars);
    take(w.rec, L.ncomp());
    take(w.node, NODE_NCOMP);
    for (int d = 0; d < 3; ++d) take(w.F[d], P.nvars);
    if (P.ns > kInlineNoiseMaxNS) {
        for (int d = 0; d < 3; ++d) {
            take(w.sA[d], P.nvars);
            take(w.sB[d], P.nvars);
        }
        take(w.rA, 1);
        take(w.rB, 1);
    }
    w.ready = true;
    return fused_rk3_step(p, P, w, cu, prim, dt, K, nullptr, nullptr, nullptr, nullptr, d_err, s);
}

}  // namespace fhdx

// ---- timing ------------------------------------------------------------------------------

namespace fhdx {

const char* const kKernelKindNames[KT_COUNT] = {"noise", "cells", "nodes", "faces", "update", "ghosts"};

hipEvent_t Timing::get()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void Timing::begin(int kind, hipStream_t s)
{
    if (!on) return;
    cur = kind;
    cur_a = get();
    (void)hipEventRecord(cur_a, s);
}

void Timing::end(hipStream_t s)
{
    if (!on || cur < 0) return;
    hipEvent_t b = get();
    (void)hipEventRecord(b, s);
    pending.push_back({cur, cur_a, b});
    cur = -1;
}

int Timing::resolve()
{
    for (const Pair& p : pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) != hipSuccess || hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess)
            return FHDX_ERR_HIP;
        total_ms[p.kind] += ms;
        launches[p.kind] += 1;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    return FHDX_OK;
}

void Timing::reset()
{
    (void)resolve();
    for (int k = 0; k < KT_COUNT; ++k) {
        total_ms[k] = 0.;
        launches[k] = 0;
    }
}

void Timing::destroy()
{
    (void)resolve();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
}

}  // namespace fhdx
