// fused_kernels.hip -- the fast RK3 step on gfx950.
//
// Per step: k_noise writes the A/B noise sets in one pass (or the caller's sets are used).
// Per stage:
//   k_cells   conservedToPrimitive + transport coefficients, once per valid cell, into a
//             ghost-padded record (u, v, w, T, p, eta, kappa, zeta, Y, X, clamped Y, chi, D);
//   k_nodes   corner velocity gradients, visccorn and the stage-weighted node noise;
//   k_face<D> all four flux families of one face direction (stochastic, diffusive, corner,
//             hyperbolic) accumulated in registers in the reference's per-component order,
//             energy fold included, written once (nvars components);
//   k_update  divergence + RK3 stage combination (+ primitive variables after stage 3).
// Only ghost cells are touched by the periodic fills.  Every value is computed with the
// reference's operation order, so the step is bit-identical to the reference-shaped path.
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

__global__ void __launch_bounds__(FB) k_fill_ghosts(Arr a, int nc, int g, int n0, int n1, int n2)
{
    ghost_body(a, nc, g, n0, n1, n2, (long)blockIdx.x * FB + threadIdx.x);
}

static int fill_ghosts(const DevParams& P, Arr a, int nc, hipStream_t s)
{
    const long G0 = P.n[0] + 2 * P.ngc, G1 = P.n[1] + 2 * P.ngc, g2 = 2L * P.ngc;
    const long n = g2 * G0 * G1 + g2 * G0 * P.n[2] + g2 * P.n[1] * P.n[2];
    k_fill_ghosts<<<grid_for(n, FB), FB, 0, s>>>(a, nc, P.ngc, P.n[0], P.n[1], P.n[2]);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

// ---- K0: the step's noise, one thread per cell (its three lower faces and its node) --------

template <int NV>
__global__ void __launch_bounds__(FB) k_noise(DevParams P, NoiseKey K, Arr A0, Arr A1, Arr A2, Arr B0, Arr B1,
                                              Arr B2, Arr rA, Arr rB)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    noise_body<NV>(P, K, A0, A1, A2, B0, B1, B2, rA, rB, i, j, k);
}

// ---- K1: cell records ------------------------------------------------------------------

template <int NS>
__global__ void __launch_bounds__(FB) k_cells(DevParams P, Arr U, Arr R, int* err)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    cell_body<NS>(P, U, R, err, i, j, k);
}

// ---- tiles: 2-D (x, y) blocks of 256 threads, one z level per block, XCD-aware order ----
//
// Workgroup b runs on XCD b % 8.  Tiles are numbered y-tile fastest, then x-tile, then z, and
// XCD x receives the contiguous range x*nb/8 .. (x+1)*nb/8, so the blocks an XCD holds at once
// cover whole z planes of one slab and their x/y neighbours share that XCD's L2.

struct TileMap {
    int tx_shift;  // tile width 1 << tx_shift (x), height 256 >> tx_shift (y)
    int ntx, nty;
};

static TileMap make_tiles(int n0, int n1)
{
    TileMap m;
    int s = 0;
    while ((1 << s) < n0 && s < 6) ++s;  // width min(64, next power of two >= n0)
    m.tx_shift = s;
    const int tx = 1 << s, ty = FB >> s;
    m.ntx = (n0 + tx - 1) / tx;
    m.nty = (n1 + ty - 1) / ty;
    return m;
}

__device__ inline bool tile_cell(const TileMap& m, int n0, int n1, int& i, int& j, int& k)
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int xcd = b & 7, r = b >> 3, per = nb >> 3, rem = nb & 7;
    const int t = xcd * per + (xcd < rem ? xcd : rem) + r;
    const int ty = t % m.nty, rest = t / m.nty;
    const int tx = rest % m.ntx;
    k = rest / m.ntx;
    i = (tx << m.tx_shift) + (threadIdx.x & ((1 << m.tx_shift) - 1));
    j = ty * (FB >> m.tx_shift) + (threadIdx.x >> m.tx_shift);
    return i < n0 && j < n1;
}

// ---- K2: nodes (flux.cpp:1627-1713) and the stochastic edge terms of each node -------------

__global__ void __launch_bounds__(FB) k_nodes(DevParams P, TileMap tm, Arr R, Arr N, Arr rA, Arr rB, double w)
{
    int i, j, k;
    if (!tile_cell(tm, P.n[0], P.n[1], i, j, k)) return;
    node_body(P, R, N, rA, rB, w, i, j, k);
}

// ---- K3: all fluxes of one face direction -------------------------------------------------

template <int NS, int D>
__global__ void __launch_bounds__(FB) k_face(DevParams P, TileMap tm, Arr Us, Arr R, Arr N, Arr A, Arr B, Arr Fo,
                                             double w, double volinv, double dtinv, double dt)
{
    int i, j, k;
    if (!tile_cell(tm, P.n[0], P.n[1], i, j, k)) return;
    face_body<NS, D>(P, Us, R, N, A, B, Fo, w, volinv, dtinv, dt, i, j, k);
}

// ---- K4: stage combination (+ primitive variables after the last stage) --------------------

template <int NS>
__global__ void __launch_bounds__(FB) k_update(DevParams P, Arr out, Arr u0, Arr us, Arr fx, Arr fy, Arr fz,
                                               double dt, int mode, Arr prim, int write_prim)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    update_body<NS>(P, out, u0, us, fx, fy, fz, dt, mode, prim, write_prim, i, j, k);
}

// ---- host side ---------------------------------------------------------------------------

void fused_free(FusedWorkspace& w)
{
    w.timing.destroy();
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
    const int g = P.ngc, nv = P.nvars;
    const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    const RecLayout L{P.ns};
    int r = 0;
    auto cell = [&](fhdx_fab& f, int nc) { if (!r) r = alloc(w, f, -g, -g, -g, n0 - 1 + g, n1 - 1 + g, n2 - 1 + g, nc); };
    auto face = [&](fhdx_fab& f, int d, int nc) {
        if (!r) r = alloc(w, f, 0, 0, 0, n0 - 1 + (d == 0), n1 - 1 + (d == 1), n2 - 1 + (d == 2), nc);
    };
    auto node = [&](fhdx_fab& f, int nc) { if (!r) r = alloc(w, f, 0, 0, 0, n0, n1, n2, nc); };
    auto per = [&](fhdx_fab& f, int nc) { if (!r) r = alloc(w, f, 0, 0, 0, n0 - 1, n1 - 1, n2 - 1, nc); };
    cell(w.cup, nv);
    cell(w.cup2, nv);
    cell(w.rec, L.ncomp());
    per(w.node, NODE_NCOMP);
    for (int d = 0; d < 3; ++d) {
        per(w.F[d], nv);
        face(w.sA[d], d, nv);
        face(w.sB[d], d, nv);
    }
    node(w.rA, 1);
    node(w.rB, 1);
    if (r) return r;
    FHDX_HIP_TRY(hipDeviceSynchronize());
    w.ready = true;
    return FHDX_OK;
}

template <int NS>
static void launch_stage(const DevParams& P, FusedWorkspace& w, Arr cu, Arr us, Arr out, Arr prim, const Arr A[3],
                         const Arr B[3], Arr rA, Arr rB, double wst, double dt, int st, int* d_err, hipStream_t s)
{
    const long ncell = (long)P.n[0] * P.n[1] * P.n[2];
    const RecLayout L{NS};
    const TileMap tm = make_tiles(P.n[0], P.n[1]);
    const unsigned ntile = (unsigned)(tm.ntx * tm.nty * P.n[2]);
    Arr R = make_arr(w.rec), N = make_arr(w.node);
    Arr F0 = make_arr(w.F[0]), F1 = make_arr(w.F[1]), F2 = make_arr(w.F[2]);
    Timing& T = w.timing;
    T.begin(KT_CELLS, s);
    k_cells<NS><<<grid_for(ncell, FB), FB, 0, s>>>(P, us, R, d_err);
    fill_ghosts(P, R, L.ncomp(), s);
    T.end(s);
    T.begin(KT_NODES, s);
    k_nodes<<<ntile, FB, 0, s>>>(P, tm, R, N, rA, rB, wst);
    T.end(s);
    const double volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
    const double dtinv = 1. / dt;
    T.begin(KT_FACE_X, s);
    k_face<NS, 0><<<ntile, FB, 0, s>>>(P, tm, us, R, N, A[0], B[0], F0, wst, volinv, dtinv, dt);
    T.end(s);
    T.begin(KT_FACE_Y, s);
    k_face<NS, 1><<<ntile, FB, 0, s>>>(P, tm, us, R, N, A[1], B[1], F1, wst, volinv, dtinv, dt);
    T.end(s);
    T.begin(KT_FACE_Z, s);
    k_face<NS, 2><<<ntile, FB, 0, s>>>(P, tm, us, R, N, A[2], B[2], F2, wst, volinv, dtinv, dt);
    T.end(s);
    T.begin(KT_UPDATE, s);
    k_update<NS><<<grid_for(ncell, FB), FB, 0, s>>>(P, out, cu, us, F0, F1, F2, dt, st, prim, st == 2);
    fill_ghosts(P, out, P.nvars, s);
    T.end(s);
}

int fused_rk3_step(const fhdx_params& p, const DevParams& P, FusedWorkspace& w, fhdx_fab cu, fhdx_fab prim,
                   double dt, const NoiseKey& K, const fhdx_fab* sA, const fhdx_fab* sB, const fhdx_fab* rA,
                   const fhdx_fab* rB, int* d_err, hipStream_t s)
{
    (void)p;
    int r = ensure(P, w);
    if (r) return r;
    Arr A[3], B[3], RA, RB;
    if (sA) {
        for (int d = 0; d < 3; ++d) {
            A[d] = make_arr(sA[d]);
            B[d] = make_arr(sB[d]);
        }
        RA = make_arr(*rA);
        RB = make_arr(*rB);
    } else {
        for (int d = 0; d < 3; ++d) {
            A[d] = make_arr(w.sA[d]);
            B[d] = make_arr(w.sB[d]);
        }
        RA = make_arr(w.rA);
        RB = make_arr(w.rB);
        const long ncell = (long)P.n[0] * P.n[1] * P.n[2];
        w.timing.begin(KT_NOISE, s);
        switch (P.ns) {
            case 2: k_noise<7><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, A[0], A[1], A[2], B[0], B[1], B[2], RA, RB); break;
            case 3: k_noise<8><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, A[0], A[1], A[2], B[0], B[1], B[2], RA, RB); break;
            case 4: k_noise<9><<<grid_for(ncell, FB), FB, 0, s>>>(P, K, A[0], A[1], A[2], B[0], B[1], B[2], RA, RB); break;
            default: return FHDX_ERR_UNSUPPORTED;
        }
        w.timing.end(s);
        FHDX_HIP_TRY(hipGetLastError());
    }
    double ws[3];
    stage_weights(ws);
    Arr U = make_arr(cu), U1 = make_arr(w.cup), U2 = make_arr(w.cup2), PR = make_arr(prim);
    w.timing.begin(KT_UPDATE, s);
    r = fill_ghosts(P, U, P.nvars, s);
    w.timing.end(s);
    if (r) return r;
    Arr in[3] = {U, U1, U2};
    Arr out[3] = {U1, U2, U};
    for (int st = 0; st < 3; ++st) {
        switch (P.ns) {
            case 2: launch_stage<2>(P, w, U, in[st], out[st], PR, A, B, RA, RB, ws[st], dt, st, d_err, s); break;
            case 3: launch_stage<3>(P, w, U, in[st], out[st], PR, A, B, RA, RB, ws[st], dt, st, d_err, s); break;
            case 4: launch_stage<4>(P, w, U, in[st], out[st], PR, A, B, RA, RB, ws[st], dt, st, d_err, s); break;
            default: return FHDX_ERR_UNSUPPORTED;
        }
        FHDX_HIP_TRY(hipGetLastError());
    }
    w.timing.begin(KT_UPDATE, s);
    r = fill_ghosts(P, PR, P.nprim, s);
    w.timing.end(s);
    return r;
}

}  // namespace fhdx

// ---- timing ------------------------------------------------------------------------------

namespace fhdx {

const char* const kKernelKindNames[KT_COUNT] = {"noise", "cells", "nodes", "face_x", "face_y", "face_z",
                                                "update"};

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
