// fused_kernels.hip -- the fast RK3 step on gfx950.
//
// Per step: one Philox pass writes the A/B noise sets (or the caller's sets are used).  Per
// stage:
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
    const int G0 = n0 + 2 * g, G1 = n1 + 2 * g;
    const long zs = 2L * g * G0 * G1, ys = 2L * g * G0 * n2, xs = 2L * g * n1 * n2;
    long t = (long)blockIdx.x * FB + threadIdx.x;
    int i, j, k;
    if (t < zs) {
        int kk = (int)(t / ((long)G0 * G1));
        long r = t % ((long)G0 * G1);
        k = kk < g ? kk - g : n2 + kk - g;
        j = (int)(r / G0) - g;
        i = (int)(r % G0) - g;
    } else if ((t -= zs) < ys) {
        k = (int)(t / (2L * g * G0));
        long r = t % (2L * g * G0);
        int jj = (int)(r / G0);
        j = jj < g ? jj - g : n1 + jj - g;
        i = (int)(r % G0) - g;
    } else if ((t -= ys) < xs) {
        k = (int)(t / (2L * g * n1));
        long r = t % (2L * g * n1);
        j = (int)(r / (2 * g));
        int ii = (int)(r % (2 * g));
        i = ii < g ? ii - g : n0 + ii - g;
    } else {
        return;
    }
    const int iv = (i + n0) % n0, jv = (j + n1) % n1, kv = (k + n2) % n2;
    for (int n = 0; n < nc; ++n) a(i, j, k, n) = a(iv, jv, kv, n);
}

static int fill_ghosts(const DevParams& P, Arr a, int nc, hipStream_t s)
{
    const long G0 = P.n[0] + 2 * P.ngc, G1 = P.n[1] + 2 * P.ngc, g2 = 2L * P.ngc;
    const long n = g2 * G0 * G1 + g2 * G0 * P.n[2] + g2 * P.n[1] * P.n[2];
    k_fill_ghosts<<<grid_for(n, FB), FB, 0, s>>>(a, nc, P.ngc, P.n[0], P.n[1], P.n[2]);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

// ---- K1: cell records ------------------------------------------------------------------

template <int NS>
__global__ void __launch_bounds__(FB) k_cells(DevParams P, Arr U, Arr R, int* err)
{
    long t = (long)blockIdx.x * FB + threadIdx.x;
    if (t >= (long)P.n[0] * P.n[1] * P.n[2]) return;
    int i, j, k;
    decode3(t, P.n[0], P.n[1], i, j, k);
    const RecLayout L{NS};
    double c[NS + 5], pr[2 * NS + 6];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) c[n] = U(i, j, k, n);
    cons_to_prim_cell<NS>(P, c, pr);
    Transport tr;
    double chi[NS], D[NS * NS];
    bool bad = transport_cell<NS>(P, pr[0], pr[4], pr[5], &pr[6], tr, chi, D);
    if (bad) atomicOr(err, 1);
    // clamped and renormalised mass fractions of the stochastic species flux (flux.cpp:343-365)
    double yn[NS], sumy = 0.;
#pragma unroll
    for (int s = 0; s < NS; ++s) yn[s] = clamp01(pr[6 + s]);
#pragma unroll
    for (int s = 0; s < NS; ++s) sumy += yn[s];
#pragma unroll
    for (int s = 0; s < NS; ++s) yn[s] /= sumy;
    R(i, j, k, RecLayout::U) = pr[1];
    R(i, j, k, RecLayout::V) = pr[2];
    R(i, j, k, RecLayout::W) = pr[3];
    R(i, j, k, RecLayout::T) = pr[4];
    R(i, j, k, RecLayout::P) = pr[5];
    R(i, j, k, RecLayout::ETA) = tr.eta;
    R(i, j, k, RecLayout::KAP) = tr.kappa;
    R(i, j, k, RecLayout::ZETA) = tr.zeta;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        R(i, j, k, RecLayout::Y + s) = pr[6 + s];
        R(i, j, k, L.X() + s) = pr[6 + NS + s];
        R(i, j, k, L.YN() + s) = yn[s];
        R(i, j, k, L.CHI() + s) = chi[s];
    }
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) R(i, j, k, L.D() + q) = D[q];
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

__device__ inline int wrap_up(int i, int n) { return i + 1 == n ? 0 : i + 1; }

// ---- K2: nodes (flux.cpp:1627-1713) and the stochastic edge terms of each node -------------

// Node-relative summation orders of the edge viscosities: cell (a,b,c) is node - 1 + (a,b,c);
// the fourth entry shifts the eta cell in z (order 2 reads eta(k-2)*T(k), flux.cpp:857-858).
__constant__ signed char c_node_order[4][8][4] = {
    {{1,0,0,0},{0,0,0,0},{1,1,0,0},{0,1,0,0},{1,0,1,0},{0,0,1,0},{1,1,1,0},{0,1,1,0}},
    {{0,1,0,0},{1,1,0,0},{0,0,0,0},{1,0,0,0},{0,1,1,0},{1,1,1,0},{0,0,1,0},{1,0,1,0}},
    {{1,1,1,0},{0,1,1,-2},{1,0,1,0},{0,0,1,-2},{1,1,0,0},{0,1,0,0},{1,0,0,0},{0,0,0,0}},
    {{0,0,1,0},{1,0,1,0},{0,1,1,0},{1,1,1,0},{0,0,0,0},{1,0,0,0},{0,1,0,0},{1,1,0,0}}};

__global__ void __launch_bounds__(FB) k_nodes(DevParams P, TileMap tm, Arr R, Arr N, Arr rA, Arr rB, double w)
{
    int i, j, k;
    if (!tile_cell(tm, P.n[0], P.n[1], i, j, k)) return;
    const int E = RecLayout::ETA, Z = RecLayout::ZETA;
    const bool visc3 = (P.visc_type == 3 || P.visc_type == -3);
    double muxp = 0.125 * (R(i, j - 1, k - 1, E) + R(i - 1, j - 1, k - 1, E) + R(i, j, k - 1, E) +
                           R(i - 1, j, k - 1, E) + R(i, j - 1, k, E) + R(i - 1, j - 1, k, E) + R(i, j, k, E) +
                           R(i - 1, j, k, E));
    double zetaxp = 0.;
    if (visc3)
        zetaxp = 0.125 * (R(i, j - 1, k - 1, Z) + R(i - 1, j - 1, k - 1, Z) + R(i, j, k - 1, Z) +
                          R(i - 1, j, k - 1, Z) + R(i, j - 1, k, Z) + R(i - 1, j - 1, k, Z) + R(i, j, k, Z) +
                          R(i - 1, j, k, Z));
    const double DX0 = P.dx[0], DX1 = P.dx[1], DX2 = P.dx[2];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int n = RecLayout::U + b;  // u, v, w
        N(i, j, k, 0 + b) = 0.25 * muxp *
                            (R(i, j - 1, k - 1, n) - R(i - 1, j - 1, k - 1, n) + R(i, j, k - 1, n) -
                             R(i - 1, j, k - 1, n) + R(i, j - 1, k, n) - R(i - 1, j - 1, k, n) + R(i, j, k, n) -
                             R(i - 1, j, k, n)) /
                            DX0;
        N(i, j, k, 3 + b) = 0.25 * muxp *
                            (R(i - 1, j, k - 1, n) - R(i - 1, j - 1, k - 1, n) + R(i, j, k - 1, n) -
                             R(i, j - 1, k - 1, n) + R(i - 1, j, k, n) - R(i - 1, j - 1, k, n) + R(i, j, k, n) -
                             R(i, j - 1, k, n)) /
                            DX1;
        N(i, j, k, 6 + b) = 0.25 * muxp *
                            (R(i - 1, j - 1, k, n) - R(i - 1, j - 1, k - 1, n) + R(i, j - 1, k, n) -
                             R(i, j - 1, k - 1, n) + R(i - 1, j, k, n) - R(i - 1, j, k - 1, n) + R(i, j, k, n) -
                             R(i, j, k - 1, n)) /
                            DX2;
    }
    const int u = RecLayout::U, v = RecLayout::V, ww = RecLayout::W;
    N(i, j, k, NODE_VC) =
        (muxp / 12. + zetaxp / 4.) *
        ((R(i, j - 1, k - 1, u) - R(i - 1, j - 1, k - 1, u)) / DX0 +
         (R(i, j, k - 1, u) - R(i - 1, j, k - 1, u)) / DX0 + (R(i, j - 1, k, u) - R(i - 1, j - 1, k, u)) / DX0 +
         (R(i, j, k, u) - R(i - 1, j, k, u)) / DX0 + (R(i - 1, j, k - 1, v) - R(i - 1, j - 1, k - 1, v)) / DX1 +
         (R(i, j, k - 1, v) - R(i, j - 1, k - 1, v)) / DX1 + (R(i - 1, j, k, v) - R(i - 1, j - 1, k, v)) / DX1 +
         (R(i, j, k, v) - R(i, j - 1, k, v)) / DX1 + (R(i - 1, j - 1, k, ww) - R(i - 1, j - 1, k - 1, ww)) / DX2 +
         (R(i, j - 1, k, ww) - R(i, j - 1, k - 1, ww)) / DX2 + (R(i - 1, j, k, ww) - R(i - 1, j, k - 1, ww)) / DX2 +
         (R(i, j, k, ww) - R(i, j, k - 1, ww)) / DX2);
    if (P.stoch_on) {
        // stage-weighted node noise (MultiFab::LinComb, timeStep.cpp:125-128)
        const double W = 1.0 * rA(i, j, k) + w * rB(i, j, k);
        const int RT = RecLayout::T;
#pragma unroll
        for (int v2 = 0; v2 < 4; ++v2) {
            double s = 0.;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const signed char* o = c_node_order[v2][q];
                const int a = i - 1 + o[0], bb = j - 1 + o[1], c = k - 1 + o[2];
                double term = R(a, bb, c + o[3], E) * R(a, bb, c, RT);
                s = (q == 0) ? term : s + term;
            }
            double mu = 0.25 * s / 3.;
            if (visc3) {
                double z = 0.;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const signed char* o = c_node_order[v2][q];
                    const int a = i - 1 + o[0], bb = j - 1 + o[1], c = k - 1 + o[2];
                    double term = R(a, bb, c + o[3], Z) * R(a, bb, c, RT);
                    z = (q == 0) ? term : z + term;
                }
                mu += 0.25 * z;
            }
            N(i, j, k, NODE_S + v2) = sqrt(mu) * W;
        }
    }
}

// ---- K3: all fluxes of one face direction -------------------------------------------------

// Sum over the 4 nodes of face (i,j,k) of component c, in visccorn order (hi,hi),(lo,hi),(hi,lo),
// (lo,lo) over the two transverse axes, or in the z-face order (hi,hi),(hi,lo),(lo,hi),(lo,lo).
// Node index n along an axis is node 0 (periodic), hence the wrapped "hi" indices.
template <int D>
__device__ inline double node_sum(const Arr& N, int c, int i, int j, int k, int ih, int jh, int kh, bool zorder,
                                  double s0, bool first)
{
    // first transverse axis: x faces y, y faces x, z faces x; second: x faces z, y faces z, z faces y
    const bool ahi[4] = {true, zorder ? true : false, zorder ? false : true, false};
    const bool bhi[4] = {true, zorder ? false : true, zorder ? true : false, false};
    double s = s0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int a = i, b = j, cc = k;
        if (D == 0) { b = ahi[q] ? jh : j; cc = bhi[q] ? kh : k; }
        if (D == 1) { a = ahi[q] ? ih : i; cc = bhi[q] ? kh : k; }
        if (D == 2) { a = ahi[q] ? ih : i; b = bhi[q] ? jh : j; }
        double v = N(a, b, cc, c);
        s = (first && q == 0) ? v : s + v;
    }
    return s;
}

template <int NS, int D>
__global__ void __launch_bounds__(FB) k_face(DevParams P, TileMap tm, Arr Us, Arr R, Arr N, Arr A, Arr B, Arr Fo,
                                             double w, double volinv, double dtinv, double dt)
{
    int i, j, k;
    if (!tile_cell(tm, P.n[0], P.n[1], i, j, k)) return;
    const int il = i - (D == 0), jl = j - (D == 1), kl = k - (D == 2);
    const int ih = wrap_up(i, P.n[0]), jh = wrap_up(j, P.n[1]), kh = wrap_up(k, P.n[2]);
    constexpr int nv = NS + 5;
    const RecLayout L{NS};
    const int RT = RecLayout::T, RE = RecLayout::ETA;
    // velocity component c (1..3) lives in record slot c-1
    auto vel = [&](int a, int b, int cc, int c) { return R(a, b, cc, RecLayout::U + c - 1); };

    double f[nv + 4];
#pragma unroll
    for (int n = 0; n < nv + 4; ++n) f[n] = 0.;

    // ---- stochastic (flux.cpp:71-1145) ----
    if (P.stoch_on) {
        double Tu = R(i, j, k, RT), Tl = R(il, jl, kl, RT);
        double mup = R(i, j, k, RE) * Tu + R(il, jl, kl, RE) * Tl;
        double kp = R(i, j, k, RecLayout::KAP) * Tu * Tu + R(il, jl, kl, RecLayout::KAP) * Tl * Tl;
        double meanT = 0.5 * (Tu + Tl);
        double wn[nv];
        wn[0] = 0.;
#pragma unroll
        for (int n = 1; n < nv; ++n) wn[n] = 1.0 * A(i, j, k, n) + w * B(i, j, k, n);
        double wi[nv];
        double fw1 = sqrt(P.k_B * mup * volinv * dtinv);
        double fw4 = sqrt(P.k_B * kp * volinv * dtinv);
        wi[0] = 0.;
        wi[1] = fw1 * wn[1];
        wi[2] = fw1 * wn[2];
        wi[3] = fw1 * wn[3];
        wi[4] = fw4 * wn[4];
        double nweight = sqrt(P.k_B * volinv * dtinv);
        // edge terms sqrt(mu)*rancorn of the pp, mp, pm, mm edges (flux.cpp:277-280, 644-647, 1011-1014)
        double sum;
        if (D == 0) {
            sum = N(i, jh, kh, NODE_S + 0);
            sum = sum + N(i, j, kh, NODE_S + 0);
            sum = sum + N(i, jh, k, NODE_S + 0);
            sum = sum + N(i, j, k, NODE_S + 0);
        } else if (D == 1) {
            sum = N(ih, j, kh, NODE_S + 0);
            sum = sum + N(i, j, kh, NODE_S + 1);
            sum = sum + N(ih, j, k, NODE_S + 0);
            sum = sum + N(i, j, k, NODE_S + 1);
        } else {
            sum = N(ih, jh, k, NODE_S + 0);
            sum = sum + N(i, jh, k, NODE_S + 1);
            sum = sum + N(ih, j, k, NODE_S + 2);
            sum = sum + N(i, j, k, NODE_S + 3);
        }
        wi[1 + D] = wi[1 + D] + 0.25 * nweight * sum;
#pragma unroll
        for (int n = 1; n < 4; ++n) f[n] = f[n] + wi[n];
        f[nv] = f[nv] + wi[4];
        constexpr int t1 = (D == 0) ? 2 : 1, t2 = (D == 2) ? 2 : 3;
        double phidiag = wi[1 + D] * (vel(il, jl, kl, 1 + D) + vel(i, j, k, 1 + D));
        double phishear = wi[t1] * (vel(il, jl, kl, t1) + vel(i, j, k, t1)) +
                          wi[t2] * (vel(il, jl, kl, t2) + vel(i, j, k, t2));
        phidiag = -0.5 * phidiag;
        phishear = -0.5 * phishear;
        f[nv + 1] = f[nv + 1] - phidiag;
        f[nv + 2] = f[nv + 2] - phishear;
        if (P.algorithm_type == 2) {
            double yy[NS], yyp[NS], DijY[NS * NS], sqD[NS * NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                wi[5 + s] = 0.;
                yy[s] = R(il, jl, kl, L.YN() + s);
                yyp[s] = R(i, j, k, L.YN() + s);
            }
            double MWmix = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                MWmix = MWmix + 0.5 * (yy[s] + yyp[s]) / P.molmass[s];
#pragma unroll
                for (int l = 0; l < NS; ++l)
                    DijY[s * NS + l] = 0.5 * (R(il, jl, kl, L.D() + l * NS + s) * yy[l] +
                                              R(i, j, k, L.D() + l * NS + s) * yyp[l] +
                                              (R(il, jl, kl, L.D() + s * NS + l) * yy[s] +
                                               R(i, j, k, L.D() + s * NS + l) * yyp[s]));
            }
#pragma unroll
            for (int s = 0; s < NS; ++s)
                if (fabs(yy[s]) + fabs(yyp[s]) <= 1.e-12)
#pragma unroll
                    for (int n = 0; n < NS; ++n) {
                        DijY[s * NS + n] = 0.;
                        DijY[n * NS + s] = 0.;
                    }
            MWmix = 1. / MWmix;
            cholesky<NS>(DijY, sqD);
            double wsp = sqrt(P.k_B * MWmix * volinv / (P.Runiv * dt));
#pragma unroll
            for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int l = 0; l <= s; ++l) wi[5 + s] = wi[5 + s] + (wsp * sqD[s * NS + l]) * wn[5 + l];
                f[5 + s] = wi[5 + s];
            }
            double soret = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double hk = P.e0[s] + P.hcp[s] * meanT;
                double ss = (hk + P.Runiv * meanT / P.molmass[s] * 0.5 *
                                      (R(il, jl, kl, L.CHI() + s) + R(i, j, k, L.CHI() + s))) *
                            wi[5 + s];
                soret = soret + ss;
            }
            f[nv + 3] = f[nv + 3] + soret;
        }
    }

    // ---- diffusive (flux.cpp:1193-1623) ----
    {
        const double half = 0.5, h = P.dx[D];
        double mup = half * (R(i, j, k, RE) + R(il, jl, kl, RE));
        double kp = half * (R(i, j, k, RecLayout::KAP) + R(il, jl, kl, RecLayout::KAP));
        double tau1 = mup * (vel(i, j, k, 1) - vel(il, jl, kl, 1)) / h;
        double tau2 = mup * (vel(i, j, k, 2) - vel(il, jl, kl, 2)) / h;
        double tau3 = mup * (vel(i, j, k, 3) - vel(il, jl, kl, 3)) / h;
        double divp = 0.;
        double s1 = vel(il, jl, kl, 1) + vel(i, j, k, 1);
        double s2 = vel(il, jl, kl, 2) + vel(i, j, k, 2);
        double s3 = vel(il, jl, kl, 3) + vel(i, j, k, 3);
        double phiflx;
        if (D == 0) phiflx = tau1 * s1 + divp * s1 + tau2 * s2 + tau3 * s3;
        else if (D == 1) phiflx = tau1 * s1 + tau2 * s2 + divp * s2 + tau3 * s3;
        else phiflx = tau1 * s1 + tau2 * s2 + tau3 * s3 + divp * s3;
        double TU = R(i, j, k, RT), TL = R(il, jl, kl, RT);
        double Qflux = kp * (TU - TL) / h;
        f[1] = f[1] - (D == 0 ? tau1 + divp : tau1);
        f[2] = f[2] - (D == 1 ? tau2 + divp : tau2);
        f[3] = f[3] - (D == 2 ? tau3 + divp : tau3);
        f[nv] = f[nv] - Qflux;
        f[nv + 1] = f[nv + 1] - (half * phiflx);
        double meanT = 0.5 * (TL + TU);
        double PU = R(i, j, k, RecLayout::P), PL = R(il, jl, kl, RecLayout::P);
        double meanP = 0.5 * (PL + PU);
        if (P.algorithm_type == 2) {
            double dk[NS], soret[NS], Fk[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double XU = R(i, j, k, L.X() + s), XL = R(il, jl, kl, L.X() + s);
                double term1 = (XU - XL) / h;
                double mX = 0.5 * (XL + XU);
                double mY = 0.5 * (R(il, jl, kl, RecLayout::Y + s) + R(i, j, k, RecLayout::Y + s));
                double term2 = (mX - mY) * (PU - PL) / h / meanP;
                dk[s] = term1 + term2;
                double ChiX;
                if (D == 2)  // flux.cpp:1561 reads chi at k and k+1
                    ChiX = 0.5 * (R(i, j, k, L.CHI() + s) * XL + R(i, j, k + 1, L.CHI() + s) * XU);
                else
                    ChiX = 0.5 * (R(il, jl, kl, L.CHI() + s) * XL + R(i, j, k, L.CHI() + s) * XU);
                soret[s] = ChiX * (TU - TL) / h / meanT;
            }
#pragma unroll
            for (int kk = 0; kk < NS; ++kk) {
                double fk = 0.;
#pragma unroll
                for (int ll = 0; ll < NS; ++ll) {
                    double Fks = half * (R(il, jl, kl, L.D() + ll * NS + kk) + R(i, j, k, L.D() + ll * NS + kk)) *
                                 (dk[ll] + soret[ll]);
                    fk = fk - Fks;
                }
                Fk[kk] = fk;
            }
            double Q5 = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double hk = P.e0[s] + P.hcp[s] * meanT;
                double csum = (D == 2) ? R(i, j, k, L.CHI() + s) + R(i, j, k, L.CHI() + s)  // flux.cpp:1604
                                       : R(il, jl, kl, L.CHI() + s) + R(i, j, k, L.CHI() + s);
                double Q5s = (hk + 0.5 * P.Runiv * meanT * csum / P.molmass[s]) * Fk[s];
                Q5 = Q5 + Q5s;
            }
            f[nv + 3] = f[nv + 3] + Q5;
#pragma unroll
            for (int s = 0; s < NS; ++s) f[5 + s] = f[5 + s] + Fk[s];
        }
    }

    // ---- corner terms (flux.cpp:1770-1980) ----
    {
        const bool zo = (D == 2);
        // x face: div pair (vy, wz), tangential (uy, uz); y: (ux, wz), (vx, vz); z: (vy, ux), (wx, wy)
        constexpr int c1 = (D == 0) ? 4 : (D == 1) ? 0 : 4;
        constexpr int c2 = (D == 0) ? 8 : (D == 1) ? 8 : 0;
        constexpr int a1 = (D == 0) ? 3 : (D == 1) ? 1 : 2;
        constexpr int a2 = (D == 0) ? 6 : (D == 1) ? 7 : 5;
        double vc4 = node_sum<D>(N, NODE_VC, i, j, k, ih, jh, kh, false, 0., true);
        double div8 = node_sum<D>(N, c1, i, j, k, ih, jh, kh, zo, 0., true);
        div8 = node_sum<D>(N, c2, i, j, k, ih, jh, kh, zo, div8, false);
        double s1 = node_sum<D>(N, a1, i, j, k, ih, jh, kh, zo, 0., true);
        double s2 = node_sum<D>(N, a2, i, j, k, ih, jh, kh, zo, 0., true);
        constexpr int cn = 1 + D;
        constexpr int tc1 = (D == 0) ? 2 : 1, tc2 = (D == 2) ? 2 : 3;
        f[cn] = f[cn] - 0.25 * vc4;
        f[cn] = f[cn] + .25 * div8;
        f[tc1] = f[tc1] - .25 * s1;
        f[tc2] = f[tc2] - .25 * s2;
        double phiflx = 0.25 * (vc4 - div8) * (vel(il, jl, kl, cn) + vel(i, j, k, cn));
        phiflx = phiflx + .25 * s1 * (vel(il, jl, kl, tc1) + vel(i, j, k, tc1));
        phiflx = phiflx + .25 * s2 * (vel(il, jl, kl, tc2) + vel(i, j, k, tc2));
        f[nv + 1] = f[nv + 1] - 0.5 * phiflx;
    }

    // ---- hyperbolic (flux.cpp:2016-2372) ----
    if (P.advection_type == 1 || P.advection_type == 2) {
        constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
        const double wgt2 = 1. / 12.;
        const double wgt1 = 0.5 + wgt2;
        double cv[NS + 5], pv[NS + 6], Yk[NS];
        if (P.advection_type == 1) {
            // primitive interpolation: rho from cons, the rest from the record
            auto pr = [&](int a, int b, int cc, int l) {
                if (l == 0) return Us(a, b, cc, 0);
                if (l <= 3) return R(a, b, cc, RecLayout::U + l - 1);
                if (l == 4) return R(a, b, cc, RecLayout::T);
                if (l == 5) return R(a, b, cc, RecLayout::P);
                return R(a, b, cc, RecLayout::Y + l - 6);
            };
#pragma unroll
            for (int l = 0; l < NS + 6; ++l)
                pv[l] = wgt1 * (pr(i, j, k, l) + pr(i - ex, j - ey, k - ez, l)) -
                        wgt2 * (pr(i - 2 * ex, j - 2 * ey, k - 2 * ez, l) + pr(i + ex, j + ey, k + ez, l));
            double rho = pv[0];
            cv[0] = rho;
#pragma unroll
            for (int n = 0; n < NS; ++n) Yk[n] = pv[6 + n];
            double ie = eos_energy<NS>(P, Yk, pv[4]);
            double vsqr = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
            cv[4] = rho * ie + 0.5 * rho * vsqr;
        } else {
#pragma unroll
            for (int l = 0; l < NS + 5; ++l)
                cv[l] = wgt1 * (Us(i, j, k, l) + Us(i - ex, j - ey, k - ez, l)) -
                        wgt2 * (Us(i - 2 * ex, j - 2 * ey, k - 2 * ez, l) + Us(i + ex, j + ey, k + ez, l));
#pragma unroll
            for (int l = 1; l < 4; ++l) pv[l] = cv[l] / cv[0];
#pragma unroll
            for (int n = 0; n < NS; ++n) Yk[n] = cv[5 + n] / cv[0];
            double vsqr = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
            double ie = cv[4] / cv[0] - 0.5 * vsqr;
            pv[4] = eos_temperature<NS>(P, ie, Yk);
            pv[5] = eos_pressure<NS>(P, Yk, cv[0], pv[4]);
        }
        const double un = pv[1 + D];
        f[0] += cv[0] * un;
        if (D == 0) {
            f[1] += cv[0] * (pv[1] * pv[1]) + pv[5];
            f[2] += cv[0] * pv[1] * pv[2];
            f[3] += cv[0] * pv[1] * pv[3];
        } else if (D == 1) {
            f[1] += cv[0] * pv[1] * pv[2];
            f[2] += cv[0] * pv[2] * pv[2] + pv[5];
            f[3] += cv[0] * pv[3] * pv[2];
        } else {
            f[1] += cv[0] * pv[1] * pv[3];
            f[2] += cv[0] * pv[2] * pv[3];
            f[3] += cv[0] * pv[3] * pv[3] + pv[5];
        }
        f[4] += un * cv[4] + pv[5] * un;
        f[4] += f[nv] + f[nv + 1] + f[nv + 2] + f[nv + 3];
        if (P.algorithm_type == 2) {
            if (P.advection_type == 1) {
#pragma unroll
                for (int n = 0; n < NS; ++n) f[5 + n] += cv[0] * pv[6 + n] * un;
            } else {
#pragma unroll
                for (int n = 0; n < NS; ++n) f[5 + n] += cv[5 + n] * un;
            }
        }
    }
#pragma unroll
    for (int n = 0; n < nv; ++n) Fo(i, j, k, n) = f[n];
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
    const int ih = wrap_up(i, P.n[0]), jh = wrap_up(j, P.n[1]), kh = wrap_up(k, P.n[2]);
    const double src = 0.0;
    double o[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) {
        // face n is face 0 of the periodic box (same inputs, same bits)
        double div = (fx(ih, j, k, n) - fx(i, j, k, n)) / P.dx[0] +
                     (fy(i, jh, k, n) - fy(i, j, k, n)) / P.dx[1] +
                     (fz(i, j, kh, n) - fz(i, j, k, n)) / P.dx[2];
        if (mode == 0) o[n] = us(i, j, k, n) - dt * div + dt * src;
        else if (mode == 1) o[n] = 0.25 * (3.0 * u0(i, j, k, n) + us(i, j, k, n) - dt * div + dt * src);
        else o[n] = (2. / 3.) * (0.5 * u0(i, j, k, n) + us(i, j, k, n) - dt * div + dt * src);
    }
    const double* g = P.grav;
    double r0 = us(i, j, k, 0), r1 = us(i, j, k, 1), r2 = us(i, j, k, 2), r3 = us(i, j, k, 3);
    if (mode == 0) {
#pragma unroll
        for (int n = 0; n < 3; ++n) o[n + 1] += dt * r0 * g[n];
        o[4] += dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
    } else if (mode == 1) {
#pragma unroll
        for (int n = 0; n < 3; ++n) o[n + 1] += 0.25 * dt * r0 * g[n];
        o[4] += 0.25 * dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
    } else {
#pragma unroll
        for (int n = 0; n < 3; ++n) o[n + 1] += 2. / 3. * dt * r0 * g[n];
        o[4] += 2. / 3. * dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
    }
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) out(i, j, k, n) = o[n];
    if (write_prim) {
        double pr[2 * NS + 6];
        cons_to_prim_cell<NS>(P, o, pr);
#pragma unroll
        for (int n = 0; n < 2 * NS + 6; ++n) prim(i, j, k, n) = pr[n];
    }
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
        w.timing.begin(KT_NOISE, s);
        r = ref_fill_noise(P, K, A, B, RA, RB, s);
        w.timing.end(s);
        if (r) return r;
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
