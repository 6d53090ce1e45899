// stage_bodies.hpp -- per-cell, per-node and per-face bodies of one RK3 stage.
//
// Shared by the single-box fast path (fused_kernels.hip) and the batched engine (engine.hip):
// the kernels there only map a thread to (instance, i, j, k) and call these.  Every body keeps
// the reference's floating-point operation order (see physics.hpp).
#pragma once

#include "fused.hpp"

namespace fhdx {


// One layout for every array of a box (a cell FAB with ngc ghost layers, component-major; face
// and node arrays use it too, their entry (i,j,k) being the lower face / node of cell (i,j,k)).
// Views made from one Grid share their index arithmetic, so a kernel computes each cell offset
// once for all arrays.
// A field of one box in its Grid layout.  F(cb, c, di, dj, dk) is component c of the cell
// (di,dj,dk) away from the cell at byte offset cb: the shift goes into the wave-uniform base
// pointer, so every access of a thread uses the same 32-bit vector offset (one VGPR) and the
// address arithmetic runs on the scalar unit.
struct Fld {
    double* p;
    long sn;
    int sy, sz;
    __device__ inline double& operator()(unsigned cb, int c, int di = 0, int dj = 0, int dk = 0) const
    {
        return *(double*)((char*)(p + (long)c * sn + di + (long)dj * sy + (long)dk * sz) + cb);
    }
    // Streaming (nontemporal) store, for arrays the next kernel reads back from HBM anyway: the
    // store does not displace the lines the writing kernel re-reads from L2.
    __device__ inline void put_nt(unsigned cb, int c, double v, int di = 0, int dj = 0, int dk = 0) const
    {
        __builtin_nontemporal_store(v, &(*this)(cb, c, di, dj, dk));
    }
};

struct Grid {
    int lo, sy, sz;
    long sn;
    __host__ __device__ Arr view(double* p) const
    {
        Arr a;
        a.p = p;
        a.lo0 = a.lo1 = a.lo2 = lo;
        a.sy = sy;
        a.sz = sz;
        a.sn = sn;
        return a;
    }
    __host__ __device__ Fld fld(double* p) const { return Fld{p, sn, sy, sz}; }
    // byte offset of cell (i,j,k) within a component (below 4 GiB: checked by the C ABI)
    __device__ unsigned off(int i, int j, int k) const
    {
        return (unsigned)((i - lo) + (j - lo) * sy + (k - lo) * sz) << 3;
    }
};

inline Grid grid_of(const DevParams& P)
{
    const int g = P.ngc;
    Grid G;
    G.lo = -g;
    G.sy = P.n[0] + 2 * g;
    G.sz = G.sy * (P.n[1] + 2 * g);
    G.sn = (long)G.sz * (P.n[2] + 2 * g);
    return G;
}

// ---- stores with periodic images ------------------------------------------------------------

// Stores value(c), c < NC, for valid cell (i,j,k) at byte offset cb and at each periodic image of
// the cell that lies in the ngc ghost layers (up to 7), so no separate ghost-fill pass is needed.
// Requires n >= 2 ngc on every axis (a cell then has at most one image per axis).
// NT: streaming stores for the cell itself (the images are few and stay ordinary stores).
template <int NC, bool NT = false, typename V>
__device__ inline void store_cell(const DevParams& P, const Fld& F, V value, int i, int j, int k, unsigned cb)
{
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (NT) F.put_nt(cb, c, value(c));
        else F(cb, c) = value(c);
    }
    const int g = P.ngc;
    const int ax = i < g ? P.n[0] : (i >= P.n[0] - g ? -P.n[0] : 0);
    const int ay = j < g ? P.n[1] : (j >= P.n[1] - g ? -P.n[1] : 0);
    const int az = k < g ? P.n[2] : (k >= P.n[2] - g ? -P.n[2] : 0);
    if ((ax | ay | az) == 0) return;
    for (int m = 1; m < 8; ++m) {
        if (((m & 1) && !ax) || ((m & 2) && !ay) || ((m & 4) && !az)) continue;
        const int d = ((m & 1) ? ax : 0) + ((m & 2) ? ay : 0) * F.sy + ((m & 4) ? az : 0) * F.sz;
        const unsigned cbi = cb + (unsigned)(d * 8);  // modulo 2^32: negative shifts wrap correctly
#pragma unroll
        for (int c = 0; c < NC; ++c) F(cbi, c) = value(c);
    }
}

// ---- periodic ghost fill: thread t of the G^3 - n^3 ghost cells ---------------------------

__device__ inline void ghost_body(const Arr& a, int nc, int g, int n0, int n1, int n2, long t)
{
    const int G0 = n0 + 2 * g, G1 = n1 + 2 * g;
    const long zs = 2L * g * G0 * G1, ys = 2L * g * G0 * n2, xs = 2L * g * n1 * n2;
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

// ---- conservedToPrimitive + transport coefficients into the cell record -------------------

// The record of cell (i,j,k) from its conserved values c[] (held in registers by the caller).
// VW: transport_type 2 (Valk-Waldmann); otherwise the type the parameters name (1 or 3).
template <int NS, bool VW = false>
__device__ inline void cell_record(const DevParams& P, const double* c, const Fld& R, int* err, int i, int j, int k,
                                   unsigned cb)
{
    const RecLayout L{NS};
    double pr[2 * NS + 6];
    cons_to_prim_cell<NS>(P, c, pr);
    Transport tr;
    double chi[NS], D[NS * NS];
    if constexpr (VW) {
        const int fl = transport_cell_vw<NS>(P, pr[0], pr[4], pr[5], &pr[6], tr, chi, D);
        if (fl) atomicOr(err, fl);
    } else {
        bool bad = transport_cell<NS>(P, pr[0], pr[4], pr[5], &pr[6], tr, chi, D);
        if (bad) atomicOr(err, 1);
    }
    // clamped and renormalised mass fractions of the stochastic species flux (flux.cpp:343-365)
    double yn[NS], sumy = 0.;
#pragma unroll
    for (int s = 0; s < NS; ++s) yn[s] = clamp01(pr[6 + s]);
#pragma unroll
    for (int s = 0; s < NS; ++s) sumy += yn[s];
#pragma unroll
    for (int s = 0; s < NS; ++s) yn[s] /= sumy;
    double rec[8 + 4 * NS + NS * NS];
    rec[RecLayout::U] = pr[1];
    rec[RecLayout::V] = pr[2];
    rec[RecLayout::W] = pr[3];
    rec[RecLayout::T] = pr[4];
    rec[RecLayout::P] = pr[5];
    rec[RecLayout::ETA] = tr.eta;
    rec[RecLayout::KAP] = tr.kappa;
    rec[RecLayout::ZETA] = tr.zeta;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        rec[RecLayout::Y + s] = pr[6 + s];
        rec[L.X() + s] = pr[6 + NS + s];
        rec[L.YN() + s] = yn[s];
        rec[L.CHI() + s] = chi[s];
    }
#pragma unroll
    for (int q = 0; q < NS * NS; ++q) rec[L.D() + q] = D[q];
    store_cell<8 + 4 * NS + NS * NS>(P, R, [&](int q) { return rec[q]; }, i, j, k, cb);
}

template <int NS, bool VW = false>
__device__ inline void cell_body(const DevParams& P, const Fld& U, const Fld& R, int* err, int i, int j, int k,
                                 unsigned cb)
{
    double c[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) c[n] = U(cb, n);
    cell_record<NS, VW>(P, c, R, err, i, j, k, cb);
}

// ---- nodes (flux.cpp:1627-1713) and the stochastic edge terms of each node -----------------

// Node-relative summation orders of the edge viscosities: cell (a,b,c) is node - 1 + (a,b,c);
// the fourth entry shifts the eta cell in z (order 2 reads eta(k-2)*T(k), flux.cpp:857-858).
struct NodeOrder {
    signed char o[4][8][4];
};
constexpr NodeOrder kNodeOrder = {{
    {{1,0,0,0},{0,0,0,0},{1,1,0,0},{0,1,0,0},{1,0,1,0},{0,0,1,0},{1,1,1,0},{0,1,1,0}},
    {{0,1,0,0},{1,1,0,0},{0,0,0,0},{1,0,0,0},{0,1,1,0},{1,1,1,0},{0,0,1,0},{1,0,1,0}},
    {{1,1,1,0},{0,1,1,-2},{1,0,1,0},{0,0,1,-2},{1,1,0,0},{0,1,0,0},{1,0,0,0},{0,0,0,0}},
    {{0,0,1,0},{1,0,1,0},{0,1,1,0},{1,1,1,0},{0,0,0,0},{1,0,0,0},{0,1,0,0},{1,1,0,0}}}};

// Cell (a,b,c) of the 2x2x2 block around node (i,j,k) is (i-1+a, j-1+b, k-1+c).
__device__ constexpr int nc8(int a, int b, int c) { return a + 2 * b + 4 * c; }

// All loads come first (into registers), all stores last: the record and node arrays are
// distinct, and keeping stores behind the loads lets the compiler load each value once.
// INL: the node noise is evaluated in place from Philox key K instead of read from rA, rB.
// Node (i,j,k) of a periodic box equals node (i,j,k) + n along each axis: the values of nodes with
// a zero index are stored at those images too, so faces read node i+1 without wrapping.
template <bool INL>
__device__ inline void node_body(const DevParams& P, const Fld& R, const Fld& N, const Fld& rA, const Fld& rB,
                                 const NoiseKey& K, double w, int i, int j, int k, unsigned cb)
{
    const int E = RecLayout::ETA, Z = RecLayout::ZETA, RT = RecLayout::T;
    const bool visc3 = (P.visc_type == 3 || P.visc_type == -3);
    double u[8], v[8], wv[8], e[8], T[8], ek2[4];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int q = nc8(a, b, c);
                u[q] = R(cb, RecLayout::U, a - 1, b - 1, c - 1);
                v[q] = R(cb, RecLayout::V, a - 1, b - 1, c - 1);
                wv[q] = R(cb, RecLayout::W, a - 1, b - 1, c - 1);
                e[q] = R(cb, E, a - 1, b - 1, c - 1);
                T[q] = R(cb, RT, a - 1, b - 1, c - 1);
            }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) ek2[a + 2 * b] = R(cb, E, a - 1, b - 1, -2);  // eta(k-2), order 2
    double rAv = 0., rBv = 0.;
    if (P.stoch_on) {
        if (INL) {
            double za, zb;
            noise_ab(K, NODE_FIELD, (uint32_t)(((long)k * P.n[1] + j) * P.n[0] + i), za, zb);
            rAv = za * P.var_mom;
            rBv = zb * P.var_mom;
        } else {
            rAv = rA(cb, 0);
            rBv = rB(cb, 0);
        }
    }

    // flux.cpp:1627-1703 (summation orders of the reference kept term by term)
    double muxp = 0.125 * (e[nc8(1, 0, 0)] + e[nc8(0, 0, 0)] + e[nc8(1, 1, 0)] + e[nc8(0, 1, 0)] +
                           e[nc8(1, 0, 1)] + e[nc8(0, 0, 1)] + e[nc8(1, 1, 1)] + e[nc8(0, 1, 1)]);
    double zetaxp = 0.;
    double z[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    if (visc3) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) z[nc8(a, b, c)] = R(cb, Z, a - 1, b - 1, c - 1);
        zetaxp = 0.125 * (z[nc8(1, 0, 0)] + z[nc8(0, 0, 0)] + z[nc8(1, 1, 0)] + z[nc8(0, 1, 0)] +
                          z[nc8(1, 0, 1)] + z[nc8(0, 0, 1)] + z[nc8(1, 1, 1)] + z[nc8(0, 1, 1)]);
    }
    const double DX0 = P.dx[0], DX1 = P.dx[1], DX2 = P.dx[2];
    double out[NODE_NCOMP];
#pragma unroll
    for (int bcomp = 0; bcomp < 3; ++bcomp) {
        const double* f = bcomp == 0 ? u : (bcomp == 1 ? v : wv);
        out[0 + bcomp] = 0.25 * muxp *
                         (f[nc8(1, 0, 0)] - f[nc8(0, 0, 0)] + f[nc8(1, 1, 0)] - f[nc8(0, 1, 0)] + f[nc8(1, 0, 1)] -
                          f[nc8(0, 0, 1)] + f[nc8(1, 1, 1)] - f[nc8(0, 1, 1)]) /
                         DX0;
        out[3 + bcomp] = 0.25 * muxp *
                         (f[nc8(0, 1, 0)] - f[nc8(0, 0, 0)] + f[nc8(1, 1, 0)] - f[nc8(1, 0, 0)] + f[nc8(0, 1, 1)] -
                          f[nc8(0, 0, 1)] + f[nc8(1, 1, 1)] - f[nc8(1, 0, 1)]) /
                         DX1;
        out[6 + bcomp] = 0.25 * muxp *
                         (f[nc8(0, 0, 1)] - f[nc8(0, 0, 0)] + f[nc8(1, 0, 1)] - f[nc8(1, 0, 0)] + f[nc8(0, 1, 1)] -
                          f[nc8(0, 1, 0)] + f[nc8(1, 1, 1)] - f[nc8(1, 1, 0)]) /
                         DX2;
    }
    // flux.cpp:1705-1711
    out[NODE_VC] =
        (muxp / 12. + zetaxp / 4.) *
        ((u[nc8(1, 0, 0)] - u[nc8(0, 0, 0)]) / DX0 + (u[nc8(1, 1, 0)] - u[nc8(0, 1, 0)]) / DX0 +
         (u[nc8(1, 0, 1)] - u[nc8(0, 0, 1)]) / DX0 + (u[nc8(1, 1, 1)] - u[nc8(0, 1, 1)]) / DX0 +
         (v[nc8(0, 1, 0)] - v[nc8(0, 0, 0)]) / DX1 + (v[nc8(1, 1, 0)] - v[nc8(1, 0, 0)]) / DX1 +
         (v[nc8(0, 1, 1)] - v[nc8(0, 0, 1)]) / DX1 + (v[nc8(1, 1, 1)] - v[nc8(1, 0, 1)]) / DX1 +
         (wv[nc8(0, 0, 1)] - wv[nc8(0, 0, 0)]) / DX2 + (wv[nc8(1, 0, 1)] - wv[nc8(1, 0, 0)]) / DX2 +
         (wv[nc8(0, 1, 1)] - wv[nc8(0, 1, 0)]) / DX2 + (wv[nc8(1, 1, 1)] - wv[nc8(1, 1, 0)]) / DX2);
    if (P.stoch_on) {
        // stage-weighted node noise (MultiFab::LinComb, timeStep.cpp:125-128)
        const double W = 1.0 * rAv + w * rBv;
#pragma unroll
        for (int v2 = 0; v2 < 4; ++v2) {
            double sm = 0.;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const signed char* o = kNodeOrder.o[v2][q];
                const int cell = nc8(o[0], o[1], o[2]);
                const double eta = o[3] == 0 ? e[cell] : ek2[o[0] + 2 * o[1]];
                const double term = eta * T[cell];
                sm = (q == 0) ? term : sm + term;
            }
            double mu = 0.25 * sm / 3.;
            if (visc3) {
                double zs = 0.;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const signed char* o = kNodeOrder.o[v2][q];
                    const int cell = nc8(o[0], o[1], o[2]);
                    // the zeta form keeps the reference's k-2 shift as well (flux.cpp:969-970)
                    const double zt = o[3] == 0 ? z[cell] : R(cb, Z, o[0] - 1, o[1] - 1, -2);
                    const double term = zt * T[cell];
                    zs = (q == 0) ? term : zs + term;
                }
                mu += 0.25 * zs;
            }
            out[NODE_S + v2] = sqrt(mu) * W;
        }
    }
    const int nout = P.stoch_on ? NODE_NCOMP : NODE_S;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        if (((m & 1) && i != 0) || ((m & 2) && j != 0) || ((m & 4) && k != 0)) continue;
        const int sx = (m & 1) ? P.n[0] : 0, sy = (m & 2) ? P.n[1] : 0, sz = (m & 4) ? P.n[2] : 0;
#pragma unroll
        for (int c = 0; c < NODE_NCOMP; ++c)
            if (c < nout) N(cb, c, sx, sy, sz) = out[c];
    }
}

// ---- all fluxes of one face direction -------------------------------------------------------

// Sum over the 4 nodes of face (i,j,k) of component c, in visccorn order (hi,hi),(lo,hi),(hi,lo),
// (lo,lo) over the two transverse axes, or in the z-face order (hi,hi),(hi,lo),(lo,hi),(lo,lo).
// Node index n along an axis is node 0 (periodic), hence the wrapped "hi" indices.
template <int D>
__device__ inline double node_sum(const Fld& N, int c, unsigned cb, bool zorder, double s0, bool first)
{
    // first transverse axis: x faces y, y faces x, z faces x; second: x faces z, y faces z, z faces y
    const bool ahi[4] = {true, zorder ? true : false, zorder ? false : true, false};
    const bool bhi[4] = {true, zorder ? false : true, zorder ? true : false, false};
    double s = s0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = ahi[q] ? 1 : 0, b = bhi[q] ? 1 : 0;
        double v = D == 0 ? N(cb, c, 0, a, b) : (D == 1 ? N(cb, c, a, 0, b) : N(cb, c, a, b, 0));
        s = (first && q == 0) ? v : s + v;
    }
    return s;
}

// The nvars face fluxes of face (i,j,k) of direction D into fo[] (energy fold included).
// INL: the face's A/B noise is evaluated in place from Philox key K (the values fhdx_fill_noise
// writes for that key); otherwise it is read from the step's arrays A, B.
template <int NS, int D, bool INL>
__device__ inline void face_flux(const DevParams& P, const Fld& Us, const Fld& R, const Fld& N, const Fld& A,
                                 const Fld& B, const NoiseKey& K, double w, double volinv, double dtinv, double dt,
                                 int i, int j, int k, unsigned cb, double* fo)
{
    // the face lies between cell (i,j,k) ("U", offset cb) and its lower neighbour ("L", -e)
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    constexpr int nv = NS + 5;
    const RecLayout L{NS};
    // record values of the two cells, loaded once and shared by the flux families below
    const double TU = R(cb, RecLayout::T), TL = R(cb, RecLayout::T, -ex, -ey, -ez);
    const double etaU = R(cb, RecLayout::ETA), etaL = R(cb, RecLayout::ETA, -ex, -ey, -ez);
    const double kapU = R(cb, RecLayout::KAP), kapL = R(cb, RecLayout::KAP, -ex, -ey, -ez);
    const double PU = R(cb, RecLayout::P), PL = R(cb, RecLayout::P, -ex, -ey, -ez);
    double vU[3], vL[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vU[c] = R(cb, RecLayout::U + c);
        vL[c] = R(cb, RecLayout::U + c, -ex, -ey, -ez);
    }
    // velocity component c (1..3)
    auto velU = [&](int c) { return vU[c - 1]; };
    auto velL = [&](int c) { return vL[c - 1]; };
    const bool species = P.algorithm_type == 2;
    double chiU[NS], chiL[NS], DU[NS * NS], DL[NS * NS];
    if (species) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            chiU[s] = R(cb, L.CHI() + s);
            chiL[s] = R(cb, L.CHI() + s, -ex, -ey, -ez);
        }
#pragma unroll
        for (int q = 0; q < NS * NS; ++q) {
            DU[q] = R(cb, L.D() + q);
            DL[q] = R(cb, L.D() + q, -ex, -ey, -ez);
        }
    }

    double f[nv + 4];
#pragma unroll
    for (int n = 0; n < nv + 4; ++n) f[n] = 0.;

    // ---- stochastic (flux.cpp:71-1145) ----
    if (P.stoch_on) {
        double Tu = TU, Tl = TL;
        double mup = etaU * Tu + etaL * Tl;
        double kp = kapU * Tu * Tu + kapL * Tl * Tl;
        double meanT = 0.5 * (Tu + Tl);
        double wn[nv];
        wn[0] = 0.;
#pragma unroll
        for (int n = 1; n < nv; ++n) {
            if (INL) {
                const uint32_t lin = (uint32_t)(((long)k * P.n[1] + j) * P.n[0] + i);
                const double sc = (n <= 3) ? P.var_mom : (n == 4) ? P.var_ener : P.var_mass;
                double za, zb;
                noise_ab(K, (uint32_t)(16 * D + n), lin, za, zb);
                const double a = za * sc, b = zb * sc;
                wn[n] = 1.0 * a + w * b;
            } else {
                wn[n] = 1.0 * A(cb, n) + w * B(cb, n);
            }
        }
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
            sum = N(cb, NODE_S + 0, 0, 1, 1);
            sum = sum + N(cb, NODE_S + 0, 0, 0, 1);
            sum = sum + N(cb, NODE_S + 0, 0, 1, 0);
            sum = sum + N(cb, NODE_S + 0);
        } else if (D == 1) {
            sum = N(cb, NODE_S + 0, 1, 0, 1);
            sum = sum + N(cb, NODE_S + 1, 0, 0, 1);
            sum = sum + N(cb, NODE_S + 0, 1, 0, 0);
            sum = sum + N(cb, NODE_S + 1);
        } else {
            sum = N(cb, NODE_S + 0, 1, 1, 0);
            sum = sum + N(cb, NODE_S + 1, 0, 1, 0);
            sum = sum + N(cb, NODE_S + 2, 1, 0, 0);
            sum = sum + N(cb, NODE_S + 3);
        }
        wi[1 + D] = wi[1 + D] + 0.25 * nweight * sum;
#pragma unroll
        for (int n = 1; n < 4; ++n) f[n] = f[n] + wi[n];
        f[nv] = f[nv] + wi[4];
        constexpr int t1 = (D == 0) ? 2 : 1, t2 = (D == 2) ? 2 : 3;
        double phidiag = wi[1 + D] * (velL(1 + D) + velU(1 + D));
        double phishear = wi[t1] * (velL(t1) + velU(t1)) +
                          wi[t2] * (velL(t2) + velU(t2));
        phidiag = -0.5 * phidiag;
        phishear = -0.5 * phishear;
        f[nv + 1] = f[nv + 1] - phidiag;
        f[nv + 2] = f[nv + 2] - phishear;
        if (species) {
            double yy[NS], yyp[NS], DijY[NS * NS], sqD[NS * NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                wi[5 + s] = 0.;
                yy[s] = R(cb, L.YN() + s, -ex, -ey, -ez);
                yyp[s] = R(cb, L.YN() + s);
            }
            double MWmix = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                MWmix = MWmix + 0.5 * (yy[s] + yyp[s]) / P.molmass[s];
#pragma unroll
                for (int l = 0; l < NS; ++l)
                    DijY[s * NS + l] = 0.5 * (DL[l * NS + s] * yy[l] + DU[l * NS + s] * yyp[l] +
                                              (DL[s * NS + l] * yy[s] + DU[s * NS + l] * yyp[s]));
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
                double ss = (hk + P.Runiv * meanT / P.molmass[s] * 0.5 * (chiL[s] + chiU[s])) * wi[5 + s];
                soret = soret + ss;
            }
            f[nv + 3] = f[nv + 3] + soret;
        }
    }

    // ---- diffusive (flux.cpp:1193-1623) ----
    {
        const double half = 0.5, h = P.dx[D];
        double mup = half * (etaU + etaL);
        double kp = half * (kapU + kapL);
        double tau1 = mup * (velU(1) - velL(1)) / h;
        double tau2 = mup * (velU(2) - velL(2)) / h;
        double tau3 = mup * (velU(3) - velL(3)) / h;
        double divp = 0.;
        double s1 = velL(1) + velU(1);
        double s2 = velL(2) + velU(2);
        double s3 = velL(3) + velU(3);
        double phiflx;
        if (D == 0) phiflx = tau1 * s1 + divp * s1 + tau2 * s2 + tau3 * s3;
        else if (D == 1) phiflx = tau1 * s1 + tau2 * s2 + divp * s2 + tau3 * s3;
        else phiflx = tau1 * s1 + tau2 * s2 + tau3 * s3 + divp * s3;
        double Qflux = kp * (TU - TL) / h;
        f[1] = f[1] - (D == 0 ? tau1 + divp : tau1);
        f[2] = f[2] - (D == 1 ? tau2 + divp : tau2);
        f[3] = f[3] - (D == 2 ? tau3 + divp : tau3);
        f[nv] = f[nv] - Qflux;
        f[nv + 1] = f[nv + 1] - (half * phiflx);
        double meanT = 0.5 * (TL + TU);
        double meanP = 0.5 * (PL + PU);
        if (species) {
            double dk[NS], soret[NS], Fk[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double XU = R(cb, L.X() + s), XL = R(cb, L.X() + s, -ex, -ey, -ez);
                double term1 = (XU - XL) / h;
                double mX = 0.5 * (XL + XU);
                double mY = 0.5 * (R(cb, RecLayout::Y + s, -ex, -ey, -ez) + R(cb, RecLayout::Y + s));
                double term2 = (mX - mY) * (PU - PL) / h / meanP;
                dk[s] = term1 + term2;
                double ChiX;
                if (D == 2)  // flux.cpp:1561 reads chi at k and k+1
                    ChiX = 0.5 * (chiU[s] * XL + R(cb, L.CHI() + s, 0, 0, 1) * XU);
                else
                    ChiX = 0.5 * (chiL[s] * XL + chiU[s] * XU);
                soret[s] = ChiX * (TU - TL) / h / meanT;
            }
#pragma unroll
            for (int kk = 0; kk < NS; ++kk) {
                double fk = 0.;
#pragma unroll
                for (int ll = 0; ll < NS; ++ll) {
                    double Fks = half * (DL[ll * NS + kk] + DU[ll * NS + kk]) * (dk[ll] + soret[ll]);
                    fk = fk - Fks;
                }
                Fk[kk] = fk;
            }
            double Q5 = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double hk = P.e0[s] + P.hcp[s] * meanT;
                double csum = (D == 2) ? chiU[s] + chiU[s]  // flux.cpp:1604
                                       : chiL[s] + chiU[s];
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
        double vc4 = node_sum<D>(N, NODE_VC, cb, false, 0., true);
        double div8 = node_sum<D>(N, c1, cb, zo, 0., true);
        div8 = node_sum<D>(N, c2, cb, zo, div8, false);
        double s1 = node_sum<D>(N, a1, cb, zo, 0., true);
        double s2 = node_sum<D>(N, a2, cb, zo, 0., true);
        constexpr int cn = 1 + D;
        constexpr int tc1 = (D == 0) ? 2 : 1, tc2 = (D == 2) ? 2 : 3;
        f[cn] = f[cn] - 0.25 * vc4;
        f[cn] = f[cn] + .25 * div8;
        f[tc1] = f[tc1] - .25 * s1;
        f[tc2] = f[tc2] - .25 * s2;
        double phiflx = 0.25 * (vc4 - div8) * (velL(cn) + velU(cn));
        phiflx = phiflx + .25 * s1 * (velL(tc1) + velU(tc1));
        phiflx = phiflx + .25 * s2 * (velL(tc2) + velU(tc2));
        f[nv + 1] = f[nv + 1] - 0.5 * phiflx;
    }

    // ---- hyperbolic (flux.cpp:2016-2372) ----
    if (P.advection_type == 1 || P.advection_type == 2) {
        const double wgt2 = 1. / 12.;
        const double wgt1 = 0.5 + wgt2;
        double cv[NS + 5], pv[NS + 6], Yk[NS];
        if (P.advection_type == 1) {
            // primitive interpolation: rho from cons, the rest from the record
            // value l of the primitive vector at cell (i,j,k) + m*e
            auto pr = [&](int m, int l) {
                if (l == 0) return Us(cb, 0, m * ex, m * ey, m * ez);
                if (l <= 3) return R(cb, RecLayout::U + l - 1, m * ex, m * ey, m * ez);
                if (l == 4) return R(cb, RecLayout::T, m * ex, m * ey, m * ez);
                if (l == 5) return R(cb, RecLayout::P, m * ex, m * ey, m * ez);
                return R(cb, RecLayout::Y + l - 6, m * ex, m * ey, m * ez);
            };
#pragma unroll
            for (int l = 0; l < NS + 6; ++l)
                pv[l] = wgt1 * (pr(0, l) + pr(-1, l)) - wgt2 * (pr(-2, l) + pr(1, l));
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
                cv[l] = wgt1 * (Us(cb, l) + Us(cb, l, -ex, -ey, -ez)) -
                        wgt2 * (Us(cb, l, -2 * ex, -2 * ey, -2 * ez) + Us(cb, l, ex, ey, ez));
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
    for (int n = 0; n < nv; ++n) fo[n] = f[n];
}

// Store the fluxes of face (i,j,k) of direction D; face 0 is also face n of the periodic box, so
// the update reads face i+1 without wrapping.
template <int NS, int D>
__device__ inline void store_face(const DevParams& P, const Fld& Fo, const double* fo, int i, int j, int k,
                                  unsigned cb)
{
    const int idx = D == 0 ? i : (D == 1 ? j : k);
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    const int n = P.n[D];
#pragma unroll
    for (int c = 0; c < NS + 5; ++c) Fo.put_nt(cb, c, fo[c]);
    if (idx == 0) {
#pragma unroll
        for (int c = 0; c < NS + 5; ++c) Fo.put_nt(cb, c, fo[c], ex * n, ey * n, ez * n);
    }
}

template <int NS, int D, bool INL>
__device__ inline void face_body(const DevParams& P, const Fld& Us, const Fld& R, const Fld& N, const Fld& A,
                                 const Fld& B, const NoiseKey& K, const Fld& Fo, double w, double volinv, double dtinv,
                                 double dt, int i, int j, int k, unsigned cb)
{
    double fo[NS + 5];
    face_flux<NS, D, INL>(P, Us, R, N, A, B, K, w, volinv, dtinv, dt, i, j, k, cb, fo);
    store_face<NS, D>(P, Fo, fo, i, j, k, cb);
}

// ---- stage combination (+ primitive variables after the last stage) -----------------------

// What a stage update needs for the chemistry source: the stage's cell records (T, p, X of the
// stage input), the noise key and weight of the stage, and the cell volume.
struct ChemStage {
    Fld R;
    NoiseKey K;
    double w, dV;
};

// div[] = (fx(i+1)-fx(i))/dx + (fy(j+1)-fy(j))/dy + (fz(k+1)-fz(k))/dz, evaluated left to right.
template <int NS>
__device__ inline void update_from_div(const DevParams& P, const Fld& out, const Fld& u0, const Fld& us,
                                       const double* div, double dt, int mode, const Fld& prim, int write_prim, int i,
                                       int j, int k, unsigned cb, const Fld* rec_next = nullptr, int* err = nullptr,
                                       const ChemStage* chem = nullptr)
{
    // source S: zero (main_driver.cpp:241-242) but for the species equations under chemistry
    double src[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) src[n] = 0.0;
    if (chem && P.chem.nreaction > 0) {
        const RecLayout L{NS};
        double X[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) X[s] = chem->R(cb, L.X() + s);
        const uint32_t lin = (uint32_t)(((long)k * P.n[1] + j) * P.n[0] + i);
        chem_source_cell<NS>(P, chem->R(cb, RecLayout::T), chem->R(cb, RecLayout::P), X,
                             [&](int m) { return chem_noise(chem->K, m, lin, chem->w); }, dt, chem->dV, &src[5]);
    }
    double o[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) {
        if (mode == 0) o[n] = us(cb, n) - dt * div[n] + dt * src[n];
        else if (mode == 1) o[n] = 0.25 * (3.0 * u0(cb, n) + us(cb, n) - dt * div[n] + dt * src[n]);
        else o[n] = (2. / 3.) * (0.5 * u0(cb, n) + us(cb, n) - dt * div[n] + dt * src[n]);
    }
    const double* g = P.grav;
    double r0 = us(cb, 0), r1 = us(cb, 1), r2 = us(cb, 2), r3 = us(cb, 3);
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
    store_cell<NS + 5, true>(P, out, [&](int n) { return o[n]; }, i, j, k, cb);
    // the next stage's record depends on this cell's new state only: computed here from registers
    if (rec_next) cell_record<NS>(P, o, *rec_next, err, i, j, k, cb);
    if (write_prim) {
        double pr[2 * NS + 6];
        cons_to_prim_cell<NS>(P, o, pr);
        store_cell<2 * NS + 6, true>(P, prim, [&](int n) { return pr[n]; }, i, j, k, cb);
    }
}

template <int NS>
__device__ inline void update_body(const DevParams& P, const Fld& out, const Fld& u0, const Fld& us, const Fld& fx,
                                   const Fld& fy, const Fld& fz, double dt, int mode, const Fld& prim, int write_prim,
                                   int i, int j, int k, unsigned cb, const Fld* rec_next = nullptr,
                                   int* err = nullptr, const ChemStage* chem = nullptr)
{
    double div[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n)
        // face n holds face 0 of the periodic box (store_face)
        div[n] = (fx(cb, n, 1, 0, 0) - fx(cb, n)) / P.dx[0] + (fy(cb, n, 0, 1, 0) - fy(cb, n)) / P.dx[1] +
                 (fz(cb, n, 0, 0, 1) - fz(cb, n)) / P.dx[2];
    update_from_div<NS>(P, out, u0, us, div, dt, mode, prim, write_prim, i, j, k, cb, rec_next, err, chem);
}

}  // namespace fhdx
