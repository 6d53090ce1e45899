// ref_kernels.hip -- the reference-shaped kernels of calculateFlux / RK3step on gfx950.
//
// One kernel per pass of the reference (transport, stochastic, diffusive, corner, corner-to-
// face, hyperbolic fluxes, stage combination, noise), each thread owning one cell, face or
// node.  These back the drop-in entry points that fill the reference's own intermediate
// arrays (fhdx_calculate_flux, fhdx_rk3_step_ref).  Arithmetic follows the reference
// operation by operation (see physics.hpp for why that is required).
#include "kernels.hpp"

namespace fhdx {

constexpr int BLK = 256;

// Edge-viscosity summation order per face direction and edge (pp, mp, pm, mm): offsets of the
// eta cell then of the T cell relative to the face index, flux.cpp:117-135 (x), 481-499 (y),
// 847-865 (z, including the eta(k-2)*T(k) terms of 857-858).
__constant__ signed char c_mu_order[3][4][8][6] = {
    {{{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,1,0,0,1,0},{-1,1,0,-1,1,0},{0,0,1,0,0,1},{-1,0,1,-1,0,1},{0,1,1,0,1,1},{-1,1,1,-1,1,1}},
     {{0,-1,0,0,-1,0},{-1,-1,0,-1,-1,0},{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,-1,1,0,-1,1},{-1,-1,1,-1,-1,1},{0,0,1,0,0,1},{-1,0,1,-1,0,1}},
     {{0,0,-1,0,0,-1},{-1,0,-1,-1,0,-1},{0,1,-1,0,1,-1},{-1,1,-1,-1,1,-1},{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,1,0,0,1,0},{-1,1,0,-1,1,0}},
     {{0,-1,-1,0,-1,-1},{-1,-1,-1,-1,-1,-1},{0,0,-1,0,0,-1},{-1,0,-1,-1,0,-1},{0,-1,0,0,-1,0},{-1,-1,0,-1,-1,0},{0,0,0,0,0,0},{-1,0,0,-1,0,0}}},
    {{{1,-1,0,1,-1,0},{0,-1,0,0,-1,0},{1,0,0,1,0,0},{0,0,0,0,0,0},{1,-1,1,1,-1,1},{0,-1,1,0,-1,1},{1,0,1,1,0,1},{0,0,1,0,0,1}},
     {{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0},{-1,0,1,-1,0,1},{0,0,1,0,0,1},{-1,-1,1,-1,-1,1},{0,-1,1,0,-1,1}},
     {{1,-1,-1,1,-1,-1},{0,-1,-1,0,-1,-1},{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,-1,0,1,-1,0},{0,-1,0,0,-1,0},{1,0,0,1,0,0},{0,0,0,0,0,0}},
     {{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1},{-1,-1,-1,-1,-1,-1},{0,-1,-1,0,-1,-1},{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0}}},
    {{{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,1,-1,1,1,-1},{0,1,-1,0,1,-1},{1,0,0,1,0,0},{0,0,0,0,0,0},{1,1,0,1,1,0},{0,1,0,0,1,0}},
     {{-1,1,-1,-1,1,-1},{0,1,-1,0,1,-1},{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1},{-1,1,0,-1,1,0},{0,1,0,0,1,0},{-1,0,0,-1,0,0},{0,0,0,0,0,0}},
     {{1,0,0,1,0,0},{0,0,-2,0,0,0},{1,-1,0,1,-1,0},{0,-1,-2,0,-1,0},{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,-1,-1,1,-1,-1},{0,-1,-1,0,-1,-1}},
     {{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0},{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,-1,-1,-1,-1},{0,-1,-1,0,-1,-1},{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1}}}};

// Node (rancorn) offsets of the pp, mp, pm, mm edges (flux.cpp:277-280, 644-647, 1011-1014).
__constant__ signed char c_rc_node[3][4][3] = {{{0, 1, 1}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0}},
                                               {{1, 0, 1}, {0, 0, 1}, {1, 0, 0}, {0, 0, 0}},
                                               {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 0}}};

// Wall faces: the four edge-viscosity cells in the ghost layer (normal offset -1 at a lo face,
// 0 at a hi face), transverse offsets per face direction and edge in the reference's order
// (flux.cpp:137-156 x, 501-520 y, 867-886 z).
__constant__ signed char c_mu_wall[3][4][4][3] = {
    {{{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}}, {{0, -1, 0}, {0, 0, 0}, {0, -1, 1}, {0, 0, 1}},
     {{0, 0, -1}, {0, 1, -1}, {0, 0, 0}, {0, 1, 0}}, {{0, -1, -1}, {0, 0, -1}, {0, -1, 0}, {0, 0, 0}}},
    {{{1, 0, 0}, {0, 0, 0}, {1, 0, 1}, {0, 0, 1}}, {{-1, 0, 0}, {0, 0, 0}, {-1, 0, 1}, {0, 0, 1}},
     {{1, 0, -1}, {0, 0, -1}, {1, 0, 0}, {0, 0, 0}}, {{-1, 0, -1}, {0, 0, -1}, {-1, 0, 0}, {0, 0, 0}}},
    {{{1, 0, 0}, {0, 0, 0}, {1, 1, 0}, {0, 1, 0}}, {{-1, 1, 0}, {0, 1, 0}, {-1, 0, 0}, {0, 0, 0}},
     {{1, 0, 0}, {0, 0, 0}, {1, -1, 0}, {0, -1, 0}}, {{-1, -1, 0}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}}}};

// Where a face of direction D stands relative to the walls: lo/hi are set on a bc_mass 1/2
// face at index 0 / n_D (the is_*_dirichlet_mass branches of flux.cpp); (wi, wj, wk) is the
// wall-side cell (the ghost at a lo face, the boundary cell at a hi face).
struct WallFace {
    bool lo, hi;
    int wi, wj, wk;
    __device__ bool any() const { return lo || hi; }
};

template <int D>
__device__ inline WallFace wall_face(const DevParams& P, int i, int j, int k)
{
    WallFace w;
    const int c = (D == 0) ? i : (D == 1) ? j : k;
    w.lo = P.any_wall && c == 0 && P.wall[2 * D];
    w.hi = P.any_wall && c == P.n[D] && P.wall[2 * D + 1];
    w.wi = i - (w.lo && D == 0);
    w.wj = j - (w.lo && D == 1);
    w.wk = k - (w.lo && D == 2);
    return w;
}

__device__ inline void face_index(long t, int fx, int fy, int& i, int& j, int& k)
{
    i = (int)(t % fx);
    long r = t / fx;
    j = (int)(r % fy);
    k = (int)(r / fy);
}

// Thread t of a launch over the launch box's cells grown by (ex, ey, ez) on the hi side (faces
// of direction d: 1 along d; nodes: 1 along every axis), as global indices (i, j, k); false
// past the end.  Wall and periodic tests keep using the domain size P.n.
__device__ inline bool box_index(const DevParams& P, long t, int ex, int ey, int ez, int& i, int& j, int& k)
{
    const int fx = P.bn[0] + ex, fy = P.bn[1] + ey, fz = P.bn[2] + ez;
    if (t >= (long)fx * fy * fz) return false;
    face_index(t, fx, fy, i, j, k);
    i += P.bo[0];
    j += P.bo[1];
    k += P.bo[2];
    return true;
}

// ---- cell kernels -------------------------------------------------------------------------

template <int NS>
__global__ void __launch_bounds__(BLK) k_c2p(DevParams P, Arr prim, Arr cons)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    int i, j, k;
    if (!box_index(P, t, 0, 0, 0, i, j, k)) return;
    double c[NS + 5], pr[2 * NS + 6];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) c[n] = cons(i, j, k, n);
    cons_to_prim_cell<NS>(P, c, pr);
#pragma unroll
    for (int n = 0; n < 2 * NS + 6; ++n) prim(i, j, k, n) = pr[n];
}

template <int NS, bool VW>
__global__ void __launch_bounds__(BLK) k_transport(DevParams P, Arr prim, Arr eta, Arr zeta,
                                                   Arr kappa, Arr chi, Arr Dm, int* err)
{
    const int g = P.ngc;
    const int gx = P.bn[0] + 2 * g, gy = P.bn[1] + 2 * g, gz = P.bn[2] + 2 * g;
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)gx * gy * gz) return;
    int i, j, k;
    face_index(t, gx, gy, i, j, k);
    i += P.bo[0] - g;
    j += P.bo[1] - g;
    k += P.bo[2] - g;
    double Y[NS], chiv[NS], Dv[NS * NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) Y[n] = prim(i, j, k, 6 + n);
    Transport tr;
    if constexpr (VW) {  // transport_type 2
        const int fl = transport_cell_vw<NS>(P, prim(i, j, k, 0), prim(i, j, k, 4), prim(i, j, k, 5), Y, tr, chiv, Dv);
        if (fl) atomicOr(err, fl);
    } else {
        bool bad = transport_cell<NS>(P, prim(i, j, k, 0), prim(i, j, k, 4), prim(i, j, k, 5), Y, tr,
                                      chiv, Dv);
        if (bad) atomicOr(err, 1);
    }
    eta(i, j, k) = tr.eta;
    zeta(i, j, k) = tr.zeta;
    kappa(i, j, k) = tr.kappa;
#pragma unroll
    for (int n = 0; n < NS; ++n) chi(i, j, k, n) = chiv[n];
#pragma unroll
    for (int n = 0; n < NS * NS; ++n) Dm(i, j, k, n) = Dv[n];
}

// Periodic ghost fill of a cell FAB: ghost (i,j,k) <- valid (i mod n, ...) along periodic
// directions; a ghost cell outside a wall has no periodic image and is left to setBC.
// blockIdx.y selects the array: the step fills the stage state and its primitive variables
// (independent arrays) in one launch.
__global__ void __launch_bounds__(BLK) k_fill_boundary(DevParams P, Arr a0, int nc0, Arr a1, int nc1)
{
    const Arr a = blockIdx.y == 0 ? a0 : a1;
    const int nc = blockIdx.y == 0 ? nc0 : nc1;
    const int g = P.ngc, n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    const int gx = n0 + 2 * g, gy = n1 + 2 * g, gz = n2 + 2 * g;
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)gx * gy * gz) return;
    int i, j, k;
    face_index(t, gx, gy, i, j, k);
    i -= g;
    j -= g;
    k -= g;
    if ((!P.periodic[0] && (i < 0 || i >= n0)) || (!P.periodic[1] && (j < 0 || j >= n1)) ||
        (!P.periodic[2] && (k < 0 || k >= n2)))
        return;
    // the image of a ghost cell, any number of periods away (boxes thinner than ngc included)
    int iv = (i % n0 + n0) % n0, jv = (j % n1 + n1) % n1, kv = (k % n2 + n2) % n2;
    if (iv == i && jv == j && kv == k) return;
    for (int n = 0; n < nc; ++n) a(i, j, k, n) = a(iv, jv, kv, n);
}

// One face of setBC (boundary.cpp:180-1036): every cell of the grown box outside face
// (D, side) takes its value from the mirror cell (-c-1 at lo, 2n-c-1 at hi) and the wall data:
// mass fractions mirrored (bc_mass 1) or fixed (2); T and p mirrored (adiabatic) or T = t_wall
// with p mirrored (isothermal); velocity: slip (normal 0, tangential mirrored) or no slip (0);
// then rho from GetDensity and the conserved state from it.
// One launch per direction covers the ghost slabs of both faces (2*ngc layers across D, the
// grown box along the other two): the lo and hi passes write disjoint cells and read only
// mirror cells inside the box, so running them together equals running them in turn.
template <int NS>
__global__ void __launch_bounds__(BLK) k_set_bc(DevParams P, Arr prim, Arr cons, int D)
{
    const int g = P.ngc, n = P.n[D];
    int ext[3] = {P.bn[0] + 2 * g, P.bn[1] + 2 * g, P.bn[2] + 2 * g};
    ext[D] = 2 * g;
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= (long)ext[0] * ext[1] * ext[2]) return;
    int c[3];
    face_index(t, ext[0], ext[1], c[0], c[1], c[2]);
    for (int d = 0; d < 3; ++d)
        if (d != D) c[d] += P.bo[d] - g;
    const int side = c[D] >= g;  // layers 0..g-1: lo ghosts -g..-1; g..2g-1: hi ghosts n..n+g-1
    // a box that does not reach this side of the domain has no wall ghosts there
    if (side ? P.bo[D] + P.bn[D] != n : P.bo[D] != 0) return;
    c[D] = side ? n + (c[D] - g) : c[D] - g;
    int m[3] = {c[0], c[1], c[2]};
    m[D] = side ? 2 * (n - 1) - c[D] + 1 : -c[D] - 1;
    const int f = 2 * D + side;
    const int bm = P.bcm[f], bt = P.bct[f], bv = P.bcv[f];
#define PR(a, q) prim(a[0], a[1], a[2], q)
#define CO(a, q) cons(a[0], a[1], a[2], q)
    if (bm == 1 && P.algorithm_type == 2)
#pragma unroll
        for (int q = 6; q < 6 + 2 * NS; ++q) PR(c, q) = PR(m, q);
    if (bm == 2 && P.algorithm_type == 2)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            PR(c, 6 + s) = P.Yw[f][s];
            PR(c, 6 + NS + s) = P.Xw[f][s];
        }
    if (bm == 3 && P.algorithm_type == 2) {
        // x reservoir (boundary.cpp:221-238, 260-288): the reservoir gas at rest; no thermal or velocity
        // branch follows (the x chain is else-if)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            PR(c, 6 + s) = P.Yw[f][s];
            PR(c, 6 + NS + s) = P.Xw[f][s];
        }
        const double rho = P.rhow[f];
        PR(c, 0) = rho;
        CO(c, 0) = rho;
        PR(c, 4) = P.tw[f];
        PR(c, 5) = P.pw[f];
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            CO(c, q) = 0.0;
            PR(c, q) = 0.0;
        }
        double Yr[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) Yr[s] = P.Yw[f][s];
        const double ie = eos_energy<NS>(P, Yr, P.tw[f]);
#pragma unroll
        for (int s = 0; s < NS; ++s) CO(c, 5 + s) = rho * Yr[s];
        CO(c, 4) = rho * ie + 0.5 * rho * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
        return;
    }
    if (bt == 1) {
        PR(c, 4) = PR(m, 4);
        PR(c, 5) = PR(m, 5);
    } else if (bt == 2) {
        PR(c, 4) = P.tw[f];
        PR(c, 5) = PR(m, 5);
    }
    if (bv != 1 && bv != 2) return;
    double v[4];
    if (bv == 1) {
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            const bool normal = (q == 1 + D);
            CO(c, q) = normal ? 0.0 : CO(m, q);
            v[q] = normal ? 0.0 : PR(m, q);
            PR(c, q) = v[q];
        }
    } else {
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            v[q] = 0.0;
            PR(c, q) = 0.0;
        }
    }
    double Y[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) Y[s] = PR(c, 6 + s);
    const double T = PR(c, 4), pt = PR(c, 5);
    double molmix = 0.;
#pragma unroll
    for (int s = 0; s < NS; ++s) molmix += Y[s] / P.molmass[s];
    molmix = 1. / molmix;
    const double rho = pt / (P.Runiv / molmix) / T;  // GetDensity, compressible_functions.H:1481-1495
    const double ie = eos_energy<NS>(P, Y, T);
    PR(c, 0) = rho;
    CO(c, 0) = rho;
    if (P.algorithm_type == 2)
#pragma unroll
        for (int s = 0; s < NS; ++s) CO(c, 5 + s) = rho * Y[s];
    if (bv == 2)
#pragma unroll
        for (int q = 1; q < 4; ++q) CO(c, q) = rho * v[q];
    CO(c, 4) = rho * ie + 0.5 * rho * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
#undef PR
#undef CO
}

__global__ void __launch_bounds__(BLK) k_fill(double* p, long n, double v)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t < n) p[t] = v;
}

// ---- stochastic face flux (flux.cpp:71-1145) -----------------------------------------------
// At a wall face (bc_mass 1/2) the face-local averages read the wall-side cell for both sides:
// 2*eta*T, 2*kappa*T^2, T, 2*w*u and the species terms of flux.cpp:88-97, 137-202, 318-327,
// 343-427 are then reproduced bit for bit (every rewrite is an exact doubling or halving).
// The edge viscosities take the four ghost-layer cells, and edges touching a tangential wall
// are scaled by sqrt(2) (slip) or 0 (no slip) (:250-280).  StochFlux (boundary.cpp:1392-1970)
// then scales the face's stochastic fluxes, applied here to the values this thread wrote.

// The step's noise of one stage: with WN the weighted values 1*A + w*B are formed here (and
// stored into ran for the face, as MultiFab::LinComb leaves them); otherwise ran holds them.
struct NoiseW {
    Arr A, B, rA, rB;
    double w;
};

template <int NS, int D, bool WN>
__device__ __forceinline__ void stoch_body(const DevParams& P, Arr prim, Arr eta, Arr zeta, Arr kappa, Arr chi,
                                           Arr Dm, Arr fl, Arr ran, Arr rc, const NoiseW& nw, double volinv,
                                           double dtinv, double dt, long t)
{
    int i, j, k;
    if (!box_index(P, t, D == 0, D == 1, D == 2, i, j, k)) return;
    if constexpr (WN) {
#pragma unroll
        for (int c = 0; c < NS + 5; ++c)
            ran(i, j, k, c) = (c < 1) ? 0.0 : 1.0 * nw.A(i, j, k, c) + nw.w * nw.B(i, j, k, c);
    }
    const WallFace W = wall_face<D>(P, i, j, k);
    // lower and upper cells of the face averages; both the wall-side cell at a wall
    const int il = W.any() ? W.wi : i - (D == 0), jl = W.any() ? W.wj : j - (D == 1),
              kl = W.any() ? W.wk : k - (D == 2);
    const int iu = W.any() ? W.wi : i, ju = W.any() ? W.wj : j, ku = W.any() ? W.wk : k;
    const int nv = NS + 5;

    double Tu = prim(iu, ju, ku, 4), Tl = prim(il, jl, kl, 4);
    double mup = eta(iu, ju, ku) * Tu + eta(il, jl, kl) * Tl;
    double kp = kappa(iu, ju, ku) * Tu * Tu + kappa(il, jl, kl) * Tl * Tl;
    double meanT = 0.5 * (Tu + Tl);
    double w[NS + 5];
    double fw1 = sqrt(P.k_B * mup * volinv * dtinv);
    double fw4 = sqrt(P.k_B * kp * volinv * dtinv);
    w[0] = 0 * ran(i, j, k, 0);
    w[1] = fw1 * ran(i, j, k, 1);
    w[2] = fw1 * ran(i, j, k, 2);
    w[3] = fw1 * ran(i, j, k, 3);
    w[4] = fw4 * ran(i, j, k, 4);
    double nweight = sqrt(P.k_B * volinv * dtinv);

    // tangential-wall factors of the four edges (pp, mp, pm, mm) over transverse axes (A, B)
    constexpr int TA = (D == 0) ? 1 : 0, TB = (D == 2) ? 1 : 2;
    double fac[4] = {1., 1., 1., 1.};
    if (P.any_wall) {
        const int idx[3] = {i, j, k};
        double flo[2] = {1., 1.}, fhi[2] = {1., 1.};
        const int ax[2] = {TA, TB};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int a = ax[q];
            const int vl = P.bcv[2 * a], vh = P.bcv[2 * a + 1];
            if ((vl == 1 || vl == 2) && idx[a] == 0) flo[q] = (vl == 1) ? sqrt(2.0) : 0.;
            if ((vh == 1 || vh == 2) && idx[a] == P.n[a] - 1) fhi[q] = (vh == 1) ? sqrt(2.0) : 0.;
        }
        fac[0] = fhi[0] * fhi[1];
        fac[1] = flo[0] * fhi[1];
        fac[2] = fhi[0] * flo[1];
        fac[3] = flo[0] * flo[1];
    }

    double sum = 0.;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double mu;
        if (W.any()) {
            const int oi = W.wi - i, oj = W.wj - j, ok = W.wk - k;  // ghost-layer shift
            double s4 = 0., z4 = 0.;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const signed char* o = c_mu_wall[D][c][q];
                const int ci = i + oi + o[0], cj = j + oj + o[1], ck = k + ok + o[2];
                const double T = prim(ci, cj, ck, 4);
                const double e = eta(ci, cj, ck) * T;
                s4 = (q == 0) ? e : s4 + e;
                if (P.visc_type == 3 || P.visc_type == -3) {
                    const double zz = zeta(ci, cj, ck) * T;
                    z4 = (q == 0) ? zz : z4 + zz;
                }
            }
            mu = 0.5 * s4 / 3.;
            if (P.visc_type == 3 || P.visc_type == -3) mu += (D == 0) ? 0.5 * z4 : 0.5 * z4 / 3.;
        } else {
            double s8 = 0.;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const signed char* o = c_mu_order[D][c][q];
                double term = eta(i + o[0], j + o[1], k + o[2]) * prim(i + o[3], j + o[4], k + o[5], 4);
                s8 = (q == 0) ? term : s8 + term;
            }
            mu = 0.25 * s8 / 3.;
            if (P.visc_type == 3 || P.visc_type == -3) {
                double z = 0.;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const signed char* o = c_mu_order[D][c][q];
                    double term = zeta(i + o[0], j + o[1], k + o[2]) * prim(i + o[3], j + o[4], k + o[5], 4);
                    z = (q == 0) ? term : z + term;
                }
                mu += 0.25 * z;
            }
        }
        const signed char* nd = c_rc_node[D][c];
        double rcv;
        if constexpr (WN) rcv = 1.0 * nw.rA(i + nd[0], j + nd[1], k + nd[2]) + nw.w * nw.rB(i + nd[0], j + nd[1], k + nd[2]);
        else rcv = rc(i + nd[0], j + nd[1], k + nd[2]);
        double term = fac[c] * sqrt(mu) * rcv;
        sum = (c == 0) ? term : sum + term;
    }
    w[1 + D] = w[1 + D] + 0.25 * nweight * sum;

    // this face's flux after the stochastic pass, every one of the nvars + 4 components: the
    // reference adds to a zeroed array (flux.cpp:24-26), so each term is 0.0 + term
    double o[NS + 9];
#pragma unroll
    for (int n = 0; n < NS + 9; ++n) o[n] = 0.0;
#pragma unroll
    for (int n = 1; n < 4; ++n) o[n] = 0.0 + w[n];
    o[nv] = 0.0 + w[4];

    constexpr int t1 = (D == 0) ? 2 : 1, t2 = (D == 2) ? 2 : 3;
    double phidiag = w[1 + D] * (prim(il, jl, kl, 1 + D) + prim(iu, ju, ku, 1 + D));
    double phishear = w[t1] * (prim(il, jl, kl, t1) + prim(iu, ju, ku, t1)) +
                      w[t2] * (prim(il, jl, kl, t2) + prim(iu, ju, ku, t2));
    phidiag = -0.5 * phidiag;
    phishear = -0.5 * phishear;
    o[nv + 1] = 0.0 - phidiag;
    o[nv + 2] = 0.0 - phishear;

    if (P.algorithm_type == 2) {
        double yy[NS], yyp[NS], DijY[NS * NS], sqD[NS * NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            w[5 + s] = 0.;
            yy[s] = clamp01(prim(il, jl, kl, 6 + s));
            yyp[s] = clamp01(prim(iu, ju, ku, 6 + s));
        }
        double sumy = 0., sumyp = 0.;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            sumy += yy[n];
            sumyp += yyp[n];
        }
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            yy[n] /= sumy;
            yyp[n] /= sumyp;
        }
        double MWmix = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            MWmix = MWmix + 0.5 * (yy[s] + yyp[s]) / P.molmass[s];
#pragma unroll
            for (int l = 0; l < NS; ++l)
                DijY[s * NS + l] = 0.5 * (Dm(il, jl, kl, l * NS + s) * yy[l] + Dm(iu, ju, ku, l * NS + s) * yyp[l] +
                                          (Dm(il, jl, kl, s * NS + l) * yy[s] + Dm(iu, ju, ku, s * NS + l) * yyp[s]));
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
            for (int l = 0; l <= s; ++l) w[5 + s] = w[5 + s] + (wsp * sqD[s * NS + l]) * ran(i, j, k, 5 + l);
            o[5 + s] = w[5 + s];
        }
        double soret = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double hk = P.e0[s] + P.hcp[s] * meanT;
            double ss = (hk + P.Runiv * meanT / P.molmass[s] * 0.5 * (chi(il, jl, kl, s) + chi(iu, ju, ku, s))) *
                        w[5 + s];
            soret = soret + ss;
        }
        o[nv + 3] = o[nv + 3] + soret;
    }

    if (P.any_wall) {  // StochFlux on the wall planes of direction D
        const int c = (D == 0) ? i : (D == 1) ? j : k;
        const int f = (c == 0) ? 2 * D : (c == P.n[D]) ? 2 * D + 1 : -1;
        if (f >= 0) {
            const int bm = P.bcm[f], bt = P.bct[f], bv = P.bcv[f];
            if (bm == 1 || bm == 2) {
                const double factor = (bm == 1) ? 0. : sqrt(2.);
                if (P.algorithm_type == 2)
#pragma unroll
                    for (int s = 0; s < NS; ++s) o[5 + s] *= factor;
                o[nv + 3] *= factor;
            }
            if (bt == 1 || bt == 2) {
                double factor = (bt == 1) ? 0. : sqrt(2.);
                if (D == 0 && bm == 3) factor = 1.0;
                o[nv] *= factor;
            }
            if (bv == 1 || bv == 2) {
                const double factor = (bv == 1) ? 0. : sqrt(2.);
                const double sqrtTwo = sqrt(2.);
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    if (n == D) {
                        o[1 + n] *= sqrtTwo;
                        o[nv + 1] *= sqrtTwo;
                        o[nv + 2] *= factor;
                    } else {
                        o[1 + n] *= factor;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int n = 0; n < NS + 9; ++n) fl(i, j, k, n) = o[n];
}

// ---- diffusive face flux (flux.cpp:1193-1623) ----------------------------------------------
// At a wall face the gradients span half a cell and the coefficients, means and the Soret
// chi*X read the wall-side cell (flux.cpp:1218-1290 and the y/z images; the z-face chi quirks
// of :1561, :1604 are not in those branches).  BCWallSpeciesFlux (boundary.cpp:1212-1390) then
// zeroes species, density and Dufour fluxes on bc_mass 1 faces; the corner pass in between
// touches none of those components, so it is applied here.

template <int NS, int D>
__device__ __forceinline__ void diff_body(const DevParams& P, Arr prim, Arr eta, Arr kappa, Arr chi, Arr Dm,
                                          Arr fl, long t)
{
    int i, j, k;
    if (!box_index(P, t, D == 0, D == 1, D == 2, i, j, k)) return;
    const int il = i - (D == 0), jl = j - (D == 1), kl = k - (D == 2);
    const WallFace W = wall_face<D>(P, i, j, k);
    const bool wall = W.any();
    // coefficient / mean cells: the two neighbours, or the wall-side cell twice
    const int ai = wall ? W.wi : il, aj = wall ? W.wj : jl, ak = wall ? W.wk : kl;
    const int bi = wall ? W.wi : i, bj = wall ? W.wj : j, bk = wall ? W.wk : k;
    const int nv = NS + 5;
    const double half = 0.5, h = wall ? 0.5 * P.dx[D] : P.dx[D];

    double mup = half * (eta(bi, bj, bk) + eta(ai, aj, ak));
    double kp = half * (kappa(bi, bj, bk) + kappa(ai, aj, ak));
    double tau1 = mup * (prim(i, j, k, 1) - prim(il, jl, kl, 1)) / h;
    double tau2 = mup * (prim(i, j, k, 2) - prim(il, jl, kl, 2)) / h;
    double tau3 = mup * (prim(i, j, k, 3) - prim(il, jl, kl, 3)) / h;
    double divp = 0.;
    double s1 = prim(ai, aj, ak, 1) + prim(bi, bj, bk, 1);
    double s2 = prim(ai, aj, ak, 2) + prim(bi, bj, bk, 2);
    double s3 = prim(ai, aj, ak, 3) + prim(bi, bj, bk, 3);
    double phiflx;
    if (D == 0) phiflx = tau1 * s1 + divp * s1 + tau2 * s2 + tau3 * s3;
    else if (D == 1) phiflx = tau1 * s1 + tau2 * s2 + divp * s2 + tau3 * s3;
    else phiflx = tau1 * s1 + tau2 * s2 + tau3 * s3 + divp * s3;
    double Qflux = kp * (prim(i, j, k, 4) - prim(il, jl, kl, 4)) / h;
    fl(i, j, k, 1) = fl(i, j, k, 1) - (D == 0 ? tau1 + divp : tau1);
    fl(i, j, k, 2) = fl(i, j, k, 2) - (D == 1 ? tau2 + divp : tau2);
    fl(i, j, k, 3) = fl(i, j, k, 3) - (D == 2 ? tau3 + divp : tau3);
    fl(i, j, k, nv) = fl(i, j, k, nv) - Qflux;
    fl(i, j, k, nv + 1) = fl(i, j, k, nv + 1) - (half * phiflx);

    double meanT = 0.5 * (prim(ai, aj, ak, 4) + prim(bi, bj, bk, 4));
    double meanP = 0.5 * (prim(ai, aj, ak, 5) + prim(bi, bj, bk, 5));
    if (P.algorithm_type == 2) {
        double dk[NS], soret[NS], Fk[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double XU = prim(i, j, k, 6 + NS + s), XL = prim(il, jl, kl, 6 + NS + s);
            double term1 = (XU - XL) / h;
            double mX = 0.5 * (prim(ai, aj, ak, 6 + NS + s) + prim(bi, bj, bk, 6 + NS + s));
            double mY = 0.5 * (prim(ai, aj, ak, 6 + s) + prim(bi, bj, bk, 6 + s));
            double term2 = (mX - mY) * (prim(i, j, k, 5) - prim(il, jl, kl, 5)) / h / meanP;
            dk[s] = term1 + term2;
            double ChiX;
            if (D == 2 && !wall)  // flux.cpp:1561 reads chi at k and k+1
                ChiX = 0.5 * (chi(i, j, k, s) * XL + chi(i, j, k + 1, s) * XU);
            else
                ChiX = 0.5 * (chi(ai, aj, ak, s) * prim(ai, aj, ak, 6 + NS + s) +
                              chi(bi, bj, bk, s) * prim(bi, bj, bk, 6 + NS + s));
            soret[s] = ChiX * (prim(i, j, k, 4) - prim(il, jl, kl, 4)) / h / meanT;
        }
#pragma unroll
        for (int kk = 0; kk < NS; ++kk) {
            double f = 0.;
#pragma unroll
            for (int ll = 0; ll < NS; ++ll) {
                double Fks = half * (Dm(ai, aj, ak, ll * NS + kk) + Dm(bi, bj, bk, ll * NS + kk)) * (dk[ll] + soret[ll]);
                f = f - Fks;
            }
            Fk[kk] = f;
        }
        double Q5 = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double hk = P.e0[s] + P.hcp[s] * meanT;
            double csum = (D == 2 && !wall) ? chi(i, j, k, s) + chi(i, j, k, s)  // flux.cpp:1604
                                            : chi(ai, aj, ak, s) + chi(bi, bj, bk, s);
            double Q5s = (hk + 0.5 * P.Runiv * meanT * csum / P.molmass[s]) * Fk[s];
            Q5 = Q5 + Q5s;
        }
        fl(i, j, k, nv + 3) = fl(i, j, k, nv + 3) + Q5;
#pragma unroll
        for (int s = 0; s < NS; ++s) fl(i, j, k, 5 + s) = fl(i, j, k, 5 + s) + Fk[s];
    }
    if (wall && P.bcm[2 * D + (W.hi ? 1 : 0)] == 1) {  // BCWallSpeciesFlux
#pragma unroll
        for (int s = 0; s < NS; ++s) fl(i, j, k, 5 + s) = 0.;
        fl(i, j, k, 0) = 0.;
        fl(i, j, k, nv + 3) = 0.;
    }
}

// ---- corner viscous terms at nodes (flux.cpp:1627-1713) ------------------------------------

__global__ void __launch_bounds__(BLK) k_corner_node(DevParams P, Arr prim, Arr eta, Arr zeta,
                                                     Arr cux, Arr cvx, Arr cwx, Arr cuy, Arr cvy,
                                                     Arr cwy, Arr cuz, Arr cvz, Arr cwz, Arr vcorn, Arr rancorn,
                                                     Arr rA, Arr rB, double w, int weigh)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    int i, j, k;
    if (!box_index(P, t, 1, 1, 1, i, j, k)) return;
    if (weigh) rancorn(i, j, k) = 1.0 * rA(i, j, k) + w * rB(i, j, k);  // the node's weighted noise
    double muxp = 0.125 * (eta(i, j - 1, k - 1) + eta(i - 1, j - 1, k - 1) + eta(i, j, k - 1) +
                           eta(i - 1, j, k - 1) + eta(i, j - 1, k) + eta(i - 1, j - 1, k) + eta(i, j, k) +
                           eta(i - 1, j, k));
    double zetaxp = 0.;
    if (P.visc_type == 3 || P.visc_type == -3)
        zetaxp = 0.125 * (zeta(i, j - 1, k - 1) + zeta(i - 1, j - 1, k - 1) + zeta(i, j, k - 1) +
                          zeta(i - 1, j, k - 1) + zeta(i, j - 1, k) + zeta(i - 1, j - 1, k) + zeta(i, j, k) +
                          zeta(i - 1, j, k));
    double DXa[3] = {P.dx[0], P.dx[1], P.dx[2]};
    if (P.any_wall) {
        // wall nodes (flux.cpp:1647-1682): half spacing across each wall met, viscosities from
        // the four ghost-layer cells of the last wall met (the checks run lo x .. hi z)
        const int nd[3] = {i, j, k};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const int d = f >> 1, side = f & 1;
            if (!P.wall[f] || nd[d] != (side ? P.n[d] : 0)) continue;
            DXa[d] = 0.5 * P.dx[d];
            const int a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
            double e4 = 0., z4 = 0.;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int c[3];
                c[d] = side ? nd[d] : nd[d] - 1;
                c[a] = nd[a] - 1 + (q >> 1);
                c[b] = nd[b] - 1 + (q & 1);
                e4 = (q == 0) ? eta(c[0], c[1], c[2]) : e4 + eta(c[0], c[1], c[2]);
                z4 = (q == 0) ? zeta(c[0], c[1], c[2]) : z4 + zeta(c[0], c[1], c[2]);
            }
            muxp = 0.25 * e4;
            zetaxp = (P.visc_type == 3 || P.visc_type == -3) ? 0.25 * z4 : 0.;
        }
    }
    const double DX0 = DXa[0], DX1 = DXa[1], DX2 = DXa[2];
    Arr gx[3] = {cux, cvx, cwx}, gy[3] = {cuy, cvy, cwy}, gz[3] = {cuz, cvz, cwz};
#pragma unroll
    for (int n = 1; n < 4; ++n) {
        gx[n - 1](i, j, k) = 0.25 * muxp *
                             (prim(i, j - 1, k - 1, n) - prim(i - 1, j - 1, k - 1, n) + prim(i, j, k - 1, n) -
                              prim(i - 1, j, k - 1, n) + prim(i, j - 1, k, n) - prim(i - 1, j - 1, k, n) +
                              prim(i, j, k, n) - prim(i - 1, j, k, n)) /
                             DX0;
        gy[n - 1](i, j, k) = 0.25 * muxp *
                             (prim(i - 1, j, k - 1, n) - prim(i - 1, j - 1, k - 1, n) + prim(i, j, k - 1, n) -
                              prim(i, j - 1, k - 1, n) + prim(i - 1, j, k, n) - prim(i - 1, j - 1, k, n) +
                              prim(i, j, k, n) - prim(i, j - 1, k, n)) /
                             DX1;
        gz[n - 1](i, j, k) = 0.25 * muxp *
                             (prim(i - 1, j - 1, k, n) - prim(i - 1, j - 1, k - 1, n) + prim(i, j - 1, k, n) -
                              prim(i, j - 1, k - 1, n) + prim(i - 1, j, k, n) - prim(i - 1, j, k - 1, n) +
                              prim(i, j, k, n) - prim(i, j, k - 1, n)) /
                             DX2;
    }
    vcorn(i, j, k) =
        (muxp / 12. + zetaxp / 4.) *
        ((prim(i, j - 1, k - 1, 1) - prim(i - 1, j - 1, k - 1, 1)) / DX0 +
         (prim(i, j, k - 1, 1) - prim(i - 1, j, k - 1, 1)) / DX0 +
         (prim(i, j - 1, k, 1) - prim(i - 1, j - 1, k, 1)) / DX0 + (prim(i, j, k, 1) - prim(i - 1, j, k, 1)) / DX0 +
         (prim(i - 1, j, k - 1, 2) - prim(i - 1, j - 1, k - 1, 2)) / DX1 +
         (prim(i, j, k - 1, 2) - prim(i, j - 1, k - 1, 2)) / DX1 +
         (prim(i - 1, j, k, 2) - prim(i - 1, j - 1, k, 2)) / DX1 + (prim(i, j, k, 2) - prim(i, j - 1, k, 2)) / DX1 +
         (prim(i - 1, j - 1, k, 3) - prim(i - 1, j - 1, k - 1, 3)) / DX2 +
         (prim(i, j - 1, k, 3) - prim(i, j - 1, k - 1, 3)) / DX2 +
         (prim(i - 1, j, k, 3) - prim(i - 1, j, k - 1, 3)) / DX2 + (prim(i, j, k, 3) - prim(i, j, k - 1, 3)) / DX2);
}

// ---- corner-to-face terms (flux.cpp:1770-1980) -------------------------------------------
// Sums over the 4 nodes of a face: visccorn in (hi,hi),(lo,hi),(hi,lo),(lo,lo) transverse
// order on every face; the corner arrays in that order on x/y faces and in
// (hi,hi),(hi,lo),(lo,hi),(lo,lo) order on z faces.

// Transverse offsets of face direction D: the first transverse axis moves (AI, AJ, 0), the
// second (0, BJ, BK): x faces (y, z), y faces (x, z), z faces (x, y).
template <int D>
struct Transverse {
    static constexpr int AI = (D != 0), AJ = (D == 0), BJ = (D == 2), BK = (D != 2);
};

// Left-to-right sum of `a` over the 4 nodes of face (i,j,k), starting from `s0` unless
// `first`; zorder selects (hi,hi),(hi,lo),(lo,hi),(lo,lo) instead of (hi,hi),(lo,hi),(hi,lo),(lo,lo).
template <int D>
__device__ inline double node4(const Arr& a, int i, int j, int k, bool zorder, double s0 = 0., bool first = true)
{
    using T = Transverse<D>;
    const int oa[4] = {1, zorder ? 1 : 0, zorder ? 0 : 1, 0};
    const int ob[4] = {1, zorder ? 0 : 1, zorder ? 1 : 0, 0};
    double s = s0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = a(i + oa[q] * T::AI, j + oa[q] * T::AJ + ob[q] * T::BJ, k + ob[q] * T::BK);
        s = (first && q == 0) ? v : s + v;
    }
    return s;
}

template <int NS, int D>
__device__ __forceinline__ void corner_face_body(const DevParams& P, Arr prim, Arr fl, Arr vcorn, Arr c1,
                                                 Arr c2, Arr ta1, Arr ta2, long t)
{
    int i, j, k;
    if (!box_index(P, t, D == 0, D == 1, D == 2, i, j, k)) return;
    const WallFace W = wall_face<D>(P, i, j, k);
    // velocity cells of the heating term: the two neighbours, or the wall-side cell twice
    // (flux.cpp:1791-1822: 0.5*x*u_wall equals 0.25*x*(u_wall + u_wall) exactly)
    const int il = W.any() ? W.wi : i - (D == 0), jl = W.any() ? W.wj : j - (D == 1),
              kl = W.any() ? W.wk : k - (D == 2);
    const int iu = W.any() ? W.wi : i, ju = W.any() ? W.wj : j, ku = W.any() ? W.wk : k;
    const int nv = NS + 5;
    const bool zo = (D == 2);
    double vc4 = node4<D>(vcorn, i, j, k, false);
    double div8 = node4<D>(c1, i, j, k, zo);
    div8 = node4<D>(c2, i, j, k, zo, div8, false);
    double s1 = node4<D>(ta1, i, j, k, zo);
    double s2 = node4<D>(ta2, i, j, k, zo);
    constexpr int cn = 1 + D;
    constexpr int tc1 = (D == 0) ? 2 : 1, tc2 = (D == 2) ? 2 : 3;
    fl(i, j, k, cn) = fl(i, j, k, cn) - 0.25 * vc4;
    fl(i, j, k, cn) = fl(i, j, k, cn) + .25 * div8;
    fl(i, j, k, tc1) = fl(i, j, k, tc1) - .25 * s1;
    fl(i, j, k, tc2) = fl(i, j, k, tc2) - .25 * s2;
    double phiflx = 0.25 * (vc4 - div8) * (prim(il, jl, kl, cn) + prim(iu, ju, ku, cn));
    phiflx = phiflx + .25 * s1 * (prim(il, jl, kl, tc1) + prim(iu, ju, ku, tc1));
    phiflx = phiflx + .25 * s2 * (prim(il, jl, kl, tc2) + prim(iu, ju, ku, tc2));
    fl(i, j, k, nv + 1) = fl(i, j, k, nv + 1) - 0.5 * phiflx;
}

// ---- hyperbolic face flux (flux.cpp:2016-2372) ---------------------------------------------

template <int NS, int D>
__device__ __forceinline__ void hyp_body(const DevParams& P, Arr cons, Arr prim, Arr fl, long t)
{
    int i, j, k;
    if (!box_index(P, t, D == 0, D == 1, D == 2, i, j, k)) return;
    constexpr int ex = (D == 0), ey = (D == 1), ez = (D == 2);
    const int nv = NS + 5;
    const double wgt2 = 1. / 12.;
    const double wgt1 = 0.5 + wgt2;
    double cv[NS + 5], pv[NS + 6], Yk[NS];
    // walls: the wall face takes the wall-side cell (flux.cpp:2028-2037); with conserved
    // interpolation the next face in uses the one-sided weights of :1994-1997 (:2190-2206)
    const WallFace W = wall_face<D>(P, i, j, k);
    const int c = (D == 0) ? i : (D == 1) ? j : k;
    const bool near_lo = P.any_wall && c == 1 && P.wall[2 * D];
    const bool near_hi = P.any_wall && c == P.n[D] - 1 && P.wall[2 * D + 1];
    if (P.advection_type == 1) {
#pragma unroll
        for (int l = 0; l < NS + 6; ++l)
            pv[l] = W.any() ? prim(W.wi, W.wj, W.wk, l)
                            : wgt1 * (prim(i, j, k, l) + prim(i - ex, j - ey, k - ez, l)) -
                                  wgt2 * (prim(i - 2 * ex, j - 2 * ey, k - 2 * ez, l) + prim(i + ex, j + ey, k + ez, l));
        double rho = pv[0];
        cv[0] = rho;
#pragma unroll
        for (int n = 0; n < NS; ++n) Yk[n] = pv[6 + n];
        double ie = eos_energy<NS>(P, Yk, pv[4]);
        double vsqr = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
        cv[4] = rho * ie + 0.5 * rho * vsqr;
#pragma unroll
        for (int n = 0; n < NS; ++n) cv[5 + n] = cv[0] * pv[6 + n];
    } else {
        const double wgta = -0.2, wgtb = 0.75, wgtc = 0.5, wgtd = -0.05;
#pragma unroll
        for (int l = 0; l < NS + 5; ++l) {
            const double cm2 = cons(i - 2 * ex, j - 2 * ey, k - 2 * ez, l), cm1 = cons(i - ex, j - ey, k - ez, l);
            const double c0 = cons(i, j, k, l), cp1 = cons(i + ex, j + ey, k + ez, l);
            if (W.lo) cv[l] = cm1;
            else if (W.hi) cv[l] = c0;
            else if (near_lo) cv[l] = wgta * cm2 + wgtb * cm1 + wgtc * c0 + wgtd * cp1;
            else if (near_hi) cv[l] = wgta * cp1 + wgtb * c0 + wgtc * cm1 + wgtd * cm2;
            else cv[l] = wgt1 * (c0 + cm1) - wgt2 * (cm2 + cp1);
        }
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
    fl(i, j, k, 0) += cv[0] * un;
    if (D == 0) {
        fl(i, j, k, 1) += cv[0] * (pv[1] * pv[1]) + pv[5];
        fl(i, j, k, 2) += cv[0] * pv[1] * pv[2];
        fl(i, j, k, 3) += cv[0] * pv[1] * pv[3];
    } else if (D == 1) {
        fl(i, j, k, 1) += cv[0] * pv[1] * pv[2];
        fl(i, j, k, 2) += cv[0] * pv[2] * pv[2] + pv[5];
        fl(i, j, k, 3) += cv[0] * pv[3] * pv[2];
    } else {
        fl(i, j, k, 1) += cv[0] * pv[1] * pv[3];
        fl(i, j, k, 2) += cv[0] * pv[2] * pv[3];
        fl(i, j, k, 3) += cv[0] * pv[3] * pv[3] + pv[5];
    }
    fl(i, j, k, 4) += un * cv[4] + pv[5] * un;
    fl(i, j, k, 4) += fl(i, j, k, nv) + fl(i, j, k, nv + 1) + fl(i, j, k, nv + 2) + fl(i, j, k, nv + 3);
    if (P.algorithm_type == 2) {
        if (P.advection_type == 1) {
#pragma unroll
            for (int n = 0; n < NS; ++n) fl(i, j, k, 5 + n) += cv[0] * pv[6 + n] * un;
        } else {
#pragma unroll
            for (int n = 0; n < NS; ++n) fl(i, j, k, 5 + n) += cv[5 + n] * un;
        }
    }
}

// ---- stage combination (timeStep.cpp:158-201, 266-306, 370-413) ----------------------------

// src: the source MultiFab (null: zero, main_driver.cpp:241-242)
template <int NS>
__global__ void __launch_bounds__(BLK) k_stage(DevParams P, Arr out, Arr u0, Arr us, Arr fx, Arr fy,
                                               Arr fz, Arr srcf, double dt, int mode)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    int i, j, k;
    if (!box_index(P, t, 0, 0, 0, i, j, k)) return;
    double o[NS + 5];
#pragma unroll
    for (int n = 0; n < NS + 5; ++n) {
        const double src = srcf.p ? srcf(i, j, k, n) : 0.0;
        double div = (fx(i + 1, j, k, n) - fx(i, j, k, n)) / P.dx[0] +
                     (fy(i, j + 1, k, n) - fy(i, j, k, n)) / P.dx[1] +
                     (fz(i, j, k + 1, n) - fz(i, j, k, n)) / P.dx[2];
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
}

// ---- chemistry source (chemistry_functions.cpp:128-193) ------------------------------------

template <int NS>
__global__ void __launch_bounds__(BLK) k_chem_source(DevParams P, Arr prim, Arr source, Arr ranchem, double dt,
                                                     double dV)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    int i, j, k;
    if (!box_index(P, t, 0, 0, 0, i, j, k)) return;
    double X[NS], src[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) X[n] = prim(i, j, k, 6 + NS + n);
    chem_source_cell<NS>(P, prim(i, j, k, 4), prim(i, j, k, 5), X, [&](int m) { return ranchem(i, j, k, m); }, dt,
                         dV, src);
#pragma unroll
    for (int n = 0; n < NS; ++n) source(i, j, k, 5 + n) = src[n];
}

// Reaction noise sets A, B (MultiFabFillRandom(ranchem_A/B, m, 1.0), timeStep.cpp:93-97).
__global__ void __launch_bounds__(BLK) k_chem_noise(DevParams P, NoiseKey K, Arr A, Arr B)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    int i, j, k;
    if (!box_index(P, t, 0, 0, 0, i, j, k)) return;
    const uint32_t lin = (uint32_t)(((long)k * P.n[1] + j) * P.n[0] + i);
    for (int m = 0; m < P.chem.nreaction; ++m) {
        double za, zb;
        noise_ab(K, CHEM_FIELD + (uint32_t)m, lin, za, zb);
        A(i, j, k, m) = za;
        B(i, j, k, m) = zb;
    }
}

// ---- noise --------------------------------------------------------------------------------

// The four weighted-noise arrays of a stage in one launch: stochFlux[d] = 1*A + w*B on comps
// 1..nvars-1 (comp 0 zeroed), rancorn likewise on its one component (MultiFab::LinComb).
struct LinSet {
    double* dst[4];
    const double* a[4];
    const double* b[4];
    long npts[4];
    int nc[4], c0[4];
};

__global__ void __launch_bounds__(BLK) k_lincomb4(LinSet L, double w)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (t >= L.npts[q]) continue;
        for (int n = 0; n < L.nc[q]; ++n) {
            const long x = t + (long)n * L.npts[q];
            L.dst[q][x] = (n < L.c0[q]) ? 0.0 : 1.0 * L.a[q][x] + w * L.b[q][x];
        }
    }
}

// stochFlux = 1*A + w*B on comps 1..nvars-1 (comp 0 zeroed), MultiFab::LinComb.
__global__ void __launch_bounds__(BLK) k_lincomb(double* dst, const double* a, const double* b, long npts,
                                                 int nc, int c0, double w)
{
    long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (t >= npts) return;
    for (int n = 0; n < nc; ++n) {
        long q = t + (long)n * npts;
        dst[q] = (n < c0) ? 0.0 : 1.0 * a[q] + w * b[q];
    }
}

// Fills every component of direction D's face noise for sets A and B: component 0 (density)
// is zero (timeStep.cpp:74-76 skips it), components 1..nvars-1 are N(0,1) * their scale.
template <int D>
__device__ __forceinline__ void noise_face_body(const DevParams& P, const NoiseKey& K, Arr A, Arr B, long t)
{
    int i, j, k;
    if (!box_index(P, t, D == 0, D == 1, D == 2, i, j, k)) return;
    const int iw = (D == 0 && i == P.n[0]) ? 0 : i;
    const int jw = (D == 1 && j == P.n[1]) ? 0 : j;
    const int kw = (D == 2 && k == P.n[2]) ? 0 : k;
    uint32_t lin = (uint32_t)(((long)kw * P.n[1] + jw) * P.n[0] + iw);
    // in a wall direction face n is a face of its own, drawn from its own field
    const int c = (D == 0) ? i : (D == 1) ? j : k;
    const uint32_t wall_hi = (c == P.n[D] && !P.periodic[D]) ? 128u : 0u;
    A(i, j, k, 0) = 0.0;
    B(i, j, k, 0) = 0.0;
    for (int comp = 1; comp < P.nvars; ++comp) {
        const double scale = (comp <= 3) ? P.var_mom : (comp == 4) ? P.var_ener : P.var_mass;
        double za, zb;
        noise_ab(K, (uint32_t)(16 * D + comp) + wall_hi, lin, za, zb);
        A(i, j, k, comp) = za * scale;
        B(i, j, k, comp) = zb * scale;
    }
}

__device__ __forceinline__ void noise_node_body(const DevParams& P, const NoiseKey& K, Arr A, Arr B, double scale,
                                                long t)
{
    int i, j, k;
    if (!box_index(P, t, 1, 1, 1, i, j, k)) return;
    const int iw = (i == P.n[0]) ? 0 : i, jw = (j == P.n[1]) ? 0 : j, kw = (k == P.n[2]) ? 0 : k;
    uint32_t lin = (uint32_t)(((long)kw * P.n[1] + jw) * P.n[0] + iw);
    // nodes on the hi plane of a wall direction are distinct: field NODE_FIELD + mask
    const uint32_t mask = ((i == P.n[0] && !P.periodic[0]) ? 1u : 0u) |
                          ((j == P.n[1] && !P.periodic[1]) ? 2u : 0u) |
                          ((k == P.n[2] && !P.periodic[2]) ? 4u : 0u);
    double za, zb;
    noise_ab(K, NODE_FIELD + mask, lin, za, zb);
    A(i, j, k) = za * scale;
    B(i, j, k) = zb * scale;
}

// One launch for a step's whole noise set (blockIdx.y: face direction 0..2, 3 = nodes); every
// value is drawn by the same body and key as in a launch of its own.
__global__ void __launch_bounds__(BLK) k_noise_all(DevParams P, NoiseKey K, Arr A0, Arr B0, Arr A1, Arr B1,
                                                   Arr A2, Arr B2, Arr rA, Arr rB)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    switch (blockIdx.y) {
        case 0: noise_face_body<0>(P, K, A0, B0, t); break;
        case 1: noise_face_body<1>(P, K, A1, B1, t); break;
        case 2: noise_face_body<2>(P, K, A2, B2, t); break;
        default: noise_node_body(P, K, rA, rB, P.var_mom, t); break;
    }
}

// ---- the face passes, one launch for all three directions ------------------------------------
// blockIdx.y is the face direction; each thread runs the same per-face body as a launch of that
// direction alone would, so the arithmetic and its order are unchanged.  A pass writes only
// the flux of its own direction and reads shared inputs, so the directions are independent.
// On small boxes (the wall decks), three directions' faces in one launch keep three times as
// many waves in flight as one direction's.

template <int NS, bool WN>
__global__ void __launch_bounds__(BLK) k_stoch3(DevParams P, Arr prim, Arr eta, Arr zeta, Arr kappa, Arr chi,
                                                Arr Dm, Arr f0, Arr f1, Arr f2, Arr r0, Arr r1, Arr r2, Arr rc,
                                                NoiseW n0, NoiseW n1, NoiseW n2, double volinv, double dtinv,
                                                double dt)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (blockIdx.y == 0)
        stoch_body<NS, 0, WN>(P, prim, eta, zeta, kappa, chi, Dm, f0, r0, rc, n0, volinv, dtinv, dt, t);
    else if (blockIdx.y == 1)
        stoch_body<NS, 1, WN>(P, prim, eta, zeta, kappa, chi, Dm, f1, r1, rc, n1, volinv, dtinv, dt, t);
    else
        stoch_body<NS, 2, WN>(P, prim, eta, zeta, kappa, chi, Dm, f2, r2, rc, n2, volinv, dtinv, dt, t);
}

template <int NS>
__global__ void __launch_bounds__(BLK) k_diff3(DevParams P, Arr prim, Arr eta, Arr kappa, Arr chi, Arr Dm,
                                               Arr f0, Arr f1, Arr f2)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (blockIdx.y == 0)
        diff_body<NS, 0>(P, prim, eta, kappa, chi, Dm, f0, t);
    else if (blockIdx.y == 1)
        diff_body<NS, 1>(P, prim, eta, kappa, chi, Dm, f1, t);
    else
        diff_body<NS, 2>(P, prim, eta, kappa, chi, Dm, f2, t);
}

// Per direction: the flux, the divergence pair and the tangential pair of corner arrays.
struct CornerFaceSet {
    Arr fl[3], c1[3], c2[3], ta1[3], ta2[3];
};

template <int NS>
__global__ void __launch_bounds__(BLK) k_corner_face3(DevParams P, Arr prim, Arr vcorn, CornerFaceSet C)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (blockIdx.y == 0)
        corner_face_body<NS, 0>(P, prim, C.fl[0], vcorn, C.c1[0], C.c2[0], C.ta1[0], C.ta2[0], t);
    else if (blockIdx.y == 1)
        corner_face_body<NS, 1>(P, prim, C.fl[1], vcorn, C.c1[1], C.c2[1], C.ta1[1], C.ta2[1], t);
    else
        corner_face_body<NS, 2>(P, prim, C.fl[2], vcorn, C.c1[2], C.c2[2], C.ta1[2], C.ta2[2], t);
}

template <int NS>
__global__ void __launch_bounds__(BLK) k_hyp3(DevParams P, Arr cons, Arr prim, Arr f0, Arr f1, Arr f2)
{
    const long t = (long)blockIdx.x * BLK + threadIdx.x;
    if (blockIdx.y == 0)
        hyp_body<NS, 0>(P, cons, prim, f0, t);
    else if (blockIdx.y == 1)
        hyp_body<NS, 1>(P, cons, prim, f1, t);
    else
        hyp_body<NS, 2>(P, cons, prim, f2, t);
}

// ---- host launchers ---------------------------------------------------------------------

#define NS_DISPATCH(NSV, CALL)                          \
    switch (NSV) {                                      \
        case 2: { constexpr int NS = 2; CALL; } break;  \
        case 3: { constexpr int NS = 3; CALL; } break;  \
        case 4: { constexpr int NS = 4; CALL; } break;  \
        case 5: { constexpr int NS = 5; CALL; } break;  \
        case 6: { constexpr int NS = 6; CALL; } break;  \
        case 7: { constexpr int NS = 7; CALL; } break;  \
        case 8: { constexpr int NS = 8; CALL; } break;  \
        default: return FHDX_ERR_UNSUPPORTED;           \
    }

// counts over the launch box (cells, faces of direction d, nodes, cells grown by ngc)
static long ncell(const DevParams& P) { return (long)P.bn[0] * P.bn[1] * P.bn[2]; }
static long nface(const DevParams& P, int d)
{
    return (long)(P.bn[0] + (d == 0)) * (P.bn[1] + (d == 1)) * (P.bn[2] + (d == 2));
}
static long nnode(const DevParams& P) { return (long)(P.bn[0] + 1) * (P.bn[1] + 1) * (P.bn[2] + 1); }
static long ngrown(const DevParams& P)
{
    const int g = P.ngc;
    return (long)(P.bn[0] + 2 * g) * (P.bn[1] + 2 * g) * (P.bn[2] + 2 * g);
}

int ref_c2p(const DevParams& P, Arr prim, Arr cons, hipStream_t s)
{
    NS_DISPATCH(P.ns, (k_c2p<NS><<<grid_for(ncell(P), BLK), BLK, 0, s>>>(P, prim, cons)));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_transport(const DevParams& P, Arr prim, Arr eta, Arr zeta, Arr kappa, Arr chi, Arr Dm,
                  int* err, hipStream_t s)
{
    if (P.transport_type == 2) {
        NS_DISPATCH(P.ns, (k_transport<NS, true><<<grid_for(ngrown(P), BLK), BLK, 0, s>>>(
                              P, prim, eta, zeta, kappa, chi, Dm, err)));
    } else {
        NS_DISPATCH(P.ns, (k_transport<NS, false><<<grid_for(ngrown(P), BLK), BLK, 0, s>>>(
                              P, prim, eta, zeta, kappa, chi, Dm, err)));
    }
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_fill_boundary(const DevParams& P, Arr a, int nc, hipStream_t s)
{
    k_fill_boundary<<<grid_for(ngrown(P), BLK), BLK, 0, s>>>(P, a, nc, a, nc);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_fill_boundary2(const DevParams& P, Arr a, int nca, Arr b, int ncb, hipStream_t s)
{
    k_fill_boundary<<<dim3(grid_for(ngrown(P), BLK), 2), BLK, 0, s>>>(P, a, nca, b, ncb);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_set_bc(const DevParams& P, Arr prim, Arr cons, hipStream_t s)
{
    if (!P.any_wall) return FHDX_OK;
    const int g = P.ngc;
    for (int d = 0; d < 3; ++d) {
        if (P.periodic[d]) continue;
        long slab = 2L * g;
        for (int e = 0; e < 3; ++e)
            if (e != d) slab *= P.bn[e] + 2 * g;
        NS_DISPATCH(P.ns, (k_set_bc<NS><<<grid_for(slab, BLK), BLK, 0, s>>>(P, prim, cons, d)));
    }
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int fill_value(double* p, long n, double v, hipStream_t s)
{
    if (n <= 0) return FHDX_OK;
    k_fill<<<grid_for(n, BLK), BLK, 0, s>>>(p, n, v);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

template <int NS>
static void flux_passes(const DevParams& P, const FluxArgs& a, hipStream_t s)
{
    const double volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
    const double dtinv = 1. / a.dt;
    long mf = nface(P, 0);
    for (int d = 1; d < 3; ++d) mf = nface(P, d) > mf ? nface(P, d) : mf;
    const dim3 grid(grid_for(mf, BLK), 3);   // every body checks its own direction's face count
    NoiseW nw[3];
    for (int d = 0; d < 3; ++d) nw[d] = NoiseW{a.nA[d], a.nB[d], a.rA, a.rB, a.w};
    if (P.stoch_on && a.weigh)
        k_stoch3<NS, true><<<grid, BLK, 0, s>>>(P, a.prim, a.eta, a.zeta, a.kappa, a.chi, a.D, a.flux[0], a.flux[1],
                                                a.flux[2], a.ran[0], a.ran[1], a.ran[2], a.rancorn, nw[0], nw[1],
                                                nw[2], volinv, dtinv, a.dt);
    else if (P.stoch_on)
        k_stoch3<NS, false><<<grid, BLK, 0, s>>>(P, a.prim, a.eta, a.zeta, a.kappa, a.chi, a.D, a.flux[0],
                                                 a.flux[1], a.flux[2], a.ran[0], a.ran[1], a.ran[2], a.rancorn,
                                                 nw[0], nw[1], nw[2], volinv, dtinv, a.dt);
    k_diff3<NS><<<grid, BLK, 0, s>>>(P, a.prim, a.eta, a.kappa, a.chi, a.D, a.flux[0], a.flux[1], a.flux[2]);
    // corn[a][b]: derivative direction a, velocity component b (cornx/corny/cornz of flux.cpp)
    const Arr(&c)[3][3] = a.corn;
    k_corner_node<<<grid_for(nnode(P), BLK), BLK, 0, s>>>(P, a.prim, a.eta, a.zeta, c[0][0], c[0][1], c[0][2],
                                                          c[1][0], c[1][1], c[1][2], c[2][0], c[2][1], c[2][2],
                                                          a.visccorn, a.rancorn, a.rA, a.rB, a.w,
                                                          (P.stoch_on && a.weigh) ? 1 : 0);
    // x face: div pair (vy, wz), tangential (uy, uz); y: (ux, wz), (vx, vz); z: (vy, ux), (wx, wy)
    CornerFaceSet C;
    for (int d = 0; d < 3; ++d) C.fl[d] = a.flux[d];
    C.c1[0] = c[1][1]; C.c2[0] = c[2][2]; C.ta1[0] = c[1][0]; C.ta2[0] = c[2][0];
    C.c1[1] = c[0][0]; C.c2[1] = c[2][2]; C.ta1[1] = c[0][1]; C.ta2[1] = c[2][1];
    C.c1[2] = c[1][1]; C.c2[2] = c[0][0]; C.ta1[2] = c[0][2]; C.ta2[2] = c[1][2];
    k_corner_face3<NS><<<grid, BLK, 0, s>>>(P, a.prim, a.visccorn, C);
    if (P.advection_type == 1 || P.advection_type == 2)
        k_hyp3<NS><<<grid, BLK, 0, s>>>(P, a.cons, a.prim, a.flux[0], a.flux[1], a.flux[2]);
}

int ref_calculate_flux(const DevParams& P, const FluxArgs& a, int flux_ncomp[3], hipStream_t s)
{
    const int nc = P.nvars + 4;
    for (int d = 0; d < 3; ++d) {
        // flux_in.setVal(0) (flux.cpp:24-26): the stochastic pass writes all nvars + 4
        // components itself; zero first only without it or for extra components
        if (P.stoch_on && flux_ncomp[d] == nc) continue;
        int r = fill_value(a.flux[d].p, a.flux[d].sn * flux_ncomp[d], 0.0, s);
        if (r) return r;
    }
    NS_DISPATCH(P.ns, flux_passes<NS>(P, a, s));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_stage(const DevParams& P, Arr out, Arr u0, Arr us, const Arr flux[3], Arr source, double dt, int mode,
              hipStream_t s)
{
    NS_DISPATCH(P.ns, (k_stage<NS><<<grid_for(ncell(P), BLK), BLK, 0, s>>>(P, out, u0, us, flux[0], flux[1],
                                                                             flux[2], source, dt, mode)));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_chem_source(const DevParams& P, Arr prim, Arr source, Arr ranchem, double dt, hipStream_t s)
{
    const double dV = P.dx[0] * P.dx[1] * P.dx[2];
    NS_DISPATCH(P.ns, (k_chem_source<NS><<<grid_for(ncell(P), BLK), BLK, 0, s>>>(P, prim, source, ranchem, dt,
                                                                                   dV)));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_fill_chem_noise(const DevParams& P, const NoiseKey& K, Arr A, Arr B, hipStream_t s)
{
    k_chem_noise<<<grid_for(ncell(P), BLK), BLK, 0, s>>>(P, K, A, B);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_lincomb4(double* const dst[4], const double* const a[4], const double* const b[4], const long npts[4],
                 const int nc[4], const int c0[4], double w, hipStream_t s)
{
    LinSet L;
    long mx = 0;
    for (int q = 0; q < 4; ++q) {
        L.dst[q] = dst[q];
        L.a[q] = a[q];
        L.b[q] = b[q];
        L.npts[q] = npts[q];
        L.nc[q] = nc[q];
        L.c0[q] = c0[q];
        mx = npts[q] > mx ? npts[q] : mx;
    }
    k_lincomb4<<<grid_for(mx, BLK), BLK, 0, s>>>(L, w);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_lincomb(double* dst, const double* a, const double* b, long npts, int nc, int c0, double w,
                hipStream_t s)
{
    k_lincomb<<<grid_for(npts, BLK), BLK, 0, s>>>(dst, a, b, npts, nc, c0, w);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int ref_fill_noise(const DevParams& P, const NoiseKey& K, const Arr A[3], const Arr B[3], Arr rA, Arr rB,
                   hipStream_t s)
{
    long mx = nnode(P);                       // the node box is the largest; each body checks its own
    for (int d = 0; d < 3; ++d) mx = nface(P, d) > mx ? nface(P, d) : mx;
    k_noise_all<<<dim3(grid_for(mx, BLK), 4), BLK, 0, s>>>(P, K, A[0], B[0], A[1], B[1], A[2], B[2], rA, rB);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

}  // namespace fhdx
