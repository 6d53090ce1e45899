// stag_kernels.hip -- the staggered compressible scheme (src_compressible_stag) on gfx950.
//
// Momentum and velocity live on faces (cumom[d], vel[d], nodal in d); density, energy and the
// species densities in cells.  One stage of RK3stepStag (timeStepStag.cpp:13-1245) is
//   transport (ref_kernels.hip) -> k_stag_cells (cells grown by 1: diagonal stresses and the
//   cell-centred momentum fluxes) -> k_stag_edges (the three edge types: off-diagonal stresses
//   and edge momentum fluxes) -> k_stag_faces<D> (mass, energy, species and the viscous-work,
//   heat and Dufour components of faceflux) -> k_stag_update (cells and the three face
//   momenta) -> ghost fill of the density -> k_stag_c2p -> ghost fill of state, momenta,
//   velocities and prim,
// one thread per cell, edge or face.  The noise of a stage (A + w B of the step's two sets) is
// evaluated in place from a counter-based Philox keyed by the wrapped global index, so duplicate
// periodic faces and edges draw one value (OverrideSync's result) and nothing is stored.  Every
// expression keeps the reference's operation order: the results equal the oracle's
// (oracle/oracle_stag.c) bit for bit.  Scope: 3-D, periodic in every direction; with chemistry
// the update adds the reaction source (timeStepStag.cpp:389-396).
#include "stag.hpp"

namespace fhdx {

constexpr int SB = 256;

// Noise fields (low byte of the counter's third word): face component c of direction d at
// 176 + 16 d + (c - 4), edge types xy, xz, yz at 224 + e, cell fields at 232 + d.
constexpr uint32_t STAG_FACE_FIELD = 176u;
constexpr uint32_t STAG_EDGE_FIELD = 224u;
constexpr uint32_t STAG_CELL_FIELD = 232u;

__device__ inline int wrapn(int i, int n) { return ((i % n) + n) % n; }

__device__ inline uint32_t wrapped_lin(const DevParams& P, int i, int j, int k)
{
    return (uint32_t)(((long)wrapn(k, P.n[2]) * P.n[1] + wrapn(j, P.n[1])) * P.n[0] + wrapn(i, P.n[0]));
}

// The edge flux array of momentum component m transported along t (m != t), in the reference's
// order edgeflux_x[0..1], edgeflux_y[0..1], edgeflux_z[0..1]: x_v, x_w, y_u, y_w, z_u, z_v.
__host__ __device__ constexpr int edge_flux_index(int m, int t)
{
    return m == 0 ? (t == 1 ? 2 : 4) : m == 1 ? (t == 0 ? 0 : 5) : (t == 0 ? 1 : 3);
}
// Edge type of the plane of two distinct axes: xy 0, xz 1, yz 2.
__host__ __device__ constexpr int edge_type(int a, int b)
{
    return (a + b == 1) ? 0 : (a + b == 2) ? 1 : 2;
}

__device__ inline void cell_of(long t, int fx, int fy, int& i, int& j, int& k)
{
    i = (int)(t % fx);
    long r = t / fx;
    j = (int)(r % fy);
    k = (int)(r / fy);
}

// ---- cells grown by 1: fluxStag.cpp:163-206, 872-924, 799-803, 1444-1449, 1838-1856 --------

__global__ void __launch_bounds__(SB) k_stag_cells(DevParams P, StagArrays A, NoiseKey K, double w, double dt)
{
    const int gx = P.n[0] + 2, gy = P.n[1] + 2, gz = P.n[2] + 2;
    const long t = (long)blockIdx.x * SB + threadIdx.x;
    if (t >= (long)gx * gy * gz) return;
    int i, j, k;
    cell_of(t, gx, gy, i, j, k);
    i -= 1;
    j -= 1;
    k -= 1;
    const double volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
    const double dtinv = 1. / dt;
    const bool visc3 = (P.visc_type == 3 || P.visc_type == -3);
    const double eta = A.eta(i, j, k), zeta = A.zeta(i, j, k);
    const double T = A.prim(i, j, k, 4);
    double ts[3] = {0., 0., 0.};
    if (P.stoch_on) {
        const double etaT = eta * T;
        const double zetaT = zeta * T;
        const double fac1 = sqrt(2.0 * P.k_B * etaT * volinv * dtinv);
        double fac2 = (-1.0 / 3.0) * sqrt(2.0 * P.k_B * etaT * volinv * dtinv);
        if (visc3)
            fac2 = sqrt(P.k_B * zetaT * volinv * dtinv / 3.0) - sqrt(2.0 * P.k_B * etaT * volinv * dtinv) / 3.0;
        const uint32_t lin = wrapped_lin(P, i, j, k);
        const double sq2 = sqrt(2.0);
        double z[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double a, b;
            noise_ab(K, STAG_CELL_FIELD + (uint32_t)c, lin, a, b);
            z[c] = 1.0 * (sq2 * a) + w * (sq2 * b);
        }
        const double traceZ = z[0] + z[1] + z[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) ts[c] = (fac1 * z[c]) + (fac2 * traceZ);
    }
    const double g[3] = {(A.vel[0](i + 1, j, k) - A.vel[0](i, j, k)) / P.dx[0],
                         (A.vel[1](i, j + 1, k) - A.vel[1](i, j, k)) / P.dx[1],
                         (A.vel[2](i, j, k + 1) - A.vel[2](i, j, k)) / P.dx[2]};
    const double div = g[0] + g[1] + g[2];
    const double p = A.prim(i, j, k, 5);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double tau = visc3 ? 2 * eta * g[c] + (zeta - 2 * eta / 3.) * div
                                 : 2 * eta * g[c] + (0.0 - 2 * eta / 3.) * div;
        A.td(i, j, k, c) = ts[c];
        A.td(i, j, k, 3 + c) = tau;
        double cf = ts[c];
        cf -= tau;
        if (P.advection_type >= 0) {
            const int e0 = (c == 0), e1 = (c == 1), e2 = (c == 2);
            const double m0 = A.mom[c](i, j, k), m1 = A.mom[c](i + e0, j + e1, k + e2);
            const double v0 = A.vel[c](i, j, k), v1 = A.vel[c](i + e0, j + e1, k + e2);
            cf += 0.25 * (m0 + m1) * (v0 + v1) + p;
        }
        A.cf[c](i, j, k) = cf;
    }
}

// ---- edges: fluxStag.cpp:209-295, 927-1022, 784-797, 1429-1441, 1767-1835 ----------------

template <int E>
__device__ inline void stag_edge(const DevParams& P, const StagArrays& A, const NoiseKey& K, double w,
                                 double dt, int i, int j, int k)
{
    constexpr int a = (E == 2) ? 1 : 0, b = (E == 0) ? 1 : 2;
    constexpr int ea0 = a == 0, ea1 = a == 1, ea2 = a == 2;
    constexpr int eb0 = b == 0, eb1 = b == 1, eb2 = b == 2;
    const int ci[4] = {i - ea0 - eb0, i - ea0, i - eb0, i};
    const int cj[4] = {j - ea1 - eb1, j - ea1, j - eb1, j};
    const int ck[4] = {k - ea2 - eb2, k - ea2, k - eb2, k};
    double ts = 0.;
    if (P.stoch_on) {
        const double volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
        const double dtinv = 1. / dt;
        double s = A.eta(ci[0], cj[0], ck[0]) * A.prim(ci[0], cj[0], ck[0], 4);
        s = s + A.eta(ci[1], cj[1], ck[1]) * A.prim(ci[1], cj[1], ck[1], 4);
        s = s + A.eta(ci[2], cj[2], ck[2]) * A.prim(ci[2], cj[2], ck[2], 4);
        s = s + A.eta(ci[3], cj[3], ck[3]) * A.prim(ci[3], cj[3], ck[3], 4);
        const double etaT = 0.25 * s;
        const double fac = sqrt(2.0 * P.k_B * etaT * volinv * dtinv);
        double za, zb;
        noise_ab(K, STAG_EDGE_FIELD + (uint32_t)E, wrapped_lin(P, i, j, k), za, zb);
        ts = fac * (1.0 * za + w * zb);
    }
    const Arr& va = A.vel[a];
    const Arr& vb = A.vel[b];
    const double da_b = (va(i, j, k) - va(i - eb0, j - eb1, k - eb2)) / P.dx[b];
    const double db_a = (vb(i, j, k) - vb(i - ea0, j - ea1, k - ea2)) / P.dx[a];
    double es = A.eta(ci[0], cj[0], ck[0]);
    es = es + A.eta(ci[1], cj[1], ck[1]);
    es = es + A.eta(ci[2], cj[2], ck[2]);
    es = es + A.eta(ci[3], cj[3], ck[3]);
    const double eta_interp = 0.25 * es;
    const double tau = eta_interp * (da_b + db_a);
    A.to[E](i, j, k, 0) = ts;
    A.to[E](i, j, k, 1) = tau;
    double fa = ts, fb = ts;  // momentum a along b, momentum b along a
    fa -= tau;
    fb -= tau;
    if (P.advection_type >= 0) {
        const Arr& ma = A.mom[a];
        const Arr& mb = A.mom[b];
        const double adv_a = 0.25 * (ma(i - eb0, j - eb1, k - eb2) + ma(i, j, k)) *
                             (vb(i - ea0, j - ea1, k - ea2) + vb(i, j, k));
        const double adv_b = 0.25 * (mb(i - ea0, j - ea1, k - ea2) + mb(i, j, k)) *
                             (va(i - eb0, j - eb1, k - eb2) + va(i, j, k));
        fa += adv_a;
        fb += adv_b;
    }
    A.ef[edge_flux_index(a, b)](i, j, k) = fa;
    A.ef[edge_flux_index(b, a)](i, j, k) = fb;
}

__global__ void __launch_bounds__(SB) k_stag_edges(DevParams P, StagArrays A, NoiseKey K, double w, double dt)
{
    const int e = blockIdx.y;
    // edge boxes: nodal along the type's two axes
    const int ex = P.n[0] + (e != 2), ey = P.n[1] + (e != 1), ez = P.n[2] + (e != 0);
    const long t = (long)blockIdx.x * SB + threadIdx.x;
    if (t >= (long)ex * ey * ez) return;
    int i, j, k;
    cell_of(t, ex, ey, i, j, k);
    if (e == 0) stag_edge<0>(P, A, K, w, dt, i, j, k);
    else if (e == 1) stag_edge<1>(P, A, K, w, dt, i, j, k);
    else stag_edge<2>(P, A, K, w, dt, i, j, k);
}

// ---- faces of direction D: fluxStag.cpp:298-780, 1026-1426, 1507-1764 --------------------

template <int NS, int D>
__global__ void __launch_bounds__(SB) k_stag_faces(DevParams P, StagArrays A, NoiseKey K, double w, double dt)
{
    constexpr int nv = NS + 5;
    constexpr int ed0 = D == 0, ed1 = D == 1, ed2 = D == 2;
    constexpr int t1 = (D == 0) ? 1 : 0, t2 = (D == 2) ? 1 : 2;
    const int fx = P.n[0] + ed0, fy = P.n[1] + ed1, fz = P.n[2] + ed2;
    const long tid = (long)blockIdx.x * SB + threadIdx.x;
    if (tid >= (long)fx * fy * fz) return;
    int i, j, k;
    cell_of(tid, fx, fy, i, j, k);
    const int li = i - ed0, lj = j - ed1, lk = k - ed2;
    const Arr& prim = A.prim;
    const Arr& cu = A.cu;
    const Arr& chi = A.chi;
    const Arr& Dm = A.D;
    const double dxd = P.dx[D];
    const double volinv = 1. / (P.dx[0] * P.dx[1] * P.dx[2]);
    const double dtinv = 1. / dt;
    const double Tu = prim(i, j, k, 4), Tl = prim(li, lj, lk, 4);
    const double vel_f = A.vel[D](i, j, k);
    double f[nv + 4];
#pragma unroll
    for (int c = 0; c < nv + 4; ++c) f[c] = 0.;
    double hk[NS];

    // stochastic part
    if (P.stoch_on) {
        const uint32_t lin = wrapped_lin(P, i, j, k);
        double wn[NS + 1];
        const double kp = A.kappa(i, j, k) * Tu * Tu + A.kappa(li, lj, lk) * Tl * Tl;
        const double meanT = 0.5 * (Tu + Tl);
        const double fw0 = sqrt(P.k_B * kp * volinv * dtinv);
        double za, zb;
        noise_ab(K, STAG_FACE_FIELD + 16u * D + 0u, lin, za, zb);
        wn[0] = fw0 * (1.0 * za + w * zb);
        f[nv] = wn[0];
        f[nv + 1] = 0.5 * vel_f * (A.td(li, lj, lk, D) + A.td(i, j, k, D));
        double vsh = 0.0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int t = q == 0 ? t1 : t2;
            const int et0 = t == 0, et1 = t == 1, et2 = t == 2;
            const Arr& vt = A.vel[t];
            const Arr& ts = A.to[edge_type(D, t)];
            const int pi = i + et0, pj = j + et1, pk = k + et2;
            vsh += 0.25 * ((vt(pi, pj, pk) + vt(pi - ed0, pj - ed1, pk - ed2)) * ts(pi, pj, pk, 0) +
                           (vt(i, j, k) + vt(li, lj, lk)) * ts(i, j, k, 0));
        }
        f[nv + 2] = vsh;
        if (P.algorithm_type == 2) {
            double yy[NS], yyp[NS], DijY[NS * NS], sqD[NS * NS];
#pragma unroll
            for (int n = 1; n < 1 + NS; ++n) wn[n] = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                yy[s] = clamp01(prim(li, lj, lk, 6 + s));
                yyp[s] = clamp01(prim(i, j, k, 6 + s));
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
                    DijY[s * NS + l] = 0.5 * (Dm(li, lj, lk, l * NS + s) * yy[l] + Dm(i, j, k, l * NS + s) * yyp[l] +
                                              (Dm(li, lj, lk, s * NS + l) * yy[s] + Dm(i, j, k, s * NS + l) * yyp[s]));
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
            double zs[NS];
#pragma unroll
            for (int l = 0; l < NS; ++l) {
                double a, b;
                noise_ab(K, STAG_FACE_FIELD + 16u * D + 1u + (uint32_t)l, lin, a, b);
                zs[l] = 1.0 * a + w * b;
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
#pragma unroll
                for (int l = 0; l <= s; ++l) {
                    const double fw = sqrt(P.k_B * MWmix * volinv / (P.Runiv * dt)) * sqD[s * NS + l];
                    wn[1 + s] = wn[1 + s] + fw * zs[l];
                }
                f[5 + s] = wn[1 + s];
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) hk[s] = P.e0[s] + P.hcp[s] * meanT;
            double soret = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s)
                soret += (hk[s] + P.Runiv * meanT / P.molmass[s] * 0.5 * (chi(li, lj, lk, s) + chi(i, j, k, s))) *
                         wn[1 + s];
            f[nv + 3] = soret;
        }
    }

    // diffusive part
    {
        const double kp = 0.5 * (A.kappa(li, lj, lk) + A.kappa(i, j, k));
        const double meanT = 0.5 * (Tl + Tu);
        const double pl = prim(li, lj, lk, 5), pu = prim(i, j, k, 5);
        const double meanP = 0.5 * (pl + pu);
        f[nv + 1] -= 0.5 * vel_f * (A.td(li, lj, lk, 3 + D) + A.td(i, j, k, 3 + D));
        double vsh = 0.0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int t = q == 0 ? t1 : t2;
            const int et0 = t == 0, et1 = t == 1, et2 = t == 2;
            const Arr& vt = A.vel[t];
            const Arr& ts = A.to[edge_type(D, t)];
            const int pi = i + et0, pj = j + et1, pk = k + et2;
            vsh -= 0.25 * ((vt(pi, pj, pk) + vt(pi - ed0, pj - ed1, pk - ed2)) * ts(pi, pj, pk, 1) +
                           (vt(i, j, k) + vt(li, lj, lk)) * ts(i, j, k, 1));
        }
        f[nv] -= kp * (Tu - Tl) / dxd;
        f[nv + 2] += vsh;
        if (P.algorithm_type == 2) {
            double dk[NS], soret[NS], Fk[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double Xl = prim(li, lj, lk, 6 + NS + s), Xu = prim(i, j, k, 6 + NS + s);
                const double Yl = prim(li, lj, lk, 6 + s), Yu = prim(i, j, k, 6 + s);
                const double term1 = (Xu - Xl) / dxd;
                const double meanXk = 0.5 * (Xl + Xu);
                const double meanYk = 0.5 * (Yl + Yu);
                const double term2 = (meanXk - meanYk) * (pu - pl) / dxd / meanP;
                dk[s] = term1 + term2;
                const double ChiX = 0.5 * (chi(li, lj, lk, s) * Xl + chi(i, j, k, s) * Xu);
                soret[s] = ChiX * (Tu - Tl) / dxd / meanT;
            }
#pragma unroll
            for (int kk = 0; kk < NS; ++kk) {
                Fk[kk] = 0.;
#pragma unroll
                for (int l = 0; l < NS; ++l) {
                    const double Fks = 0.5 * (Dm(li, lj, lk, l * NS + kk) + Dm(i, j, k, l * NS + kk)) * (dk[l] + soret[l]);
                    Fk[kk] -= Fks;
                }
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) hk[s] = P.e0[s] + P.hcp[s] * meanT;
            double Q5 = 0.;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // z faces add chi(i,j,k) twice (fluxStag.cpp:1398)
                const double chil = (D == 2) ? chi(i, j, k, s) : chi(li, lj, lk, s);
                Q5 += (hk[s] + 0.5 * P.Runiv * meanT * (chil + chi(i, j, k, s)) / P.molmass[s]) * Fk[s];
            }
            f[nv + 3] += Q5;
#pragma unroll
            for (int s = 0; s < NS; ++s) f[5 + s] += Fk[s];
        }
    }

    // hyperbolic part
    if (P.advection_type >= 0) {
        const double mom = A.mom[D](i, j, k);
        f[0] += mom;
        const double meanT = 0.5 * (Tl + Tu);
        const double meanRho = 0.5 * (prim(li, lj, lk, 0) + prim(i, j, k, 0));
        const double meanP = 0.5 * (prim(li, lj, lk, 5) + prim(i, j, k, 5));
#pragma unroll
        for (int s = 0; s < NS; ++s) hk[s] = P.e0[s] + P.hcp[s] * meanT;
        const double meanE = 0.5 * (cu(li, lj, lk, 4) + cu(i, j, k, 4)) / meanRho;
        if (P.advection_type == 1) {
            f[4] += mom * (meanE + (meanP / meanRho));
        } else if (P.advection_type == 2) {
            f[4] += 0.5 * (cu(li, lj, lk, 4) + cu(i, j, k, 4)) * vel_f + 0.5 * (prim(li, lj, lk, 5) + prim(i, j, k, 5)) * vel_f;
        }
        if (P.algorithm_type == 2) {
#pragma unroll
            for (int n = 0; n < NS; ++n) {
                const double Yk = 0.5 * (prim(li, lj, lk, 6 + n) + prim(i, j, k, 6 + n));
                if (P.advection_type == 0 || P.advection_type == 1)
                    f[5 + n] += Yk * mom;
                else if (P.advection_type == 2)
                    f[5 + n] += 0.5 * (cu(li, lj, lk, 5 + n) + cu(i, j, k, 5 + n)) * vel_f;
                if (P.advection_type == 0) f[4] += mom * Yk * hk[n];
            }
        }
        if (P.advection_type == 0) {
            const Arr& mx = A.mom[0];
            const Arr& my = A.mom[1];
            const Arr& mz = A.mom[2];
            double keP = 0., keM = 0., s;
            s = mx(i + 1, j, k) + mx(i, j, k);
            keP += s * s;
            s = my(i, j + 1, k) + my(i, j, k);
            keP += s * s;
            s = mz(i, j, k + 1) + mz(i, j, k);
            keP += s * s;
            keP *= (0.125 / cu(i, j, k, 0) / cu(i, j, k, 0));
            s = mx(li + 1, lj, lk) + mx(li, lj, lk);
            keM += s * s;
            s = my(li, lj + 1, lk) + my(li, lj, lk);
            keM += s * s;
            s = mz(li, lj, lk + 1) + mz(li, lj, lk);
            keM += s * s;
            // z faces divide by the density of cell (i, j-1, k) (fluxStag.cpp:1754)
            const double rm = (D == 2) ? cu(i, j - 1, k, 0) : cu(li, lj, lk, 0);
            keM *= (0.125 / rm / rm);
            f[4] += mom * 0.5 * (keP + keM);
        }
    }
    f[4] += f[nv] + f[nv + 1] + f[nv + 2] + f[nv + 3];
#pragma unroll
    for (int c = 0; c < nv + 4; ++c) A.ff[D](i, j, k, c) = f[c];
}

// ---- stage update: timeStepStag.cpp:470-538, 771-857, 1106-1199 --------------------------

template <int NS>
__global__ void __launch_bounds__(SB) k_stag_update(DevParams P, StagUpdate U, int st, double dt)
{
    constexpr int nv = NS + 5;
    const int part = blockIdx.y;  // 0 cells, 1..3 faces of direction part-1
    const long tid = (long)blockIdx.x * SB + threadIdx.x;
    const double* dx = P.dx;
    const double* grav = P.grav;
    if (part == 0) {
        const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
        if (tid >= (long)n0 * n1 * n2) return;
        int i, j, k;
        cell_of(tid, n0, n1, i, j, k);
        double src[nv];
#pragma unroll
        for (int n = 0; n < nv; ++n) src[n] = 0.0;
        if (U.chem) {  // compute_compressible_chemistry_source_CLE from the stage's prim
            const double dV = P.dx[0] * P.dx[1] * P.dx[2];
            double X[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) X[s] = U.prim(i, j, k, 6 + NS + s);
            const uint32_t lin = (uint32_t)(((long)k * n1 + j) * n0 + i);
            chem_source_cell<NS>(P, U.prim(i, j, k, 4), U.prim(i, j, k, 5), X,
                                 [&](int m) { return chem_noise(U.K, m, lin, U.w); }, dt, dV, &src[5]);
            if (U.ranchem.p)
                for (int m = 0; m < P.chem.nreaction; ++m) U.ranchem(i, j, k, m) = chem_noise(U.K, m, lin, U.w);
            if (U.source.p)
#pragma unroll
                for (int n = 5; n < nv; ++n) U.source(i, j, k, n) = src[n];
        }
        // Components 1-3 (cell-centred momentum) are skipped: the reference computes them here
        // and conservedToPrimitiveStag overwrites them (and FillBoundary their ghosts) before
        // anything reads them, so the result does not depend on them.
#pragma unroll
        for (int n = 0; n < nv; ++n) {
            if (n >= 1 && n <= 3) continue;
            const double div = (U.ff[0](i + 1, j, k, n) - U.ff[0](i, j, k, n)) / dx[0] +
                               (U.ff[1](i, j + 1, k, n) - U.ff[1](i, j, k, n)) / dx[1] +
                               (U.ff[2](i, j, k + 1, n) - U.ff[2](i, j, k, n)) / dx[2];
            double v;
            if (st == 0)
                v = U.u0(i, j, k, n) - dt * div + dt * src[n];
            else if (st == 1)
                v = 0.25 * (3.0 * U.u0(i, j, k, n) + U.us(i, j, k, n) - dt * div + dt * src[n]);
            else
                v = (2. / 3.) * (0.5 * U.u0(i, j, k, n) + U.us(i, j, k, n) - dt * div + dt * src[n]);
            U.out(i, j, k, n) = v;
        }
        const double gs = grav[0] * (U.ms[0](i + 1, j, k) + U.ms[0](i, j, k)) +
                          grav[1] * (U.ms[1](i, j + 1, k) + U.ms[1](i, j, k)) +
                          grav[2] * (U.ms[2](i, j, k + 1) + U.ms[2](i, j, k));
        double e = U.out(i, j, k, 4);
        if (st == 0)
            e += 0.5 * dt * gs;
        else if (st == 1)
            e += 0.5 * 0.25 * dt * gs;
        else
            e += 0.5 * (2. / 3.) * dt * gs;
        U.out(i, j, k, 4) = e;
        return;
    }
    const int d = part - 1;
    const int ed0 = d == 0, ed1 = d == 1, ed2 = d == 2;
    const int fx = P.n[0] + ed0, fy = P.n[1] + ed1, fz = P.n[2] + ed2;
    if (tid >= (long)fx * fy * fz) return;
    int i, j, k;
    cell_of(tid, fx, fy, i, j, k);
    double c[3];  // flux differences along x, y, z
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        if (t == d) {
            c[t] = U.cf[d](i, j, k) - U.cf[d](i - ed0, j - ed1, k - ed2);
        } else {
            const Arr& e = U.ef[edge_flux_index(d, t)];
            c[t] = e(i + (t == 0), j + (t == 1), k + (t == 2)) - e(i, j, k);
        }
    }
    const double rs = U.us(i - ed0, j - ed1, k - ed2, 0) + U.us(i, j, k, 0);
    const double m0 = U.m0[d](i, j, k);
    double v;
    if (st == 0)
        v = m0 - dt * c[0] / dx[0] - dt * c[1] / dx[1] - dt * c[2] / dx[2] + 0.5 * dt * grav[d] * rs;
    else if (st == 1)
        v = 0.25 * 3.0 * m0 + 0.25 * U.ms[d](i, j, k) - 0.25 * dt * c[0] / dx[0] - 0.25 * dt * c[1] / dx[1] -
            0.25 * dt * c[2] / dx[2] + 0.5 * 0.25 * dt * grav[d] * rs;
    else
        v = (2. / 3.) * (0.5 * m0 + U.ms[d](i, j, k)) - (2. / 3.) * dt * c[0] / dx[0] - (2. / 3.) * dt * c[1] / dx[1] -
            (2. / 3.) * dt * c[2] / dx[2] + 0.5 * (2. / 3.) * dt * grav[d] * rs;
    U.outm[d](i, j, k) = v;
}

// ---- conservedToPrimitiveStag (conservedPrimitiveConversionsStag.cpp:58-140) ---------------

// TR: 0 conversion only; 1 also the transport coefficients of the parameters' type (1 or 3) and
// 2 of transport_type 2, from the prim values just computed (the same inputs k_transport reads).
template <int NS, int TR>
__global__ void __launch_bounds__(SB) k_stag_c2p(DevParams P, Arr prim, Arr vx, Arr vy, Arr vz, Arr cons, Arr mx,
                                                 Arr my, Arr mz, StagTransport tc)
{
    const int part = blockIdx.y;  // 0 cells, 1..3 face velocities of direction part-1
    const long tid = (long)blockIdx.x * SB + threadIdx.x;
    if (part > 0) {
        const int d = part - 1;
        const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
        const int fx = P.n[0] + e0, fy = P.n[1] + e1, fz = P.n[2] + e2;
        if (tid >= (long)fx * fy * fz) return;
        int i, j, k;
        cell_of(tid, fx, fy, i, j, k);
        const Arr& m = d == 0 ? mx : d == 1 ? my : mz;
        const Arr& v = d == 0 ? vx : d == 1 ? vy : vz;
        v(i, j, k) = 2 * m(i, j, k) / (cons(i, j, k, 0) + cons(i - e0, j - e1, k - e2, 0));
        return;
    }
    const int n0 = P.n[0], n1 = P.n[1], n2 = P.n[2];
    if (tid >= (long)n0 * n1 * n2) return;
    int i, j, k;
    cell_of(tid, n0, n1, i, j, k);
    const double rho = cons(i, j, k, 0);
    // the face velocities of this cell, as the face pass computes them
    const double vxl = 2 * mx(i, j, k) / (cons(i, j, k, 0) + cons(i - 1, j, k, 0));
    const double vxh = 2 * mx(i + 1, j, k) / (cons(i + 1, j, k, 0) + cons(i, j, k, 0));
    const double vyl = 2 * my(i, j, k) / (cons(i, j, k, 0) + cons(i, j - 1, k, 0));
    const double vyh = 2 * my(i, j + 1, k) / (cons(i, j + 1, k, 0) + cons(i, j, k, 0));
    const double vzl = 2 * mz(i, j, k) / (cons(i, j, k, 0) + cons(i, j, k - 1, 0));
    const double vzh = 2 * mz(i, j, k + 1) / (cons(i, j, k + 1, 0) + cons(i, j, k, 0));
    prim(i, j, k, 0) = rho;
    prim(i, j, k, 1) = 0.5 * (vxl + vxh);
    prim(i, j, k, 2) = 0.5 * (vyl + vyh);
    prim(i, j, k, 3) = 0.5 * (vzl + vzh);
    cons(i, j, k, 1) = 0.5 * (mx(i, j, k) + mx(i + 1, j, k));
    cons(i, j, k, 2) = 0.5 * (my(i, j, k) + my(i, j + 1, k));
    cons(i, j, k, 3) = 0.5 * (mz(i, j, k) + mz(i, j, k + 1));
    const double sx = mx(i + 1, j, k) + mx(i, j, k);
    const double sy = my(i, j + 1, k) + my(i, j, k);
    const double sz = mz(i, j, k + 1) + mz(i, j, k);
    double kin = 0.;
    kin += sx * sx;
    kin += sy * sy;
    kin += sz * sz;
    kin *= (0.125 / rho);
    const double intenergy = (cons(i, j, k, 4) - kin) / rho;
    double Yk[NS], Yf[NS];
    double sumYk = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        Yk[n] = cons(i, j, k, 5 + n) / rho;
        Yf[n] = clamp01(Yk[n]);
        sumYk += Yf[n];
    }
#pragma unroll
    for (int n = 0; n < NS; ++n) Yf[n] /= sumYk;
    const double T = eos_temperature<NS>(P, intenergy, Yf);
    prim(i, j, k, 4) = T;
    double molmix = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) molmix += Yk[n] / P.molmass[n];
    molmix = 1. / molmix;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        prim(i, j, k, 6 + n) = Yk[n];
        prim(i, j, k, 6 + NS + n) = Yk[n] * molmix / P.molmass[n];
    }
    const double pres = eos_pressure<NS>(P, Yk, rho, T);
    prim(i, j, k, 5) = pres;
    if constexpr (TR != 0) {  // calculateTransportCoeffs of this cell (transCoeffs.cpp:37-88)
        Transport tr;
        double chiv[NS], Dv[NS * NS];
        if constexpr (TR == 2) {
            const int fl = transport_cell_vw<NS>(P, rho, T, pres, Yk, tr, chiv, Dv);
            if (fl) atomicOr(tc.err, fl);
        } else {
            if (transport_cell<NS>(P, rho, T, pres, Yk, tr, chiv, Dv)) atomicOr(tc.err, 1);
        }
        tc.eta(i, j, k) = tr.eta;
        tc.zeta(i, j, k) = tr.zeta;
        tc.kappa(i, j, k) = tr.kappa;
#pragma unroll
        for (int n = 0; n < NS; ++n) tc.chi(i, j, k, n) = chiv[n];
#pragma unroll
        for (int n = 0; n < NS * NS; ++n) tc.D(i, j, k, n) = Dv[n];
    }
}

// ---- periodic ghost fill of cell and face FABs ----------------------------------------------

// blockIdx.y selects an array.  The threads enumerate only the ghost points: the two z slabs,
// then the two y slabs between them, then the two x slabs; each takes the periodic image.
__global__ void __launch_bounds__(SB) k_stag_fill(DevParams P, FillList L)
{
    const FillItem& it = L.item[blockIdx.y];
    const int g = L.g;
    int V[3], G[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        V[d] = P.n[d] + (d == it.nodal);
        G[d] = V[d] + 2 * g;
    }
    const long nz_slab = 2L * g * G[1] * G[0], ny_slab = 2L * V[2] * g * G[0], nx_slab = 2L * V[2] * V[1] * g;
    long t = (long)blockIdx.x * SB + threadIdx.x;
    int i, j, k;
    if (t < nz_slab) {
        const long per = (long)g * G[1] * G[0];
        const int side = t >= per;
        const long r = t - side * per;
        i = (int)(r % G[0]) - g;
        j = (int)((r / G[0]) % G[1]) - g;
        const int kk = (int)(r / ((long)G[0] * G[1]));
        k = side ? V[2] + kk : kk - g;
    } else if ((t -= nz_slab) < ny_slab) {
        const long per = (long)V[2] * g * G[0];
        const int side = t >= per;
        const long r = t - side * per;
        i = (int)(r % G[0]) - g;
        const int jj = (int)((r / G[0]) % g);
        k = (int)(r / ((long)G[0] * g));
        j = side ? V[1] + jj : jj - g;
    } else if ((t -= ny_slab) < nx_slab) {
        const long per = (long)V[2] * V[1] * g;
        const int side = t >= per;
        const long r = t - side * per;
        const int ii = (int)(r % g);
        j = (int)((r / g) % V[1]);
        k = (int)(r / ((long)g * V[1]));
        i = side ? V[0] + ii : ii - g;
    } else {
        return;
    }
    int v[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        while (v[d] < 0) v[d] += P.n[d];
        while (v[d] > V[d] - 1) v[d] -= P.n[d];
    }
    for (int c = it.c0; c < it.c0 + it.nc; ++c) it.a(i, j, k, c) = it.a(v[0], v[1], v[2], c);
}

// ---- host launchers -----------------------------------------------------------------------

#define STAG_NS_DISPATCH(NSV, CALL)                     \
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

static long ghost_points(const DevParams& P, int nodal, int g)
{
    long V = 1, G = 1;
    for (int d = 0; d < 3; ++d) {
        const long v = P.n[d] + (d == nodal);
        V *= v;
        G *= v + 2 * g;
    }
    return G - V;
}

int stag_fill(const DevParams& P, const FillList& L, int count, hipStream_t s)
{
    if (count <= 0) return FHDX_OK;
    long mx = 0;
    for (int q = 0; q < count; ++q) mx = std::max(mx, ghost_points(P, L.item[q].nodal, L.g));
    k_stag_fill<<<dim3(grid_for(mx, SB), count), SB, 0, s>>>(P, L);
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int stag_flux(const DevParams& P, const StagArrays& A, const NoiseKey& K, double w, double dt, hipStream_t s)
{
    const long ngrown1 = (long)(P.n[0] + 2) * (P.n[1] + 2) * (P.n[2] + 2);
    k_stag_cells<<<grid_for(ngrown1, SB), SB, 0, s>>>(P, A, K, w, dt);
    long ne = 0;
    for (int e = 0; e < 3; ++e)
        ne = std::max(ne, (long)(P.n[0] + (e != 2)) * (P.n[1] + (e != 1)) * (P.n[2] + (e != 0)));
    k_stag_edges<<<dim3(grid_for(ne, SB), 3), SB, 0, s>>>(P, A, K, w, dt);
    const long nf[3] = {(long)(P.n[0] + 1) * P.n[1] * P.n[2], (long)P.n[0] * (P.n[1] + 1) * P.n[2],
                        (long)P.n[0] * P.n[1] * (P.n[2] + 1)};
    STAG_NS_DISPATCH(P.ns, (k_stag_faces<NS, 0><<<grid_for(nf[0], SB), SB, 0, s>>>(P, A, K, w, dt),
                            k_stag_faces<NS, 1><<<grid_for(nf[1], SB), SB, 0, s>>>(P, A, K, w, dt),
                            k_stag_faces<NS, 2><<<grid_for(nf[2], SB), SB, 0, s>>>(P, A, K, w, dt)));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int stag_update(const DevParams& P, const StagUpdate& U, int st, double dt, hipStream_t s)
{
    long mx = (long)P.n[0] * P.n[1] * P.n[2];
    for (int d = 0; d < 3; ++d) mx = std::max(mx, (long)(P.n[0] + (d == 0)) * (P.n[1] + (d == 1)) * (P.n[2] + (d == 2)));
    STAG_NS_DISPATCH(P.ns, (k_stag_update<NS><<<dim3(grid_for(mx, SB), 4), SB, 0, s>>>(P, U, st, dt)));
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

int stag_c2p(const DevParams& P, Arr prim, const Arr vel[3], Arr cons, const Arr mom[3], const StagTransport* tr,
             hipStream_t s)
{
    long mx = (long)P.n[0] * P.n[1] * P.n[2];
    for (int d = 0; d < 3; ++d) mx = std::max(mx, (long)(P.n[0] + (d == 0)) * (P.n[1] + (d == 1)) * (P.n[2] + (d == 2)));
    const dim3 grid(grid_for(mx, SB), 4);
    const StagTransport tc = tr ? *tr : StagTransport{};
    if (!tr) {
        STAG_NS_DISPATCH(P.ns, (k_stag_c2p<NS, 0><<<grid, SB, 0, s>>>(P, prim, vel[0], vel[1], vel[2], cons, mom[0],
                                                                       mom[1], mom[2], tc)));
    } else if (P.transport_type == 2) {
        STAG_NS_DISPATCH(P.ns, (k_stag_c2p<NS, 2><<<grid, SB, 0, s>>>(P, prim, vel[0], vel[1], vel[2], cons, mom[0],
                                                                       mom[1], mom[2], tc)));
    } else {
        STAG_NS_DISPATCH(P.ns, (k_stag_c2p<NS, 1><<<grid, SB, 0, s>>>(P, prim, vel[0], vel[1], vel[2], cons, mom[0],
                                                                       mom[1], mom[2], tc)));
    }
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

}  // namespace fhdx
