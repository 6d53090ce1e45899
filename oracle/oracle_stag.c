/*
 * oracle_stag.c -- FP64 CPU restatement of FHDeX's staggered compressible RK3 step
 * (src_compressible_stag) on a fully periodic box.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP staggered path, never the product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * Momentum lives on faces (cumom[d], nodal in d), velocities likewise (vel[d]); density, energy
 * and the species densities live in cells (cu).  Each routine keeps the reference's operation
 * order (left-to-right sums, its parenthesisation, divisions where it divides); build with
 * -ffp-contract=off.  Scope: the 3-D branch (do_1D = do_2D = 0), periodic in every direction,
 * no reservoir, membrane, turbulence forcing or chemistry.  In that scope setBCStag,
 * BCMassTempPress, BCMomNormal, BCMomTrans, BCRhoRhoE, StochFluxStag and
 * BCWallReservoirFluxStag change nothing (their branches need bc codes other than -1,
 * boundaryStag.cpp:268-300, 840-1289, 1550-1663, 2086-2130), so a stage ends with the periodic
 * FillBoundary calls alone.
 *
 * Functions and the reference code they follow (src_compressible_stag/):
 *   c2p_stag     conservedPrimitiveConversionsStag.cpp:4-144 (+ GetTemperature, GetMolfrac,
 *                GetPressureGas of src_compressible/compressible_functions.H:138-199)
 *   stag_cell    fluxStag.cpp:163-206 (diagonal stochastic stress), :872-924 (diagonal
 *                stress), :799-803, 1444-1449, 1838-1856 (cell-centred momentum fluxes)
 *   stag_edge    fluxStag.cpp:209-295 (off-diagonal stochastic stress), :927-1022
 *                (off-diagonal stress), :784-797, 1429-1441, 1767-1835 (edge momentum fluxes)
 *   stag_face    fluxStag.cpp:298-780 (stochastic face fluxes), :1026-1426 (diffusive),
 *                :1507-1764 (hyperbolic: mass, energy, species)
 *   stag_update  timeStepStag.cpp:470-538 (stage 1), 771-857 (stage 2), 1106-1199 (stage 3)
 *   rk3_step     timeStepStag.cpp:13-1245 (noise weights :237, 568, 901; LinComb of the
 *                A/B sets before each calculateFluxStag)
 *   noise        replaces the MultiFabFillRandomNormal calls of timeStepStag.cpp:171-223
 *                (Philox4x32-10 + Box-Muller keyed by the wrapped global index, which gives
 *                the same value to duplicate periodic faces and edges, as OverrideSync does)
 *
 * Pinned in tests/test_oracle_stag.py: exact discrete conservation of mass, momentum (summed
 * over the n distinct faces of each direction), energy and species with noise on; a uniform
 * state at rest with the noise off stays put bit for bit; a transverse velocity wave decays at
 * nu * k_eff^2 and a sound wave oscillates at c_s * k_eff with the staggered grid's effective
 * wavenumber 2 sin(k dx / 2) / dx; cyclic relabelling of the axes maps the fluxes of a
 * perturbed state onto each other (to rounding) except where the reference's z-face quirks
 * break it; and the equilibrium variances of the noisy step against fluctuation theory.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

/* Noise fields of the staggered scheme (the low byte of the Philox counter's third word): face
 * components 4..nvars-1 of direction d at 176 + 16 d + (c - 4); the three edge fields (xy, xz,
 * yz) at 224 + e; the three cell fields at 232 + d.  Disjoint from the collocated step's. */
#define STAG_FACE_FIELD 176u
#define STAG_EDGE_FIELD 224u
#define STAG_CELL_FIELD 232u

/* Edge types: 0 xy (nodal in x and y), 1 xz, 2 yz; their axes (a < b). */
static const int EDGE_AX[3][2] = {{0, 1}, {0, 2}, {1, 2}};
/* The edge flux array of momentum component m transported along direction t (m != t), in the
 * reference's order edgeflux_x[0..1], edgeflux_y[0..1], edgeflux_z[0..1]: x_v, x_w, y_u, y_w,
 * z_u, z_v (fluxStag.cpp:101-106). */
static int edge_flux_index(int m, int t)
{
    static const int T[3][3] = {{-1, 2, 4}, {0, -1, 5}, {1, 3, -1}};
    return T[m][t];
}
/* The edge type of the plane of two distinct axes. */
static int edge_type(int a, int b)
{
    if (a > b) { int t = a; a = b; b = t; }
    return (a == 0) ? (b == 1 ? 0 : 1) : 2;
}

typedef struct {
    const fhdx_params* P;
    int NS, nv;
    double dt, volinv, dtinv;
    fab_t cu, prim, eta, zeta, kappa, chi, D;
    fab_t mom[3], vel[3];
    fab_t ff[3];  /* face fluxes, nvars + 4 comps */
    fab_t ef[6];  /* edge fluxes */
    fab_t cf[3];  /* cell-centred fluxes, one ghost layer */
    fab_t sf[3];  /* weighted face noise (components 4..nvars-1) */
    fab_t se[3];  /* weighted edge noise: xy, xz, yz */
    fab_t sc[3];  /* weighted cell noise, one ghost layer */
    fab_t td;     /* cells grown by 1: tau_stoch xx yy zz, tau xx yy zz */
    fab_t to[3];  /* edges: tau_stoch, tau of each edge type */
} stag_args;

/* ---- periodic ghost fill of a cell (nodal = -1) or face (nodal = d) FAB ----------------- */

static void fill_ghosts(const fhdx_params* P, fab_t f, int nodal)
{
    const int n[3] = {P->n_cells[0], P->n_cells[1], P->n_cells[2]};
#pragma omp parallel for schedule(static)
    for (int k = f.lo2; k <= f.hi2; ++k)
        for (int j = f.lo1; j <= f.hi1; ++j)
            for (int i = f.lo0; i <= f.hi0; ++i) {
                const int c[3] = {i, j, k};
                int v[3], ghost = 0;
                for (int d = 0; d < 3; ++d) {
                    const int hi = n[d] - 1 + (d == nodal);
                    v[d] = c[d];
                    while (v[d] < 0) v[d] += n[d];
                    while (v[d] > hi) v[d] -= n[d];
                    ghost |= (v[d] != c[d]);
                }
                if (!ghost) continue;
                for (int q = 0; q < f.nc; ++q) AT(f, i, j, k, q) = AT(f, v[0], v[1], v[2], q);
            }
}

/* ---- conservedToPrimitiveStag (conservedPrimitiveConversionsStag.cpp:4-144) ------------- */

static void c2p_stag_valid(const fhdx_params* P, fab_t prim, const fab_t* vel, fab_t cons,
                           const fab_t* mom)
{
    const int NS = P->nspecies;
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    /* face velocities (valid faces 0..n along the face direction) */
    for (int d = 0; d < 3; ++d) {
        const int e0 = (d == 0), e1 = (d == 1), e2 = (d == 2);
        fab_t v = vel[d], m = mom[d];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n2 + e2; ++k)
            for (int j = 0; j < n1 + e1; ++j)
                for (int i = 0; i < n0 + e0; ++i)
                    AT(v, i, j, k, 0) = 2 * AT(m, i, j, k, 0) /
                                        (AT(cons, i, j, k, 0) + AT(cons, i - e0, j - e1, k - e2, 0));
    }
    const fab_t mx = mom[0], my = mom[1], mz = mom[2], vx = vel[0], vy = vel[1], vz = vel[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                double Xk[MS] = {0.}, Yk[MS], Yf[MS];
                const double rho = AT(cons, i, j, k, 0);
                AT(prim, i, j, k, 0) = rho;
                AT(prim, i, j, k, 1) = 0.5 * (AT(vx, i, j, k, 0) + AT(vx, i + 1, j, k, 0));
                AT(prim, i, j, k, 2) = 0.5 * (AT(vy, i, j, k, 0) + AT(vy, i, j + 1, k, 0));
                AT(prim, i, j, k, 3) = 0.5 * (AT(vz, i, j, k, 0) + AT(vz, i, j, k + 1, 0));
                AT(cons, i, j, k, 1) = 0.5 * (AT(mx, i, j, k, 0) + AT(mx, i + 1, j, k, 0));
                AT(cons, i, j, k, 2) = 0.5 * (AT(my, i, j, k, 0) + AT(my, i, j + 1, k, 0));
                AT(cons, i, j, k, 3) = 0.5 * (AT(mz, i, j, k, 0) + AT(mz, i, j, k + 1, 0));
                const double sx = AT(mx, i + 1, j, k, 0) + AT(mx, i, j, k, 0);
                const double sy = AT(my, i, j + 1, k, 0) + AT(my, i, j, k, 0);
                const double sz = AT(mz, i, j, k + 1, 0) + AT(mz, i, j, k, 0);
                double kin = 0.;
                kin += sx * sx;
                kin += sy * sy;
                kin += sz * sz;
                kin *= (0.125 / rho);
                const double intenergy = (AT(cons, i, j, k, 4) - kin) / rho;
                double sumYk = 0.;
                for (int n = 0; n < NS; ++n) {
                    Yk[n] = AT(cons, i, j, k, 5 + n) / rho;
                    Yf[n] = clamp01(Yk[n]);
                    sumYk += Yf[n];
                }
                for (int n = 0; n < NS; ++n) Yf[n] /= sumYk;
                const double T = get_temperature(P, intenergy, Yf);
                AT(prim, i, j, k, 4) = T;
                get_molfrac(P, Yk, Xk);
                for (int n = 0; n < NS; ++n) {
                    AT(prim, i, j, k, 6 + n) = Yk[n];
                    AT(prim, i, j, k, 6 + NS + n) = Xk[n];
                }
                AT(prim, i, j, k, 5) = get_pressure(P, Yk, rho, T);
            }
}

/* The end of a stage (timeStepStag.cpp:515-538): FillBoundary of the conserved state and
 * momentum, conservedToPrimitiveStag, FillBoundary of velocity, prim and cons. */
static void stage_end(const fhdx_params* P, fab_t prim, const fab_t* vel, fab_t cons,
                      const fab_t* mom)
{
    for (int d = 0; d < 3; ++d) fill_ghosts(P, mom[d], d);
    fill_ghosts(P, cons, -1);
    c2p_stag_valid(P, prim, vel, cons, mom);
    for (int d = 0; d < 3; ++d) fill_ghosts(P, vel[d], d);
    fill_ghosts(P, prim, -1);
    fill_ghosts(P, cons, -1);
}

int orc_c2p_stag(const fhdx_params* P, fhdx_fab prim, const fhdx_fab vel[3], fhdx_fab cu,
                 const fhdx_fab cumom[3])
{
    fab_t v[3], m[3];
    for (int d = 0; d < 3; ++d) {
        v[d] = F(vel[d]);
        m[d] = F(cumom[d]);
    }
    stage_end(P, F(prim), v, F(cu), m);
    return FHDX_OK;
}

/* ---- calculateFluxStag ------------------------------------------------------------------- */

/* One cell of the box grown by 1: diagonal stochastic stress (fluxStag.cpp:163-206), diagonal
 * stress (:872-924) and the cell-centred momentum fluxes cenflux (:799-803, 1444-1449,
 * 1838-1856: tau_stoch - tau + (momentum advection + p)). */
static void stag_cell(const stag_args* A, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const double* dx = P->dx;
    const int visc3 = abs(P->visc_type) == 3;
    const double eta = AT(A->eta, i, j, k, 0), zeta = AT(A->zeta, i, j, k, 0);
    const double T = AT(A->prim, i, j, k, 4);
    double ts[3] = {0., 0., 0.};
    if (P->stoch_stress_form == 1) {
        const double etaT = eta * T;
        const double zetaT = zeta * T;
        const double fac1 = sqrt(2.0 * P->k_B * etaT * A->volinv * A->dtinv);
        double fac2 = (-1.0 / 3.0) * sqrt(2.0 * P->k_B * etaT * A->volinv * A->dtinv);
        if (visc3)
            fac2 = sqrt(P->k_B * zetaT * A->volinv * A->dtinv / 3.0) -
                   sqrt(2.0 * P->k_B * etaT * A->volinv * A->dtinv) / 3.0;
        const double zx = AT(A->sc[0], i, j, k, 0), zy = AT(A->sc[1], i, j, k, 0),
                     zz = AT(A->sc[2], i, j, k, 0);
        const double traceZ = zx + zy + zz;
        ts[0] = (fac1 * zx) + (fac2 * traceZ);
        ts[1] = (fac1 * zy) + (fac2 * traceZ);
        ts[2] = (fac1 * zz) + (fac2 * traceZ);
    }
    const double g[3] = {(AT(A->vel[0], i + 1, j, k, 0) - AT(A->vel[0], i, j, k, 0)) / dx[0],
                         (AT(A->vel[1], i, j + 1, k, 0) - AT(A->vel[1], i, j, k, 0)) / dx[1],
                         (AT(A->vel[2], i, j, k + 1, 0) - AT(A->vel[2], i, j, k, 0)) / dx[2]};
    const double div = g[0] + g[1] + g[2];
    double tau[3];
    for (int c = 0; c < 3; ++c)
        tau[c] = visc3 ? 2 * eta * g[c] + (zeta - 2 * eta / 3.) * div
                       : 2 * eta * g[c] + (0.0 - 2 * eta / 3.) * div;
    const double p = AT(A->prim, i, j, k, 5);
    for (int c = 0; c < 3; ++c) {
        AT(A->td, i, j, k, c) = ts[c];
        AT(A->td, i, j, k, 3 + c) = tau[c];
        double cf = ts[c];
        cf -= tau[c];
        if (P->advection_type >= 0) {
            const int e0 = (c == 0), e1 = (c == 1), e2 = (c == 2);
            const double m0 = AT(A->mom[c], i, j, k, 0), m1 = AT(A->mom[c], i + e0, j + e1, k + e2, 0);
            const double v0 = AT(A->vel[c], i, j, k, 0), v1 = AT(A->vel[c], i + e0, j + e1, k + e2, 0);
            cf += 0.25 * (m0 + m1) * (v0 + v1) + p;
        }
        AT(A->cf[c], i, j, k, 0) = cf;
    }
}

/* One edge of type e (axes a < b): off-diagonal stochastic stress (fluxStag.cpp:209-295),
 * off-diagonal stress (:927-1022) and the two edge momentum fluxes (tau_stoch - tau +
 * advection, :784-797, 1429-1441, 1767-1835). */
static void stag_edge(const stag_args* A, int e, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const int a = EDGE_AX[e][0], b = EDGE_AX[e][1];
    const int ea[3] = {a == 0, a == 1, a == 2}, eb[3] = {b == 0, b == 1, b == 2};
    /* the four cells around the edge, in the reference's order: (-a,-b), (-a,0), (0,-b), (0,0) */
    const int ci[4] = {i - ea[0] - eb[0], i - ea[0], i - eb[0], i};
    const int cj[4] = {j - ea[1] - eb[1], j - ea[1], j - eb[1], j};
    const int ck[4] = {k - ea[2] - eb[2], k - ea[2], k - eb[2], k};
    double ts = 0.;
    if (P->stoch_stress_form == 1) {
        double s = AT(A->eta, ci[0], cj[0], ck[0], 0) * AT(A->prim, ci[0], cj[0], ck[0], 4);
        s = s + AT(A->eta, ci[1], cj[1], ck[1], 0) * AT(A->prim, ci[1], cj[1], ck[1], 4);
        s = s + AT(A->eta, ci[2], cj[2], ck[2], 0) * AT(A->prim, ci[2], cj[2], ck[2], 4);
        s = s + AT(A->eta, ci[3], cj[3], ck[3], 0) * AT(A->prim, ci[3], cj[3], ck[3], 4);
        const double etaT = 0.25 * s;
        const double fac = sqrt(2.0 * P->k_B * etaT * A->volinv * A->dtinv);
        ts = fac * AT(A->se[e], i, j, k, 0);
    }
    const fab_t va = A->vel[a], vb = A->vel[b], ma = A->mom[a], mb = A->mom[b];
    const double da_b = (AT(va, i, j, k, 0) - AT(va, i - eb[0], j - eb[1], k - eb[2], 0)) / P->dx[b];
    const double db_a = (AT(vb, i, j, k, 0) - AT(vb, i - ea[0], j - ea[1], k - ea[2], 0)) / P->dx[a];
    double es = AT(A->eta, ci[0], cj[0], ck[0], 0);
    es = es + AT(A->eta, ci[1], cj[1], ck[1], 0);
    es = es + AT(A->eta, ci[2], cj[2], ck[2], 0);
    es = es + AT(A->eta, ci[3], cj[3], ck[3], 0);
    const double eta_interp = 0.25 * es;
    const double tau = eta_interp * (da_b + db_a);
    AT(A->to[e], i, j, k, 0) = ts;
    AT(A->to[e], i, j, k, 1) = tau;
    double fa = ts, fb = ts;  /* momentum a along b, momentum b along a */
    fa -= tau;
    fb -= tau;
    if (P->advection_type >= 0) {
        const double adv_a = 0.25 * (AT(ma, i - eb[0], j - eb[1], k - eb[2], 0) + AT(ma, i, j, k, 0)) *
                             (AT(vb, i - ea[0], j - ea[1], k - ea[2], 0) + AT(vb, i, j, k, 0));
        const double adv_b = 0.25 * (AT(mb, i - ea[0], j - ea[1], k - ea[2], 0) + AT(mb, i, j, k, 0)) *
                             (AT(va, i - eb[0], j - eb[1], k - eb[2], 0) + AT(va, i, j, k, 0));
        fa += adv_a;
        fb += adv_b;
    }
    AT(A->ef[edge_flux_index(a, b)], i, j, k, 0) = fa;
    AT(A->ef[edge_flux_index(b, a)], i, j, k, 0) = fb;
}

/* One face of direction d (lower cell l = face - e_d, upper cell u = face): the stochastic
 * (fluxStag.cpp:298-780), diffusive (:1026-1426) and hyperbolic (:1507-1764) parts, every
 * component of faceflux in the reference's order. */
static void stag_face(const stag_args* A, int d, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    const int ed[3] = {d == 0, d == 1, d == 2};
    const int li = i - ed[0], lj = j - ed[1], lk = k - ed[2];
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int tr[2] = {t1, t2};
    const fab_t *prim = &A->prim, *cu = &A->cu, *chi = &A->chi, *Dm = &A->D;
    const fab_t *vel_d = &A->vel[d], *mom_d = &A->mom[d];
    const double dxd = P->dx[d];
    const double Tu = AT(*prim, i, j, k, 4), Tl = AT(*prim, li, lj, lk, 4);
    double f[MS + 9];
    for (int c = 0; c < nv + 4; ++c) f[c] = 0.;
    double hk[MS];

    /* stochastic part */
    if (P->stoch_stress_form == 1) {
        double w[MS + 1];
        const double kp = AT(A->kappa, i, j, k, 0) * Tu * Tu + AT(A->kappa, li, lj, lk, 0) * Tl * Tl;
        const double meanT = 0.5 * (Tu + Tl);
        const double fw0 = sqrt(P->k_B * kp * A->volinv * A->dtinv);
        w[0] = fw0 * AT(A->sf[d], i, j, k, 4);
        f[nv] = w[0];
        f[nv + 1] = 0.5 * AT(*vel_d, i, j, k, 0) * (AT(A->td, li, lj, lk, d) + AT(A->td, i, j, k, d));
        double vsh = 0.0;
        for (int q = 0; q < 2; ++q) {
            const int t = tr[q];
            const int et[3] = {t == 0, t == 1, t == 2};
            const fab_t *vt = &A->vel[t], *ts = &A->to[edge_type(d, t)];
            const int pi = i + et[0], pj = j + et[1], pk = k + et[2];
            vsh += 0.25 * ((AT(*vt, pi, pj, pk, 0) + AT(*vt, pi - ed[0], pj - ed[1], pk - ed[2], 0)) *
                               AT(*ts, pi, pj, pk, 0) +
                           (AT(*vt, i, j, k, 0) + AT(*vt, li, lj, lk, 0)) * AT(*ts, i, j, k, 0));
        }
        f[nv + 2] = vsh;
        if (P->algorithm_type == 2) {
            double yy[MS], yyp[MS], DijY[MS * MS], sqD[MS * MS];
            for (int n = 1; n < 1 + NS; ++n) w[n] = 0.;
            for (int s = 0; s < NS; ++s) {
                yy[s] = clamp01(AT(*prim, li, lj, lk, 6 + s));
                yyp[s] = clamp01(AT(*prim, i, j, k, 6 + s));
            }
            double sumy = 0., sumyp = 0.;
            for (int n = 0; n < NS; ++n) {
                sumy += yy[n];
                sumyp += yyp[n];
            }
            for (int n = 0; n < NS; ++n) {
                yy[n] /= sumy;
                yyp[n] /= sumyp;
            }
            double MWmix = 0.;
            for (int s = 0; s < NS; ++s) {
                MWmix = MWmix + 0.5 * (yy[s] + yyp[s]) / P->molmass[s];
                for (int l = 0; l < NS; ++l)
                    DijY[s * NS + l] = 0.5 * (AT(*Dm, li, lj, lk, l * NS + s) * yy[l] +
                                              AT(*Dm, i, j, k, l * NS + s) * yyp[l] +
                                              (AT(*Dm, li, lj, lk, s * NS + l) * yy[s] +
                                               AT(*Dm, i, j, k, s * NS + l) * yyp[s]));
            }
            for (int s = 0; s < NS; ++s)
                if (fabs(yy[s]) + fabs(yyp[s]) <= 1.e-12)
                    for (int n = 0; n < NS; ++n) {
                        DijY[s * NS + n] = 0.;
                        DijY[n * NS + s] = 0.;
                    }
            MWmix = 1. / MWmix;
            cholesky(NS, DijY, sqD);
            for (int s = 0; s < NS; ++s) {
                for (int l = 0; l <= s; ++l) {
                    const double fw = sqrt(P->k_B * MWmix * A->volinv / (P->Runiv * A->dt)) * sqD[s * NS + l];
                    w[1 + s] = w[1 + s] + fw * AT(A->sf[d], i, j, k, 5 + l);
                }
                f[5 + s] = w[1 + s];
            }
            for (int s = 0; s < NS; ++s) hk[s] = P->e0[s] + P->hcp[s] * meanT;
            double soret = 0.;
            for (int s = 0; s < NS; ++s)
                soret += (hk[s] + P->Runiv * meanT / P->molmass[s] * 0.5 *
                                      (AT(*chi, li, lj, lk, s) + AT(*chi, i, j, k, s))) *
                         w[1 + s];
            f[nv + 3] = soret;
        }
    }

    /* diffusive part */
    {
        const double kp = 0.5 * (AT(A->kappa, li, lj, lk, 0) + AT(A->kappa, i, j, k, 0));
        const double meanT = 0.5 * (Tl + Tu);
        const double pl = AT(*prim, li, lj, lk, 5), pu = AT(*prim, i, j, k, 5);
        const double meanP = 0.5 * (pl + pu);
        f[nv + 1] -= 0.5 * AT(*vel_d, i, j, k, 0) * (AT(A->td, li, lj, lk, 3 + d) + AT(A->td, i, j, k, 3 + d));
        double vsh = 0.0;
        for (int q = 0; q < 2; ++q) {
            const int t = tr[q];
            const int et[3] = {t == 0, t == 1, t == 2};
            const fab_t *vt = &A->vel[t], *ts = &A->to[edge_type(d, t)];
            const int pi = i + et[0], pj = j + et[1], pk = k + et[2];
            vsh -= 0.25 * ((AT(*vt, pi, pj, pk, 0) + AT(*vt, pi - ed[0], pj - ed[1], pk - ed[2], 0)) *
                               AT(*ts, pi, pj, pk, 1) +
                           (AT(*vt, i, j, k, 0) + AT(*vt, li, lj, lk, 0)) * AT(*ts, i, j, k, 1));
        }
        f[nv] -= kp * (Tu - Tl) / dxd;
        f[nv + 2] += vsh;
        if (P->algorithm_type == 2) {
            double dk[MS], soret[MS], Fk[MS];
            for (int s = 0; s < NS; ++s) {
                const double Xl = AT(*prim, li, lj, lk, 6 + NS + s), Xu = AT(*prim, i, j, k, 6 + NS + s);
                const double Yl = AT(*prim, li, lj, lk, 6 + s), Yu = AT(*prim, i, j, k, 6 + s);
                const double term1 = (Xu - Xl) / dxd;
                const double meanXk = 0.5 * (Xl + Xu);
                const double meanYk = 0.5 * (Yl + Yu);
                const double term2 = (meanXk - meanYk) * (pu - pl) / dxd / meanP;
                dk[s] = term1 + term2;
                const double ChiX = 0.5 * (AT(*chi, li, lj, lk, s) * Xl + AT(*chi, i, j, k, s) * Xu);
                soret[s] = ChiX * (Tu - Tl) / dxd / meanT;
            }
            for (int kk = 0; kk < NS; ++kk) {
                Fk[kk] = 0.;
                for (int l = 0; l < NS; ++l) {
                    const double Fks = 0.5 * (AT(*Dm, li, lj, lk, l * NS + kk) + AT(*Dm, i, j, k, l * NS + kk)) *
                                       (dk[l] + soret[l]);
                    Fk[kk] -= Fks;
                }
            }
            for (int s = 0; s < NS; ++s) hk[s] = P->e0[s] + P->hcp[s] * meanT;
            double Q5 = 0.;
            for (int s = 0; s < NS; ++s) {
                /* z faces add chi(i,j,k) twice (fluxStag.cpp:1398) */
                const double chil = (d == 2) ? AT(*chi, i, j, k, s) : AT(*chi, li, lj, lk, s);
                Q5 += (hk[s] + 0.5 * P->Runiv * meanT * (chil + AT(*chi, i, j, k, s)) / P->molmass[s]) * Fk[s];
            }
            f[nv + 3] += Q5;
            for (int s = 0; s < NS; ++s) f[5 + s] += Fk[s];
        }
    }

    /* hyperbolic part */
    if (P->advection_type >= 0) {
        const double mom = AT(*mom_d, i, j, k, 0), vel = AT(*vel_d, i, j, k, 0);
        f[0] += mom;
        const double meanT = 0.5 * (Tl + Tu);
        const double meanRho = 0.5 * (AT(*prim, li, lj, lk, 0) + AT(*prim, i, j, k, 0));
        const double meanP = 0.5 * (AT(*prim, li, lj, lk, 5) + AT(*prim, i, j, k, 5));
        for (int s = 0; s < NS; ++s) hk[s] = P->e0[s] + P->hcp[s] * meanT;
        const double meanE = 0.5 * (AT(*cu, li, lj, lk, 4) + AT(*cu, i, j, k, 4)) / meanRho;
        if (P->advection_type == 1) {
            f[4] += mom * (meanE + (meanP / meanRho));
        } else if (P->advection_type == 2) {
            f[4] += 0.5 * (AT(*cu, li, lj, lk, 4) + AT(*cu, i, j, k, 4)) * vel +
                    0.5 * (AT(*prim, li, lj, lk, 5) + AT(*prim, i, j, k, 5)) * vel;
        }
        if (P->algorithm_type == 2) {
            for (int n = 0; n < NS; ++n) {
                const double Yk = 0.5 * (AT(*prim, li, lj, lk, 6 + n) + AT(*prim, i, j, k, 6 + n));
                if (P->advection_type == 0 || P->advection_type == 1)
                    f[5 + n] += Yk * mom;
                else if (P->advection_type == 2)
                    f[5 + n] += 0.5 * (AT(*cu, li, lj, lk, 5 + n) + AT(*cu, i, j, k, 5 + n)) * vel;
                if (P->advection_type == 0) f[4] += mom * Yk * hk[n];
            }
        }
        if (P->advection_type == 0) {
            const fab_t *mx = &A->mom[0], *my = &A->mom[1], *mz = &A->mom[2];
            double keP = 0., keM = 0.;
            double s;
            s = AT(*mx, i + 1, j, k, 0) + AT(*mx, i, j, k, 0);
            keP += s * s;
            s = AT(*my, i, j + 1, k, 0) + AT(*my, i, j, k, 0);
            keP += s * s;
            s = AT(*mz, i, j, k + 1, 0) + AT(*mz, i, j, k, 0);
            keP += s * s;
            keP *= (0.125 / AT(*cu, i, j, k, 0) / AT(*cu, i, j, k, 0));
            s = AT(*mx, li + 1, lj, lk, 0) + AT(*mx, li, lj, lk, 0);
            keM += s * s;
            s = AT(*my, li, lj + 1, lk, 0) + AT(*my, li, lj, lk, 0);
            keM += s * s;
            s = AT(*mz, li, lj, lk + 1, 0) + AT(*mz, li, lj, lk, 0);
            keM += s * s;
            /* z faces divide by the density of cell (i, j-1, k) (fluxStag.cpp:1754) */
            const double rm = (d == 2) ? AT(*cu, i, j - 1, k, 0) : AT(*cu, li, lj, lk, 0);
            keM *= (0.125 / rm / rm);
            f[4] += mom * 0.5 * (keP + keM);
        }
    }
    f[4] += f[nv] + f[nv + 1] + f[nv + 2] + f[nv + 3];
    for (int c = 0; c < nv + 4; ++c) AT(A->ff[d], i, j, k, c) = f[c];
}

static void flux_stag(const stag_args* A)
{
    const fhdx_params* P = A->P;
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
#pragma omp parallel for schedule(static)
    for (int k = -1; k <= n2; ++k)
        for (int j = -1; j <= n1; ++j)
            for (int i = -1; i <= n0; ++i) stag_cell(A, i, j, k);
    for (int e = 0; e < 3; ++e) {
        const int a = EDGE_AX[e][0], b = EDGE_AX[e][1];
        const int x1 = (a == 0 || b == 0), y1 = (a == 1 || b == 1), z1 = (a == 2 || b == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n2 + z1; ++k)
            for (int j = 0; j < n1 + y1; ++j)
                for (int i = 0; i < n0 + x1; ++i) stag_edge(A, e, i, j, k);
    }
    for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n2 + (d == 2); ++k)
            for (int j = 0; j < n1 + (d == 1); ++j)
                for (int i = 0; i < n0 + (d == 0); ++i) stag_face(A, d, i, j, k);
    }
}

/* ---- host arrays -------------------------------------------------------------------------- */

static fhdx_fab mkfab(int l0, int l1, int l2, int h0, int h1, int h2, int nc)
{
    fhdx_fab f;
    memset(&f, 0, sizeof(f));
    f.lo[0] = l0; f.lo[1] = l1; f.lo[2] = l2;
    f.hi[0] = h0; f.hi[1] = h1; f.hi[2] = h2;
    f.ncomp = nc;
    f.p = (double*)calloc((size_t)(h0 - l0 + 1) * (h1 - l1 + 1) * (h2 - l2 + 1) * nc, sizeof(double));
    return f;
}

static fhdx_fab mk_cell(const fhdx_params* P, int g, int nc)
{
    return mkfab(-g, -g, -g, P->n_cells[0] - 1 + g, P->n_cells[1] - 1 + g, P->n_cells[2] - 1 + g, nc);
}

/* face FAB nodal in d with g ghost layers */
static fhdx_fab mk_face(const fhdx_params* P, int d, int g, int nc)
{
    return mkfab(-g, -g, -g, P->n_cells[0] - 1 + g + (d == 0), P->n_cells[1] - 1 + g + (d == 1),
                 P->n_cells[2] - 1 + g + (d == 2), nc);
}

/* edge FAB of type e (nodal along both its axes), no ghosts */
static fhdx_fab mk_edge(const fhdx_params* P, int e, int nc)
{
    const int a = EDGE_AX[e][0], b = EDGE_AX[e][1];
    int h[3];
    for (int d = 0; d < 3; ++d) h[d] = P->n_cells[d] - 1 + (d == a || d == b);
    return mkfab(0, 0, 0, h[0], h[1], h[2], nc);
}

/* ---- noise ------------------------------------------------------------------------------ */

static void normal_at(const uint32_t key[2], uint32_t step_lo, uint32_t fhi, uint32_t instance,
                      uint32_t field, uint32_t lin, double* z0, double* z1)
{
    const uint32_t ctr[4] = {lin, step_lo, field | fhi, instance};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    orc_normal_pair(r, z0, z1);
}

static uint32_t wrapped_lin(const fhdx_params* P, int i, int j, int k)
{
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const int a = ((i % n0) + n0) % n0, b = ((j % n1) + n1) % n1, c = ((k % n2) + n2) % n2;
    return (uint32_t)(((long)c * n1 + b) * n0 + a);
}

int orc_fill_noise_stag(const fhdx_params* P, uint64_t seed, uint64_t step, uint32_t instance,
                        const fhdx_fab faceA[3], const fhdx_fab faceB[3], const fhdx_fab edgeA[3],
                        const fhdx_fab edgeB[3], const fhdx_fab cenA[3], const fhdx_fab cenB[3])
{
    const int nv = nvars_of(P);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t slo = (uint32_t)step, fhi = ((uint32_t)(step >> 32)) << 8;
    const double sq2 = sqrt(2.0);
    for (int d = 0; d < 3; ++d) {
        fab_t a = F(faceA[d]), b = F(faceB[d]);
        for (int c = 4; c < nv; ++c) {
            const uint32_t field = STAG_FACE_FIELD + 16u * (uint32_t)d + (uint32_t)(c - 4);
#pragma omp parallel for schedule(static)
            for (int k = a.lo2; k <= a.hi2; ++k)
                for (int j = a.lo1; j <= a.hi1; ++j)
                    for (int i = a.lo0; i <= a.hi0; ++i) {
                        double z0, z1;
                        normal_at(key, slo, fhi, instance, field, wrapped_lin(P, i, j, k), &z0, &z1);
                        AT(a, i, j, k, c) = z0;
                        AT(b, i, j, k, c) = z1;
                    }
        }
    }
    for (int e = 0; e < 3; ++e) {
        fab_t a = F(edgeA[e]), b = F(edgeB[e]);
#pragma omp parallel for schedule(static)
        for (int k = a.lo2; k <= a.hi2; ++k)
            for (int j = a.lo1; j <= a.hi1; ++j)
                for (int i = a.lo0; i <= a.hi0; ++i) {
                    double z0, z1;
                    normal_at(key, slo, fhi, instance, STAG_EDGE_FIELD + (uint32_t)e, wrapped_lin(P, i, j, k),
                              &z0, &z1);
                    AT(a, i, j, k, 0) = z0;
                    AT(b, i, j, k, 0) = z1;
                }
    }
    for (int d = 0; d < 3; ++d) {
        fab_t a = F(cenA[d]), b = F(cenB[d]);
#pragma omp parallel for schedule(static)
        for (int k = a.lo2; k <= a.hi2; ++k)
            for (int j = a.lo1; j <= a.hi1; ++j)
                for (int i = a.lo0; i <= a.hi0; ++i) {
                    double z0, z1;
                    normal_at(key, slo, fhi, instance, STAG_CELL_FIELD + (uint32_t)d, wrapped_lin(P, i, j, k),
                              &z0, &z1);
                    /* variance 2 (timeStepStag.cpp:211-216): standard deviation sqrt(2) */
                    AT(a, i, j, k, 0) = sq2 * z0;
                    AT(b, i, j, k, 0) = sq2 * z1;
                }
    }
    return FHDX_OK;
}

/* dst = 1.0 A + w B over components c0 .. c0+nc-1 of every point (MultiFab::LinComb). */
static void lincomb(fhdx_fab dst_, fhdx_fab a_, fhdx_fab b_, int c0, int nc, double w)
{
    fab_t dst = F(dst_), a = F(a_), b = F(b_);
    for (int c = c0; c < c0 + nc; ++c) {
        double *pd = dst.p + (long)c * dst.sn, *pa = a.p + (long)c * a.sn, *pb = b.p + (long)c * b.sn;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < dst.sn; ++q) pd[q] = 1.0 * pa[q] + w * pb[q];
    }
}

/* ---- the stage update (timeStepStag.cpp:470-538, 771-857, 1106-1199) ------------------- */

/* stage 0: out = u0 - dt div F; 1: out = 1/4 (3 u0 + us - dt div F); 2: out = 2/3 (1/2 u0 + us
 * - dt div F), with the momenta updated from the cell-centred and edge fluxes, the gravity terms
 * and dt * source (source NULL: zero, no chemistry). */
static void stag_update(const fhdx_params* P, int st, fab_t out, fab_t* outm, fab_t u0, fab_t* m0,
                        fab_t us, fab_t* ms, const fab_t* ff, const fab_t* ef, const fab_t* cf,
                        const fab_t* source, double dt)
{
    const int nv = nvars_of(P);
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const double* dx = P->dx;
    const double* grav = P->grav;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                for (int n = 0; n < nv; ++n) {
                    const double src = source ? AT(*source, i, j, k, n) : 0.0;
                    const double div = (AT(ff[0], i + 1, j, k, n) - AT(ff[0], i, j, k, n)) / dx[0] +
                                       (AT(ff[1], i, j + 1, k, n) - AT(ff[1], i, j, k, n)) / dx[1] +
                                       (AT(ff[2], i, j, k + 1, n) - AT(ff[2], i, j, k, n)) / dx[2];
                    double v;
                    if (st == 0)
                        v = AT(u0, i, j, k, n) - dt * div + dt * src;
                    else if (st == 1)
                        v = 0.25 * (3.0 * AT(u0, i, j, k, n) + AT(us, i, j, k, n) - dt * div + dt * src);
                    else
                        v = (2. / 3.) * (0.5 * AT(u0, i, j, k, n) + AT(us, i, j, k, n) - dt * div + dt * src);
                    AT(out, i, j, k, n) = v;
                }
                const double gs = grav[0] * (AT(ms[0], i + 1, j, k, 0) + AT(ms[0], i, j, k, 0)) +
                                  grav[1] * (AT(ms[1], i, j + 1, k, 0) + AT(ms[1], i, j, k, 0)) +
                                  grav[2] * (AT(ms[2], i, j, k + 1, 0) + AT(ms[2], i, j, k, 0));
                if (st == 0)
                    AT(out, i, j, k, 4) += 0.5 * dt * gs;
                else if (st == 1)
                    AT(out, i, j, k, 4) += 0.5 * 0.25 * dt * gs;
                else
                    AT(out, i, j, k, 4) += 0.5 * (2. / 3.) * dt * gs;
            }
    for (int d = 0; d < 3; ++d) {
        const int ed[3] = {d == 0, d == 1, d == 2};
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n2 + ed[2]; ++k)
            for (int j = 0; j < n1 + ed[1]; ++j)
                for (int i = 0; i < n0 + ed[0]; ++i) {
                    double c[3];  /* flux differences along x, y, z */
                    for (int t = 0; t < 3; ++t) {
                        if (t == d) {
                            c[t] = AT(cf[d], i, j, k, 0) - AT(cf[d], i - ed[0], j - ed[1], k - ed[2], 0);
                        } else {
                            const fab_t e = ef[edge_flux_index(d, t)];
                            c[t] = AT(e, i + (t == 0), j + (t == 1), k + (t == 2), 0) - AT(e, i, j, k, 0);
                        }
                    }
                    const double rs = AT(us, i - ed[0], j - ed[1], k - ed[2], 0) + AT(us, i, j, k, 0);
                    double v;
                    if (st == 0)
                        v = AT(m0[d], i, j, k, 0) - dt * c[0] / dx[0] - dt * c[1] / dx[1] - dt * c[2] / dx[2] +
                            0.5 * dt * grav[d] * rs;
                    else if (st == 1)
                        v = 0.25 * 3.0 * AT(m0[d], i, j, k, 0) + 0.25 * AT(ms[d], i, j, k, 0) -
                            0.25 * dt * c[0] / dx[0] - 0.25 * dt * c[1] / dx[1] - 0.25 * dt * c[2] / dx[2] +
                            0.5 * 0.25 * dt * grav[d] * rs;
                    else
                        v = (2. / 3.) * (0.5 * AT(m0[d], i, j, k, 0) + AT(ms[d], i, j, k, 0)) -
                            (2. / 3.) * dt * c[0] / dx[0] - (2. / 3.) * dt * c[1] / dx[1] -
                            (2. / 3.) * dt * c[2] / dx[2] + 0.5 * (2. / 3.) * dt * grav[d] * rs;
                    AT(outm[d], i, j, k, 0) = v;
                }
    }
}

/* ---- RK3stepStag ------------------------------------------------------------------------ */

typedef struct {
    fhdx_fab cup, cup2, cupm[3], cup2m[3];
    fhdx_fab sf[3], se[3], sc[3];
    fhdx_fab td, to[3];
} stag_scratch;

static void scratch_alloc(const fhdx_params* P, stag_scratch* S)
{
    const int g = P->ngc, nv = nvars_of(P);
    S->cup = mk_cell(P, g, nv);
    S->cup2 = mk_cell(P, g, nv);
    for (int d = 0; d < 3; ++d) {
        S->cupm[d] = mk_face(P, d, g, 1);
        S->cup2m[d] = mk_face(P, d, g, 1);
        S->sf[d] = mk_face(P, d, 0, nv);
        S->se[d] = mk_edge(P, d, 1);
        S->sc[d] = mk_cell(P, 1, 1);
        S->to[d] = mk_edge(P, d, 2);
    }
    S->td = mk_cell(P, 1, 6);
}

static void scratch_free(stag_scratch* S)
{
    free(S->cup.p);
    free(S->cup2.p);
    for (int d = 0; d < 3; ++d) {
        free(S->cupm[d].p);
        free(S->cup2m[d].p);
        free(S->sf[d].p);
        free(S->se[d].p);
        free(S->sc[d].p);
        free(S->to[d].p);
    }
    free(S->td.p);
}

/* Chemistry of a staggered step (timeStepStag.cpp:389-396 and the stage-2/3 forms): the
 * reactions, the caller's source (nvars comps) and ranchem (nreaction) and the reaction noise
 * sets A, B of the step. */
typedef struct {
    const fhdx_chemistry* chem;
    fhdx_fab source, ranchem, A, B;
} stag_chem;

static int rk3_step_stag(const fhdx_params* P, const stag_chem* ch, stag_scratch* S, fhdx_fab cu,
                         const fhdx_fab cumom[3],
                         fhdx_fab prim, const fhdx_fab vel[3], fhdx_fab eta, fhdx_fab zeta,
                         fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab ff[3],
                         const fhdx_fab ef[6], const fhdx_fab cf[3], const fhdx_fab faceA[3],
                         const fhdx_fab faceB[3], const fhdx_fab edgeA[3], const fhdx_fab edgeB[3],
                         const fhdx_fab cenA[3], const fhdx_fab cenB[3], double dt)
{
    const int nv = nvars_of(P);
    const double sq2 = sqrt(2.0), sq3 = sqrt(3.0);
    const double wts[3] = {(2.0 * sq2 + 1.0 * sq3) / 5.0, (-4.0 * sq2 + 3.0 * sq3) / 5.0,
                           (1.0 * sq2 - 2.0 * sq3) / 10.0};
    stag_args A;
    memset(&A, 0, sizeof(A));
    A.P = P;
    A.NS = P->nspecies;
    A.nv = nv;
    A.dt = dt;
    A.volinv = 1. / (P->dx[0] * P->dx[1] * P->dx[2]);
    A.dtinv = 1. / dt;
    A.prim = F(prim);
    A.eta = F(eta);
    A.zeta = F(zeta);
    A.kappa = F(kappa);
    A.chi = F(chi);
    A.D = F(D);
    for (int d = 0; d < 3; ++d) {
        A.vel[d] = F(vel[d]);
        A.ff[d] = F(ff[d]);
        A.cf[d] = F(cf[d]);
        A.sf[d] = F(S->sf[d]);
        A.se[d] = F(S->se[d]);
        A.sc[d] = F(S->sc[d]);
        A.to[d] = F(S->to[d]);
    }
    for (int e = 0; e < 6; ++e) A.ef[e] = F(ef[e]);
    A.td = F(S->td);
    fab_t u0 = F(cu), m0[3], m1[3], m2[3], vv[3];
    for (int d = 0; d < 3; ++d) {
        m0[d] = F(cumom[d]);
        m1[d] = F(S->cupm[d]);
        m2[d] = F(S->cup2m[d]);
        vv[d] = F(vel[d]);
    }
    fab_t c1 = F(S->cup), c2 = F(S->cup2);
    for (int st = 0; st < 3; ++st) {
        const double w = wts[st];
        for (int d = 0; d < 3; ++d) {
            lincomb(S->sf[d], faceA[d], faceB[d], 4, nv - 4, w);
            lincomb(S->se[d], edgeA[d], edgeB[d], 0, 1, w);
            lincomb(S->sc[d], cenA[d], cenB[d], 0, 1, w);
        }
        int rc = orc_transport(P, prim, eta, zeta, kappa, chi, D);
        if (rc != FHDX_OK) return rc;
        fab_t us = (st == 0) ? u0 : (st == 1) ? c1 : c2;
        fab_t* ms = (st == 0) ? m0 : (st == 1) ? m1 : m2;
        fab_t out = (st == 0) ? c1 : (st == 1) ? c2 : u0;
        fab_t* outm = (st == 0) ? m1 : (st == 1) ? m2 : m0;
        A.cu = us;
        for (int d = 0; d < 3; ++d) A.mom[d] = ms[d];
        flux_stag(&A);
        fab_t src;
        if (ch) { /* ranchem = 1*A + w*B, then the source from this stage's prim */
            lincomb(ch->ranchem, ch->A, ch->B, 0, ch->chem->nreaction, w);
            rc = orc_chem_source(P, ch->chem, prim, ch->source, ch->ranchem, dt);
            if (rc != FHDX_OK) return rc;
            src = F(ch->source);
        }
        stag_update(P, st, out, outm, u0, m0, us, ms, A.ff, A.ef, A.cf, ch ? &src : NULL, dt);
        stage_end(P, A.prim, vv, out, outm);
    }
    return FHDX_OK;
}

int orc_rk3_step_stag(const fhdx_params* P, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim,
                      const fhdx_fab vel[3], fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa,
                      fhdx_fab chi, fhdx_fab D, const fhdx_fab faceflux[3],
                      const fhdx_fab edgeflux[6], const fhdx_fab cenflux[3],
                      const fhdx_fab faceA[3], const fhdx_fab faceB[3], const fhdx_fab edgeA[3],
                      const fhdx_fab edgeB[3], const fhdx_fab cenA[3], const fhdx_fab cenB[3],
                      double dt)
{
    stag_scratch S;
    scratch_alloc(P, &S);
    int rc = rk3_step_stag(P, NULL, &S, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, faceflux, edgeflux,
                           cenflux, faceA, faceB, edgeA, edgeB, cenA, cenB, dt);
    scratch_free(&S);
    return rc;
}

int orc_flux_stag(const fhdx_params* P, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim,
                  const fhdx_fab vel[3], fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi,
                  fhdx_fab D, const fhdx_fab faceflux[3], const fhdx_fab edgeflux[6],
                  const fhdx_fab cenflux[3], const fhdx_fab stochface[3],
                  const fhdx_fab stochedge[3], const fhdx_fab stochcen[3], double dt)
{
    stag_args A;
    memset(&A, 0, sizeof(A));
    A.P = P;
    A.NS = P->nspecies;
    A.nv = nvars_of(P);
    A.dt = dt;
    A.volinv = 1. / (P->dx[0] * P->dx[1] * P->dx[2]);
    A.dtinv = 1. / dt;
    A.cu = F(cu);
    A.prim = F(prim);
    A.eta = F(eta);
    A.zeta = F(zeta);
    A.kappa = F(kappa);
    A.chi = F(chi);
    A.D = F(D);
    fhdx_fab td = mk_cell(P, 1, 6), to[3];
    for (int d = 0; d < 3; ++d) {
        to[d] = mk_edge(P, d, 2);
        A.mom[d] = F(cumom[d]);
        A.vel[d] = F(vel[d]);
        A.ff[d] = F(faceflux[d]);
        A.cf[d] = F(cenflux[d]);
        A.sf[d] = F(stochface[d]);
        A.se[d] = F(stochedge[d]);
        A.sc[d] = F(stochcen[d]);
        A.to[d] = F(to[d]);
    }
    for (int e = 0; e < 6; ++e) A.ef[e] = F(edgeflux[e]);
    A.td = F(td);
    flux_stag(&A);
    free(td.p);
    for (int d = 0; d < 3; ++d) free(to[d].p);
    return FHDX_OK;
}

static int advance_stag(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu,
                        const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3], int nsteps,
                        double dt, uint64_t seed, uint64_t step0, uint32_t instance)
{
    const int g = P->ngc, NS = P->nspecies, nv = nvars_of(P);
    fhdx_fab eta = mk_cell(P, g, 1), zeta = mk_cell(P, g, 1), kappa = mk_cell(P, g, 1);
    fhdx_fab chi = mk_cell(P, g, NS), D = mk_cell(P, g, NS * NS);
    fhdx_fab ff[3], ef[6], cf[3], fA[3], fB[3], eA[3], eB[3], cA[3], cB[3];
    for (int d = 0; d < 3; ++d) {
        ff[d] = mk_face(P, d, 0, nv + 4);
        cf[d] = mk_cell(P, 1, 1);
        fA[d] = mk_face(P, d, 0, nv);
        fB[d] = mk_face(P, d, 0, nv);
        eA[d] = mk_edge(P, d, 1);
        eB[d] = mk_edge(P, d, 1);
        cA[d] = mk_cell(P, 1, 1);
        cB[d] = mk_cell(P, 1, 1);
    }
    /* edgeflux_x[0] xy, [1] xz; _y[0] xy, [1] yz; _z[0] xz, [1] yz */
    const int et[6] = {0, 1, 0, 2, 1, 2};
    for (int e = 0; e < 6; ++e) ef[e] = mk_edge(P, et[e], 1);
    stag_scratch S;
    scratch_alloc(P, &S);
    stag_chem ch;
    memset(&ch, 0, sizeof(ch));
    if (chem) {
        const int nr = chem->nreaction;
        ch.chem = chem;
        ch.source = mk_cell(P, g, nv);
        ch.ranchem = mk_cell(P, g, nr);
        ch.A = mk_cell(P, g, nr);
        ch.B = mk_cell(P, g, nr);
    }
    int rc = FHDX_OK;
    for (int s = 0; s < nsteps && rc == FHDX_OK; ++s) {
        orc_fill_noise_stag(P, seed, step0 + (uint64_t)s, instance, fA, fB, eA, eB, cA, cB);
        if (chem) orc_fill_chem_noise(P, seed, step0 + (uint64_t)s, instance, ch.A, ch.B);
        rc = rk3_step_stag(P, chem ? &ch : NULL, &S, cu, cumom, prim, vel, eta, zeta, kappa, chi, D, ff, ef, cf,
                           fA, fB, eA, eB, cA, cB, dt);
    }
    scratch_free(&S);
    if (chem) {
        free(ch.source.p);
        free(ch.ranchem.p);
        free(ch.A.p);
        free(ch.B.p);
    }
    free(eta.p);
    free(zeta.p);
    free(kappa.p);
    free(chi.p);
    free(D.p);
    for (int d = 0; d < 3; ++d) {
        free(ff[d].p);
        free(cf[d].p);
        free(fA[d].p);
        free(fB[d].p);
        free(eA[d].p);
        free(eB[d].p);
        free(cA[d].p);
        free(cB[d].p);
    }
    for (int e = 0; e < 6; ++e) free(ef[e].p);
    return rc;
}

int orc_advance_stag(const fhdx_params* P, fhdx_fab cu, const fhdx_fab cumom[3], fhdx_fab prim,
                     const fhdx_fab vel[3], int nsteps, double dt, uint64_t seed, uint64_t step0,
                     uint32_t instance)
{
    return advance_stag(P, NULL, cu, cumom, prim, vel, nsteps, dt, seed, step0, instance);
}

int orc_advance_stag_chem(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu,
                          const fhdx_fab cumom[3], fhdx_fab prim, const fhdx_fab vel[3], int nsteps,
                          double dt, uint64_t seed, uint64_t step0, uint32_t instance)
{
    return advance_stag(P, chem, cu, cumom, prim, vel, nsteps, dt, seed, step0, instance);
}
