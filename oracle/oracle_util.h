/*
 * oracle_util.h -- helpers shared by the oracle's translation units (oracle.c, oracle_stag.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Host views of fhdx_fab arrays, the clamps of
 * amrex::max/min, the EOS helpers of compressible_functions.H and CholeskyDecomp.
 */
#ifndef FHDX_ORACLE_UTIL_H
#define FHDX_ORACLE_UTIL_H

#include "oracle.h"

#include <math.h>

#define MS FHDX_MAX_SPECIES

typedef struct {
    double* p;
    int lo0, lo1, lo2;
    int hi0, hi1, hi2;
    long sy, sz, sn;
    int nc;
} fab_t;

static inline fab_t F(fhdx_fab f)
{
    fab_t r;
    r.p = f.p;
    r.lo0 = f.lo[0]; r.lo1 = f.lo[1]; r.lo2 = f.lo[2];
    r.hi0 = f.hi[0]; r.hi1 = f.hi[1]; r.hi2 = f.hi[2];
    long nx = f.hi[0] - f.lo[0] + 1, ny = f.hi[1] - f.lo[1] + 1, nz = f.hi[2] - f.lo[2] + 1;
    r.sy = nx;
    r.sz = nx * ny;
    r.sn = nx * ny * nz;
    r.nc = f.ncomp;
    return r;
}

#define AT(f, i, j, k, n) \
    ((f).p[((long)(i) - (f).lo0) + ((long)(j) - (f).lo1) * (f).sy + \
           ((long)(k) - (f).lo2) * (f).sz + (long)(n) * (f).sn])

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
/* amrex::max(0., amrex::min(1., y)) */
static inline double clamp01(double y) { return dmax(0., dmin(1., y)); }

static inline int nvars_of(const fhdx_params* P) { return 5 + P->nspecies; }

/* ------------------------------------------------------------------------------------- */
/* EOS helpers, compressible_functions.H:138-199, 1497-1520                               */
/* ------------------------------------------------------------------------------------- */

static inline double get_temperature(const fhdx_params* P, double energy, const double* Yk)
{
    double cvmix = 0., e0mix = 0.;
    for (int n = 0; n < P->nspecies; ++n) {
        cvmix += Yk[n] * P->hcv[n];
        e0mix += Yk[n] * P->e0[n];
    }
    return (energy - e0mix) / cvmix;
}

static inline double get_pressure(const fhdx_params* P, const double* Yk, double rho, double T)
{
    double molmix = 0.;
    for (int n = 0; n < P->nspecies; ++n) molmix += Yk[n] / P->molmass[n];
    molmix = 1. / molmix;
    return rho * (P->Runiv / molmix) * T;
}

static inline void get_molfrac(const fhdx_params* P, const double* Yk, double* Xk)
{
    double molmix = 0.;
    for (int n = 0; n < P->nspecies; ++n) molmix += Yk[n] / P->molmass[n];
    molmix = 1. / molmix;
    for (int n = 0; n < P->nspecies; ++n) Xk[n] = Yk[n] * molmix / P->molmass[n];
}

static inline double get_energy(const fhdx_params* P, const double* Yk, double T)
{
    double cvmix = 0., e0mix = 0.;
    for (int n = 0; n < P->nspecies; ++n) {
        cvmix += Yk[n] * P->hcv[n];
        e0mix += Yk[n] * P->e0[n];
    }
    return e0mix + T * cvmix;
}

/* CholeskyDecomp (compressible_functions.H:460-509).  Only the lower-triangle entries and the
 * pivots feed the result; the reference's writes into the upper triangle never do. */
static inline void cholesky(int NS, double* a, double* sqda)
{
    double p[MS];
    for (int i = 0; i < NS; ++i) {
        int ising = 0;
        for (int j = i; j < NS; ++j) {
            double sum1 = a[i * NS + j];
            for (int k = i - 1; k >= 0; --k) sum1 = sum1 - a[i * NS + k] * a[j * NS + k];
            if (i == j) {
                if (sum1 <= 0.) {
                    p[i] = 0.;
                    ising = 1;
                } else {
                    p[i] = sqrt(sum1);
                }
            }
            if (ising == 0) a[j * NS + i] = sum1 / p[i];
            else a[j * NS + i] = 0.;
        }
    }
    for (int i = 0; i < NS * NS; ++i) sqda[i] = 0.;
    for (int i = 0; i < NS; ++i) {
        for (int j = i - 1; j >= 0; --j) sqda[i * NS + j] = a[i * NS + j];
        sqda[i * NS + i] = p[i];
    }
}

#endif
