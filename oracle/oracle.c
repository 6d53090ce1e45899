/*
 * oracle.c -- FP64 CPU restatement of FHDeX's compressible RK3 step (src_compressible).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * Each routine restates one reference routine and keeps its floating-point operation order
 * (left-to-right sums, the same parenthesisation, divisions where the reference divides),
 * because the species-noise Cholesky factors a matrix that is singular in exact arithmetic:
 * its last pivot is rounding noise, so any reordering shows up at ~1e-11 in rho*Y_k, j and
 * rho*E.  Build with -ffp-contract=off.  The wall branches (every is_*_dirichlet_mass flag of
 * flux.cpp:29-34, setBC, StochFlux, BCWallSpeciesFlux) and the periodic ones are both here.
 *
 * Functions and the reference code they follow:
 *   cons_to_prim     conservedPrimitiveConversions.cpp:48-89, compressible_functions.H:138-199
 *   transport        transCoeffs.cpp:37-88, compressible_functions.H:1153-1352 (GIO),
 *                    :647-778 (D_GIO1), :1357-1478 (HCB binary, transport_type 3),
 *                    :221-458 and :782-1150 (Valk-Waldmann, transport_type 2)
 *   stoch_face       flux.cpp:71-1145 (+ CholeskyDecomp compressible_functions.H:460-509)
 *   diff_face        flux.cpp:1193-1623
 *   corner_node      flux.cpp:1627-1713
 *   corner_face      flux.cpp:1770-1980
 *   hyp_face         flux.cpp:2016-2175 (advection_type 1), :2180-2372 (advection_type 2)
 *   rk3_step         timeStep.cpp:11-427 (LinComb of the noise sets at :118-128 etc.)
 *   fill_boundary    MultiFab::FillBoundary(periodicity) (periodic directions only)
 *   set_bc           boundary.cpp:172-1209; stoch_flux_walls :1392-1970; wall_species_flux
 *                    :1212-1390
 *   init_cons        compressible_functions.cpp:103-413 (prob_types 1-4, 100)
 *   noise            replaces MultiFabFillRandom.cpp:5-31 (Philox4x32-10 + Box-Muller)
 *   chem_source      src_chemistry/chemistry_functions.cpp:128-193
 *   evaluate_stats   src_compressible/stats.cpp:5-363
 *
 * How each is pinned is tested under tests/test_oracle*.py (Random123 known answers, closed
 * forms, invariants, textbook first approximations of the mixture transport).  Parity unpinned:
 * for three or more species, D_GIO1's mixture-averaged preconditioner changes D only within the
 * truncation of its three-term series (<= 2e-3 against the exact Stefan-Maxwell solution), so
 * that sub-expression is pinned to the truncation level, not to the bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_util.h"

/* ------------------------------------------------------------------------------------- */
/* conservedToPrimitive                                                                    */
/* ------------------------------------------------------------------------------------- */

static void c2p_cell(const fhdx_params* P, const double* c, double* pr)
{
    const int NS = P->nspecies;
    double Yk[MS], Yf[MS], Xk[MS];
    pr[0] = c[0];
    pr[1] = c[1] / c[0];
    pr[2] = c[2] / c[0];
    pr[3] = c[3] / c[0];
    double vsqr = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
    double intenergy = c[4] / c[0] - 0.5 * vsqr;
    double sumYk = 0.;
    for (int n = 0; n < NS; ++n) {
        Yk[n] = c[5 + n] / c[0];
        Yf[n] = clamp01(Yk[n]);
        sumYk += Yf[n];
    }
    for (int n = 0; n < NS; ++n) Yf[n] /= sumYk;
    pr[4] = get_temperature(P, intenergy, Yf);
    get_molfrac(P, Yk, Xk);
    for (int n = 0; n < NS; ++n) {
        pr[6 + n] = Yk[n];
        pr[6 + NS + n] = Xk[n];
    }
    pr[5] = get_pressure(P, Yk, pr[0], pr[4]);
}

int orc_cons_to_prim(const fhdx_params* P, fhdx_fab prim_, fhdx_fab cons_)
{
    fab_t prim = F(prim_), cons = F(cons_);
    const int nx = P->n_cells[0], ny = P->n_cells[1], nz = P->n_cells[2];
    const int nv = nvars_of(P), npv = 6 + 2 * P->nspecies;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double c[MS + 5], pr[2 * MS + 6];
                for (int n = 0; n < nv; ++n) c[n] = AT(cons, i, j, k, n);
                c2p_cell(P, c, pr);
                for (int n = 0; n < npv; ++n) AT(prim, i, j, k, n) = pr[n];
            }
    return FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* Transport: IdealMixtureTransportGIO + D_GIO1                                            */
/* ------------------------------------------------------------------------------------- */

/* Gaussian elimination without pivoting followed by back substitution, as written at
 * compressible_functions.H:1254-1271. */
static void gio_solve(int NS, double amat[MS][MS], double* Bi)
{
    for (int j = 0; j < NS - 1; ++j)
        for (int i = j + 1; i < NS; ++i) {
            double fact1 = amat[i][j] / amat[j][j];
            for (int k = 0; k < NS; ++k) amat[i][k] -= fact1 * amat[j][k];
            Bi[i] -= fact1 * Bi[j];
        }
    Bi[NS - 1] /= amat[NS - 1][NS - 1];
    for (int i = NS - 2; i > -1; --i) {
        double sum1 = 0.0;
        for (int j = NS - 1; j > i; --j) sum1 += amat[i][j] * Bi[j];
        Bi[i] = (Bi[i] - sum1) / amat[i][i];
    }
}

/* D_GIO1, compressible_functions.H:647-778; out[j*NS+i] = Diff_ij * Ykp_i */
static void d_gio1(int NS, double Dbin[MS][MS], const double* Ykp, const double* Xkp,
                   double* out)
{
    double Di[MS], Minv[MS], Mmat[MS];
    double Pm[MS][MS], Zm[MS][MS], Jm[MS][MS], PJ[MS][MS], m1[MS][MS], m2[MS][MS];
    double Diff[MS][MS];
    for (int i = 0; i < NS; ++i) {
        double term2 = 0.;
        for (int j = 0; j < NS; ++j)
            if (j != i) term2 = term2 + Xkp[j] / Dbin[i][j];
        Di[i] = (1. - Ykp[i]) / term2;
    }
    for (int i = 0; i < NS; ++i) {
        Mmat[i] = Xkp[i] / Di[i];
        Minv[i] = Di[i] / Xkp[i];
    }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            Pm[i][j] = -Ykp[j];
            if (i == j) Pm[i][j] = Pm[i][j] + 1.;
        }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            double Delta;
            if (i == j) {
                double term1 = 0.;
                for (int k = 0; k < NS; ++k)
                    if (k != i) term1 = term1 + Xkp[i] * Xkp[k] / Dbin[i][k];
                Delta = term1;
            } else {
                Delta = -Xkp[i] * Xkp[j] / Dbin[i][j];
            }
            Zm[i][j] = -Delta;
        }
    for (int i = 0; i < NS; ++i) Zm[i][i] = Zm[i][i] + Mmat[i];
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) Jm[i][j] = Minv[i] * Zm[i][j];
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            PJ[i][j] = 0.;
            for (int k = 0; k < NS; ++k) PJ[i][j] = PJ[i][j] + Pm[i][k] * Jm[k][j];
        }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            double scr = 0.;
            for (int k = 0; k < NS; ++k) scr = scr + Pm[i][k] * Minv[k] * Pm[j][k];
            m2[i][j] = scr;
            Diff[i][j] = scr;
        }
    for (int ll = 0; ll < 3; ++ll) {
        for (int i = 0; i < NS; ++i)
            for (int j = 0; j < NS; ++j) {
                double scr = 0.;
                for (int k = 0; k < NS; ++k) scr = scr + PJ[i][k] * Diff[k][j];
                m1[i][j] = scr + m2[i][j];
            }
        for (int i = 0; i < NS; ++i)
            for (int j = 0; j < NS; ++j) Diff[i][j] = m1[i][j];
    }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) out[j * NS + i] = Diff[i][j] * Ykp[i];
}

/* IdealMixtureTransportGIO (compressible_functions.H:1153-1352) followed by the rho scaling
 * (transCoeffs.cpp:76-80) and bulk viscosity (:84-86).  Returns 1 if a mass fraction is outside
 * (0,1) (the reference aborts at transCoeffs.cpp:45-46). */
/* IdealMixtureTransportHCBBin (compressible_functions.H:1357-1478): Hirschfelder-Curtiss-Bird
 * first approximations for a binary mixture of hard spheres with the correction factors 1.016
 * (viscosity), 1.025 (conductivity), 1.133 (diffusion) and 1.299 (thermal diffusion); Xk from
 * GetMolfrac of the clamped mass fractions (transCoeffs.cpp:60-61).  diff_ij[4] and chitil[2]
 * out (diff_ij before the density factor). */
static void hcb_bin(const fhdx_params* P, double density, double temperature, const double* Xk,
                    double* eta, double* kappa, double* zeta, double* diff_ij, double* chitil)
{
    const double pi = 3.1415926535897932;
    const double kB = P->k_B;
    double mm[2], xxtr[2], yytr[2];
    for (int n = 0; n < 2; ++n) mm[n] = P->molmass[n] * (kB / P->Runiv);
    for (int ii = 0; ii < 2; ++ii) {
        double sumXk = 0.;
        for (int jj = 0; jj < 2; ++jj) sumXk += Xk[jj];
        xxtr[ii] = Xk[ii] + 1.e-15 * (sumXk / 2.0 - Xk[ii]);
    }
    double MWmix = 0.;
    for (int ii = 0; ii < 2; ++ii) MWmix += xxtr[ii] * P->molmass[ii];
    for (int ii = 0; ii < 2; ++ii) yytr[ii] = P->molmass[ii] / MWmix * xxtr[ii];
    const double m0 = mm[0], m1 = mm[1];
    const double d0 = P->diameter[0], d1 = P->diameter[1];
    const double mx = 2.0 * m0 * m1 / (m0 + m1);
    const double dx = 0.5 * (d0 + d1);
    const double conc = xxtr[1], onc = xxtr[0];
    const double sm = (m0 + m1) * (m0 + m1), dm = (m0 - m1) * (m0 - m1);
    /* viscosity */
    const double eta0 = (5.0 / (16.0 * d0 * d0)) * sqrt(m0 * kB * temperature / pi);
    const double eta1 = (5.0 / (16.0 * d1 * d1)) * sqrt(m1 * kB * temperature / pi);
    const double etax = (5.0 / (16.0 * dx * dx)) * sqrt(mx * kB * temperature / pi);
    const double Xeta = (onc * onc / eta0) + (2.0 * conc * onc / etax) + (conc * conc / eta1);
    const double Yeta = (3.0 / 5.0) * ((onc * onc * m0) / (eta0 * m1) +
                                       (2.0 * conc * onc * etax * (m0 + m1) * (m0 + m1)) / (4.0 * eta0 * eta1 * m0 * m1) +
                                       (conc * conc * m1) / (eta1 * m0));
    const double Zeta = (3.0 / 5.0) * ((onc * onc * m0 / m1) +
                                       2.0 * conc * onc * ((sm / (4.0 * m0 * m1)) * ((etax / eta0) + (etax / eta1)) - 1.0) +
                                       conc * conc * m1 / m0);
    *eta = 1.016 * (1.0 + Zeta) / (Xeta + Yeta);
    /* thermal conductivity */
    const double kappa0 = 15.0 * kB * eta0 / (4.0 * m0);
    const double kappa1 = 15.0 * kB * eta1 / (4.0 * m1);
    const double kappax = 15.0 * kB * etax / (4.0 * mx);
    const double U0 = 4.0 / 15.0 - 17.0 * m0 / (60.0 * m1) + dm / (2.0 * m0 * m1);
    const double U1 = 4.0 / 15.0 - 17.0 * m1 / (60.0 * m0) + dm / (2.0 * m0 * m1);
    const double UY = 4.0 * (m0 + m1) * (m0 + m1) * kappax * kappax / (15.0 * 4.0 * m0 * m1 * kappa0 * kappa1) -
                      17.0 / 60.0 + 13.0 * (m0 - m1) * (m0 - m1) / (32.0 * m0 * m1);
    const double UZ = (4.0 / 15.0) * ((sm / (4.0 * m0 * m1)) * ((kappax / kappa0) + (kappax / kappa1)) - 1.0) - 17.0 / 60.0;
    const double Xkappa = onc * onc / kappa0 + 2.0 * conc * onc / kappax + conc * conc / kappa1;
    const double Ykappa = onc * onc * U0 / kappa0 + 2.0 * conc * onc * UY / kappax + conc * conc * U1 / kappa1;
    const double Zkappa = onc * onc * U0 + 2.0 * conc * onc * UZ + conc * conc * U1;
    *kappa = 1.025 * (1.0 + Zkappa) / (Xkappa + Ykappa);
    *zeta = 0.0;
    /* diffusion coefficient and thermal diffusion ratio */
    const double diff = 6.0 * 1.133 * etax / (5.0 * density);
    const double S0 = (m0 + m1) * kappax / (2.0 * m1 * kappa0) - 15 * (m1 - m0) / (8.0 * m0) - 1.0;
    const double S1 = (m0 + m1) * kappax / (2.0 * m0 * kappa1) - 15 * (m0 - m1) / (8.0 * m1) - 1.0;
    double thermal_diff = 1.299 * conc * onc * (S1 * conc - S0 * onc) / (6.0 * kappax * (Xkappa + Ykappa));
    const double mR = m0 / m1;
    const double d0t = mR * diff * yytr[1] * yytr[1] * yytr[1] / (yytr[0] * xxtr[1] * xxtr[1]);
    const double d1t = mR * diff * yytr[1] * yytr[1] / (xxtr[1] * xxtr[1]);
    diff_ij[0] = 1.0 * yytr[0] * d0t;
    diff_ij[1] = -1.0 * yytr[0] * d0t;
    diff_ij[2] = -1.0 * yytr[0] * d1t;
    diff_ij[3] = 1.0 * yytr[0] * d1t;
    thermal_diff = thermal_diff * mR * yytr[1] * yytr[1] / (xxtr[1] * xxtr[1]);
    chitil[0] = -1.0 * xxtr[1] * xxtr[1] * thermal_diff / (mR * xxtr[0] * yytr[1] * yytr[1]);
    chitil[1] = 1.0 * xxtr[1] * thermal_diff / (mR * yytr[1] * yytr[1]);
}

/* ------------------------------------------------------------------------------------- */
/* transport_type 2: IdealMixtureTransportVW (compressible_functions.H:1047-1150)           */
/* ------------------------------------------------------------------------------------- */
/* The reference's pow(a, 2.) is evaluated as a*a (the correctly rounded square, which libm's pow
 * returns) and pow(a, 3.) as (a*a)*a (within an ulp of libm's).  Matrices are flat with rows
 * `stride` apart, as the reference keeps them. */
static inline double vw_sq(double a) { return a * a; }
static inline double vw_cube(double a) { return (a * a) * a; }

/* Decomp and Decomp2 (H:221-320): LU factorisation with partial pivoting; ip holds 1-based row
 * numbers.  Returns 1 where the reference aborts ("singular matrix in decomp"). */
static int vw_decomp(int neq, int stride, double* a, int* ip)
{
    for (int n = 0; n < neq; ++n) ip[n] = n + 1;
    for (int i = 0; i < neq - 1; ++i) {
        int k = i + 1;
        double apm = fabs(a[(ip[i] - 1) * stride + i]);
        for (int j = i + 1; j < neq; ++j) {
            const double apn = fabs(a[(ip[j] - 1) * stride + i]);
            if (apm >= apn) continue;
            apm = apn;
            k = j + 1;
        }
        const int piv = ip[k - 1];
        ip[k - 1] = ip[i];
        ip[i] = piv;
        if (apm < 1.e-30) return 1;
        for (int l = i + 1; l < neq; ++l) {
            const int n = ip[l];
            const double amult = a[(n - 1) * stride + i] / a[(piv - 1) * stride + i];
            a[(n - 1) * stride + i] = amult;
            if (fabs(amult) < 1.e-60) continue;
            for (int c = i + 1; c < neq; ++c)
                a[(n - 1) * stride + c] = a[(n - 1) * stride + c] - amult * a[(piv - 1) * stride + c];
        }
    }
    return fabs(a[(ip[neq - 1] - 1) * stride + neq - 1]) < 1.e-30;
}

/* Solve and Solve2 (H:322-402): forward and back substitution through the permutation, then the
 * permutation undone in place cycle by cycle (ip is reset to the identity). */
static void vw_solve(int neq, int stride, const double* a, double* b, int* ip)
{
    for (int l = 1; l < neq; ++l) {
        const int n = ip[l];
        for (int k = 0; k < l; ++k) b[n - 1] = b[n - 1] - a[(n - 1) * stride + k] * b[ip[k] - 1];
    }
    b[ip[neq - 1] - 1] = b[ip[neq - 1] - 1] / a[(ip[neq - 1] - 1) * stride + neq - 1];
    for (int l = 0; l < neq - 1; ++l) {
        const int j = neq - (l + 1);
        const int n = ip[j - 1];
        for (int k = j; k < neq; ++k) b[n - 1] = b[n - 1] - a[(n - 1) * stride + k] * b[ip[k] - 1];
        b[n - 1] = b[n - 1] / a[(n - 1) * stride + j - 1];
    }
    for (int n = 0; n < neq; ++n)
        while (ip[n] != n + 1) {
            const int j = ip[n];
            const double scr = b[j - 1];
            ip[n] = ip[j - 1];
            b[j - 1] = b[ip[j - 1] - 1];
            b[ip[j - 1] - 1] = scr;
            ip[j - 1] = j;
        }
}

/* Decompnp / Solvenp (H:404-458): the same without pivoting. */
static int vw_decomp_np(int neq, int ndim, double* a)
{
    for (int i = 0; i < neq - 1; ++i)
        for (int l = i + 1; l < neq; ++l) {
            const double amult = a[l * ndim + i] / a[i * ndim + i];
            a[l * ndim + i] = amult;
            if (fabs(amult) < 1.e-60) continue;
            for (int k = i + 1; k < neq; ++k) a[l * ndim + k] = a[l * ndim + k] - amult * a[i * ndim + k];
        }
    return fabs(a[(neq - 1) * ndim + neq - 1]) < 1.e-30;
}

static void vw_solve_np(int neq, int ndim, const double* a, double* b)
{
    for (int l = 1; l < neq; ++l)
        for (int k = 0; k < l; ++k) b[l] = b[l] - a[l * ndim + k] * b[k];
    b[neq - 1] = b[neq - 1] / a[(neq - 1) * ndim + neq - 1];
    for (int l = 0; l < neq - 1; ++l) {
        const int j = neq - (l + 1);
        for (int k = j; k < neq; ++k) b[j - 1] = b[j - 1] - a[(j - 1) * ndim + k] * b[k];
        b[j - 1] = b[j - 1] / a[(j - 1) * ndim + j - 1];
    }
}

/* ViscLin (H:969-1045): the first Sonine approximation of the mixture viscosity, HCB 7.4-56,
 * 62, 63. */
static int vw_visc_lin(int NS, const double* om, const double* Ykp, double T, double rho, const double* mk,
                       double kB, double* eta)
{
    double nk[MS], bS[MS], diag[MS], Q[MS * MS];
    int ip[MS];
    for (int i = 0; i < NS; ++i) nk[i] = rho * Ykp[i] / mk[i];
    for (int i = 0; i < NS; ++i) {
        diag[i] = 0.;
        for (int k = 0; k < NS; ++k)
            diag[i] = diag[i] + nk[k] * mk[k] / vw_sq(mk[i] + mk[k]) *
                                    (5. * mk[i] * om[i * NS + k] + 3. * mk[k] * om[i * NS + k]);
    }
    for (int i = 0; i < NS; ++i) {
        for (int j = 0; j < NS; ++j) {
            double s1 = -nk[j] * mk[j] / vw_sq(mk[i] + mk[j]) * 2. * mk[j] * om[i * NS + j];
            if (i == j) s1 = s1 + diag[i];
            Q[i * NS + j] = -(16. / 15.) * (mk[i] / mk[j]) * s1;
        }
        bS[i] = -1.;
    }
    if (vw_decomp(NS, NS, Q, ip)) return 1;
    vw_solve(NS, NS, Q, bS, ip);
    double s = 0.;
    for (int i = 0; i < NS; ++i) s = s + nk[i] * bS[i];
    *eta = 0.5 * kB * T * s;
    return 0;
}

/* LambdaLin (H:782-912): the 2NS-term Sonine system for the thermal conductivity and the
 * thermal diffusion coefficients (HCB 7.4-50..54), lambda' (7.4-33) and lambda (7.4-65). */
static int vw_lambda_lin(int NS, const double* Dbin, const double* om, const double* Ykp, double T, double rho,
                         const double* mk, double kB, double* lam)
{
    const int S = 2 * NS;
    double nk[MS], beta[MS], d1[MS], d2[MS], d3[MS], d4[MS], DT[MS], QQ[4 * MS * MS], aS[2 * MS];
    int ip[2 * MS];
    for (int i = 0; i < NS; ++i) {
        nk[i] = rho * Ykp[i] / mk[i];
        d1[i] = d2[i] = d3[i] = d4[i] = 0.;
    }
    for (int i = 0; i < NS; ++i)
        for (int k = 0; k < NS; ++k) {
            const double o = om[i * NS + k], s = mk[i] + mk[k];
            d1[i] = d1[i] + nk[k] * mk[k] / s * nk[i] * mk[i] * o;
            if (k != i) d2[i] = d2[i] + nk[k] * mk[k] / s * o;
            d3[i] = d3[i] + nk[i] * nk[k] * vw_sq(mk[k]) / vw_sq(s) * .5 * o;
            d4[i] = d4[i] + nk[i] * nk[k] * mk[k] / vw_cube(s) *
                                ((5. / 4. * (6. * vw_sq(mk[i]) + 5. * vw_sq(mk[k])) * o - 3. * vw_sq(mk[k]) * o) +
                                 4. * mk[i] * mk[k] * o);
        }
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            const double o = om[i * NS + j], s = mk[i] + mk[j];
            double s00 = -nk[j] * mk[j] * d2[i] - nk[j] * mk[j] / s * nk[i] * mk[i] * o;
            double s01 = -nk[i] * nk[j] * vw_sq(mk[j]) / vw_sq(s) * .5 * o;
            /* the off-diagonal [11] term as the reference writes it: m_j throughout */
            double s11 = nk[i] * nk[j] * mk[j] / vw_cube(s) *
                         (-(5. / 4. * (6. * vw_sq(mk[j]) + 5. * vw_sq(mk[j])) * o - 3. * vw_sq(mk[j]) * o) +
                          4. * mk[j] * mk[j] * o);
            if (i == j) {
                s00 = s00 + d1[i];
                s01 = s01 + d3[i];
                s11 = s11 + d4[i];
            }
            const double ratm = mk[i] / mk[j], sqratm = sqrt(ratm), scr = ratm * sqratm;
            QQ[i * S + j] = 8. * sqratm / mk[i] * s00;
            QQ[i * S + j + NS] = -8. * scr * s01;
            QQ[(i + NS) * S + j] = -8. * sqratm * s01;
            QQ[(i + NS) * S + j + NS] = 8. * ratm * sqratm * s11;
        }
    for (int i = 0; i < NS; ++i) {
        aS[i] = 0.;
        beta[i] = sqrt(2. * kB * T / mk[i]);
        aS[i + NS] = -(15. / 4.) * nk[i] * beta[i];
    }
    if (vw_decomp(S, S, QQ, ip)) return 1;
    vw_solve(S, S, QQ, aS, ip);
    for (int i = 0; i < NS; ++i) DT[i] = 0.5 * nk[i] * beta[i] * mk[i] * aS[i];
    double s1 = 0.;
    for (int i = 0; i < NS; ++i) s1 = s1 + nk[i] * beta[i] * aS[NS + i];
    const double lp = -5. / 4. * kB * s1;
    s1 = 0.;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j)
            s1 = s1 + nk[i] * nk[j] / Dbin[i * NS + j] * vw_sq(DT[i] / (nk[i] * mk[i]) - DT[j] / (nk[j] * mk[j]));
    double ntotal = 0.;
    for (int i = 0; i < NS; ++i) ntotal += nk[i];
    *lam = lp - 0.5 * kB / ntotal * s1;
    return 0;
}

/* ThermalDiff (H:914-967): Valk's thermal diffusion ratios (kT into chitil). */
static int vw_thermal_diff(int NS, const double* a1, const double* a2, const double* ab, const double* Xkp,
                           double sqrtT, const double* mk, double* kT)
{
    double A[MS * MS], AA[MS], al[MS * MS];
    for (int i = 0; i < NS; ++i) {
        for (int j = 0; j < NS; ++j) A[i * NS + j] = Xkp[j] * a2[i * NS + j];
        AA[i] = 15. / 4.;
    }
    for (int i = 0; i < NS; ++i) {
        double s = 0.;
        for (int k = 0; k < NS; ++k) s = s + Xkp[k] * a1[i * NS + k];
        A[i * NS + i] = A[i * NS + i] + s;
    }
    if (vw_decomp_np(NS, NS, A)) return 1;
    vw_solve_np(NS, NS, A, AA);
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) al[i * NS + j] = ab[i * NS + j] * (AA[i] / mk[i] - AA[j] / mk[j]) / sqrtT;
    for (int k = 0; k < NS; ++k) {
        double s = 0.;
        for (int j = 0; j < NS; ++j) s = s + Xkp[j] * al[k * NS + j];
        kT[k] = s;
    }
    return 0;
}

/* IdealMixtureTransportVW (H:1047-1150): Xk from GetMolfrac of the clamped mass fractions
 * (transCoeffs.cpp:60-61); eta, kappa, zeta = 0, diff_ij (out[j*NS+i] before the density factor)
 * and chitil.  Returns 1 where one of the linear solves meets a singular matrix (the reference
 * aborts there). */
static int vw_transport(const fhdx_params* P, double density, double T, double pressure, const double* Xk,
                        double* eta, double* kappa, double* zeta, double* diff, double* chi)
{
    const int NS = P->nspecies;
    const double kB = P->k_B, pi = 3.1415926535897932;
    double mk[MS], xx[MS], yy[MS];
    double Dflat[MS * MS], om[MS * MS], a1[MS * MS], a2[MS * MS], ab[MS * MS], Dbin[MS][MS];
    for (int n = 0; n < NS; ++n) mk[n] = P->molmass[n] * (kB / P->Runiv);
    for (int ii = 0; ii < NS; ++ii) {
        double sumXk = 0.;
        for (int jj = 0; jj < NS; ++jj) sumXk += Xk[jj];
        xx[ii] = Xk[ii] + 1.e-15 * (sumXk / NS - Xk[ii]);
    }
    double MWmix = 0.;
    for (int ii = 0; ii < NS; ++ii) MWmix += xx[ii] * P->molmass[ii];
    for (int ii = 0; ii < NS; ++ii) yy[ii] = P->molmass[ii] / MWmix * xx[ii];
    const double Fijstar = -27. / 5.;
    const double sqrtT = sqrt(T);
    *eta = *kappa = 0.; /* what a singular solve leaves (the reference aborts there) */
    for (int n = 0; n < NS; ++n) chi[n] = 0.;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            const double diamat = 0.5 * (P->diameter[i] + P->diameter[j]);
            const double mi = mk[i], mj = mk[j];
            const double mu = mi * mj / (mi + mj);
            const double Fij = (6.0 * mi * mi + 13.0 / 5.0 * mj * mj + 16.0 / 5.0 * mi * mj) / vw_sq(mi + mj);
            const double s11bar = sqrt(kB / (2.0 * pi * mu)) * pi * vw_sq(diamat);
            ab[i * NS + j] = 8.0 / (3.0 * kB) * mu * mu * (-.5 * s11bar);
            Dflat[i * NS + j] = 3.0 / 16.0 * sqrt(2.0 * pi * vw_cube(kB) * (mi + mj) / mi / mj) /
                                (pi * vw_sq(diamat)) * T * sqrtT / pressure;
            Dbin[i][j] = Dflat[i * NS + j];
            om[i * NS + j] = sqrt(pi * kB / (2.0 * mu)) * vw_sq(diamat) * sqrtT;
            a1[i * NS + j] = (5.0 / (kB) * mi * mj / (mi + mj) * Fij * s11bar) / sqrtT;
            a2[i * NS + j] = (5.0 / (kB) * mi * mj * mi * mj / (vw_cube(mi + mj)) * Fijstar * s11bar) / sqrtT;
        }
    int bad = vw_visc_lin(NS, om, yy, T, density, mk, kB, eta);
    *zeta = 0.;
    bad |= vw_lambda_lin(NS, Dflat, om, yy, T, density, mk, kB, kappa);
    d_gio1(NS, Dbin, yy, xx, diff); /* D_GIO (H:513-646) is D_GIO1's arithmetic on a flat Dbin */
    bad |= vw_thermal_diff(NS, a1, a2, ab, xx, sqrtT, mk, chi);
    return bad;
}

static int transport_cell(const fhdx_params* P, const double* pr, double* eta, double* zeta,
                          double* kappa, double* chi, double* D)
{
    const int NS = P->nspecies;
    double Yk[MS];
    double sumYk = 0.;
    int bad = 0;
    for (int n = 0; n < NS; ++n) {
        double y = pr[6 + n];
        if (y <= 0.0) bad = 1;
        if (y >= 1.0) bad = 1;
        Yk[n] = clamp01(y);
        sumYk += Yk[n];
    }
    for (int n = 0; n < NS; ++n) Yk[n] /= sumYk;

    if (P->transport_type == 3) { /* Hirschfelder-Curtiss-Bird, binary (transCoeffs.cpp:71-75) */
        double Xf[MS];
        get_molfrac(P, Yk, Xf);
        hcb_bin(P, pr[0], pr[4], Xf, eta, kappa, zeta, D, chi);
        for (int n = 0; n < 4; ++n) D[n] *= pr[0];
        if (abs(P->visc_type) == 3) *zeta = P->zeta_ratio * *eta;
        return bad;
    }
    if (P->transport_type == 2) { /* Valk-Waldmann (transCoeffs.cpp:66-69) */
        double Xf[MS];
        get_molfrac(P, Yk, Xf);
        if (vw_transport(P, pr[0], pr[4], pr[5], Xf, eta, kappa, zeta, D, chi)) bad |= 2;
        for (int n = 0; n < NS * NS; ++n) D[n] *= pr[0];
        if (abs(P->visc_type) == 3) *zeta = P->zeta_ratio * *eta;
        return bad;
    }

    const double T = pr[4], p = pr[5];
    const double pi = 3.1415926535897932;
    double mm[MS], Xk[MS], Xkp[MS], Ykp[MS], eta1[MS], rhs[MS], Bi[MS];
    double mu[MS][MS], diam[MS][MS], Dbin[MS][MS], amat[MS][MS];
    for (int n = 0; n < NS; ++n) mm[n] = P->molmass[n] * (P->k_B / P->Runiv);
    double mbar = 0.0;
    for (int n = 0; n < NS; ++n) mbar += Yk[n] / mm[n];
    mbar = 1.0 / mbar;
    double Xksum = 0.0;
    for (int n = 0; n < NS; ++n) {
        Xk[n] = Yk[n] * mbar / mm[n];
        Xksum += Xk[n];
    }
    for (int n = 0; n < NS; ++n) Xkp[n] = Xk[n] + 1.0e-16 * (Xksum / (double)NS - Xk[n]);
    for (int n = 0; n < NS; ++n) Ykp[n] = Xkp[n] * mm[n] / mbar;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            diam[i][j] = 0.5 * (P->diameter[i] + P->diameter[j]);
            mu[i][j] = mm[i] * mm[j] / (mm[i] + mm[j]);
        }
    const double kT = P->k_B * T;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j)
            Dbin[i][j] = (3.0 / 16.0) * sqrt(2.0 * pi * kT * kT * kT / mu[i][j]) /
                         (p * pi * diam[i][j] * diam[i][j]);
    const double AKL = 1.0, BKL = 1.0, CKL = 1.0;

    /* viscosity */
    for (int i = 0; i < NS; ++i)
        eta1[i] = 5.0 / (16.0 * diam[i][i] * diam[i][i]) * sqrt(mm[i] * P->k_B * T / pi);
    for (int i = 0; i < NS; ++i) {
        rhs[i] = Xkp[i];
        Bi[i] = rhs[i];
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = Xkp[i] * Xkp[i] / eta1[i];
                for (int k = 0; k < NS; ++k)
                    if (k != i)
                        amat[i][i] += 2.0 * kT * Xkp[i] * Xkp[k] /
                                      (p * Dbin[i][k] * (mm[i] + mm[k])) *
                                      (1.0 + 0.6 * mm[k] / mm[i] * AKL);
            } else {
                amat[i][j] = 2.0 * kT * Xkp[i] * Xkp[j] / (p * Dbin[i][j] * (mm[i] + mm[j])) *
                             (0.6 * AKL - 1.0);
            }
        }
    }
    gio_solve(NS, amat, Bi);
    double e = 0.0;
    for (int i = 0; i < NS; ++i) e += Xkp[i] * Bi[i];
    *eta = e;
    *zeta = 0.;

    /* thermal conductivity */
    for (int i = 0; i < NS; ++i) {
        rhs[i] = 2.5 * Xkp[i];
        Bi[i] = rhs[i];
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = Xkp[i] * Xkp[i] / Dbin[i][j] * 2.0 * AKL;
                for (int k = 0; k < NS; ++k)
                    if (k != i)
                        amat[i][i] += Xkp[i] * Xkp[k] / Dbin[i][k] * mu[i][k] / (mm[i] + mm[k]) *
                                      (7.5 * mm[i] / mm[k] + 6.25 * mm[k] / mm[i] -
                                       3.0 * mm[k] / mm[i] * BKL + 4.0 * AKL);
            } else {
                amat[i][j] = -1.0 * Xkp[i] * Xkp[j] / Dbin[i][j] * mu[i][j] / (mm[i] + mm[j]) *
                             (55.0 / 4.0 - 3.0 * BKL - 4.0 * AKL);
            }
        }
    }
    gio_solve(NS, amat, Bi);
    double kap = 0.0;
    for (int i = 0; i < NS; ++i) kap += rhs[i] * Bi[i];
    kap *= p / T;
    *kappa = kap;

    /* thermal diffusion ratios */
    for (int i = 0; i < NS; ++i) {
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = 0.0;
                for (int k = 0; k < NS; ++k)
                    if (k != i)
                        amat[i][i] -= 0.5 * Xkp[k] / Dbin[i][k] * mm[k] / (mm[i] + mm[k]) *
                                      (6.0 * CKL - 5.0);
            } else {
                amat[i][j] =
                    0.5 * Xkp[j] / Dbin[i][j] * mm[i] / (mm[i] + mm[j]) * (6.0 * CKL - 5.0);
            }
        }
    }
    for (int i = 0; i < NS; ++i) {
        double c = 0.0;
        for (int j = 0; j < NS; ++j) c += amat[i][j] * Bi[j];
        chi[i] = c;
    }

    d_gio1(NS, Dbin, Ykp, Xkp, D);
    for (int kk = 0; kk < NS; ++kk)
        for (int ll = 0; ll < NS; ++ll) D[kk * NS + ll] *= pr[0];
    if (abs(P->visc_type) == 3) *zeta = P->zeta_ratio * *eta;
    return bad;
}

int orc_transport(const fhdx_params* P, fhdx_fab prim_, fhdx_fab eta_, fhdx_fab zeta_,
                  fhdx_fab kappa_, fhdx_fab chi_, fhdx_fab D_)
{
    fab_t prim = F(prim_), eta = F(eta_), zeta = F(zeta_), kappa = F(kappa_), chi = F(chi_),
          D = F(D_);
    const int NS = P->nspecies, npv = 6 + 2 * NS, g = P->ngc;
    const int nx = P->n_cells[0], ny = P->n_cells[1], nz = P->n_cells[2];
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int k = -g; k < nz + g; ++k)
        for (int j = -g; j < ny + g; ++j)
            for (int i = -g; i < nx + g; ++i) {
                double pr[2 * MS + 6], c[MS], d[MS * MS], e, z, kp;
                for (int n = 0; n < npv; ++n) pr[n] = AT(prim, i, j, k, n);
                bad |= transport_cell(P, pr, &e, &z, &kp, c, d);
                AT(eta, i, j, k, 0) = e;
                AT(zeta, i, j, k, 0) = z;
                AT(kappa, i, j, k, 0) = kp;
                for (int n = 0; n < NS; ++n) AT(chi, i, j, k, n) = c[n];
                for (int n = 0; n < NS * NS; ++n) AT(D, i, j, k, n) = d[n];
            }
    return (bad & 1) ? FHDX_ERR_MASSFRAC : (bad & 2) ? FHDX_ERR_SINGULAR : FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* Periodic ghost fill                                                                     */
/* ------------------------------------------------------------------------------------- */

static int wrap(int i, int n) { return ((i % n) + n) % n; }

/* A direction is periodic when its bc_vel is -1 (SetupBC, boundary.cpp:3-12). */
static int is_periodic(const fhdx_params* P, int d) { return P->bc_vel_lo[d] == -1; }
/* is_{lo,hi}_*_dirichlet_mass (flux.cpp:29-34): a wall face unless periodic or reservoir. */
static int dir_mass(int code) { return code != 3 && code != -1; }
static int wall_lo(const fhdx_params* P, int d) { return dir_mass(P->bc_mass_lo[d]); }
static int wall_hi(const fhdx_params* P, int d) { return dir_mass(P->bc_mass_hi[d]); }

/* FillBoundary(periodicity): a ghost cell whose periodic image is a valid cell takes that
 * image; ghost cells outside a wall keep what they hold (setBC writes them). */
int orc_fill_boundary(const fhdx_params* P, fhdx_fab mf_)
{
    fab_t f = F(mf_);
    const int n[3] = {P->n_cells[0], P->n_cells[1], P->n_cells[2]};
    const int per[3] = {is_periodic(P, 0), is_periodic(P, 1), is_periodic(P, 2)};
#pragma omp parallel for schedule(static)
    for (int k = f.lo2; k <= f.hi2; ++k)
        for (int j = f.lo1; j <= f.hi1; ++j)
            for (int i = f.lo0; i <= f.hi0; ++i) {
                const int c[3] = {i, j, k};
                int v[3], inside = 1;
                for (int d = 0; d < 3; ++d) {
                    if (per[d]) v[d] = wrap(c[d], n[d]);
                    else {
                        v[d] = c[d];
                        if (c[d] < 0 || c[d] >= n[d]) inside = 0;
                    }
                }
                if (!inside || (v[0] == i && v[1] == j && v[2] == k)) continue;
                for (int q = 0; q < f.nc; ++q) AT(f, i, j, k, q) = AT(f, v[0], v[1], v[2], q);
            }
    return FHDX_OK;
}

/* GetDensity (compressible_functions.H:1481-1495). */
static double get_density(const fhdx_params* P, double pres, double T, const double* Y)
{
    double molmix = 0.;
    for (int n = 0; n < P->nspecies; ++n) molmix += Y[n] / P->molmass[n];
    molmix = 1. / molmix;
    return pres / (P->Runiv / molmix) / T;
}

/* One ghost cell of setBC's face (d, side) (boundary.cpp:180-1036; mass-fraction branches
 * :185-209, thermal :230-254, slip :283-339, no slip :340-378 for lo x and their mirror images
 * for the other faces).  The mirror of ghost index c is -c-1 at a lo face and 2n-c-1 at a hi
 * face.  bc_mass 3 (reservoir) is accepted on x faces only (the parameter check). */
static void set_bc_cell(const fhdx_params* P, fab_t* prim, fab_t* cons, int d, int side, int i,
                        int j, int k)
{
    const int NS = P->nspecies, np = 6 + 2 * NS;
    const int n = P->n_cells[d], f = 2 * d + side;
    int c[3] = {i, j, k}, m[3] = {i, j, k};
    m[d] = side ? 2 * (n - 1) - c[d] + 1 : -c[d] - 1;
    const int bm = side ? P->bc_mass_hi[d] : P->bc_mass_lo[d];
    const int bt = side ? P->bc_therm_hi[d] : P->bc_therm_lo[d];
    const int bv = side ? P->bc_vel_hi[d] : P->bc_vel_lo[d];
    const double tw = side ? P->t_hi[d] : P->t_lo[d];
#define PR(a, q) AT(*prim, a[0], a[1], a[2], q)
#define CO(a, q) AT(*cons, a[0], a[1], a[2], q)
    if (bm == 1 && P->algorithm_type == 2)
        for (int q = 6; q < np; ++q) PR(c, q) = PR(m, q);
    if (bm == 2 && P->algorithm_type == 2)
        for (int s = 0; s < NS; ++s) {
            PR(c, 6 + s) = P->bc_Yk[f][s];
            PR(c, 6 + NS + s) = P->bc_Xk[f][s];
        }
    if (bm == 3 && P->algorithm_type == 2) { /* x reservoir, boundary.cpp:221-238 and 260-288 */
        const double rw = side ? P->rho_hi[d] : P->rho_lo[d], pw = side ? P->p_hi[d] : P->p_lo[d];
        for (int s = 0; s < NS; ++s) {
            PR(c, 6 + s) = P->bc_Yk[f][s];
            PR(c, 6 + NS + s) = P->bc_Xk[f][s];
            PR(c, 0) = rw;
            CO(c, 0) = rw;
            PR(c, 4) = tw;
            PR(c, 5) = pw;
        }
        /* the momentum block of a reservoir: gas at rest, energy from rho and T */
        for (int q = 1; q < 4; ++q) {
            CO(c, q) = 0.0;
            PR(c, q) = 0.0;
        }
        double Yr[MS];
        for (int s = 0; s < NS; ++s) Yr[s] = PR(c, 6 + s);
        const double temp = PR(c, 4), rho = PR(c, 0);
        const double ie = get_energy(P, Yr, temp);
        for (int s = 0; s < NS; ++s) CO(c, 5 + s) = rho * PR(c, 6 + s);
        CO(c, 4) = rho * ie + 0.5 * rho * (PR(c, 1) * PR(c, 1) + PR(c, 2) * PR(c, 2) + PR(c, 3) * PR(c, 3));
        return;
    }
    if (bt == 1) {
        PR(c, 4) = PR(m, 4);
        PR(c, 5) = PR(m, 5);
    } else if (bt == 2) {
        PR(c, 4) = tw;
        PR(c, 5) = PR(m, 5);
    }
    if (bv != 1 && bv != 2) return;
    if (bv == 1) { /* slip: normal velocity 0, tangential mirrored (cons and prim) */
        for (int q = 1; q < 4; ++q) {
            CO(c, q) = (q == 1 + d) ? 0.0 : CO(m, q);
            PR(c, q) = (q == 1 + d) ? 0.0 : PR(m, q);
        }
    } else {
        PR(c, 1) = 0.0;
        PR(c, 2) = 0.0;
        PR(c, 3) = 0.0;
    }
    double Y[MS];
    for (int s = 0; s < NS; ++s) Y[s] = PR(c, 6 + s);
    const double T = PR(c, 4), pt = PR(c, 5);
    const double rho = get_density(P, pt, T, Y);
    const double ie = get_energy(P, Y, T);
    PR(c, 0) = rho;
    CO(c, 0) = rho;
    if (P->algorithm_type == 2)
        for (int s = 0; s < NS; ++s) CO(c, 5 + s) = rho * PR(c, 6 + s);
    if (bv == 2)
        for (int q = 1; q < 4; ++q) CO(c, q) = rho * PR(c, q);
    CO(c, 4) = rho * ie + 0.5 * rho * (PR(c, 1) * PR(c, 1) + PR(c, 2) * PR(c, 2) + PR(c, 3) * PR(c, 3));
#undef PR
#undef CO
}

/* setBC (boundary.cpp:172-1209): faces lo x, hi x, lo y, hi y, lo z, hi z in turn, each over
 * every cell of the grown box outside that face; a periodic direction is skipped. */
int orc_set_bc(const fhdx_params* P, fhdx_fab prim_, fhdx_fab cons_)
{
    fab_t prim = F(prim_), cons = F(cons_);
    for (int d = 0; d < 3; ++d) {
        if (is_periodic(P, d)) continue;
        const int n = P->n_cells[d];
        for (int side = 0; side < 2; ++side) {
#pragma omp parallel for schedule(static)
            for (int k = prim.lo2; k <= prim.hi2; ++k)
                for (int j = prim.lo1; j <= prim.hi1; ++j)
                    for (int i = prim.lo0; i <= prim.hi0; ++i) {
                        const int c = (d == 0) ? i : (d == 1) ? j : k;
                        if (side == 0 ? c >= 0 : c < n) continue;
                        set_bc_cell(P, &prim, &cons, d, side, i, j, k);
                    }
        }
    }
    return FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* calculateFlux                                                                           */
/* ------------------------------------------------------------------------------------- */

/* Offsets relative to the face index (i,j,k) of the 8 cells summed, in the reference's
 * order, for the four edge viscosities of a face (pp, mp, pm, mm), x/y/z faces:
 * flux.cpp:117-135 (x), :481-499 (y), :847-865 (z).  Each entry: eta offset then T offset
 * (they differ only in the z-face pm quirk of flux.cpp:857-858). */
static const int MU_ORDER[3][4][8][6] = {
    { /* x */
      {{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,1,0,0,1,0},{-1,1,0,-1,1,0},{0,0,1,0,0,1},{-1,0,1,-1,0,1},{0,1,1,0,1,1},{-1,1,1,-1,1,1}},
      {{0,-1,0,0,-1,0},{-1,-1,0,-1,-1,0},{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,-1,1,0,-1,1},{-1,-1,1,-1,-1,1},{0,0,1,0,0,1},{-1,0,1,-1,0,1}},
      {{0,0,-1,0,0,-1},{-1,0,-1,-1,0,-1},{0,1,-1,0,1,-1},{-1,1,-1,-1,1,-1},{0,0,0,0,0,0},{-1,0,0,-1,0,0},{0,1,0,0,1,0},{-1,1,0,-1,1,0}},
      {{0,-1,-1,0,-1,-1},{-1,-1,-1,-1,-1,-1},{0,0,-1,0,0,-1},{-1,0,-1,-1,0,-1},{0,-1,0,0,-1,0},{-1,-1,0,-1,-1,0},{0,0,0,0,0,0},{-1,0,0,-1,0,0}} },
    { /* y */
      {{1,-1,0,1,-1,0},{0,-1,0,0,-1,0},{1,0,0,1,0,0},{0,0,0,0,0,0},{1,-1,1,1,-1,1},{0,-1,1,0,-1,1},{1,0,1,1,0,1},{0,0,1,0,0,1}},
      {{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0},{-1,0,1,-1,0,1},{0,0,1,0,0,1},{-1,-1,1,-1,-1,1},{0,-1,1,0,-1,1}},
      {{1,-1,-1,1,-1,-1},{0,-1,-1,0,-1,-1},{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,-1,0,1,-1,0},{0,-1,0,0,-1,0},{1,0,0,1,0,0},{0,0,0,0,0,0}},
      {{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1},{-1,-1,-1,-1,-1,-1},{0,-1,-1,0,-1,-1},{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0}} },
    { /* z */
      {{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,1,-1,1,1,-1},{0,1,-1,0,1,-1},{1,0,0,1,0,0},{0,0,0,0,0,0},{1,1,0,1,1,0},{0,1,0,0,1,0}},
      {{-1,1,-1,-1,1,-1},{0,1,-1,0,1,-1},{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1},{-1,1,0,-1,1,0},{0,1,0,0,1,0},{-1,0,0,-1,0,0},{0,0,0,0,0,0}},
      {{1,0,0,1,0,0},{0,0,-2,0,0,0},{1,-1,0,1,-1,0},{0,-1,-2,0,-1,0},{1,0,-1,1,0,-1},{0,0,-1,0,0,-1},{1,-1,-1,1,-1,-1},{0,-1,-1,0,-1,-1}},
      {{-1,-1,0,-1,-1,0},{0,-1,0,0,-1,0},{-1,0,0,-1,0,0},{0,0,0,0,0,0},{-1,-1,-1,-1,-1,-1},{0,-1,-1,0,-1,-1},{-1,0,-1,-1,0,-1},{0,0,-1,0,0,-1}} }
};
/* node (rancorn) offsets of the pp, mp, pm, mm edges: flux.cpp:277-280, 644-647, 1011-1014 */
static const int RC_NODE[3][4][3] = {
    {{0,1,1},{0,0,1},{0,1,0},{0,0,0}},
    {{1,0,1},{0,0,1},{1,0,0},{0,0,0}},
    {{1,1,0},{0,1,0},{1,0,0},{0,0,0}}
};

typedef struct {
    const fhdx_params* P;
    fab_t cons, prim, eta, zeta, kappa, chi, D;
    fab_t flux[3], ran[3], corn[3][3], visccorn, rancorn;
    double dt;
    int nv, NS;
} flux_args;

/* Cells of the four edge viscosities at a wall face, all in the ghost layer (normal offset -1
 * at a lo face, 0 at a hi face), in the reference's summation order: transverse offsets per
 * face direction and edge (pp, mp, pm, mm), flux.cpp:137-156 (x), 501-520 (y), 867-886 (z). */
static const int MU_WALL[3][4][4][3] = {
    { {{0,0,0},{0,1,0},{0,0,1},{0,1,1}}, {{0,-1,0},{0,0,0},{0,-1,1},{0,0,1}},
      {{0,0,-1},{0,1,-1},{0,0,0},{0,1,0}}, {{0,-1,-1},{0,0,-1},{0,-1,0},{0,0,0}} },
    { {{1,0,0},{0,0,0},{1,0,1},{0,0,1}}, {{-1,0,0},{0,0,0},{-1,0,1},{0,0,1}},
      {{1,0,-1},{0,0,-1},{1,0,0},{0,0,0}}, {{-1,0,-1},{0,0,-1},{-1,0,0},{0,0,0}} },
    { {{1,0,0},{0,0,0},{1,1,0},{0,1,0}}, {{-1,1,0},{0,1,0},{-1,0,0},{0,0,0}},
      {{1,0,0},{0,0,0},{1,-1,0},{0,-1,0}}, {{-1,-1,0},{0,-1,0},{-1,0,0},{0,0,0}} }
};

/* The two transverse axes of a face direction: x faces (y, z), y faces (x, z), z faces (x, y). */
static void transverse_axes(int d, int* a, int* b)
{
    *a = (d == 0) ? 1 : 0;
    *b = (d == 2) ? 1 : 2;
}

/* Stochastic flux of one face (flux.cpp:71-1145), with the wall branches of a bc_mass 1/2 face
 * (:88-97, :137-202 for x and their y/z images) and the tangential-wall factors of the
 * divergence noise (:250-280). */
static void stoch_face(const flux_args* A, int d, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const int il = i - di, jl = j - dj, kl = k - dk; /* lower cell; upper cell is (i,j,k) */
    const fab_t *eta = &A->eta, *zeta = &A->zeta, *prim = &A->prim, *kappa = &A->kappa;
    const fab_t *ran = &A->ran[d], *rc = &A->rancorn;
    fab_t* fl = (fab_t*)&A->flux[d];
    const double volinv = 1. / (P->dx[0] * P->dx[1] * P->dx[2]);
    const double dtinv = 1. / A->dt;
    const int idx[3] = {i, j, k};
    const int lo_w = idx[d] == 0 && wall_lo(P, d);
    const int hi_w = idx[d] == P->n_cells[d] && wall_hi(P, d);

    double Tu = AT(*prim, i, j, k, 4), Tl = AT(*prim, il, jl, kl, 4);
    double mup = AT(*eta, i, j, k, 0) * Tu + AT(*eta, il, jl, kl, 0) * Tl;
    double kp = AT(*kappa, i, j, k, 0) * Tu * Tu + AT(*kappa, il, jl, kl, 0) * Tl * Tl;
    double meanT = 0.5 * (Tu + Tl);
    if (lo_w) {
        mup = 2.0 * AT(*eta, il, jl, kl, 0) * Tl;
        kp = 2.0 * AT(*kappa, il, jl, kl, 0) * Tl * Tl;
        meanT = Tl;
    }
    if (hi_w) {
        mup = 2.0 * AT(*eta, i, j, k, 0) * Tu;
        kp = 2.0 * AT(*kappa, i, j, k, 0) * Tu * Tu;
        meanT = Tu;
    }
    double fw[MS + 5], w[MS + 5];
    fw[0] = 0;
    fw[1] = sqrt(P->k_B * mup * volinv * dtinv);
    fw[2] = fw[1];
    fw[3] = fw[1];
    fw[4] = sqrt(P->k_B * kp * volinv * dtinv);
    for (int n = 0; n < 5; ++n) w[n] = fw[n] * AT(*ran, i, j, k, n);
    double nweight = sqrt(P->k_B * volinv * dtinv);

    double mus[4];
    for (int c = 0; c < 4; ++c) {
        if (lo_w || hi_w) {
            const int off = lo_w ? -1 : 0;
            double s = 0., z = 0.;
            for (int t = 0; t < 4; ++t) {
                const int* o = MU_WALL[d][c][t];
                const int ci = i + o[0] + off * di, cj = j + o[1] + off * dj, ck = k + o[2] + off * dk;
                double e = AT(*eta, ci, cj, ck, 0) * AT(*prim, ci, cj, ck, 4);
                double zz = AT(*zeta, ci, cj, ck, 0) * AT(*prim, ci, cj, ck, 4);
                s = (t == 0) ? e : s + e;
                z = (t == 0) ? zz : z + zz;
            }
            mus[c] = 0.5 * s / 3.;
            if (abs(P->visc_type) == 3) /* x walls do not divide by 3, y and z walls do */
                mus[c] += (d == 0) ? 0.5 * z : 0.5 * z / 3.;
            continue;
        }
        const int(*o)[6] = MU_ORDER[d][c];
        double s = AT(*eta, i + o[0][0], j + o[0][1], k + o[0][2], 0) *
                   AT(*prim, i + o[0][3], j + o[0][4], k + o[0][5], 4);
        for (int t = 1; t < 8; ++t)
            s = s + AT(*eta, i + o[t][0], j + o[t][1], k + o[t][2], 0) *
                        AT(*prim, i + o[t][3], j + o[t][4], k + o[t][5], 4);
        mus[c] = 0.25 * s / 3.;
        if (abs(P->visc_type) == 3) {
            double z = AT(*zeta, i + o[0][0], j + o[0][1], k + o[0][2], 0) *
                       AT(*prim, i + o[0][3], j + o[0][4], k + o[0][5], 4);
            for (int t = 1; t < 8; ++t)
                z = z + AT(*zeta, i + o[t][0], j + o[t][1], k + o[t][2], 0) *
                            AT(*prim, i + o[t][3], j + o[t][4], k + o[t][5], 4);
            mus[c] += 0.25 * z;
        }
    }
    /* tangential walls: sqrt(2) (slip) or 0 (no slip) on the edges that touch them */
    int ta, tb;
    transverse_axes(d, &ta, &tb);
    double f_lo[2] = {1., 1.}, f_hi[2] = {1., 1.};
    const int ax[2] = {ta, tb};
    for (int q = 0; q < 2; ++q) {
        const int t = ax[q];
        if ((P->bc_vel_lo[t] == 1 || P->bc_vel_lo[t] == 2) && idx[t] == 0)
            f_lo[q] = (P->bc_vel_lo[t] == 1) ? sqrt(2.0) : 0.;
        if ((P->bc_vel_hi[t] == 1 || P->bc_vel_hi[t] == 2) && idx[t] == P->n_cells[t] - 1)
            f_hi[q] = (P->bc_vel_hi[t] == 1) ? sqrt(2.0) : 0.;
    }
    const double fac[4] = {f_hi[0] * f_hi[1], f_lo[0] * f_hi[1], f_hi[0] * f_lo[1], f_lo[0] * f_lo[1]};
    double sum = 0.;
    for (int c = 0; c < 4; ++c) {
        const int* q = RC_NODE[d][c];
        double term = fac[c] * sqrt(mus[c]) * AT(*rc, i + q[0], j + q[1], k + q[2], 0);
        sum = (c == 0) ? term : sum + term;
    }
    w[1 + d] = w[1 + d] + 0.25 * nweight * sum;

    for (int n = 1; n < 4; ++n) AT(*fl, i, j, k, n) = AT(*fl, i, j, k, n) + w[n];
    AT(*fl, i, j, k, nv) = AT(*fl, i, j, k, nv) + w[4];

    /* viscous heating: the normal component first, the two tangential ones after */
    const int t1 = (d == 0) ? 2 : 1, t2 = (d == 2) ? 2 : 3;
    double phidiag = w[1 + d] * (AT(*prim, il, jl, kl, 1 + d) + AT(*prim, i, j, k, 1 + d));
    double phishear = w[t1] * (AT(*prim, il, jl, kl, t1) + AT(*prim, i, j, k, t1)) +
                      w[t2] * (AT(*prim, il, jl, kl, t2) + AT(*prim, i, j, k, t2));
    if (lo_w || hi_w) {
        const int wi = lo_w ? il : i, wj = lo_w ? jl : j, wk = lo_w ? kl : k;
        phidiag = 2.0 * w[1 + d] * AT(*prim, wi, wj, wk, 1 + d);
        phishear = 2.0 * w[t1] * AT(*prim, wi, wj, wk, t1) + 2.0 * w[t2] * AT(*prim, wi, wj, wk, t2);
    }
    phidiag = -0.5 * phidiag;
    phishear = -0.5 * phishear;
    AT(*fl, i, j, k, nv + 1) = AT(*fl, i, j, k, nv + 1) - phidiag;
    AT(*fl, i, j, k, nv + 2) = AT(*fl, i, j, k, nv + 2) - phishear;

    if (P->algorithm_type == 2) {
        /* at a wall both sides read the ghost cell (lo) or the boundary cell (hi) */
        const int Li = hi_w ? i : il, Lj = hi_w ? j : jl, Lk = hi_w ? k : kl;
        const int Ui = lo_w ? il : i, Uj = lo_w ? jl : j, Uk = lo_w ? kl : k;
        double yy[MS], yyp[MS], DijY[MS * MS], sqD[MS * MS], hk[MS];
        for (int n = 5; n < 5 + NS; ++n) w[n] = 0.;
        for (int s = 0; s < NS; ++s) {
            yy[s] = clamp01(AT(*prim, Li, Lj, Lk, 6 + s));
            yyp[s] = clamp01(AT(*prim, Ui, Uj, Uk, 6 + s));
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
                DijY[s * NS + l] = 0.5 * (AT(A->D, Li, Lj, Lk, l * NS + s) * yy[l] +
                                          AT(A->D, Ui, Uj, Uk, l * NS + s) * yyp[l] +
                                          (AT(A->D, Li, Lj, Lk, s * NS + l) * yy[s] +
                                           AT(A->D, Ui, Uj, Uk, s * NS + l) * yyp[s]));
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
                fw[5 + l] = sqrt(P->k_B * MWmix * volinv / (P->Runiv * A->dt)) * sqD[s * NS + l];
                w[5 + s] = w[5 + s] + fw[5 + l] * AT(*ran, i, j, k, 5 + l);
            }
            AT(*fl, i, j, k, 5 + s) = w[5 + s];
        }
        for (int s = 0; s < NS; ++s) hk[s] = P->e0[s] + P->hcp[s] * meanT;
        double soret = 0.;
        for (int s = 0; s < NS; ++s) {
            double ss;
            if (lo_w || hi_w) /* flux.cpp:422-427 */
                ss = (hk[s] + P->Runiv * meanT / P->molmass[s] * AT(A->chi, Li, Lj, Lk, s)) * w[5 + s];
            else
                ss = (hk[s] + P->Runiv * meanT / P->molmass[s] * 0.5 *
                                  (AT(A->chi, il, jl, kl, s) + AT(A->chi, i, j, k, s))) *
                     w[5 + s];
            soret = soret + ss;
        }
        AT(*fl, i, j, k, nv + 3) = AT(*fl, i, j, k, nv + 3) + soret;
    }
}

/* Diffusive flux of one face (flux.cpp:1193-1623), with the one-sided branches of a bc_mass
 * 1/2 wall face. */
static void diff_face(const flux_args* A, int d, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    const int il = i - (d == 0), jl = j - (d == 1), kl = k - (d == 2);
    const fab_t *prim = &A->prim, *eta = &A->eta, *kappa = &A->kappa, *chi = &A->chi;
    fab_t* fl = (fab_t*)&A->flux[d];
    const double half = 0.5, h = P->dx[d];
    const int idx[3] = {i, j, k};
    const int lo_w = idx[d] == 0 && wall_lo(P, d);
    const int hi_w = idx[d] == P->n_cells[d] && wall_hi(P, d);
    const int wall = lo_w || hi_w;
    /* the wall-side cell: the ghost at a lo face, the boundary cell at a hi face */
    const int wi = lo_w ? il : i, wj = lo_w ? jl : j, wk = lo_w ? kl : k;
#define PU(n) AT(*prim, i, j, k, n)
#define PL(n) AT(*prim, il, jl, kl, n)
#define PW(n) AT(*prim, wi, wj, wk, n)
    double mup = half * (AT(*eta, i, j, k, 0) + AT(*eta, il, jl, kl, 0));
    double kp = half * (AT(*kappa, i, j, k, 0) + AT(*kappa, il, jl, kl, 0));
    double tau[4];
    for (int c = 1; c < 4; ++c) tau[c] = mup * (PU(c) - PL(c)) / h;
    double divp = 0.;
    double s1 = PL(1) + PU(1), s2 = PL(2) + PU(2), s3 = PL(3) + PU(3);
    double phiflx;
    if (d == 0) phiflx = tau[1] * s1 + divp * s1 + tau[2] * s2 + tau[3] * s3;
    else if (d == 1) phiflx = tau[1] * s1 + tau[2] * s2 + divp * s2 + tau[3] * s3;
    else phiflx = tau[1] * s1 + tau[2] * s2 + tau[3] * s3 + divp * s3;
    double Qflux = kp * (PU(4) - PL(4)) / h;
    if (wall) { /* flux.cpp:1218-1241 (x), 1361-1384 (y), 1506-1529 (z) */
        mup = AT(*eta, wi, wj, wk, 0);
        kp = AT(*kappa, wi, wj, wk, 0);
        for (int c = 1; c < 4; ++c) tau[c] = mup * (PU(c) - PL(c)) / (0.5 * h);
        if (d == 0) phiflx = 2.0 * (tau[1] * PW(1) + divp * PW(1) + tau[2] * PW(2) + tau[3] * PW(3));
        else if (d == 1) phiflx = 2.0 * (tau[1] * PW(1) + tau[2] * PW(2) + divp * PW(2) + tau[3] * PW(3));
        else phiflx = 2.0 * (tau[1] * PW(1) + tau[2] * PW(2) + tau[3] * PW(3) + divp * PW(3));
        Qflux = kp * (PU(4) - PL(4)) / (0.5 * h);
    }
    for (int c = 1; c < 4; ++c) {
        double t = (c == 1 + d) ? (tau[c] + divp) : tau[c];
        AT(*fl, i, j, k, c) = AT(*fl, i, j, k, c) - t;
    }
    AT(*fl, i, j, k, nv) = AT(*fl, i, j, k, nv) - Qflux;
    AT(*fl, i, j, k, nv + 1) = AT(*fl, i, j, k, nv + 1) - (half * phiflx);

    double meanT = 0.5 * (PL(4) + PU(4));
    double meanP = 0.5 * (PL(5) + PU(5));
    if (wall) {
        meanT = PW(4);
        meanP = PW(5);
    }
    if (P->algorithm_type == 2) {
        double dk[MS], soret[MS], Fk[MS], hk[MS];
        for (int s = 0; s < NS; ++s) {
            double term1 = (PU(6 + NS + s) - PL(6 + NS + s)) / h;
            double mX = 0.5 * (PL(6 + NS + s) + PU(6 + NS + s));
            double mY = 0.5 * (PL(6 + s) + PU(6 + s));
            double term2 = (mX - mY) * (PU(5) - PL(5)) / h / meanP;
            dk[s] = term1 + term2;
            double ChiX;
            if (d == 2) /* flux.cpp:1561 reads chi at k and k+1 */
                ChiX = 0.5 * (AT(*chi, i, j, k, s) * PL(6 + NS + s) +
                              AT(*chi, i, j, k + 1, s) * PU(6 + NS + s));
            else
                ChiX = 0.5 * (AT(*chi, il, jl, kl, s) * PL(6 + NS + s) +
                              AT(*chi, i, j, k, s) * PU(6 + NS + s));
            soret[s] = ChiX * (PU(4) - PL(4)) / h / meanT;
            if (wall) { /* flux.cpp:1271-1290: half-cell gradients, wall-side values */
                term1 = (PU(6 + NS + s) - PL(6 + NS + s)) / (0.5 * h);
                mX = PW(6 + NS + s);
                mY = PW(6 + s);
                term2 = (mX - mY) * (PU(5) - PL(5)) / (0.5 * h) / meanP;
                dk[s] = term1 + term2;
                ChiX = AT(*chi, wi, wj, wk, s) * PW(6 + NS + s);
                soret[s] = ChiX * (PU(4) - PL(4)) / (0.5 * h) / meanT;
            }
        }
        for (int kk = 0; kk < NS; ++kk) {
            Fk[kk] = 0.;
            for (int ll = 0; ll < NS; ++ll) {
                double Fks = half * (AT(A->D, il, jl, kl, ll * NS + kk) +
                                     AT(A->D, i, j, k, ll * NS + kk)) *
                             (dk[ll] + soret[ll]);
                if (wall) Fks = AT(A->D, wi, wj, wk, ll * NS + kk) * (dk[ll] + soret[ll]);
                Fk[kk] = Fk[kk] - Fks;
            }
        }
        for (int s = 0; s < NS; ++s) hk[s] = P->e0[s] + P->hcp[s] * meanT;
        double Q5 = 0.;
        for (int s = 0; s < NS; ++s) {
            double csum = (d == 2) /* flux.cpp:1604 sums chi(i,j,k) twice */
                              ? AT(*chi, i, j, k, s) + AT(*chi, i, j, k, s)
                              : AT(*chi, il, jl, kl, s) + AT(*chi, i, j, k, s);
            double Q5s = (hk[s] + 0.5 * P->Runiv * meanT * csum / P->molmass[s]) * Fk[s];
            if (wall) Q5s = (hk[s] + P->Runiv * meanT * AT(*chi, wi, wj, wk, s) / P->molmass[s]) * Fk[s];
            Q5 = Q5 + Q5s;
        }
        AT(*fl, i, j, k, nv + 3) = AT(*fl, i, j, k, nv + 3) + Q5;
        for (int s = 0; s < NS; ++s) AT(*fl, i, j, k, 5 + s) = AT(*fl, i, j, k, 5 + s) + Fk[s];
    }
#undef PU
#undef PL
#undef PW
}

/* Corner (nodal) viscous terms at node (i,j,k) (flux.cpp:1627-1713). */
static void corner_node(const flux_args* A, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const fab_t *eta = &A->eta, *zeta = &A->zeta, *pr = &A->prim;
#define E(a, b, c) AT(*eta, a, b, c, 0)
#define Z(a, b, c) AT(*zeta, a, b, c, 0)
#define V(a, b, c, n) AT(*pr, a, b, c, n)
    double muxp = 0.125 * (E(i, j - 1, k - 1) + E(i - 1, j - 1, k - 1) + E(i, j, k - 1) +
                           E(i - 1, j, k - 1) + E(i, j - 1, k) + E(i - 1, j - 1, k) + E(i, j, k) +
                           E(i - 1, j, k));
    double zetaxp;
    if (abs(P->visc_type) == 3)
        zetaxp = 0.125 * (Z(i, j - 1, k - 1) + Z(i - 1, j - 1, k - 1) + Z(i, j, k - 1) +
                          Z(i - 1, j, k - 1) + Z(i, j - 1, k) + Z(i - 1, j - 1, k) + Z(i, j, k) +
                          Z(i - 1, j, k));
    else
        zetaxp = 0.;
    double DXa[3] = {P->dx[0], P->dx[1], P->dx[2]};
    /* wall nodes (flux.cpp:1647-1682): half spacing across the wall and the four ghost-layer
     * cells for the viscosities; the checks run in turn, so the last wall met sets mu */
    {
        const int nd[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                if (!(side ? (nd[d] == P->n_cells[d] && wall_hi(P, d)) : (nd[d] == 0 && wall_lo(P, d))))
                    continue;
                DXa[d] = 0.5 * P->dx[d];
                /* the other two axes, in (i, j, k) order; ghost layer index g on axis d */
                const int g = side ? nd[d] : nd[d] - 1;
                const int a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
                double e4 = 0., z4 = 0.;
                for (int t = 0; t < 4; ++t) {
                    int c[3];
                    c[d] = g;
                    c[a] = nd[a] - 1 + (t >> 1);
                    c[b] = nd[b] - 1 + (t & 1);
                    e4 = (t == 0) ? E(c[0], c[1], c[2]) : e4 + E(c[0], c[1], c[2]);
                    z4 = (t == 0) ? Z(c[0], c[1], c[2]) : z4 + Z(c[0], c[1], c[2]);
                }
                muxp = 0.25 * e4;
                zetaxp = (abs(P->visc_type) == 3) ? 0.25 * z4 : 0.;
            }
    }
    const double DX0 = DXa[0], DX1 = DXa[1], DX2 = DXa[2];
    double gx[4], gy[4], gz[4]; /* index 1..3: u, v, w */
    for (int n = 1; n < 4; ++n) {
        gx[n] = 0.25 * muxp *
                (V(i, j - 1, k - 1, n) - V(i - 1, j - 1, k - 1, n) + V(i, j, k - 1, n) -
                 V(i - 1, j, k - 1, n) + V(i, j - 1, k, n) - V(i - 1, j - 1, k, n) + V(i, j, k, n) -
                 V(i - 1, j, k, n)) /
                DX0;
        gy[n] = 0.25 * muxp *
                (V(i - 1, j, k - 1, n) - V(i - 1, j - 1, k - 1, n) + V(i, j, k - 1, n) -
                 V(i, j - 1, k - 1, n) + V(i - 1, j, k, n) - V(i - 1, j - 1, k, n) + V(i, j, k, n) -
                 V(i, j - 1, k, n)) /
                DX1;
        gz[n] = 0.25 * muxp *
                (V(i - 1, j - 1, k, n) - V(i - 1, j - 1, k - 1, n) + V(i, j - 1, k, n) -
                 V(i, j - 1, k - 1, n) + V(i - 1, j, k, n) - V(i - 1, j, k - 1, n) + V(i, j, k, n) -
                 V(i, j, k - 1, n)) /
                DX2;
    }
    /* corn[a][b]: a = derivative direction x/y/z, b = component u/v/w (cornx_in[b] holds
     * d/dx of component b, flux.cpp:1173-1181) */
    for (int n = 1; n < 4; ++n) {
        AT(A->corn[0][n - 1], i, j, k, 0) = gx[n];
        AT(A->corn[1][n - 1], i, j, k, 0) = gy[n];
        AT(A->corn[2][n - 1], i, j, k, 0) = gz[n];
    }
    AT(A->visccorn, i, j, k, 0) =
        (muxp / 12. + zetaxp / 4.) *
        ((V(i, j - 1, k - 1, 1) - V(i - 1, j - 1, k - 1, 1)) / DX0 +
         (V(i, j, k - 1, 1) - V(i - 1, j, k - 1, 1)) / DX0 +
         (V(i, j - 1, k, 1) - V(i - 1, j - 1, k, 1)) / DX0 + (V(i, j, k, 1) - V(i - 1, j, k, 1)) / DX0 +
         (V(i - 1, j, k - 1, 2) - V(i - 1, j - 1, k - 1, 2)) / DX1 +
         (V(i, j, k - 1, 2) - V(i, j - 1, k - 1, 2)) / DX1 +
         (V(i - 1, j, k, 2) - V(i - 1, j - 1, k, 2)) / DX1 + (V(i, j, k, 2) - V(i, j - 1, k, 2)) / DX1 +
         (V(i - 1, j - 1, k, 3) - V(i - 1, j - 1, k - 1, 3)) / DX2 +
         (V(i, j - 1, k, 3) - V(i, j - 1, k - 1, 3)) / DX2 +
         (V(i - 1, j, k, 3) - V(i - 1, j, k - 1, 3)) / DX2 + (V(i, j, k, 3) - V(i, j, k - 1, 3)) / DX2);
#undef E
#undef Z
#undef V
}

/* Corner-to-face terms (flux.cpp:1770-1980).  Node offsets (transverse a,b) summed in the
 * reference order: visccorn always (hi,hi),(lo,hi),(hi,lo),(lo,lo); the corner arrays in the
 * same order on x/y faces and (hi,hi),(hi,lo),(lo,hi),(lo,lo) on z faces. */
static void corner_face(const flux_args* A, int d, int i, int j, int k)
{
    const int nv = A->nv;
    const fab_t* pr = &A->prim;
    fab_t* fl = (fab_t*)&A->flux[d];
    const int il = i - (d == 0), jl = j - (d == 1), kl = k - (d == 2);
    /* transverse unit offsets */
    int ta[3], tb[3];
    if (d == 0) { ta[0] = 0; ta[1] = 1; ta[2] = 0; tb[0] = 0; tb[1] = 0; tb[2] = 1; }
    else if (d == 1) { ta[0] = 1; ta[1] = 0; ta[2] = 0; tb[0] = 0; tb[1] = 0; tb[2] = 1; }
    else { ta[0] = 1; ta[1] = 0; ta[2] = 0; tb[0] = 0; tb[1] = 1; tb[2] = 0; }
    static const int vc_ord[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
    static const int zc_ord[4][2] = {{1, 1}, {1, 0}, {0, 1}, {0, 0}};
    const int(*cord)[2] = (d == 2) ? zc_ord : vc_ord;
#define NODE(f, o, q) AT(f, i + (o)[q][0] * ta[0] + (o)[q][1] * tb[0], \
                         j + (o)[q][0] * ta[1] + (o)[q][1] * tb[1], \
                         k + (o)[q][0] * ta[2] + (o)[q][1] * tb[2], 0)
    double vc4 = NODE(A->visccorn, vc_ord, 0);
    for (int q = 1; q < 4; ++q) vc4 = vc4 + NODE(A->visccorn, vc_ord, q);
    /* normal momentum: transverse divergence pair; x: (vy, wz); y: (ux, wz); z: (vy, ux) */
    const fab_t *c1, *c2, *ta1, *ta2;
    int tc1, tc2; /* tangential momentum components */
    if (d == 0) { c1 = &A->corn[1][1]; c2 = &A->corn[2][2]; ta1 = &A->corn[1][0]; ta2 = &A->corn[2][0]; tc1 = 2; tc2 = 3; }
    else if (d == 1) { c1 = &A->corn[0][0]; c2 = &A->corn[2][2]; ta1 = &A->corn[0][1]; ta2 = &A->corn[2][1]; tc1 = 1; tc2 = 3; }
    else { c1 = &A->corn[1][1]; c2 = &A->corn[0][0]; ta1 = &A->corn[0][2]; ta2 = &A->corn[1][2]; tc1 = 1; tc2 = 2; }
    double div8 = NODE(*c1, cord, 0);
    for (int q = 1; q < 4; ++q) div8 = div8 + NODE(*c1, cord, q);
    for (int q = 0; q < 4; ++q) div8 = div8 + NODE(*c2, cord, q);
    double s1 = NODE(*ta1, cord, 0), s2 = NODE(*ta2, cord, 0);
    for (int q = 1; q < 4; ++q) {
        s1 = s1 + NODE(*ta1, cord, q);
        s2 = s2 + NODE(*ta2, cord, q);
    }
#undef NODE
    const int cn = 1 + d;
    AT(*fl, i, j, k, cn) = AT(*fl, i, j, k, cn) - 0.25 * vc4;
    AT(*fl, i, j, k, cn) = AT(*fl, i, j, k, cn) + .25 * div8;
    AT(*fl, i, j, k, tc1) = AT(*fl, i, j, k, tc1) - .25 * s1;
    AT(*fl, i, j, k, tc2) = AT(*fl, i, j, k, tc2) - .25 * s2;
    double phiflx = 0.25 * (vc4 - div8) * (AT(*pr, il, jl, kl, cn) + AT(*pr, i, j, k, cn));
    phiflx = phiflx + .25 * s1 * (AT(*pr, il, jl, kl, tc1) + AT(*pr, i, j, k, tc1));
    phiflx = phiflx + .25 * s2 * (AT(*pr, il, jl, kl, tc2) + AT(*pr, i, j, k, tc2));
    const int idx[3] = {i, j, k};
    const int lo_w = idx[d] == 0 && wall_lo(A->P, d);
    const int hi_w = idx[d] == A->P->n_cells[d] && wall_hi(A->P, d);
    if (lo_w || hi_w) { /* flux.cpp:1791-1822: the wall-side velocities only */
        const int wi = lo_w ? il : i, wj = lo_w ? jl : j, wk = lo_w ? kl : k;
        phiflx = 0.5 * (vc4 - div8) * AT(*pr, wi, wj, wk, cn);
        phiflx = phiflx + .5 * s1 * AT(*pr, wi, wj, wk, tc1);
        phiflx = phiflx + .5 * s2 * AT(*pr, wi, wj, wk, tc2);
    }
    AT(*fl, i, j, k, nv + 1) = AT(*fl, i, j, k, nv + 1) - 0.5 * phiflx;
}

/* Hyperbolic flux of one face (flux.cpp:2016-2372), with the wall-face branches. */
static void hyp_face(const flux_args* A, int d, int i, int j, int k)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    const int ex = (d == 0), ey = (d == 1), ez = (d == 2);
    fab_t* fl = (fab_t*)&A->flux[d];
    const double wgt2 = 1. / 12.;
    const double wgt1 = 0.5 + wgt2;
    /* boundary interpolation weights (flux.cpp:1991-1997) */
    const double wgta = -0.2, wgtb = 0.75, wgtc = 0.5, wgtd = -0.05;
    const int idx[3] = {i, j, k}, n = P->n_cells[d];
    const int wl = wall_lo(P, d), wh = wall_hi(P, d);
    double cv[MS + 5], pv[MS + 6], Yk[MS];
    if (P->advection_type == 1) {
        for (int l = 0; l < NS + 6; ++l)
            pv[l] = wgt1 * (AT(A->prim, i, j, k, l) + AT(A->prim, i - ex, j - ey, k - ez, l)) -
                    wgt2 * (AT(A->prim, i - 2 * ex, j - 2 * ey, k - 2 * ez, l) +
                            AT(A->prim, i + ex, j + ey, k + ez, l));
        if (idx[d] == 0 && wl) /* flux.cpp:2028-2037: the wall face takes the ghost value */
            for (int l = 0; l < NS + 6; ++l) pv[l] = AT(A->prim, i - ex, j - ey, k - ez, l);
        if (idx[d] == n && wh)
            for (int l = 0; l < NS + 6; ++l) pv[l] = AT(A->prim, i, j, k, l);
        double temp = pv[4], rho = pv[0];
        cv[0] = rho;
        for (int n = 0; n < NS; ++n) Yk[n] = pv[6 + n];
        double ie = get_energy(P, Yk, temp);
        double vsqr = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
        cv[4] = rho * ie + 0.5 * rho * vsqr;
    } else {
        for (int l = 0; l < NS + 5; ++l)
            cv[l] = wgt1 * (AT(A->cons, i, j, k, l) + AT(A->cons, i - ex, j - ey, k - ez, l)) -
                    wgt2 * (AT(A->cons, i - 2 * ex, j - 2 * ey, k - 2 * ez, l) +
                            AT(A->cons, i + ex, j + ey, k + ez, l));
        /* flux.cpp:2190-2206: the wall face takes the ghost value, the next face in a one-sided
         * four-point interpolation */
#define CS(o, l) AT(A->cons, i + (o) * ex, j + (o) * ey, k + (o) * ez, l)
        if (idx[d] == 0 && wl)
            for (int l = 0; l < NS + 5; ++l) cv[l] = CS(-1, l);
        else if (idx[d] == 1 && wl)
            for (int l = 0; l < NS + 5; ++l)
                cv[l] = wgta * CS(-2, l) + wgtb * CS(-1, l) + wgtc * CS(0, l) + wgtd * CS(1, l);
        if (idx[d] == n && wh)
            for (int l = 0; l < NS + 5; ++l) cv[l] = CS(0, l);
        else if (idx[d] == n - 1 && wh)
            for (int l = 0; l < NS + 5; ++l)
                cv[l] = wgta * CS(1, l) + wgtb * CS(0, l) + wgtc * CS(-1, l) + wgtd * CS(-2, l);
#undef CS
        for (int l = 1; l < 4; ++l) pv[l] = cv[l] / cv[0];
        for (int n = 0; n < NS; ++n) Yk[n] = cv[5 + n] / cv[0];
        double vsqr = pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3];
        double ie = cv[4] / cv[0] - 0.5 * vsqr;
        pv[4] = get_temperature(P, ie, Yk);
        pv[5] = get_pressure(P, Yk, cv[0], pv[4]);
    }
    const double un = pv[1 + d];
    AT(*fl, i, j, k, 0) += cv[0] * un;
    if (d == 0) {
        AT(*fl, i, j, k, 1) += cv[0] * (pv[1] * pv[1]) + pv[5];
        AT(*fl, i, j, k, 2) += cv[0] * pv[1] * pv[2];
        AT(*fl, i, j, k, 3) += cv[0] * pv[1] * pv[3];
    } else if (d == 1) {
        AT(*fl, i, j, k, 1) += cv[0] * pv[1] * pv[2];
        AT(*fl, i, j, k, 2) += cv[0] * pv[2] * pv[2] + pv[5];
        AT(*fl, i, j, k, 3) += cv[0] * pv[3] * pv[2];
    } else {
        AT(*fl, i, j, k, 1) += cv[0] * pv[1] * pv[3];
        AT(*fl, i, j, k, 2) += cv[0] * pv[2] * pv[3];
        AT(*fl, i, j, k, 3) += cv[0] * pv[3] * pv[3] + pv[5];
    }
    AT(*fl, i, j, k, 4) += un * cv[4] + pv[5] * un;
    AT(*fl, i, j, k, 4) += AT(*fl, i, j, k, nv) + AT(*fl, i, j, k, nv + 1) +
                           AT(*fl, i, j, k, nv + 2) + AT(*fl, i, j, k, nv + 3);
    if (P->algorithm_type == 2) {
        if (P->advection_type == 1)
            for (int n = 0; n < NS; ++n) AT(*fl, i, j, k, 5 + n) += cv[0] * pv[6 + n] * un;
        else
            for (int n = 0; n < NS; ++n) AT(*fl, i, j, k, 5 + n) += cv[5 + n] * un;
    }
}

/* Visits every face of the wall plane (d, side): index 0 (lo) or n_d (hi) along d. */
#define FOR_WALL_FACES(P, d, side, BODY)                                                     \
    do {                                                                                     \
        const int n0_ = (P)->n_cells[0], n1_ = (P)->n_cells[1], n2_ = (P)->n_cells[2];       \
        const int fix_ = (side) ? (P)->n_cells[d] : 0;                                       \
        for (int k = ((d) == 2 ? fix_ : 0); k <= ((d) == 2 ? fix_ : n2_ - 1); ++k)           \
            for (int j = ((d) == 1 ? fix_ : 0); j <= ((d) == 1 ? fix_ : n1_ - 1); ++j)       \
                for (int i = ((d) == 0 ? fix_ : 0); i <= ((d) == 0 ? fix_ : n0_ - 1); ++i) { \
                    BODY;                                                                    \
                }                                                                            \
    } while (0)

/* StochFlux (boundary.cpp:1392-1970): on wall faces the stochastic fluxes are scaled.  Mass
 * walls: species and Dufour by 0 (bc_mass 1) or sqrt(2) (bc_mass 2); thermal walls: the heat
 * flux by 0 (adiabatic) or sqrt(2) (isothermal); velocity walls: normal momentum and the
 * diagonal viscous heating by sqrt(2), tangential momentum and the shear heating by 0 (slip)
 * or sqrt(2) (no slip). */
static void stoch_flux_walls(const flux_args* A)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            const int bm = side ? P->bc_mass_hi[d] : P->bc_mass_lo[d];
            if (bm != 1 && bm != 2) continue;
            const double factor = (bm == 1) ? 0. : sqrt(2.);
            fab_t* fl = (fab_t*)&A->flux[d];
            FOR_WALL_FACES(P, d, side, {
                for (int n = 0; n < NS; ++n) AT(*fl, i, j, k, n + 5) *= factor;
                AT(*fl, i, j, k, nv + 3) *= factor;
            });
        }
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            const int bt = side ? P->bc_therm_hi[d] : P->bc_therm_lo[d];
            if (bt != 1 && bt != 2) continue;
            double factor = (bt == 1) ? 0. : sqrt(2.);
            /* a reservoir leaves the heat flux alone, written for x faces only (:1593, :1620) */
            if (d == 0 && (side ? P->bc_mass_hi[0] : P->bc_mass_lo[0]) == 3) factor = 1.0;
            fab_t* fl = (fab_t*)&A->flux[d];
            FOR_WALL_FACES(P, d, side, { AT(*fl, i, j, k, nv) *= factor; });
        }
    const double sqrtTwo = sqrt(2.);
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            const int bv = side ? P->bc_vel_hi[d] : P->bc_vel_lo[d];
            if (bv != 1 && bv != 2) continue;
            const double factor = (bv == 1) ? 0. : sqrt(2.);
            fab_t* fl = (fab_t*)&A->flux[d];
            FOR_WALL_FACES(P, d, side, {
                for (int n = 0; n < 3; ++n) {
                    if (n == d) {
                        AT(*fl, i, j, k, 1 + n) *= sqrtTwo;
                        AT(*fl, i, j, k, nv + 1) *= sqrtTwo;
                        AT(*fl, i, j, k, nv + 2) *= factor;
                    } else {
                        AT(*fl, i, j, k, 1 + n) *= factor;
                    }
                }
            });
        }
}

/* BCWallSpeciesFlux (boundary.cpp:1212-1390): species, density and Dufour fluxes are zero on
 * bc_mass 1 faces. */
static void wall_species_flux(const flux_args* A)
{
    const fhdx_params* P = A->P;
    const int NS = A->NS, nv = A->nv;
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            if ((side ? P->bc_mass_hi[d] : P->bc_mass_lo[d]) != 1) continue;
            fab_t* fl = (fab_t*)&A->flux[d];
            FOR_WALL_FACES(P, d, side, {
                for (int n = 0; n < NS; ++n) AT(*fl, i, j, k, n + 5) = 0.;
                AT(*fl, i, j, k, 0) = 0.;
                AT(*fl, i, j, k, nv + 3) = 0.;
            });
        }
}

int orc_calculate_flux(const fhdx_params* P, fhdx_fab cons, fhdx_fab prim, fhdx_fab eta,
                       fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi, fhdx_fab D,
                       const fhdx_fab flux[3], const fhdx_fab stochFlux[3],
                       const fhdx_fab cornx[3], const fhdx_fab corny[3],
                       const fhdx_fab cornz[3], fhdx_fab visccorn, fhdx_fab rancorn,
                       double dt)
{
    flux_args A;
    A.P = P;
    A.cons = F(cons); A.prim = F(prim); A.eta = F(eta); A.zeta = F(zeta);
    A.kappa = F(kappa); A.chi = F(chi); A.D = F(D);
    for (int d = 0; d < 3; ++d) {
        A.flux[d] = F(flux[d]);
        A.ran[d] = F(stochFlux[d]);
        A.corn[0][d] = F(cornx[d]);
        A.corn[1][d] = F(corny[d]);
        A.corn[2][d] = F(cornz[d]);
    }
    A.visccorn = F(visccorn);
    A.rancorn = F(rancorn);
    A.dt = dt;
    A.NS = P->nspecies;
    A.nv = nvars_of(P);
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const int nc = A.nv + 4;

    /* flux_in[d].setVal(0) (flux.cpp:24-26) */
    for (int d = 0; d < 3; ++d) {
        long npts = A.flux[d].sn * nc;
        memset(A.flux[d].p, 0, sizeof(double) * npts);
    }
    /* each pass visits faces in order; faces are independent, so threads split over k */
    if (P->stoch_stress_form == 1)
        for (int d = 0; d < 3; ++d) {
            const int kx = n0 + (d == 0), ky = n1 + (d == 1), kz = n2 + (d == 2);
#pragma omp parallel for schedule(static)
            for (int k = 0; k < kz; ++k)
                for (int j = 0; j < ky; ++j)
                    for (int i = 0; i < kx; ++i) stoch_face(&A, d, i, j, k);
        }
    if (P->stoch_stress_form == 1) stoch_flux_walls(&A); /* flux.cpp:1149 */
    for (int d = 0; d < 3; ++d) {
        const int kx = n0 + (d == 0), ky = n1 + (d == 1), kz = n2 + (d == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < kz; ++k)
            for (int j = 0; j < ky; ++j)
                for (int i = 0; i < kx; ++i) diff_face(&A, d, i, j, k);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= n2; ++k)
        for (int j = 0; j <= n1; ++j)
            for (int i = 0; i <= n0; ++i) corner_node(&A, i, j, k);
    for (int d = 0; d < 3; ++d) {
        const int kx = n0 + (d == 0), ky = n1 + (d == 1), kz = n2 + (d == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < kz; ++k)
            for (int j = 0; j < ky; ++j)
                for (int i = 0; i < kx; ++i) corner_face(&A, d, i, j, k);
    }
    wall_species_flux(&A); /* flux.cpp:1985 */
    if (P->advection_type == 1 || P->advection_type == 2)
        for (int d = 0; d < 3; ++d) {
            const int kx = n0 + (d == 0), ky = n1 + (d == 1), kz = n2 + (d == 2);
#pragma omp parallel for schedule(static)
            for (int k = 0; k < kz; ++k)
                for (int j = 0; j < ky; ++j)
                    for (int i = 0; i < kx; ++i) hyp_face(&A, d, i, j, k);
        }
    return FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* RK3 step (timeStep.cpp:11-427)                                                          */
/* ------------------------------------------------------------------------------------- */

/* stochFlux = 1*A + w*B over comps 1..nvars-1, rancorn likewise (MultiFab::LinComb). */
static void weight_noise(const fhdx_params* P, const fhdx_fab stochFlux[3], fhdx_fab rancorn,
                         const fhdx_fab sA[3], const fhdx_fab sB[3], fhdx_fab rA, fhdx_fab rB,
                         double w)
{
    const int nv = nvars_of(P);
    for (int d = 0; d < 3; ++d) {
        fab_t s = F(stochFlux[d]), a = F(sA[d]), b = F(sB[d]);
        memset(s.p, 0, sizeof(double) * s.sn * s.nc);
        for (int n = 1; n < nv; ++n)
            for (long q = 0; q < s.sn; ++q)
                s.p[q + n * s.sn] = 1.0 * a.p[q + n * a.sn] + w * b.p[q + n * b.sn];
    }
    fab_t r = F(rancorn), a = F(rA), b = F(rB);
    for (long q = 0; q < r.sn; ++q) r.p[q] = 1.0 * a.p[q] + w * b.p[q];
}

/* stage combination: mode 0 -> out = U - dt*div + dt*S;
 * mode 1 -> out = 0.25*(3*U0 + U - dt*div + dt*S); mode 2 -> out = 2/3*(0.5*U0 + U - ...). */
static void stage_update(const fhdx_params* P, fhdx_fab out_, fhdx_fab u0_, fhdx_fab us_,
                         const fhdx_fab flux[3], const fhdx_fab* source, double dt, int mode)
{
    fab_t out = F(out_), u0 = F(u0_), us = F(us_);
    fab_t sf = source ? F(*source) : us;
    fab_t fx = F(flux[0]), fy = F(flux[1]), fz = F(flux[2]);
    const int nv = nvars_of(P);
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const double dx0 = P->dx[0], dx1 = P->dx[1], dx2 = P->dx[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                double o[MS + 5];
                for (int n = 0; n < nv; ++n) {
                    /* the source MultiFab, zero without chemistry (main_driver.cpp:241-242) */
                    const double src = source ? AT(sf, i, j, k, n) : 0.0;
                    double div = (AT(fx, i + 1, j, k, n) - AT(fx, i, j, k, n)) / dx0 +
                                 (AT(fy, i, j + 1, k, n) - AT(fy, i, j, k, n)) / dx1 +
                                 (AT(fz, i, j, k + 1, n) - AT(fz, i, j, k, n)) / dx2;
                    if (mode == 0)
                        o[n] = AT(us, i, j, k, n) - dt * div + dt * src;
                    else if (mode == 1)
                        o[n] = 0.25 * (3.0 * AT(u0, i, j, k, n) + AT(us, i, j, k, n) - dt * div +
                                       dt * src);
                    else
                        o[n] = (2. / 3.) * (0.5 * AT(u0, i, j, k, n) + AT(us, i, j, k, n) -
                                            dt * div + dt * src);
                }
                /* gravity, using the previous stage's state (timeStep.cpp:190-201 etc.) */
                const double *g = P->grav;
                double r0 = AT(us, i, j, k, 0), r1 = AT(us, i, j, k, 1), r2 = AT(us, i, j, k, 2),
                       r3 = AT(us, i, j, k, 3);
                if (mode == 0) {
                    for (int n = 0; n < 3; ++n) o[n + 1] += dt * r0 * g[n];
                    o[4] += dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
                } else if (mode == 1) {
                    for (int n = 0; n < 3; ++n) o[n + 1] += 0.25 * dt * r0 * g[n];
                    o[4] += 0.25 * dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
                } else {
                    for (int n = 0; n < 3; ++n) o[n + 1] += 2. / 3. * dt * r0 * g[n];
                    o[4] += 2. / 3. * dt * (g[0] * r1 + g[1] * r2 + g[2] * r3);
                }
                for (int n = 0; n < nv; ++n) AT(out, i, j, k, n) = o[n];
            }
}

static int rk3_step_impl(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu,
                         fhdx_fab cup, fhdx_fab cup2, fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta,
                         fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab flux[3],
                         const fhdx_fab stochFlux[3], const fhdx_fab cornx[3],
                         const fhdx_fab corny[3], const fhdx_fab cornz[3], fhdx_fab visccorn,
                         fhdx_fab rancorn, const fhdx_fab stochA[3], const fhdx_fab stochB[3],
                         fhdx_fab rancornA, fhdx_fab rancornB, const fhdx_fab* ranchemA,
                         const fhdx_fab* ranchemB, const fhdx_fab* ranchem, const fhdx_fab* source,
                         double dt)
{
    const double sq2 = sqrt(2.0), sq3 = sqrt(3.0);
    const double w1 = (2.0 * sq2 + 1.0 * sq3) / 5.0;
    const double w2 = (-4.0 * sq2 + 3.0 * sq3) / 5.0;
    const double w3 = (1.0 * sq2 - 2.0 * sq3) / 10.0;
    const double ws[3] = {w1, w2, w3};
    fhdx_fab in[3] = {cu, cup, cup2};
    fhdx_fab outs[3] = {cup, cup2, cu};
    int err = FHDX_OK;
    for (int s = 0; s < 3; ++s) {
        int e = orc_transport(P, prim, eta, zeta, kappa, chi, D);
        if (e) err = e;
        weight_noise(P, stochFlux, rancorn, stochA, stochB, rancornA, rancornB, ws[s]);
        orc_calculate_flux(P, in[s], prim, eta, zeta, kappa, chi, D, flux, stochFlux, cornx,
                           corny, cornz, visccorn, rancorn, dt);
        if (chem) { /* timeStep.cpp:135-143: ranchem = 1*A + w*B, source from this stage's prim */
            fab_t r = F(*ranchem), a = F(*ranchemA), b = F(*ranchemB);
            for (int m = 0; m < chem->nreaction; ++m)
                for (long q = 0; q < r.sn; ++q)
                    r.p[q + m * r.sn] = 1.0 * a.p[q + m * a.sn] + ws[s] * b.p[q + m * b.sn];
            orc_chem_source(P, chem, prim, *source, *ranchem, dt);
        }
        stage_update(P, outs[s], cu, in[s], flux, chem ? source : NULL, dt, s);
        orc_cons_to_prim(P, prim, outs[s]);
        orc_fill_boundary(P, outs[s]);
        orc_fill_boundary(P, prim);
        orc_set_bc(P, prim, outs[s]); /* timeStep.cpp:209, 314, 425 */
    }
    return err;
}

int orc_rk3_step(const fhdx_params* P, fhdx_fab cu, fhdx_fab cup, fhdx_fab cup2,
                 fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta, fhdx_fab kappa, fhdx_fab chi,
                 fhdx_fab D, const fhdx_fab flux[3], const fhdx_fab stochFlux[3],
                 const fhdx_fab cornx[3], const fhdx_fab corny[3], const fhdx_fab cornz[3],
                 fhdx_fab visccorn, fhdx_fab rancorn, const fhdx_fab stochA[3],
                 const fhdx_fab stochB[3], fhdx_fab rancornA, fhdx_fab rancornB, double dt)
{
    return rk3_step_impl(P, NULL, cu, cup, cup2, prim, eta, zeta, kappa, chi, D, flux, stochFlux,
                         cornx, corny, cornz, visccorn, rancorn, stochA, stochB, rancornA,
                         rancornB, NULL, NULL, NULL, NULL, dt);
}

int orc_rk3_step_chem(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu,
                      fhdx_fab cup, fhdx_fab cup2, fhdx_fab prim, fhdx_fab eta, fhdx_fab zeta,
                      fhdx_fab kappa, fhdx_fab chi, fhdx_fab D, const fhdx_fab flux[3],
                      const fhdx_fab stochFlux[3], const fhdx_fab cornx[3],
                      const fhdx_fab corny[3], const fhdx_fab cornz[3], fhdx_fab visccorn,
                      fhdx_fab rancorn, const fhdx_fab stochA[3], const fhdx_fab stochB[3],
                      fhdx_fab rancornA, fhdx_fab rancornB, fhdx_fab ranchemA, fhdx_fab ranchemB,
                      fhdx_fab ranchem, fhdx_fab source, double dt)
{
    if (!chem || chem->nreaction != P->nreaction || chem->nreaction < 1) return FHDX_ERR_ARG;
    return rk3_step_impl(P, chem, cu, cup, cup2, prim, eta, zeta, kappa, chi, D, flux, stochFlux,
                         cornx, corny, cornz, visccorn, rancorn, stochA, stochB, rancornA,
                         rancornB, &ranchemA, &ranchemB, &ranchem, &source, dt);
}

/* ------------------------------------------------------------------------------------- */
/* Noise: Philox4x32-10 (Salmon et al., SC'11) and a Box-Muller normal pair               */
/* ------------------------------------------------------------------------------------- */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Box-Muller with elementary functions built from IEEE +, -, *, /, fma and sqrt, in the same
 * operation order as the device generator (fhdex_amd/csrc/physics.hpp), so host and device
 * noise agree bit for bit.  ln x, x in (0,1]: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * ln x = e ln2 + 2 atanh(s), s = (m-1)/(m+1), atanh series to s^23. */
static double bm_log(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &mb, sizeof m);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double sr = (m - 1.0) / (m + 1.0);
    const double z = sr * sr;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    p = fma(p, z, 1.0);
    const double lm = 2.0 * sr * p;
    const double de = (double)e;
    return de * 0.6931467056274414 + fma(de, 4.7493250390316726e-07, lm);
}

/* sin(pi t), cos(pi t), t in [0,2): t = n/2 + r, |r| <= 1/4; Taylor series in r, quadrant n mod 4. */
static void bm_sincospi(double t, double* sn, double* cs)
{
    const double n = rint(2.0 * t);
    const double r = t - 0.5 * n;
    const double r2 = r * r;
    double ps = -2.2948428997269873e-08;
    ps = fma(ps, r2, 7.952054001475513e-07);
    ps = fma(ps, r2, -2.1915353447830217e-05);
    ps = fma(ps, r2, 0.00046630280576761255);
    ps = fma(ps, r2, -0.0073704309457143504);
    ps = fma(ps, r2, 0.08214588661112823);
    ps = fma(ps, r2, -0.5992645293207921);
    ps = fma(ps, r2, 2.5501640398773455);
    ps = fma(ps, r2, -5.16771278004997);
    ps = fma(ps, r2, 3.141592653589793);
    const double S = r * ps;
    double pc = -1.3878952462213771e-07;
    pc = fma(pc, r2, 4.303069587032947e-06);
    pc = fma(pc, r2, -0.0001046381049248457);
    pc = fma(pc, r2, 0.0019295743094039231);
    pc = fma(pc, r2, -0.02580689139001406);
    pc = fma(pc, r2, 0.2353306303588932);
    pc = fma(pc, r2, -1.3352627688545895);
    pc = fma(pc, r2, 4.0587121264167685);
    pc = fma(pc, r2, -4.934802200544679);
    const double C = fma(pc, r2, 1.0);
    /* quadrant q = n mod 4: (sin, cos) = (S, C), (C, -S), (-S, -C), (-C, S) */
    const int q = ((int)n) & 3;
    const double a = (q & 1) ? C : S, b = (q & 1) ? S : C;
    *sn = (q & 2) ? -a : a;
    *cs = ((q + 1) & 2) ? -b : b;
}

/* u1 in (0,1] from words 0,1; u2 in [0,1) from words 2,3 (53 bits each);
 * z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2). */
void orc_normal_pair(const uint32_t r[4], double* z0, double* z1)
{
    const double two_m53 = 1.1102230246251565e-16; /* 2^-53 */
    uint64_t a = (((uint64_t)r[0] << 32) | r[1]) >> 11;
    uint64_t b = (((uint64_t)r[2] << 32) | r[3]) >> 11;
    double u1 = ((double)a + 0.5) * two_m53;
    double u2 = (double)b * two_m53;
    double rad = sqrt(-2.0 * bm_log(u1));
    double s, c;
    bm_sincospi(2.0 * u2, &s, &c);
    *z0 = rad * c;
    *z1 = rad * s;
}

int orc_fill_noise(const fhdx_params* P, uint64_t seed, uint64_t step, uint32_t instance,
                   const fhdx_fab stochA[3], const fhdx_fab stochB[3], fhdx_fab rancornA,
                   fhdx_fab rancornB)
{
    const int nv = nvars_of(P);
    const int n[3] = {P->n_cells[0], P->n_cells[1], P->n_cells[2]};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t step_lo = (uint32_t)step, step_hi = (uint32_t)(step >> 32);
    for (int d = 0; d < 3; ++d) {
        fab_t a = F(stochA[d]), b = F(stochB[d]);
        memset(a.p, 0, sizeof(double) * a.sn * a.nc);
        memset(b.p, 0, sizeof(double) * b.sn * b.nc);
        for (int c = 1; c < nv; ++c) {
            double vc = (c <= 3) ? P->variance_coef_mom
                        : (c == 4) ? P->variance_coef_ener : P->variance_coef_mass;
            double sc = sqrt(vc * vc);
            uint32_t field = (uint32_t)(d * 16 + c);
#pragma omp parallel for schedule(static)
            for (int k = 0; k <= a.hi2; ++k)
                for (int j = 0; j <= a.hi1; ++j)
                    for (int i = 0; i <= a.hi0; ++i) {
                        uint32_t lin = (uint32_t)(((long)(k % n[2]) * n[1] + (j % n[1])) * n[0] +
                                                  (i % n[0]));
                        /* a wall direction's face n is its own face, with its own stream */
                        const int fi = (d == 0) ? i : (d == 1) ? j : k;
                        const uint32_t fld = (fi == n[d] && !is_periodic(P, d)) ? field + 128u : field;
                        uint32_t ctr[4] = {lin, step_lo, fld | (step_hi << 8), instance};
                        uint32_t r[4];
                        double z0, z1;
                        orc_philox4x32_10(ctr, key, r);
                        orc_normal_pair(r, &z0, &z1);
                        AT(a, i, j, k, c) = z0 * sc;
                        AT(b, i, j, k, c) = z1 * sc;
                    }
        }
    }
    fab_t a = F(rancornA), b = F(rancornB);
    const double sc = sqrt(P->variance_coef_mom * P->variance_coef_mom);
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= a.hi2; ++k)
        for (int j = 0; j <= a.hi1; ++j)
            for (int i = 0; i <= a.hi0; ++i) {
                uint32_t lin =
                    (uint32_t)(((long)(k % n[2]) * n[1] + (j % n[1])) * n[0] + (i % n[0]));
                /* nodes on the hi plane of a wall direction are distinct: field 48 + mask */
                const int nd[3] = {i, j, k};
                uint32_t mask = 0;
                for (int d = 0; d < 3; ++d)
                    if (nd[d] == n[d] && !is_periodic(P, d)) mask |= 1u << d;
                uint32_t ctr[4] = {lin, step_lo, (48u + mask) | (step_hi << 8), instance};
                uint32_t r[4];
                double z0, z1;
                orc_philox4x32_10(ctr, key, r);
                orc_normal_pair(r, &z0, &z1);
                AT(a, i, j, k, 0) = z0 * sc;
                AT(b, i, j, k, 0) = z1 * sc;
            }
    return FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* Initial state and a self-allocating driver                                              */
/* ------------------------------------------------------------------------------------- */

int orc_init_cons(const fhdx_params* P, fhdx_fab cu_, int prob_type, double rho0,
                  const double* rhobar, double T_init)
{
    fab_t cu = F(cu_);
    const int NS = P->nspecies, nv = nvars_of(P);
    double massvec[MS];
    for (int n = 0; n < NS; ++n) massvec[n] = rhobar[n];
    double ie = get_energy(P, massvec, T_init);
    for (long q = 0; q < cu.sn * cu.nc; ++q) cu.p[q] = 0.0;
    for (long q = 0; q < cu.sn; ++q) {
        cu.p[q] = rho0;
        cu.p[q + 1 * cu.sn] = 0;
        cu.p[q + 2 * cu.sn] = 0;
        cu.p[q + 3 * cu.sn] = 0;
        cu.p[q + 4 * cu.sn] = rho0 * ie;
        for (int n = 0; n < NS; ++n) cu.p[q + (5 + n) * cu.sn] = rho0 * rhobar[n];
    }
    if (prob_type == 100) {
        const double pi = acos(-1.);
        const double Ly = P->dx[1] * P->n_cells[1];
        for (int k = 0; k < P->n_cells[2]; ++k)
            for (int j = 0; j < P->n_cells[1]; ++j)
                for (int i = 0; i < P->n_cells[0]; ++i) {
                    double y = (j + 0.5) * P->dx[1];
                    double Yk[MS];
                    for (int l = 0; l < NS; ++l) Yk[l] = AT(cu, i, j, k, 5 + l) / AT(cu, i, j, k, 0);
                    AT(cu, i, j, k, 0) = rho0 + 0.1 * rho0 * sin(2. * pi * y / Ly);
                    for (int l = 0; l < NS; ++l) AT(cu, i, j, k, 5 + l) = AT(cu, i, j, k, 0) * Yk[l];
                }
        orc_fill_boundary(P, cu_);
    } else if (prob_type == 2 || prob_type == 3) {
        /* Valid cells only (the ParallelFor over the tile box); ghost cells keep the uniform
         * values until FillBoundary / setBC.  The domain is [0, n dx] (prob_lo = 0). */
        if (prob_type == 2 && NS != 4) return FHDX_ERR_UNSUPPORTED;
        const double zc = P->n_cells[2] * P->dx[2] / 2.;
        const double Ly = P->n_cells[1] * P->dx[1];
        for (int k = 0; k < P->n_cells[2]; ++k)
            for (int j = 0; j < P->n_cells[1]; ++j)
                for (int i = 0; i < P->n_cells[0]; ++i) {
                    const double posy = (j + 0.5) * P->dx[1];
                    const double posz = (k + 0.5) * P->dx[2];
                    double mv[MS] = {0};
                    if (prob_type == 2) {
                        /* Rayleigh-Taylor: two layers of equal mean molar mass, density
                         * exponential in z at hydrostatic equilibrium (:172-214) */
                        const int top = posz - zc >= 0.;
                        mv[0] = top ? 0.4 : 0.1;
                        mv[1] = top ? 0.4 : 0.1;
                        mv[2] = top ? 0.1 : 0.4;
                        mv[3] = top ? 0.1 : 0.4;
                        double molmix = 0.;
                        for (int l = 0; l < NS; ++l) molmix += mv[l] / P->molmass[l];
                        molmix = 1. / molmix;
                        const double pamb = AT(cu, i, j, k, 0) * (P->Runiv / molmix) * T_init;
                        molmix = 0.;
                        for (int l = 0; l < NS; ++l) molmix = molmix + mv[l] / P->molmass[l];
                        molmix = 1.0 / molmix;
                        const double rgasmix = P->Runiv / molmix;
                        const double alpha = P->grav[2] / (rgasmix * T_init);
                        AT(cu, i, j, k, 0) = pamb * exp(alpha * posz) / (rgasmix * T_init);
                        for (int l = 0; l < NS; ++l) AT(cu, i, j, k, 5 + l) = AT(cu, i, j, k, 0) * mv[l];
                    } else {
                        /* diffusion barrier: mass fractions linear in y between the y walls'
                         * compositions (:216-227) */
                        for (int l = 0; l < NS; ++l) {
                            const double Ygrad = (P->bc_Yk[3][l] - P->bc_Yk[2][l]) / Ly;
                            mv[l] = Ygrad * posy + P->bc_Yk[2][l];
                            AT(cu, i, j, k, 5 + l) = AT(cu, i, j, k, 0) * mv[l];
                        }
                    }
                    const double e = get_energy(P, mv, T_init);
                    const double r = AT(cu, i, j, k, 0), m1 = AT(cu, i, j, k, 1),
                                 m2 = AT(cu, i, j, k, 2), m3 = AT(cu, i, j, k, 3);
                    AT(cu, i, j, k, 4) = r * e + 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r;
                }
    } else if (prob_type == 4) {
        /* Taylor-Green vortex (:228-262): two species, 50/50; x, y, z measured from the
         * domain's lower corner, every direction scaled by the x extent Lf; rho from the
         * pressure field at T_init, rho E = p/(5/3 - 1) + kinetic energy. */
        if (NS != 2) return FHDX_ERR_UNSUPPORTED;
        const double pi = acos(-1.);
        const double Lf = P->n_cells[0] * P->dx[0];
        const double avogadro = P->Runiv / P->k_B;
        const double velscale = 9237.;
        const double pscale = 884.147e3;
        const double rhoscale = P->molmass[0] / avogadro * pscale / (P->k_B * T_init);
        for (int k = 0; k < P->n_cells[2]; ++k)
            for (int j = 0; j < P->n_cells[1]; ++j)
                for (int i = 0; i < P->n_cells[0]; ++i) {
                    const double x = (i + 0.5) * P->dx[0];
                    const double y = (j + 0.5) * P->dx[1];
                    const double z = (k + 0.5) * P->dx[2];
                    const double pres = pscale + rhoscale * velscale * velscale * cos(2. * pi * x / Lf) *
                                                     cos(4. * pi * y / Lf) * (cos(4. * pi * z / Lf) + 2.);
                    const double r = (P->molmass[0] / avogadro) * pres / (P->k_B * T_init);
                    const double m1 = velscale * r * sin(2. * pi * x / Lf) * cos(2. * pi * y / Lf) *
                                      cos(2. * pi * z / Lf);
                    const double m2 = -velscale * r * cos(2. * pi * x / Lf) * sin(2. * pi * y / Lf) *
                                      cos(2. * pi * z / Lf);
                    const double m3 = 0.;
                    AT(cu, i, j, k, 0) = r;
                    AT(cu, i, j, k, 1) = m1;
                    AT(cu, i, j, k, 2) = m2;
                    AT(cu, i, j, k, 3) = m3;
                    AT(cu, i, j, k, 4) = pres / (5. / 3. - 1.) + 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r;
                    AT(cu, i, j, k, 5) = 0.5 * r;
                    AT(cu, i, j, k, 6) = 0.5 * r;
                }
    } else if (prob_type != 1) {
        return FHDX_ERR_UNSUPPORTED;
    }
    (void)nv;
    return FHDX_OK;
}

/* ------------------------------------------------------------------------------------- */
/* Chemistry source (src_chemistry/chemistry_functions.cpp:128-193)                          */
/* ------------------------------------------------------------------------------------- */

/* e^x and x^y from IEEE operations only, the same sequence as the device (physics.hpp:
 * chem_exp, chem_pow): x = k ln2 + r, |r| <= ln2/2, Taylor series of e^r to r^14, times 2^k;
 * x^y = e^(y ln x) with the Box-Muller logarithm; y = 0 gives 1, y = 1 gives x. */
double orc_chem_exp(double x)
{
    if (x == 0.0) return 1.0;
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    const double r = fma(-k, 4.7493250390316726e-07, fma(-k, 0.6931467056274414, x));
    double p = 1.1470745597729725e-11;
    p = fma(p, r, 1.6059043836821613e-10);
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const int ki = (int)k, k1 = ki / 2, k2 = ki - k1;
    const uint64_t b1 = (uint64_t)(k1 + 1023) << 52, b2 = (uint64_t)(k2 + 1023) << 52;
    double s1, s2;
    memcpy(&s1, &b1, sizeof s1);
    memcpy(&s2, &b2, sizeof s2);
    return p * s1 * s2;
}

double orc_chem_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
    if (x < 0.0) return NAN;
    return orc_chem_exp(y * bm_log(x));
}

/* pow(c, s) for an integer stoichiometric coefficient s >= 0, by repeated multiplication */
static double chem_ipow(double c, int s)
{
    if (s == 0) return 1.0;
    double r = c;
    for (int q = 1; q < s; ++q) r = r * c;
    return r;
}

int orc_chem_source(const fhdx_params* P, const fhdx_chemistry* C, fhdx_fab prim_,
                    fhdx_fab source_, fhdx_fab ranchem_, double dt)
{
    if (C->reaction_type != 0 && C->reaction_type != 1) return FHDX_ERR_UNSUPPORTED;
    if (!(C->T0_chem > 0.)) return FHDX_ERR_ARG;
    fab_t prim = F(prim_), source = F(source_), ranchem = F(ranchem_);
    const int NS = P->nspecies, nr = C->nreaction;
    const double dV = P->dx[0] * P->dx[1] * P->dx[2];
    const double Runiv = P->Runiv;
    const double Navo = P->Runiv / P->k_B; /* avogadro (common_functions.cpp:775) */
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const double pres = AT(prim, i, j, k, 5);
                const double T = AT(prim, i, j, k, 4);
                const double T0 = C->T0_chem;
                const double ctot = pres / Runiv / T;
                double ck[MS], rate[FHDX_MAX_REACTION], src[MS];
                for (int n = 0; n < NS; ++n) ck[n] = AT(prim, i, j, k, 6 + NS + n) * ctot;
                for (int m = 0; m < nr; ++m) {
                    rate[m] = C->rate_const[m];
                    rate[m] *= orc_chem_exp(-C->alpha_param[m] / Runiv * (1 / T - 1 / T0));
                    rate[m] *= orc_chem_pow(T / T0, C->beta_param[m]);
                    for (int n = 0; n < NS; ++n) rate[m] *= chem_ipow(ck[n], C->stoich_R[m][n]);
                }
                for (int n = 0; n < NS; ++n) src[n] = 0.;
                for (int m = 0; m < nr; ++m) {
                    rate[m] = (0. < rate[m]) ? rate[m] : 0.; /* std::max(0., rate) */
                    const double W = AT(ranchem, i, j, k, m) / sqrt(dt * dV);
                    for (int n = 0; n < NS; ++n) {
                        const int pr = C->stoich_P[m][n] - C->stoich_R[m][n];
                        src[n] += P->molmass[n] * pr * rate[m];
                        if (C->reaction_type == 1)
                            src[n] += P->molmass[n] * pr * sqrt(rate[m] / Navo) * W;
                    }
                }
                for (int n = 0; n < NS; ++n) AT(source, i, j, k, 5 + n) = src[n];
            }
    return FHDX_OK;
}

/* Reaction noise sets A and B (nreaction comps, valid cells) of key (seed, step, instance):
 * field 64 + m, the linear cell index as the counter, variance 1 (timeStep.cpp:93-97). */
int orc_fill_chem_noise(const fhdx_params* P, uint64_t seed, uint64_t step, uint32_t instance,
                        fhdx_fab A_, fhdx_fab B_)
{
    fab_t A = F(A_), B = F(B_);
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t fhi = ((uint32_t)(step >> 32)) << 8;
    for (int m = 0; m < P->nreaction; ++m)
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) {
                    const uint32_t lin = (uint32_t)(((long)k * n1 + j) * n0 + i);
                    const uint32_t ctr[4] = {lin, (uint32_t)step, (64u + (uint32_t)m) | fhi, instance};
                    uint32_t r[4];
                    double z0, z1;
                    orc_philox4x32_10(ctr, key, r);
                    orc_normal_pair(r, &z0, &z1);
                    AT(A, i, j, k, m) = z0;
                    AT(B, i, j, k, m) = z1;
                }
    return FHDX_OK;
}

/* evaluateStats (src_compressible/stats.cpp:5-363): running means and variances of cons and
 * prim, the yz-slice averages at every x, the values of the cross_cell slice (miscVals) and the
 * spatial cross-correlations against that slice.  Serial loops in the reference's k, j, i order,
 * so every sum is accumulated in the reference's order. */
int orc_evaluate_stats(const fhdx_params* P, fhdx_fab cons_, fhdx_fab consMean_, fhdx_fab consVar_,
                       fhdx_fab prim_, fhdx_fab primMean_, fhdx_fab primVar_, fhdx_fab spatialCross_,
                       fhdx_fab miscStats_, double* miscVals, int steps, int cross_cell)
{
    fab_t cu = F(cons_), cumeans = F(consMean_), cuvars = F(consVar_), prim = F(prim_);
    fab_t primmeans = F(primMean_), primvars = F(primVar_), spatialcross = F(spatialCross_);
    fab_t miscstats = F(miscStats_);
    const int NS = P->nspecies, nvars = nvars_of(P), nprimvars = 6 + 2 * NS;
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    if (steps < 1 || cross_cell < 0 || cross_cell >= n0) return FHDX_ERR_ARG;
    const double stepsminusone = steps - 1.;
    const double stepsinv = 1. / steps;
    const int n_cells_yz = n1 * n2;
    double fracvec[MS];

    /* evaluate_means (stats.cpp:45-88) */
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                for (int l = 0; l < nvars; ++l)
                    AT(cumeans, i, j, k, l) = (AT(cumeans, i, j, k, l) * stepsminusone + AT(cu, i, j, k, l)) * stepsinv;
                double densitymeaninv = 1.0 / AT(cumeans, i, j, k, 0);
                for (int l = 5; l < nvars; ++l) fracvec[l - 5] = AT(cumeans, i, j, k, l) * densitymeaninv;
                AT(primmeans, i, j, k, 0) = AT(cumeans, i, j, k, 0);
                AT(primmeans, i, j, k, 1) = AT(cumeans, i, j, k, 1) * densitymeaninv;
                AT(primmeans, i, j, k, 2) = AT(cumeans, i, j, k, 2) * densitymeaninv;
                AT(primmeans, i, j, k, 3) = AT(cumeans, i, j, k, 3) * densitymeaninv;
                double vsqr = AT(primmeans, i, j, k, 1) * AT(primmeans, i, j, k, 1) +
                              AT(primmeans, i, j, k, 2) * AT(primmeans, i, j, k, 2) +
                              AT(primmeans, i, j, k, 3) * AT(primmeans, i, j, k, 3);
                double intenergy = AT(cumeans, i, j, k, 4) / AT(cumeans, i, j, k, 0) - 0.5 * vsqr;
                AT(primmeans, i, j, k, 4) = get_temperature(P, intenergy, fracvec);
                AT(primmeans, i, j, k, 5) = get_pressure(P, fracvec, AT(cumeans, i, j, k, 0), AT(primmeans, i, j, k, 4));
            }

    /* the cross_cell slice (stats.cpp:96-150) */
    for (int q = 0; q < 20; ++q) miscVals[q] = 0.;
    {
        const int x = cross_cell;
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j) {
                miscVals[0] += AT(cumeans, x, j, k, 1);
                miscVals[1] += AT(cu, x, j, k, 1);
                miscVals[2] += AT(primmeans, x, j, k, 1);
                miscVals[3] += AT(cumeans, x, j, k, 0);
                miscVals[4] += AT(cu, x, j, k, 0);
                miscVals[5] += AT(prim, x, j, k, 1);
                miscVals[6] += AT(cu, x, j, k, 4);
                miscVals[7] += AT(cumeans, x, j, k, 4);
                miscVals[8] += AT(cu, x, j, k, 2);
                miscVals[9] += AT(cumeans, x, j, k, 2);
                miscVals[10] += AT(cu, x, j, k, 3);
                miscVals[11] += AT(cumeans, x, j, k, 3);
                double cv = 0;
                for (int l = 0; l < NS; ++l) cv = cv + P->hcv[l] * AT(cumeans, x, j, k, 5 + l) / AT(cumeans, x, j, k, 0);
                miscVals[12] += cv;
                miscVals[13] += AT(primmeans, x, j, k, 4);
                miscVals[14] += AT(primmeans, x, j, k, 2);
                miscVals[15] += AT(primmeans, x, j, k, 3);
                miscVals[16] += AT(prim, x, j, k, 4);
            }
    }
    for (int q = 0; q < 20; ++q) miscVals[q] /= n_cells_yz;

    /* yz averages at every x (stats.cpp:156-210) */
    const int nstats = 19;
    double* yz = (double*)calloc((size_t)n0 * nstats, sizeof(double));
    if (!yz) return FHDX_ERR_NOMEM;
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                double* Y = yz + (long)i * nstats;
                Y[0] += AT(cu, i, j, k, 0);
                Y[1] += AT(cumeans, i, j, k, 0);
                Y[2] += AT(cu, i, j, k, 4);
                Y[3] += AT(cumeans, i, j, k, 4);
                Y[4] += AT(cu, i, j, k, 1);
                Y[5] += AT(cumeans, i, j, k, 1);
                Y[6] += AT(cu, i, j, k, 2);
                Y[7] += AT(cumeans, i, j, k, 2);
                Y[8] += AT(cu, i, j, k, 3);
                Y[9] += AT(cumeans, i, j, k, 3);
                Y[10] += AT(prim, i, j, k, 1);
                Y[11] += AT(primmeans, i, j, k, 1);
                Y[12] += AT(prim, i, j, k, 2);
                Y[13] += AT(primmeans, i, j, k, 2);
                Y[14] += AT(prim, i, j, k, 3);
                Y[15] += AT(primmeans, i, j, k, 3);
                double cv = 0;
                for (int l = 0; l < NS; ++l) cv = cv + P->hcv[l] * AT(cumeans, i, j, k, 5 + l) / AT(cumeans, i, j, k, 0);
                Y[16] += cv;
                Y[17] += AT(prim, i, j, k, 4);
                Y[18] += AT(primmeans, i, j, k, 4);
            }
    for (long q = 0; q < (long)n0 * nstats; ++q) yz[q] /= n_cells_yz;

    /* variances and correlations (stats.cpp:215-360) */
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const double* Y = yz + (long)i * nstats;
                double densitymeaninv = 1.0 / AT(cumeans, i, j, k, 0);
                double delrho = AT(cu, i, j, k, 0) - AT(cumeans, i, j, k, 0);
                double delpx = AT(cu, i, j, k, 1) - AT(cumeans, i, j, k, 1);
                double delpy = AT(cu, i, j, k, 2) - AT(cumeans, i, j, k, 2);
                double delpz = AT(cu, i, j, k, 3) - AT(cumeans, i, j, k, 3);
                double delenergy = AT(cu, i, j, k, 4) - AT(cumeans, i, j, k, 4);
                double delrhoS = Y[0] - Y[1];
                double delES = Y[3] - Y[2];
                double delrhoSstar = miscVals[4] - miscVals[3];
                double delpxSstar = miscVals[1] - miscVals[0];
                AT(cuvars, i, j, k, 0) = (AT(cuvars, i, j, k, 0) * stepsminusone + delrho * delrho) * stepsinv;
                AT(cuvars, i, j, k, 1) = (AT(cuvars, i, j, k, 1) * stepsminusone + delpx * delpx) * stepsinv;
                AT(cuvars, i, j, k, 2) = (AT(cuvars, i, j, k, 2) * stepsminusone + delpy * delpy) * stepsinv;
                AT(cuvars, i, j, k, 3) = (AT(cuvars, i, j, k, 3) * stepsminusone + delpz * delpz) * stepsinv;
                AT(cuvars, i, j, k, 4) = (AT(cuvars, i, j, k, 4) * stepsminusone + delenergy * delenergy) * stepsinv;
                double delvelx = (delpx - AT(primmeans, i, j, k, 1) * delrho) * densitymeaninv;
                double delvely = (delpy - AT(primmeans, i, j, k, 2) * delrho) * densitymeaninv;
                double delvelz = (delpz - AT(primmeans, i, j, k, 3) * delrho) * densitymeaninv;
                AT(primvars, i, j, k, 0) = AT(cuvars, i, j, k, 0);
                AT(primvars, i, j, k, 1) = (AT(primvars, i, j, k, 1) * stepsminusone + delvelx * delvelx) * stepsinv;
                AT(primvars, i, j, k, 2) = (AT(primvars, i, j, k, 2) * stepsminusone + delvely * delvely) * stepsinv;
                AT(primvars, i, j, k, 3) = (AT(primvars, i, j, k, 3) * stepsminusone + delvelz * delvelz) * stepsinv;
                double delg = AT(primmeans, i, j, k, 1) * delpx + AT(primmeans, i, j, k, 2) * delpy +
                              AT(primmeans, i, j, k, 3) * delpz;
                AT(primvars, i, j, k, nprimvars) =
                    (AT(primvars, i, j, k, nprimvars) * stepsminusone + delg * delg) * stepsinv;
                AT(primvars, i, j, k, nprimvars + 1) =
                    (AT(primvars, i, j, k, nprimvars + 1) * stepsminusone + delg * delenergy) * stepsinv;
                AT(primvars, i, j, k, nprimvars + 2) =
                    (AT(primvars, i, j, k, nprimvars + 2) * stepsminusone + delrho * delenergy) * stepsinv;
                AT(primvars, i, j, k, nprimvars + 3) =
                    (AT(primvars, i, j, k, nprimvars + 3) * stepsminusone + delrho * delg) * stepsinv;
                double delT = AT(prim, i, j, k, 4) - AT(primmeans, i, j, k, 4);
                AT(primvars, i, j, k, 4) = (AT(primvars, i, j, k, 4) * stepsminusone + delT * delT) * stepsinv;
                AT(miscstats, i, j, k, 0) = (AT(miscstats, i, j, k, 0) * stepsminusone + miscVals[1] * Y[0]) * stepsinv;
                double delpdelrho = AT(miscstats, i, j, k, 0) - miscVals[0] * AT(cumeans, i, j, k, 0);
                AT(miscstats, i, j, k, 1) = (AT(miscstats, i, j, k, 1) * stepsminusone + delrhoS * delrhoSstar) * stepsinv;
                AT(miscstats, i, j, k, 2) = (AT(miscstats, i, j, k, 2) * stepsminusone + miscVals[16] * Y[17]) * stepsinv;
                AT(miscstats, i, j, k, 3) = (AT(miscstats, i, j, k, 3) * stepsminusone + miscVals[16] * Y[0]) * stepsinv;
                AT(miscstats, i, j, k, 4) = (AT(miscstats, i, j, k, 4) * stepsminusone + delrhoS * delpxSstar) * stepsinv;
                AT(miscstats, i, j, k, 5) = (AT(miscstats, i, j, k, 5) * stepsminusone + delES * delrhoSstar) * stepsinv;
                AT(spatialcross, i, j, k, 0) = miscVals[13];
                AT(spatialcross, i, j, k, 1) = Y[18];
                AT(spatialcross, i, j, k, 2) = AT(miscstats, i, j, k, 2);
                AT(spatialcross, i, j, k, 3) = AT(miscstats, i, j, k, 2) - Y[18] * miscVals[13];
                AT(spatialcross, i, j, k, 4) = AT(miscstats, i, j, k, 3) - Y[1] * miscVals[13];
                if (miscVals[3] == 0.)
                    AT(spatialcross, i, j, k, 5) = 0.;
                else
                    AT(spatialcross, i, j, k, 5) = (delpdelrho - miscVals[2] * AT(miscstats, i, j, k, 1)) / miscVals[3];
                AT(spatialcross, i, j, k, 6) = AT(miscstats, i, j, k, 4);
                AT(spatialcross, i, j, k, 7) = AT(miscstats, i, j, k, 5);
            }
    free(yz);
    return FHDX_OK;
}

long orc_noise_doubles_per_step(const fhdx_params* P)
{
    const long n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    const long nv = nvars_of(P);
    long faces = ((n0 + 1) * n1 * n2 + n0 * (n1 + 1) * n2 + n0 * n1 * (n2 + 1)) * nv;
    long nodes = (n0 + 1) * (n1 + 1) * (n2 + 1);
    return 2 * (faces + nodes);
}

static fhdx_fab mkfab(int l0, int l1, int l2, int h0, int h1, int h2, int nc)
{
    fhdx_fab f;
    f.lo[0] = l0; f.lo[1] = l1; f.lo[2] = l2;
    f.hi[0] = h0; f.hi[1] = h1; f.hi[2] = h2;
    f.ncomp = nc;
    f.pad = 0;
    long n = (long)(h0 - l0 + 1) * (h1 - l1 + 1) * (h2 - l2 + 1) * nc;
    f.p = (double*)calloc((size_t)n, sizeof(double));
    return f;
}

static int advance_impl(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu,
                        fhdx_fab prim, int nsteps, double dt, const double* noise, uint64_t seed,
                        uint64_t step0, uint32_t instance)
{
    const int NS = P->nspecies, nv = nvars_of(P), g = P->ngc;
    const int n0 = P->n_cells[0], n1 = P->n_cells[1], n2 = P->n_cells[2];
    fhdx_fab cup = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nv);
    fhdx_fab cup2 = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nv);
    fhdx_fab eta = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, 1);
    fhdx_fab zeta = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, 1);
    fhdx_fab kappa = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, 1);
    fhdx_fab chi = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, NS);
    fhdx_fab D = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, NS * NS);
    fhdx_fab flux[3], sf[3], sA[3], sB[3], cx[3], cy[3], cz[3];
    for (int d = 0; d < 3; ++d) {
        int h0 = n0 - 1 + (d == 0), h1 = n1 - 1 + (d == 1), h2 = n2 - 1 + (d == 2);
        flux[d] = mkfab(0, 0, 0, h0, h1, h2, nv + 4);
        sf[d] = mkfab(0, 0, 0, h0, h1, h2, nv);
        sA[d] = mkfab(0, 0, 0, h0, h1, h2, nv);
        sB[d] = mkfab(0, 0, 0, h0, h1, h2, nv);
        cx[d] = mkfab(0, 0, 0, n0, n1, n2, 1);
        cy[d] = mkfab(0, 0, 0, n0, n1, n2, 1);
        cz[d] = mkfab(0, 0, 0, n0, n1, n2, 1);
    }
    fhdx_fab vcorn = mkfab(0, 0, 0, n0, n1, n2, 1);
    fhdx_fab rc = mkfab(0, 0, 0, n0, n1, n2, 1);
    fhdx_fab rA = mkfab(0, 0, 0, n0, n1, n2, 1);
    fhdx_fab rB = mkfab(0, 0, 0, n0, n1, n2, 1);
    const int nr = chem ? chem->nreaction : 1;
    fhdx_fab cA = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nr);
    fhdx_fab cB = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nr);
    fhdx_fab cR = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nr);
    fhdx_fab src = mkfab(-g, -g, -g, n0 + g - 1, n1 + g - 1, n2 + g - 1, nv);

    orc_fill_boundary(P, cu);
    orc_cons_to_prim(P, prim, cu);
    orc_fill_boundary(P, prim);
    orc_set_bc(P, prim, cu); /* main_driver.cpp:535-542 */
    int err = FHDX_OK;
    const long per = orc_noise_doubles_per_step(P);
    for (int s = 0; s < nsteps; ++s) {
        if (noise) {
            const double* q = noise + per * s;
            fhdx_fab* sets[2] = {sA, sB};
            fhdx_fab nodes[2] = {rA, rB};
            for (int set = 0; set < 2; ++set) {
                for (int d = 0; d < 3; ++d) {
                    fab_t f = F(sets[set][d]);
                    memcpy(f.p, q, sizeof(double) * f.sn * f.nc);
                    q += f.sn * f.nc;
                }
                fab_t f = F(nodes[set]);
                memcpy(f.p, q, sizeof(double) * f.sn);
                q += f.sn;
            }
        } else {
            orc_fill_noise(P, seed, step0 + (uint64_t)s, instance, sA, sB, rA, rB);
        }
        int e;
        if (chem) {
            orc_fill_chem_noise(P, seed, step0 + (uint64_t)s, instance, cA, cB);
            e = orc_rk3_step_chem(P, chem, cu, cup, cup2, prim, eta, zeta, kappa, chi, D, flux, sf,
                                  cx, cy, cz, vcorn, rc, sA, sB, rA, rB, cA, cB, cR, src, dt);
        } else {
            e = orc_rk3_step(P, cu, cup, cup2, prim, eta, zeta, kappa, chi, D, flux, sf, cx, cy,
                             cz, vcorn, rc, sA, sB, rA, rB, dt);
        }
        if (e) err = e;
    }
    free(cA.p); free(cB.p); free(cR.p); free(src.p);
    free(cup.p); free(cup2.p); free(eta.p); free(zeta.p); free(kappa.p); free(chi.p); free(D.p);
    for (int d = 0; d < 3; ++d) {
        free(flux[d].p); free(sf[d].p); free(sA[d].p); free(sB[d].p);
        free(cx[d].p); free(cy[d].p); free(cz[d].p);
    }
    free(vcorn.p); free(rc.p); free(rA.p); free(rB.p);
    return err;
}

int orc_advance(const fhdx_params* P, fhdx_fab cu, fhdx_fab prim, int nsteps, double dt,
                const double* noise, uint64_t seed, uint64_t step0, uint32_t instance)
{
    return advance_impl(P, NULL, cu, prim, nsteps, dt, noise, seed, step0, instance);
}

int orc_advance_chem(const fhdx_params* P, const fhdx_chemistry* chem, fhdx_fab cu, fhdx_fab prim,
                     int nsteps, double dt, uint64_t seed, uint64_t step0, uint32_t instance)
{
    if (!chem || chem->nreaction != P->nreaction || chem->nreaction < 1) return FHDX_ERR_ARG;
    return advance_impl(P, chem, cu, prim, nsteps, dt, NULL, seed, step0, instance);
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
