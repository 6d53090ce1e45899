// physics.hpp -- device-side FP64 physics of the compressible FHD step (gfx950).
//
// Every routine keeps the reference's floating-point operation order (left-to-right sums,
// the same parenthesisation, a division wherever the reference divides) and the library is
// compiled with -ffp-contract=off, so results are bit-identical to a faithful CPU evaluation
// of the reference formulas.  That matters: the species-noise Cholesky factors a matrix that
// is singular in exact arithmetic (its last pivot is rounding noise), so any reordering moves
// rho*Y_k, j and rho*E by ~1e-11 relative per step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fhdx.h"

namespace fhdx {

constexpr int MS = FHDX_MAX_SPECIES;

// Kernel-argument copy of the run parameters plus per-call constants.
constexpr int MR = 8;  // reactions (FHDX_MAX_REACTION)

// Reaction data of the chemistry source (chemistry_namespace.H:1-39); nreaction 0: none.
struct DevChem {
    int nreaction, reaction_type;
    signed char R[MR][8], PR[MR][8];  // stoich_coeffs_R and P - R (chemistry_functions.cpp:86-88)
    double rate[MR], alpha[MR], beta[MR];
    double T0, Navo;                  // T0_chem; avogadro = Runiv / k_B (common_functions.cpp:775)
};

struct DevParams {
    int ns, nvars, nprim, ngc;
    int n[3];       // domain cells per direction
    // The valid box a launch covers: lower corner bo and extent bn (global cell indices).  The
    // whole domain unless the arrays are one FAB of a MultiFab split into several boxes.
    int bo[3], bn[3];
    int algorithm_type, advection_type, visc_type, stoch_on;
    int transport_type;  // 1 Giovangigli, 3 Hirschfelder-Curtiss-Bird binary
    double dx[3];
    double molmass[MS], diameter[MS], hcv[MS], hcp[MS], e0[MS];
    double k_B, Runiv, zeta_ratio;
    double grav[3];
    double var_mom, var_ener, var_mass;  // sqrt(coef*coef), the MultiFabFillRandom scale
    DevChem chem;
    // Walls (face f = 2*direction + side).  periodic[d]: bc_vel -1; wall[f]: bc_mass 1 or 2
    // (is_*_dirichlet_mass, flux.cpp:29-34); bcv/bct/bcm: bc_vel/therm/mass codes.
    int any_wall;
    int periodic[3];
    int wall[6], bcv[6], bct[6], bcm[6];
    double tw[6], pw[6], rhow[6];  // wall / reservoir temperature; reservoir pressure, density
    double Yw[6][MS], Xw[6][MS];
};

// Strided view of one FAB (AMReX Array4 layout: x fastest, then y, z, component).
struct Arr {
    double* p;
    int lo0, lo1, lo2;
    int sy, sz;
    long sn;
    // Component base (wave-uniform) plus a 32-bit byte offset within the component: device
    // loads and stores then take the scalar-base + 32-bit vector-offset form instead of a
    // 64-bit vector address per access.  A component of a FAB must stay below 4 GiB (the
    // C ABI checks it).
    __host__ __device__ inline double& operator()(int i, int j, int k, int c = 0) const
    {
        const unsigned off = (unsigned)((i - lo0) + (j - lo1) * sy + (k - lo2) * sz);
        return *(double*)((char*)(p + (long)c * sn) + (off << 3));
    }
};

__host__ __device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double dmin(double a, double b) { return a < b ? a : b; }
// amrex::max(0., amrex::min(1., y))
__host__ __device__ inline double clamp01(double y) { return dmax(0., dmin(1., y)); }

// ---- EOS (compressible_functions.H:138-199, 1497-1520) --------------------------------

template <int NS>
__device__ inline double eos_temperature(const DevParams& P, double energy, const double* Yk)
{
    double cvmix = 0., e0mix = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        cvmix += Yk[n] * P.hcv[n];
        e0mix += Yk[n] * P.e0[n];
    }
    return (energy - e0mix) / cvmix;
}

template <int NS>
__device__ inline double eos_pressure(const DevParams& P, const double* Yk, double rho, double T)
{
    double molmix = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) molmix += Yk[n] / P.molmass[n];
    molmix = 1. / molmix;
    return rho * (P.Runiv / molmix) * T;
}

template <int NS>
__device__ inline double eos_energy(const DevParams& P, const double* Yk, double T)
{
    double cvmix = 0., e0mix = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        cvmix += Yk[n] * P.hcv[n];
        e0mix += Yk[n] * P.e0[n];
    }
    return e0mix + T * cvmix;
}

// conservedToPrimitive for one cell (conservedPrimitiveConversions.cpp:57-88).
// c[0..nvars), pr[0..nprim).
template <int NS>
__device__ inline void cons_to_prim_cell(const DevParams& P, const double* c, double* pr)
{
    double Yk[NS], Yf[NS];
    pr[0] = c[0];
    pr[1] = c[1] / c[0];
    pr[2] = c[2] / c[0];
    pr[3] = c[3] / c[0];
    double vsqr = pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3];
    double intenergy = c[4] / c[0] - 0.5 * vsqr;
    double sumYk = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        Yk[n] = c[5 + n] / c[0];
        Yf[n] = clamp01(Yk[n]);
        sumYk += Yf[n];
    }
#pragma unroll
    for (int n = 0; n < NS; ++n) Yf[n] /= sumYk;
    pr[4] = eos_temperature<NS>(P, intenergy, Yf);
    // GetMolfrac of the raw mass fractions
    double molmix = 0.;
#pragma unroll
    for (int n = 0; n < NS; ++n) molmix += Yk[n] / P.molmass[n];
    molmix = 1. / molmix;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        pr[6 + n] = Yk[n];
        pr[6 + NS + n] = Yk[n] * molmix / P.molmass[n];
    }
    pr[5] = eos_pressure<NS>(P, Yk, pr[0], pr[4]);
}

// ---- transport: IdealMixtureTransportGIO + D_GIO1 (compressible_functions.H:647-1352) ----

// Gaussian elimination without pivoting and back substitution (H:1254-1271).
template <int NS>
__device__ inline void gio_solve(double (&a)[NS][NS], double* Bi)
{
#pragma unroll
    for (int j = 0; j < NS - 1; ++j)
#pragma unroll
        for (int i = j + 1; i < NS; ++i) {
            double fact1 = a[i][j] / a[j][j];
#pragma unroll
            for (int k = 0; k < NS; ++k) a[i][k] -= fact1 * a[j][k];
            Bi[i] -= fact1 * Bi[j];
        }
    Bi[NS - 1] /= a[NS - 1][NS - 1];
#pragma unroll
    for (int i = NS - 2; i > -1; --i) {
        double sum1 = 0.0;
#pragma unroll
        for (int j = NS - 1; j > i; --j) sum1 += a[i][j] * Bi[j];
        Bi[i] = (Bi[i] - sum1) / a[i][i];
    }
}

struct Transport {
    double eta, zeta, kappa;
};

// IdealMixtureTransportHCBBin (compressible_functions.H:1357-1478): Hirschfelder-Curtiss-Bird
// first approximations for a binary mixture of hard spheres with the correction factors 1.016
// (viscosity), 1.025 (conductivity), 1.133 (diffusion) and 1.299 (thermal diffusion).  Xk are the
// mole fractions of the clamped mass fractions (transCoeffs.cpp:60-61); D[4] before the density
// factor, chi[2].
__device__ inline void hcb_binary(const DevParams& P, double density, double temperature, const double* Xk,
                                  Transport& tr, double* chi, double* D)
{
    const double pi = 3.1415926535897932;
    const double kB = P.k_B;
    const double m0 = P.molmass[0] * (kB / P.Runiv), m1 = P.molmass[1] * (kB / P.Runiv);
    double xxtr[2], yytr[2];
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
        double sumXk = 0.;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) sumXk += Xk[jj];
        xxtr[ii] = Xk[ii] + 1.e-15 * (sumXk / 2.0 - Xk[ii]);
    }
    double MWmix = 0.;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) MWmix += xxtr[ii] * P.molmass[ii];
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) yytr[ii] = P.molmass[ii] / MWmix * xxtr[ii];
    const double d0 = P.diameter[0], d1 = P.diameter[1];
    const double mx = 2.0 * m0 * m1 / (m0 + m1);
    const double dx = 0.5 * (d0 + d1);
    const double conc = xxtr[1], onc = xxtr[0];
    const double sm = (m0 + m1) * (m0 + m1), dm = (m0 - m1) * (m0 - m1);
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
    tr.eta = 1.016 * (1.0 + Zeta) / (Xeta + Yeta);
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
    tr.kappa = 1.025 * (1.0 + Zkappa) / (Xkappa + Ykappa);
    tr.zeta = 0.0;
    const double diff = 6.0 * 1.133 * etax / (5.0 * density);
    const double S0 = (m0 + m1) * kappax / (2.0 * m1 * kappa0) - 15 * (m1 - m0) / (8.0 * m0) - 1.0;
    const double S1 = (m0 + m1) * kappax / (2.0 * m0 * kappa1) - 15 * (m0 - m1) / (8.0 * m1) - 1.0;
    double thermal_diff = 1.299 * conc * onc * (S1 * conc - S0 * onc) / (6.0 * kappax * (Xkappa + Ykappa));
    const double mR = m0 / m1;
    const double d0t = mR * diff * yytr[1] * yytr[1] * yytr[1] / (yytr[0] * xxtr[1] * xxtr[1]);
    const double d1t = mR * diff * yytr[1] * yytr[1] / (xxtr[1] * xxtr[1]);
    D[0] = 1.0 * yytr[0] * d0t;
    D[1] = -1.0 * yytr[0] * d0t;
    D[2] = -1.0 * yytr[0] * d1t;
    D[3] = 1.0 * yytr[0] * d1t;
    thermal_diff = thermal_diff * mR * yytr[1] * yytr[1] / (xxtr[1] * xxtr[1]);
    chi[0] = -1.0 * xxtr[1] * xxtr[1] * thermal_diff / (mR * xxtr[0] * yytr[1] * yytr[1]);
    chi[1] = 1.0 * xxtr[1] * thermal_diff / (mR * yytr[1] * yytr[1]);
}

// D_GIO1 (compressible_functions.H:647-778; D_GIO at :513-646 is the same arithmetic): the
// three-term series of the multicomponent diffusion matrix, D[j*NS+i] = Diff_ij Ykp_i.
template <int NS>
__device__ __forceinline__ void d_gio_series(const double (&Dbin)[NS][NS], const double* Ykp, const double* Xkp,
                                             double* D)
{
    double Di[NS], Minv[NS], Mmat[NS];
    double Pm[NS][NS], Zm[NS][NS], PJ[NS][NS], m1[NS][NS], m2[NS][NS], Diff[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        double term2 = 0.;
#pragma unroll
        for (int j = 0; j < NS; ++j)
            if (j != i) term2 = term2 + Xkp[j] / Dbin[i][j];
        Di[i] = (1. - Ykp[i]) / term2;
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        Mmat[i] = Xkp[i] / Di[i];
        Minv[i] = Di[i] / Xkp[i];
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            Pm[i][j] = -Ykp[j];
            if (i == j) Pm[i][j] = Pm[i][j] + 1.;
        }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            double Delta;
            if (i == j) {
                double term1 = 0.;
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    if (k != i) term1 = term1 + Xkp[i] * Xkp[k] / Dbin[i][k];
                Delta = term1;
            } else {
                Delta = -Xkp[i] * Xkp[j] / Dbin[i][j];
            }
            Zm[i][j] = -Delta;
        }
#pragma unroll
    for (int i = 0; i < NS; ++i) Zm[i][i] = Zm[i][i] + Mmat[i];
    // Jmat = Minv_i * Zmat_ij, PJ = P * J
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            double s = 0.;
#pragma unroll
            for (int k = 0; k < NS; ++k) s = s + Pm[i][k] * (Minv[k] * Zm[k][j]);
            PJ[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            double scr = 0.;
#pragma unroll
            for (int k = 0; k < NS; ++k) scr = scr + Pm[i][k] * Minv[k] * Pm[j][k];
            m2[i][j] = scr;
            Diff[i][j] = scr;
        }
#pragma unroll
    for (int ll = 0; ll < 3; ++ll) {
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                double scr = 0.;
#pragma unroll
                for (int k = 0; k < NS; ++k) scr = scr + PJ[i][k] * Diff[k][j];
                m1[i][j] = scr + m2[i][j];
            }
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) Diff[i][j] = m1[i][j];
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) D[j * NS + i] = Diff[i][j] * Ykp[i];
}

// ---- transport_type 2: IdealMixtureTransportVW (compressible_functions.H:1047-1150) ----------
// The reference's pow(a, 2.) is evaluated as a*a (the correctly rounded square libm's pow
// returns) and pow(a, 3.) as (a*a)*a (within an ulp of libm's), the same sequence as the oracle.
// Matrices are flat with rows `stride` apart; the pivot rows are chosen at run time, so these
// arrays live in private memory.  The functions return 1 where the reference aborts with
// "singular matrix in decomp".
__device__ inline double vw_sq(double a) { return a * a; }
__device__ inline double vw_cube(double a) { return (a * a) * a; }

// Decomp / Decomp2 (H:221-320): LU with partial pivoting, 1-based row numbers in ip.
__device__ inline int vw_decomp(int neq, int stride, double* a, int* ip)
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

// Solve / Solve2 (H:322-402); the closing loop undoes the permutation cycle by cycle (each pass
// fixes one more row, so it ends after at most neq passes for any permutation in ip).
__device__ inline void vw_solve(int neq, int stride, const double* a, double* b, int* ip)
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

// Decompnp / Solvenp (H:404-458): without pivoting.
template <int N>
__device__ inline int vw_decomp_np(double* a)
{
    for (int i = 0; i < N - 1; ++i)
        for (int l = i + 1; l < N; ++l) {
            const double amult = a[l * N + i] / a[i * N + i];
            a[l * N + i] = amult;
            if (fabs(amult) < 1.e-60) continue;
            for (int k = i + 1; k < N; ++k) a[l * N + k] = a[l * N + k] - amult * a[i * N + k];
        }
    return fabs(a[(N - 1) * N + N - 1]) < 1.e-30;
}

template <int N>
__device__ inline void vw_solve_np(const double* a, double* b)
{
    for (int l = 1; l < N; ++l)
        for (int k = 0; k < l; ++k) b[l] = b[l] - a[l * N + k] * b[k];
    b[N - 1] = b[N - 1] / a[(N - 1) * N + N - 1];
    for (int l = 0; l < N - 1; ++l) {
        const int j = N - (l + 1);
        for (int k = j; k < N; ++k) b[j - 1] = b[j - 1] - a[(j - 1) * N + k] * b[k];
        b[j - 1] = b[j - 1] / a[(j - 1) * N + j - 1];
    }
}

// ViscLin (H:969-1045): the first Sonine approximation of the viscosity (HCB 7.4-56, 62, 63).
template <int NS>
__device__ inline int vw_visc_lin(const double* om, const double* Ykp, double T, double rho, const double* mk,
                                  double kB, double& eta)
{
    double nk[NS], bS[NS], diag[NS], Q[NS * NS];
    int ip[NS];
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
    eta = 0.5 * kB * T * s;
    return 0;
}

// LambdaLin (H:782-912): the 2NS Sonine system for conductivity and thermal diffusion (HCB
// 7.4-50..54), lambda' (7.4-33) and lambda (7.4-65).
template <int NS>
__device__ inline int vw_lambda_lin(const double* Dbin, const double* om, const double* Ykp, double T, double rho,
                                    const double* mk, double kB, double& lam)
{
    constexpr int S = 2 * NS;
    double nk[NS], beta[NS], d1[NS], d2[NS], d3[NS], d4[NS], DT[NS], QQ[S * S], aS[S];
    int ip[S];
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
            // the off-diagonal [11] term as the reference writes it: m_j throughout
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
    lam = lp - 0.5 * kB / ntotal * s1;
    return 0;
}

// ThermalDiff (H:914-967): Valk's thermal diffusion ratios into chi.
template <int NS>
__device__ inline int vw_thermal_diff(const double* a1, const double* a2, const double* ab, const double* Xkp,
                                      double sqrtT, const double* mk, double* kT)
{
    double A[NS * NS], AA[NS], al[NS * NS];
    for (int i = 0; i < NS; ++i) {
        for (int j = 0; j < NS; ++j) A[i * NS + j] = Xkp[j] * a2[i * NS + j];
        AA[i] = 15. / 4.;
    }
    for (int i = 0; i < NS; ++i) {
        double s = 0.;
        for (int k = 0; k < NS; ++k) s = s + Xkp[k] * a1[i * NS + k];
        A[i * NS + i] = A[i * NS + i] + s;
    }
    if (vw_decomp_np<NS>(A)) return 1;
    vw_solve_np<NS>(A, AA);
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) al[i * NS + j] = ab[i * NS + j] * (AA[i] / mk[i] - AA[j] / mk[j]) / sqrtT;
    for (int k = 0; k < NS; ++k) {
        double s = 0.;
        for (int j = 0; j < NS; ++j) s = s + Xkp[j] * al[k * NS + j];
        kT[k] = s;
    }
    return 0;
}

// IdealMixtureTransportVW (H:1047-1150).  Xk: mole fractions of the clamped mass fractions
// (transCoeffs.cpp:60-61).  D[j*NS+i] before the density factor.  Returns 1 on a singular matrix.
template <int NS>
__device__ inline int vw_transport(const DevParams& P, double density, double T, double pressure, const double* Xk,
                                   Transport& tr, double* chi, double* D)
{
    const double kB = P.k_B, pi = 3.1415926535897932;
    double mk[NS], xx[NS], yy[NS];
    double Dflat[NS * NS], om[NS * NS], a1[NS * NS], a2[NS * NS], ab[NS * NS], Dbin[NS][NS];
    for (int n = 0; n < NS; ++n) mk[n] = P.molmass[n] * (kB / P.Runiv);
    for (int ii = 0; ii < NS; ++ii) {
        double sumXk = 0.;
        for (int jj = 0; jj < NS; ++jj) sumXk += Xk[jj];
        xx[ii] = Xk[ii] + 1.e-15 * (sumXk / NS - Xk[ii]);
    }
    double MWmix = 0.;
    for (int ii = 0; ii < NS; ++ii) MWmix += xx[ii] * P.molmass[ii];
    for (int ii = 0; ii < NS; ++ii) yy[ii] = P.molmass[ii] / MWmix * xx[ii];
    const double Fijstar = -27. / 5.;
    const double sqrtT = sqrt(T);
    tr.eta = tr.kappa = 0.;  // what a singular solve leaves (the reference aborts there)
    for (int n = 0; n < NS; ++n) chi[n] = 0.;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            const double diamat = 0.5 * (P.diameter[i] + P.diameter[j]);
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
    int bad = vw_visc_lin<NS>(om, yy, T, density, mk, kB, tr.eta);
    tr.zeta = 0.;
    bad |= vw_lambda_lin<NS>(Dflat, om, yy, T, density, mk, kB, tr.kappa);
    d_gio_series<NS>(Dbin, yy, xx, D);  // D_GIO (H:513-646)
    bad |= vw_thermal_diff<NS>(a1, a2, ab, xx, sqrtT, mk, chi);
    return bad;
}

// transport_type 2 for one cell (transCoeffs.cpp:37-88 with :66-69).  Returns flags: bit 0 a mass
// fraction outside (0,1), bit 1 a singular matrix in the Valk-Waldmann solves (the reference
// aborts on either).
template <int NS>
__device__ inline int transport_cell_vw(const DevParams& P, double rho, double T, double p, const double* Yraw,
                                        Transport& tr, double* chi, double* D)
{
    double Yk[NS], Xf[NS];
    double sumYk = 0.;
    int flags = 0;
    for (int n = 0; n < NS; ++n) {
        const double y = Yraw[n];
        if ((y <= 0.0) | (y >= 1.0)) flags = 1;
        Yk[n] = clamp01(y);
        sumYk += Yk[n];
    }
    for (int n = 0; n < NS; ++n) Yk[n] /= sumYk;
    double molmix = 0.;
    for (int n = 0; n < NS; ++n) molmix += Yk[n] / P.molmass[n];
    molmix = 1. / molmix;
    for (int n = 0; n < NS; ++n) Xf[n] = Yk[n] * molmix / P.molmass[n];
    if (vw_transport<NS>(P, rho, T, p, Xf, tr, chi, D)) flags |= 2;
    for (int n = 0; n < NS * NS; ++n) D[n] *= rho;
    if (P.visc_type == 3 || P.visc_type == -3) tr.zeta = P.zeta_ratio * tr.eta;
    return flags;
}

// Transport coefficients of one cell (transCoeffs.cpp:37-88).  chi[NS], D[NS*NS] out.
// Returns true if a mass fraction lies outside (0,1) (the reference aborts there).
template <int NS>
__device__ inline bool transport_cell(const DevParams& P, double rho, double T, double p,
                                      const double* Yraw, Transport& tr, double* chi, double* D)
{
    double Yk[NS];
    double sumYk = 0.;
    bool bad = false;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        double y = Yraw[n];
        bad |= (y <= 0.0) | (y >= 1.0);
        Yk[n] = clamp01(y);
        sumYk += Yk[n];
    }
#pragma unroll
    for (int n = 0; n < NS; ++n) Yk[n] /= sumYk;

    if constexpr (NS == 2) {
        if (P.transport_type == 3) {  // Hirschfelder-Curtiss-Bird, binary (transCoeffs.cpp:71-75)
            double molmix = 0., Xf[2];
#pragma unroll
            for (int n = 0; n < 2; ++n) molmix += Yk[n] / P.molmass[n];
            molmix = 1. / molmix;
#pragma unroll
            for (int n = 0; n < 2; ++n) Xf[n] = Yk[n] * molmix / P.molmass[n];
            hcb_binary(P, rho, T, Xf, tr, chi, D);
#pragma unroll
            for (int n = 0; n < 4; ++n) D[n] *= rho;
            if (P.visc_type == 3 || P.visc_type == -3) tr.zeta = P.zeta_ratio * tr.eta;
            return bad;
        }
    }

    const double pi = 3.1415926535897932;
    double mm[NS], Xk[NS], Xkp[NS], Ykp[NS], eta1[NS], rhs[NS], Bi[NS];
    double mu[NS][NS], diam[NS][NS], Dbin[NS][NS], amat[NS][NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) mm[n] = P.molmass[n] * (P.k_B / P.Runiv);
    double mbar = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) mbar += Yk[n] / mm[n];
    mbar = 1.0 / mbar;
    double Xksum = 0.0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        Xk[n] = Yk[n] * mbar / mm[n];
        Xksum += Xk[n];
    }
#pragma unroll
    for (int n = 0; n < NS; ++n) Xkp[n] = Xk[n] + 1.0e-16 * (Xksum / double(NS) - Xk[n]);
#pragma unroll
    for (int n = 0; n < NS; ++n) Ykp[n] = Xkp[n] * mm[n] / mbar;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            diam[i][j] = 0.5 * (P.diameter[i] + P.diameter[j]);
            mu[i][j] = mm[i] * mm[j] / (mm[i] + mm[j]);
        }
    const double kT = P.k_B * T;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j)
            Dbin[i][j] = (3.0 / 16.0) * sqrt(2.0 * pi * kT * kT * kT / mu[i][j]) /
                         (p * pi * diam[i][j] * diam[i][j]);
    const double AKL = 1.0, BKL = 1.0, CKL = 1.0;

    // viscosity
#pragma unroll
    for (int i = 0; i < NS; ++i)
        eta1[i] = 5.0 / (16.0 * diam[i][i] * diam[i][i]) * sqrt(mm[i] * P.k_B * T / pi);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        rhs[i] = Xkp[i];
        Bi[i] = rhs[i];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = Xkp[i] * Xkp[i] / eta1[i];
#pragma unroll
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
    gio_solve<NS>(amat, Bi);
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) e += Xkp[i] * Bi[i];
    tr.eta = e;
    tr.zeta = 0.;

    // thermal conductivity
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        rhs[i] = 2.5 * Xkp[i];
        Bi[i] = rhs[i];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = Xkp[i] * Xkp[i] / Dbin[i][j] * 2.0 * AKL;
#pragma unroll
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
    gio_solve<NS>(amat, Bi);
    double kap = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) kap += rhs[i] * Bi[i];
    kap *= p / T;
    tr.kappa = kap;

    // thermal diffusion ratios
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (i == j) {
                amat[i][i] = 0.0;
#pragma unroll
                for (int k = 0; k < NS; ++k)
                    if (k != i)
                        amat[i][i] -= 0.5 * Xkp[k] / Dbin[i][k] * mm[k] / (mm[i] + mm[k]) *
                                      (6.0 * CKL - 5.0);
            } else {
                amat[i][j] = 0.5 * Xkp[j] / Dbin[i][j] * mm[i] / (mm[i] + mm[j]) * (6.0 * CKL - 5.0);
            }
        }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        double c = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) c += amat[i][j] * Bi[j];
        chi[i] = c;
    }

    d_gio_series<NS>(Dbin, Ykp, Xkp, D);  // D_GIO1 (H:647-778)
#pragma unroll
    for (int n = 0; n < NS * NS; ++n) D[n] *= rho;
    if (P.visc_type == 3 || P.visc_type == -3) tr.zeta = P.zeta_ratio * tr.eta;
    return bad;
}

// CholeskyDecomp (compressible_functions.H:460-509): lower factor in sqda (row-major NSxNS).
// Entries the reference writes above the diagonal never feed the result and are skipped.
template <int NS>
__device__ inline void cholesky(double* a, double* sqda)
{
    double p[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        bool ising = false;
#pragma unroll
        for (int j = i; j < NS; ++j) {
            double sum1 = a[i * NS + j];
#pragma unroll
            for (int k = i - 1; k >= 0; --k) sum1 = sum1 - a[i * NS + k] * a[j * NS + k];
            if (i == j) {
                if (sum1 <= 0.) {
                    p[i] = 0.;
                    ising = true;
                } else {
                    p[i] = sqrt(sum1);
                }
            }
            a[j * NS + i] = ising ? 0. : sum1 / p[i];
        }
    }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) sqda[i * NS + j] = (j < i) ? a[i * NS + j] : (j == i ? p[i] : 0.);
}

// ---- counter-based noise: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller ------------

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 o;
        o.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        o.y = (uint32_t)p1;
        o.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        o.w = (uint32_t)p0;
        c = o;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Box-Muller with its own elementary functions, written with IEEE +, -, *, /, fma and sqrt only,
// so that the oracle (oracle.c:orc_normal_pair) evaluates the identical operation sequence and
// gets the same bits on the host.  Accuracy: both series are truncated below 2e-18 relative.

// ln x for x in (0, 1]: x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln x = e ln2 + 2 atanh(s),
// s = (m-1)/(m+1) (|s| <= 0.1716), atanh(s)/s = sum_k s^(2k)/(2k+1), k = 0..11.
__device__ inline double bm_log(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
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
    // ln2 = LN2_HI + LN2_LO; LN2_HI has 21 significant bits, so de * LN2_HI is exact
    return de * 0.6931467056274414 + fma(de, 4.7493250390316726e-07, lm);
}

// sin(pi t) and cos(pi t) for t in [0, 2): t = n/2 + r with |r| <= 1/4 (exact), Taylor series of
// sin(pi r) to r^19 and cos(pi r) to r^18, then the quadrant n mod 4.
__device__ inline void bm_sincospi(double t, double& sn, double& cs)
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
    // quadrant q = n mod 4: (sin, cos) = (S, C), (C, -S), (-S, -C), (-C, S)
    const int q = ((int)n) & 3;
    const double a = (q & 1) ? C : S, b = (q & 1) ? S : C;
    sn = (q & 2) ? -a : a;
    cs = ((q + 1) & 2) ? -b : b;
}

// Two N(0,1) values from one Philox block: 53-bit u1 in (0,1], u2 in [0,1).
__device__ inline void normal_pair(U4 r, double& z0, double& z1)
{
    const double two_m53 = 1.1102230246251565e-16;
    uint64_t a = ((((uint64_t)r.x) << 32) | r.y) >> 11;
    uint64_t b = ((((uint64_t)r.z) << 32) | r.w) >> 11;
    double u1 = ((double)a + 0.5) * two_m53;
    double u2 = (double)b * two_m53;
    double rad = sqrt(-2.0 * bm_log(u1));
    double s, c;
    bm_sincospi(2.0 * u2, s, c);
    z0 = rad * c;
    z1 = rad * s;
}

// Noise keys: field = 16*dir + comp for face components 1..nvars-1, 48 for the node field,
// 64 + m for reaction m.
constexpr uint32_t NODE_FIELD = 48u;
constexpr uint32_t CHEM_FIELD = 64u;

struct NoiseKey {
    uint32_t k0, k1;     // seed
    uint32_t step_lo;    // step
    uint32_t field_hi;   // (step >> 32) << 8
    uint32_t instance;
};

// A/B pair of the noise value at wrapped linear index `lin` of `field`.
__device__ inline void noise_ab(const NoiseKey& K, uint32_t field, uint32_t lin, double& a,
                                double& b)
{
    U4 c{lin, K.step_lo, field | K.field_hi, K.instance};
    normal_pair(philox4x32_10(c, K.k0, K.k1), a, b);
}


// ---- chemistry source (chemistry_functions.cpp:128-193) -------------------------------------
//
// exp and pow are evaluated here (and identically in the oracle) with IEEE operations only, so
// host and device agree bit for bit.  e^x: x = k ln2 + r with |r| <= ln2/2, Taylor series of e^r
// to r^14 (truncation below 1e-19 relative), times 2^k.  x^y = e^(y ln x) with the Box-Muller
// logarithm above (valid for any positive normal x); y = 0 gives 1 and y = 1 gives x exactly.
__device__ inline double chem_exp(double x)
{
    if (x == 0.0) return 1.0;
    if (x > 709.0) return __longlong_as_double(0x7ff0000000000000ll);
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
    // times 2^k as two exact power-of-two factors (each within the normal range)
    const int ki = (int)k, k1 = ki / 2, k2 = ki - k1;
    const double s1 = __longlong_as_double((long long)(k1 + 1023) << 52);
    const double s2 = __longlong_as_double((long long)(k2 + 1023) << 52);
    return p * s1 * s2;
}

__device__ inline double chem_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    if (x == 0.0) return y > 0.0 ? 0.0 : __longlong_as_double(0x7ff0000000000000ll);
    if (x < 0.0) return __longlong_as_double(0x7ff8000000000000ll);
    return chem_exp(y * bm_log(x));
}

// c^s for an integer stoichiometric coefficient s >= 0 (pow(ck, stoich) of :167).
__device__ inline double chem_ipow(double c, int s)
{
    if (s == 0) return 1.0;
    double r = c;
    for (int q = 1; q < s; ++q) r = r * c;
    return r;
}

// Species source of one cell (chemistry_functions.cpp:147-190), operation order kept: T, p
// and X_n of the cell, ranchem(m) the weighted noise of reaction m, dV the cell volume.
template <int NS, typename Noise>
__device__ inline void chem_source_cell(const DevParams& P, double T, double pres, const double* X,
                                        Noise ranchem, double dt, double dV, double* src)
{
    const DevChem& C = P.chem;
    const double T0 = C.T0;
    const double ctot = pres / P.Runiv / T;
    const double Navo = C.Navo;
    double ck[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) ck[n] = X[n] * ctot;
#pragma unroll
    for (int n = 0; n < NS; ++n) src[n] = 0.;
    for (int m = 0; m < C.nreaction; ++m) {
        double rate = C.rate[m];
        rate *= chem_exp(-C.alpha[m] / P.Runiv * (1 / T - 1 / T0));
        rate *= chem_pow(T / T0, C.beta[m]);
#pragma unroll
        for (int n = 0; n < NS; ++n) rate *= chem_ipow(ck[n], C.R[m][n]);
        rate = (0. < rate) ? rate : 0.;  // std::max(0., rate)
        const double W = ranchem(m) / sqrt(dt * dV);
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            src[n] += P.molmass[n] * C.PR[m][n] * rate;
            if (C.reaction_type == 1) src[n] += P.molmass[n] * C.PR[m][n] * sqrt(rate / Navo) * W;
        }
    }
}

// The weighted reaction noise A + w B of reaction m at linear cell index lin (as the stochastic
// fluxes: MultiFabFillRandom with variance 1, then LinComb, timeStep.cpp:93-97, 135-140).
__device__ inline double chem_noise(const NoiseKey& K, int m, uint32_t lin, double w)
{
    double za, zb;
    noise_ab(K, CHEM_FIELD + (uint32_t)m, lin, za, zb);
    return 1.0 * za + w * zb;
}

}  // namespace fhdx
