// stats.hip -- evaluateStats (src_compressible/stats.cpp:5-363) on the device.
//
// Three passes over the valid box of one FAB (valid cells only, so any boundary conditions):
//   k_stats_means  running means of cons and the primitive variables of the mean state
//                  (stats.cpp:45-88), per cell;
//   k_stats_yz_*   the 19 yz-slice sums at every x (stats.cpp:156-210): each thread sums one
//                  contiguous part of the (j,k) pairs of one x index, then the parts are added
//                  in order (a fixed order: reproducible run to run; it differs from the
//                  reference's serial k-j order only by rounding); the cross_cell slice values
//                  miscVals (stats.cpp:96-150) are the same sums at x = cross_cell;
//   k_stats_vars   running variances and cross-correlations (stats.cpp:215-360), per cell.
// Per-cell arithmetic keeps the reference's operation order.
#include "kernels.hpp"

namespace fhdx {

constexpr int SB = 256;
constexpr int NSTAT = 19;

// mean cv of a cell: sum_l hcv_l * rhoY_l / rho over the means (stats.cpp:123-126, 190-193)
template <int NS>
__device__ inline double mean_cv(const DevParams& P, const Arr& cm, int i, int j, int k)
{
    double cv = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cv = cv + P.hcv[l] * cm(i, j, k, 5 + l) / cm(i, j, k, 0);
    return cv;
}

template <int NS>
__global__ void __launch_bounds__(SB) k_stats_means(DevParams P, Arr cu, Arr cm, Arr pm, double stepsminusone,
                                                    double stepsinv)
{
    const long t = (long)blockIdx.x * SB + threadIdx.x;
    const int n0 = P.n[0], n1 = P.n[1];
    if (t >= (long)n0 * n1 * P.n[2]) return;
    const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((long)n0 * n1));
    constexpr int nvars = NS + 5;
    double m[nvars];
#pragma unroll
    for (int l = 0; l < nvars; ++l) {
        m[l] = (cm(i, j, k, l) * stepsminusone + cu(i, j, k, l)) * stepsinv;
        cm(i, j, k, l) = m[l];
    }
    const double densitymeaninv = 1.0 / m[0];
    double fracvec[NS];
#pragma unroll
    for (int l = 5; l < nvars; ++l) fracvec[l - 5] = m[l] * densitymeaninv;
    const double p1 = m[1] * densitymeaninv, p2 = m[2] * densitymeaninv, p3 = m[3] * densitymeaninv;
    const double vsqr = p1 * p1 + p2 * p2 + p3 * p3;
    const double intenergy = m[4] / m[0] - 0.5 * vsqr;
    const double T = eos_temperature<NS>(P, intenergy, fracvec);
    pm(i, j, k, 0) = m[0];
    pm(i, j, k, 1) = p1;
    pm(i, j, k, 2) = p2;
    pm(i, j, k, 3) = p3;
    pm(i, j, k, 4) = T;
    pm(i, j, k, 5) = eos_pressure<NS>(P, fracvec, m[0], T);
}

// Part p of the yz sums at x index i: the contiguous range of (j,k) pairs (q = j + n1 k)
// [nyz p / nparts, nyz (p+1) / nparts), summed in q order.  64 threads of a block cover 64
// consecutive i (coalesced loads), 4 parts per block.
template <int NS>
__global__ void __launch_bounds__(SB) k_stats_yz_part(DevParams P, Arr cu, Arr cm, Arr prim, Arr pm, double* part,
                                                      int nparts)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int p = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int n0 = P.n[0], n1 = P.n[1];
    if (i >= n0 || p >= nparts) return;
    const long nyz = (long)n1 * P.n[2];
    const long q0 = nyz * p / nparts, q1 = nyz * (p + 1) / nparts;
    double s[NSTAT];
#pragma unroll
    for (int q = 0; q < NSTAT; ++q) s[q] = 0.;
    for (long q = q0; q < q1; ++q) {
        const int j = (int)(q % n1), k = (int)(q / n1);
        s[0] += cu(i, j, k, 0);
        s[1] += cm(i, j, k, 0);
        s[2] += cu(i, j, k, 4);
        s[3] += cm(i, j, k, 4);
        s[4] += cu(i, j, k, 1);
        s[5] += cm(i, j, k, 1);
        s[6] += cu(i, j, k, 2);
        s[7] += cm(i, j, k, 2);
        s[8] += cu(i, j, k, 3);
        s[9] += cm(i, j, k, 3);
        s[10] += prim(i, j, k, 1);
        s[11] += pm(i, j, k, 1);
        s[12] += prim(i, j, k, 2);
        s[13] += pm(i, j, k, 2);
        s[14] += prim(i, j, k, 3);
        s[15] += pm(i, j, k, 3);
        s[16] += mean_cv<NS>(P, cm, i, j, k);
        s[17] += prim(i, j, k, 4);
        s[18] += pm(i, j, k, 4);
    }
    double* out = part + ((long)p * n0 + i) * NSTAT;
#pragma unroll
    for (int q = 0; q < NSTAT; ++q) out[q] = s[q];
}

// yz[i][s] = (sum over parts in order) / nyz
__global__ void __launch_bounds__(SB) k_stats_yz_sum(const double* part, int nparts, int n0, long nyz, double* yz)
{
    const long t = (long)blockIdx.x * SB + threadIdx.x;
    if (t >= (long)n0 * NSTAT) return;
    double sum = 0.;
    for (int p = 0; p < nparts; ++p) sum += part[(long)p * n0 * NSTAT + t];
    yz[t] = sum / (double)nyz;
}

// miscVals (stats.cpp:96-150) are yz sums of the cross_cell slice: index of each in the yz row.
__constant__ int c_misc_from_yz[17] = {5, 4, 11, 1, 0, 10, 2, 3, 6, 7, 8, 9, 16, 18, 13, 15, 17};

__global__ void k_stats_misc(const double* yz, int cross_cell, double* misc)
{
    const int q = threadIdx.x;
    if (q < 17) misc[q] = yz[(long)cross_cell * NSTAT + c_misc_from_yz[q]];
    else if (q < 20) misc[q] = 0.;
}

template <int NS>
__global__ void __launch_bounds__(SB) k_stats_vars(DevParams P, Arr cu, Arr cm, Arr cv, Arr prim, Arr pm, Arr pv,
                                                   Arr sc, Arr ms, const double* yz, const double* miscVals,
                                                   double stepsminusone, double stepsinv)
{
    const long t = (long)blockIdx.x * SB + threadIdx.x;
    const int n0 = P.n[0], n1 = P.n[1];
    if (t >= (long)n0 * n1 * P.n[2]) return;
    const int i = (int)(t % n0), j = (int)((t / n0) % n1), k = (int)(t / ((long)n0 * n1));
    const int nprimvars = P.nprim;
    const double* Y = yz + (long)i * NSTAT;
    const double densitymeaninv = 1.0 / cm(i, j, k, 0);
    const double delrho = cu(i, j, k, 0) - cm(i, j, k, 0);
    const double delpx = cu(i, j, k, 1) - cm(i, j, k, 1);
    const double delpy = cu(i, j, k, 2) - cm(i, j, k, 2);
    const double delpz = cu(i, j, k, 3) - cm(i, j, k, 3);
    const double delenergy = cu(i, j, k, 4) - cm(i, j, k, 4);
    const double delrhoS = Y[0] - Y[1];
    const double delES = Y[3] - Y[2];
    const double delrhoSstar = miscVals[4] - miscVals[3];
    const double delpxSstar = miscVals[1] - miscVals[0];
    const double cv0 = (cv(i, j, k, 0) * stepsminusone + delrho * delrho) * stepsinv;
    cv(i, j, k, 0) = cv0;
    cv(i, j, k, 1) = (cv(i, j, k, 1) * stepsminusone + delpx * delpx) * stepsinv;
    cv(i, j, k, 2) = (cv(i, j, k, 2) * stepsminusone + delpy * delpy) * stepsinv;
    cv(i, j, k, 3) = (cv(i, j, k, 3) * stepsminusone + delpz * delpz) * stepsinv;
    cv(i, j, k, 4) = (cv(i, j, k, 4) * stepsminusone + delenergy * delenergy) * stepsinv;
    const double pm1 = pm(i, j, k, 1), pm2 = pm(i, j, k, 2), pm3 = pm(i, j, k, 3);
    const double delvelx = (delpx - pm1 * delrho) * densitymeaninv;
    const double delvely = (delpy - pm2 * delrho) * densitymeaninv;
    const double delvelz = (delpz - pm3 * delrho) * densitymeaninv;
    pv(i, j, k, 0) = cv0;
    pv(i, j, k, 1) = (pv(i, j, k, 1) * stepsminusone + delvelx * delvelx) * stepsinv;
    pv(i, j, k, 2) = (pv(i, j, k, 2) * stepsminusone + delvely * delvely) * stepsinv;
    pv(i, j, k, 3) = (pv(i, j, k, 3) * stepsminusone + delvelz * delvelz) * stepsinv;
    const double delg = pm1 * delpx + pm2 * delpy + pm3 * delpz;
    pv(i, j, k, nprimvars) = (pv(i, j, k, nprimvars) * stepsminusone + delg * delg) * stepsinv;
    pv(i, j, k, nprimvars + 1) = (pv(i, j, k, nprimvars + 1) * stepsminusone + delg * delenergy) * stepsinv;
    pv(i, j, k, nprimvars + 2) = (pv(i, j, k, nprimvars + 2) * stepsminusone + delrho * delenergy) * stepsinv;
    pv(i, j, k, nprimvars + 3) = (pv(i, j, k, nprimvars + 3) * stepsminusone + delrho * delg) * stepsinv;
    const double delT = prim(i, j, k, 4) - pm(i, j, k, 4);
    pv(i, j, k, 4) = (pv(i, j, k, 4) * stepsminusone + delT * delT) * stepsinv;
    const double ms0 = (ms(i, j, k, 0) * stepsminusone + miscVals[1] * Y[0]) * stepsinv;
    ms(i, j, k, 0) = ms0;
    const double delpdelrho = ms0 - miscVals[0] * cm(i, j, k, 0);
    const double ms1 = (ms(i, j, k, 1) * stepsminusone + delrhoS * delrhoSstar) * stepsinv;
    const double ms2 = (ms(i, j, k, 2) * stepsminusone + miscVals[16] * Y[17]) * stepsinv;
    const double ms3 = (ms(i, j, k, 3) * stepsminusone + miscVals[16] * Y[0]) * stepsinv;
    const double ms4 = (ms(i, j, k, 4) * stepsminusone + delrhoS * delpxSstar) * stepsinv;
    const double ms5 = (ms(i, j, k, 5) * stepsminusone + delES * delrhoSstar) * stepsinv;
    ms(i, j, k, 1) = ms1;
    ms(i, j, k, 2) = ms2;
    ms(i, j, k, 3) = ms3;
    ms(i, j, k, 4) = ms4;
    ms(i, j, k, 5) = ms5;
    sc(i, j, k, 0) = miscVals[13];
    sc(i, j, k, 1) = Y[18];
    sc(i, j, k, 2) = ms2;
    sc(i, j, k, 3) = ms2 - Y[18] * miscVals[13];
    sc(i, j, k, 4) = ms3 - Y[1] * miscVals[13];
    sc(i, j, k, 5) = (miscVals[3] == 0.) ? 0. : (delpdelrho - miscVals[2] * ms1) / miscVals[3];
    sc(i, j, k, 6) = ms4;
    sc(i, j, k, 7) = ms5;
}

template <int NS>
static void stats_passes(const DevParams& P, const StatsArgs& a, hipStream_t s)
{
    const long ncell = (long)P.n[0] * P.n[1] * P.n[2];
    const double stepsminusone = a.steps - 1.;
    const double stepsinv = 1. / a.steps;
    k_stats_means<NS><<<grid_for(ncell, SB), SB, 0, s>>>(P, a.cu, a.cm, a.pm, stepsminusone, stepsinv);
    const long nyz = (long)P.n[1] * P.n[2];
    dim3 g((P.n[0] + 63) / 64, (a.nparts + 3) / 4);
    k_stats_yz_part<NS><<<g, SB, 0, s>>>(P, a.cu, a.cm, a.prim, a.pm, a.part, a.nparts);
    k_stats_yz_sum<<<grid_for((long)P.n[0] * NSTAT, SB), SB, 0, s>>>(a.part, a.nparts, P.n[0], nyz, a.yz);
    k_stats_misc<<<1, 32, 0, s>>>(a.yz, a.cross_cell, a.misc);
    k_stats_vars<NS><<<grid_for(ncell, SB), SB, 0, s>>>(P, a.cu, a.cm, a.cv, a.prim, a.pm, a.pv, a.sc, a.ms, a.yz,
                                                        a.misc, stepsminusone, stepsinv);
}

int stats_launch(const DevParams& P, const StatsArgs& a, hipStream_t s)
{
    switch (P.ns) {
        case 2: stats_passes<2>(P, a, s); break;
        case 3: stats_passes<3>(P, a, s); break;
        case 4: stats_passes<4>(P, a, s); break;
        case 5: stats_passes<5>(P, a, s); break;
        case 6: stats_passes<6>(P, a, s); break;
        case 7: stats_passes<7>(P, a, s); break;
        case 8: stats_passes<8>(P, a, s); break;
        default: return FHDX_ERR_UNSUPPORTED;
    }
    FHDX_HIP_TRY(hipGetLastError());
    return FHDX_OK;
}

}  // namespace fhdx
