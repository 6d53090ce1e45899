// common.hpp -- host-side plumbing shared by the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "physics.hpp"

namespace fhdx {

// Sets the thread-local error text returned by fhdx_last_error().
void set_error(const std::string& msg);

#define FHDX_HIP_TRY(expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::fhdx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return FHDX_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

inline Arr make_arr(const fhdx_fab& f)
{
    Arr a;
    a.p = f.p;
    a.lo0 = f.lo[0];
    a.lo1 = f.lo[1];
    a.lo2 = f.lo[2];
    int nx = f.hi[0] - f.lo[0] + 1, ny = f.hi[1] - f.lo[1] + 1, nz = f.hi[2] - f.lo[2] + 1;
    a.sy = nx;
    a.sz = nx * ny;
    a.sn = (long)nx * ny * nz;
    return a;
}

inline DevParams make_dev_params(const fhdx_params& p)
{
    DevParams d{};
    d.ns = p.nspecies;
    d.nvars = 5 + p.nspecies;
    d.nprim = 6 + 2 * p.nspecies;
    d.ngc = p.ngc;
    for (int i = 0; i < 3; ++i) {
        d.n[i] = p.n_cells[i];
        d.bo[i] = 0;
        d.bn[i] = p.n_cells[i];
        d.dx[i] = p.dx[i];
        d.grav[i] = p.grav[i];
    }
    d.algorithm_type = p.algorithm_type;
    d.advection_type = p.advection_type;
    d.visc_type = p.visc_type;
    d.stoch_on = (p.stoch_stress_form == 1);
    d.transport_type = p.transport_type;
    for (int n = 0; n < MS; ++n) {
        d.molmass[n] = p.molmass[n];
        d.diameter[n] = p.diameter[n];
        d.hcv[n] = p.hcv[n];
        d.hcp[n] = p.hcp[n];
        d.e0[n] = p.e0[n];
    }
    d.k_B = p.k_B;
    d.Runiv = p.Runiv;
    d.zeta_ratio = p.zeta_ratio;
    d.var_mom = std::sqrt(p.variance_coef_mom * p.variance_coef_mom);
    d.var_ener = std::sqrt(p.variance_coef_ener * p.variance_coef_ener);
    d.var_mass = std::sqrt(p.variance_coef_mass * p.variance_coef_mass);
    d.any_wall = 0;
    for (int i = 0; i < 3; ++i) {
        d.periodic[i] = (p.bc_vel_lo[i] == -1);
        if (!d.periodic[i]) d.any_wall = 1;
        for (int side = 0; side < 2; ++side) {
            const int f = 2 * i + side;
            d.bcv[f] = side ? p.bc_vel_hi[i] : p.bc_vel_lo[i];
            d.bct[f] = side ? p.bc_therm_hi[i] : p.bc_therm_lo[i];
            d.bcm[f] = side ? p.bc_mass_hi[i] : p.bc_mass_lo[i];
            d.wall[f] = (d.bcm[f] != 3 && d.bcm[f] != -1);
            d.tw[f] = side ? p.t_hi[i] : p.t_lo[i];
            d.pw[f] = side ? p.p_hi[i] : p.p_lo[i];
            d.rhow[f] = side ? p.rho_hi[i] : p.rho_lo[i];
            for (int n = 0; n < MS; ++n) {
                d.Yw[f][n] = p.bc_Yk[f][n];
                d.Xw[f][n] = p.bc_Xk[f][n];
            }
        }
    }
    return d;
}

inline NoiseKey make_noise_key(uint64_t seed, uint64_t step, uint32_t instance)
{
    NoiseKey k;
    k.k0 = (uint32_t)seed;
    k.k1 = (uint32_t)(seed >> 32);
    k.step_lo = (uint32_t)step;
    k.field_hi = ((uint32_t)(step >> 32)) << 8;
    k.instance = instance;
    return k;
}

// RK3 stochastic weights of the three stages (timeStep.cpp:109, 218, 323).
inline void stage_weights(double w[3])
{
    const double sq2 = std::sqrt(2.0), sq3 = std::sqrt(3.0);
    w[0] = (2.0 * sq2 + 1.0 * sq3) / 5.0;
    w[1] = (-4.0 * sq2 + 3.0 * sq3) / 5.0;
    w[2] = (1.0 * sq2 - 2.0 * sq3) / 10.0;
}

inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }

// Boxes the fused kernels serve: at least 4 and at least 2 ngc cells per direction, so every
// ghost cell is the image of exactly one valid cell and the tiles cover whole planes.
inline bool fused_box_ok(const fhdx_params& p)
{
    for (int d = 0; d < 3; ++d)
        if (p.n_cells[d] < 4 || 2 * p.ngc > p.n_cells[d]) return false;
    return true;
}

}  // namespace fhdx
