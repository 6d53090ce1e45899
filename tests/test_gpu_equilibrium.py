"""Equilibrium fluctuations of the noisy RK3 step against fluctuation theory.

A periodic argon box started uniform reaches thermal equilibrium; the variances of the cell
values of rho, j_x, j_y, j_z and rho E (over cells and snapshots) must match

    Var rho   = rho m / dV
    Var j_d   = rho k_B T / dV
    Var rhoE  = (15/4) rho (k_B T)^2 / (m dV)       (monatomic, e0 = 0)

with m = M k_B / R the molecular mass and T the mean temperature of the sampled states (the
thermal kinetic energy the fluctuations take up lowers T by ~1/N of a cell's molecules).  These
are the linearised (Gaussian) values; they hold when a cell holds many molecules: 1.4e7 argon
atoms at dx = 8e-5 cm, 1.4e4 at the 32^3 argon deck's dx = 8e-6 cm.  At 8e-5 cm the bar is 1%.  At
8e-6 cm the CFL-0.1 step has a 10x larger diffusive number, and Var(rho E) then carries a
time-step bias of the scheme, second order in dt: +1.1% at dt, +0.3% at dt/2, 0.0% at dt/4.
That case runs at dt (held to 1.5%) and at dt/2 (held to 1%, with the rho E excess at least
halved).
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def theory(cfg, T):
    p = cfg.params
    kB, R = p.k_B, p.Runiv
    m = p.molmass[0] * kB / R
    dV = p.dx[0] * p.dx[1] * p.dx[2]
    rho = cfg.rho0
    return np.array([rho * m / dV, rho * kB * T / dV, rho * kB * T / dV, rho * kB * T / dV,
                     3.75 * rho * (kB * T) ** 2 / (m * dV)])


def sample(cfg, seed, burn, nsnap, every):
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(E.uniform_cons(cfg)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    g = p.ngc
    n0, n1, n2 = p.n_cells
    step = 0
    for _ in range(burn):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, step)
        step += 1
    acc = torch.zeros(5, dtype=torch.float64, device="cuda")
    tsum = torch.zeros((), dtype=torch.float64, device="cuda")
    for _ in range(nsnap):
        for _ in range(every):
            fc.rk3_step(ctx, cu, prim, cfg.dt, seed, step)
            step += 1
        v = cu.t[:5, g:g + n2, g:g + n1, g:g + n0].reshape(5, -1)
        acc += v.var(dim=1)
        tsum += prim.t[4, g:g + n2, g:g + n1, g:g + n0].mean()
    ctx.check()
    return (acc / nsnap).cpu().numpy(), float(tsum.cpu()) / nsnap


def argon(dx, n):
    cfg = P.argon_gas(n)
    p = cfg.params
    for d in range(3):
        p.dx[d] = dx
    cfg.dt = P.cfl_dt(p, cfg.rhobar, cfg.T_init, P.ARGON_CFL)
    return cfg


def measure(cfg, div=1):
    cfg.dt = cfg.dt / div
    say(f"case equilibrium argon {tuple(cfg.params.n_cells)} dx={cfg.params.dx[0]} dt={cfg.dt:.3e}")
    got, T = sample(cfg, seed=2, burn=2000 * div, nsnap=100, every=20 * div)
    want = theory(cfg, T)
    say(f"  mean T {T:.4f} K (start {cfg.T_init} K)")
    ratio = got / want
    for name, g, w, r in zip(["rho", "jx", "jy", "jz", "rhoE"], got, want, ratio):
        say(f"  Var {name}: {g:.5e} theory {w:.5e} ratio {r:.4f}")
    return ratio


def test_equilibrium_variances():
    ratio = measure(argon(8e-5, 32))
    assert np.all(np.abs(ratio - 1) < 0.01), ratio


def test_equilibrium_variances_c1_cells_and_dt_bias():
    """At the 32^3 argon deck's 8e-6 cm cells: rho and j within 1 % at the CFL-0.1 dt; Var rho E
    carries the scheme's time-step bias (the diffusion number is 100x that at 8e-5 cm), which
    at least halves when dt is halved (second order in dt), and at dt/2 all five are within 1 %."""
    r1 = measure(argon(8e-6, 32))
    r2 = measure(argon(8e-6, 32), div=2)
    assert np.all(np.abs(r1[:4] - 1) < 0.01), r1
    assert abs(r1[4] - 1) < 0.015, r1
    assert np.all(np.abs(r2 - 1) < 0.01), r2
    assert abs(r2[4] - 1) <= 0.5 * abs(r1[4] - 1), (r1[4], r2[4])


def test_mixture_equilibrium_variances():
    """Four species (the He/Ne/Ar/Kr mixture on a 32^3 box of 3.2e-5 cm cells, dt 1.6e-11 s): at
    equilibrium each partial density fluctuates independently, Var rho_k = rho_k m_k / dV with
    m_k = M_k k_B / R; Var rho is their sum; Var j_d = rho k_B T / dV; Var rho E =
    sum_k e_k^2 Var rho_k + rho c_v k_B T^2 / dV with e_k = e0_k + c_v,k T.  All nine within 1 %.
    The cells hold ~28000 molecules of the heaviest species: with few molecules per cell (~440
    at the 8e-6 cm cells of the 128^3 benchmark) the reference scheme's variances fall short of
    these values, a finite-number effect the oracle shows on the CPU
    (test_oracle_stats.py::test_mixture_variance_shortfall_is_a_finite_number_effect)."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    cfg = P.four_species(32)
    cfg.params = P.make_params(4, (32, 32, 32), 3.2e-5, [4.0026, 20.1797, 39.9480, 83.8000],
                               [2.18e-8, 2.58e-8, 3.63e-8, 4.16e-8])
    cfg.dt = 1.6e-11
    p = cfg.params
    say(f"case equilibrium four species 32^3 dx=3.2e-5 dt={cfg.dt:.1e}")
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(E.uniform_cons(cfg)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    g = p.ngc
    n0, n1, n2 = p.n_cells
    step = 0
    for _ in range(2000):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 9, step)
        step += 1
    acc = torch.zeros(p.nvars, dtype=torch.float64, device="cuda")
    tsum = torch.zeros((), dtype=torch.float64, device="cuda")
    nsnap = 100
    for _ in range(nsnap):
        for _ in range(20):
            fc.rk3_step(ctx, cu, prim, cfg.dt, 9, step)
            step += 1
        v = cu.t[:, g:g + n2, g:g + n1, g:g + n0].reshape(p.nvars, -1)
        acc += v.var(dim=1)
        tsum += prim.t[4, g:g + n2, g:g + n1, g:g + n0].mean()
    ctx.check()
    got = (acc / nsnap).cpu().numpy()
    T = float(tsum.cpu()) / nsnap
    kB, R = p.k_B, p.Runiv
    dV = p.dx[0] * p.dx[1] * p.dx[2]
    rho = cfg.rho0
    ns = p.nspecies
    rk = [rho * cfg.rhobar[k] for k in range(ns)]
    mk = [p.molmass[k] * kB / R for k in range(ns)]
    var_k = [rk[k] * mk[k] / dV for k in range(ns)]
    cv = sum(cfg.rhobar[k] * p.hcv[k] for k in range(ns))
    ek = [p.e0[k] + p.hcv[k] * T for k in range(ns)]
    want = np.array([sum(var_k)] + [rho * kB * T / dV] * 3 +
                    [sum(ek[k] ** 2 * var_k[k] for k in range(ns)) + rho * cv * kB * T * T / dV] + var_k)
    names = ["rho", "jx", "jy", "jz", "rhoE"] + [f"rhoY{k}" for k in range(ns)]
    ratio = got / want
    say(f"  mean T {T:.4f} K")
    for name, gv, w, r in zip(names, got, want, ratio):
        say(f"  Var {name}: {gv:.5e} theory {w:.5e} ratio {r:.4f}")
    assert np.all(np.abs(ratio - 1) < 0.01), ratio
