"""The benchmark shapes themselves, checked through properties that do not depend on size.

* argon 256^3 and four species 128^3, 200 noisy steps each from the uniform start: every flux
  (hyperbolic, viscous, heat, species, stochastic) enters in divergence form on a periodic box,
  so the totals of rho, rho*u, rho*E and rho*Y_k stay at their start values up to rounding;
  the state stays finite and the error flags stay clear.
* four species 128^3: the fused step equals the reference-shaped step bit for bit.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def totals(torch, cu, p):
    g = p.ngc
    v = cu.t[:, g:-g, g:-g, g:-g]
    return v.sum(dim=(1, 2, 3)).cpu().numpy(), v.abs().sum(dim=(1, 2, 3)).cpu().numpy()


@pytest.mark.parametrize("name,mk", [("argon 256^3", lambda: P.argon_gas(256, 200)),
                                     ("four species 128^3", lambda: P.four_species(128, 200))])
def test_totals_conserved_over_200_steps(name, mk):
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    cfg = mk()
    p = cfg.params
    say(f"case full size {name} x {cfg.nsteps}")
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(E.uniform_cons(cfg)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    s0, _ = totals(torch, cu, p)
    for step in range(cfg.nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 77, step)
    torch.cuda.synchronize()
    ctx.check()
    assert bool(torch.isfinite(cu.t).all())
    s1, a1 = totals(torch, cu, p)
    for c in range(p.nvars):
        scale = a1[c]
        drift = abs(s1[c] - s0[c]) / scale
        say(f"  component {c}: total {s0[c]:.6e} -> {s1[c]:.6e}, drift {drift:.2e} of sum|.|")
        assert drift <= 1e-12, c
    # the noise did act: momentum fluctuates cell by cell
    assert a1[1] > 0 and a1[2] > 0 and a1[3] > 0


def test_fused_equals_reference_shaped_four_species_128_cubed(orc):
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    from tests.helpers import perturbed_state
    cfg = P.four_species(128)
    p = cfg.params
    say("case fused vs reference-shaped, four species 128^3 x 2 steps")
    st = perturbed_state(p, cfg, orc, seed=5, amp=0.01, vel=300.0)
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    cu = fc.MultiFab(torch.from_numpy(st.cu.copy()).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for step in range(2):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=3, step=step)
        fc.rk3_step(ctx, cu, prim, cfg.dt, 3, step)
    ctx.check()
    assert torch.equal(ds.cu.t, cu.t)
    assert torch.equal(ds.prim.t, prim.t)


def test_taylor_green_deck_matches_oracle_and_keeps_totals(orc):
    """exec/compressible/inputs_tg at its own size (64^3, two argon species, prob_type 4):
    three noisy fused steps from the vortex equal the oracle bit for bit; over the deck's 100
    steps the totals of rho, rho u, rho E and rho Y_k stay at their start values."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    cfg = P.taylor_green(64)
    p = cfg.params
    say(f"case Taylor-Green {tuple(p.n_cells)} x 3 vs oracle, then x {cfg.nsteps}")
    cu0 = E.initial_cons(cfg)
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(cu0.copy()).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for step in range(3):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, step)
    ctx.check()
    ref = cu0.copy()
    rprim = np.zeros((p.nprimvars,) + ref.shape[1:])
    assert orc.advance(p, ref, rprim, 3, cfg.dt, seed=1, step0=0) == 0
    got = cu.t.cpu().numpy()
    say(f"  exact={np.array_equal(got, ref)}")
    assert np.array_equal(got, ref)
    cu = fc.MultiFab(torch.from_numpy(cu0.copy()).cuda(), (-p.ngc,) * 3)
    s0, _ = totals(torch, cu, p)
    for step in range(cfg.nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, step)
    ctx.check()
    s1, a1 = totals(torch, cu, p)
    for c in range(p.nvars):
        assert abs(s1[c] - s0[c]) <= 1e-12 * a1[c], c


def test_deterministic_wave_decays_at_the_thermal_diffusion_rate(orc):
    """exec/compressible/inputs_test_deterministic (argon 4 x 80 x 4, noise off, prob_type 100):
    an isobaric density wave.  100 fused steps equal the oracle bit for bit.  Its entropy
    part q = rho' - p'/c^2 (the density not carried by sound) relaxes by heat conduction at
    lambda = kappa k^2 / (rho c_p), with kappa the hard-sphere closed form
    (75/64)(k_B/d^2) sqrt(k_B T / (pi m)) at 300 K (GIO gives it for two identical species),
    c_p = 5R/2M and k = 2 pi / L_y: lambda = 1.161e8 /s.  After 16000 steps (1.6e-8 s, 1.9
    e-folds) the measured mean rate is within 3 % of lambda (the oracle gives 0.5 %; the
    residual sound the wave sheds makes the running rate swing by +-4 % before that).  A
    wrong factor in kappa, c_p or the heat-flux stencil moves it by far more."""
    import math
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    cfg = P.deterministic_wave()
    p = cfg.params
    g = p.ngc
    say(f"case deterministic wave {tuple(p.n_cells)} x 100 vs oracle, then x 16000")
    cu0 = E.initial_cons(cfg)
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(cu0.copy()).cuda(), (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for step in range(100):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, step)
    ctx.check()
    ref = cu0.copy()
    rprim = np.zeros((p.nprimvars,) + ref.shape[1:])
    assert orc.advance(p, ref, rprim, 100, cfg.dt) == 0
    assert np.array_equal(cu.t.cpu().numpy(), ref)

    n1 = p.n_cells[1]
    k = 2 * math.pi / (n1 * p.dx[1])
    s = np.sin(k * (np.arange(n1) + 0.5) * p.dx[1])

    def q_amp(pr):
        r = pr[0, g:-g, g:-g, g:-g].mean(axis=(0, 2))
        pp = pr[5, g:-g, g:-g, g:-g].mean(axis=(0, 2))
        c2 = 5.0 / 3.0 * pp.mean() / r.mean()
        return 2 * (((r - r.mean()) - (pp - pp.mean()) / c2) * s).mean()

    cu = fc.MultiFab(torch.from_numpy(cu0.copy()).cuda(), (-g,) * 3)
    fc.conservedToPrimitive(ctx, prim, cu)
    q0 = q_amp(prim.t.cpu().numpy())
    nsteps = 16000
    for step in range(nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, step)
    ctx.check()
    q1 = q_amp(prim.t.cpu().numpy())
    M, kB, R, d, T = p.molmass[0], p.k_B, p.Runiv, p.diameter[0], cfg.T_init
    kappa = 75.0 / 64.0 * kB / d ** 2 * math.sqrt(kB * T / (math.pi * M * kB / R))
    lam = kappa / (cfg.rho0 * 2.5 * R / M) * k * k
    rate = -math.log(q1 / q0) / (nsteps * cfg.dt)
    say(f"  q/q0 = {q1 / q0:.5f}, rate {rate:.4e} /s, closed form {lam:.4e} /s, ratio {rate / lam:.4f}")
    assert abs(rate / lam - 1) < 0.03
