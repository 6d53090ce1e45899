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
