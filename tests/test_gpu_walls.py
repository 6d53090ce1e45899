"""Walls on the GPU against the oracle, through the C ABI: setBC ghosts, the wall-aware noise,
calculateFlux (one-sided face branches, StochFlux, BCWallSpeciesFlux) with injected noise, and
RK3 steps of the reference-shaped path with Philox noise, on the reference's wall decks
(exec/compressible/inputs_regression_{slip_wall_adiabatic, noslip_reservoir_isothermal,
temp_grad, RT_3d}, inputs_giantfluct_3d) at reduced and at full size, a bulk-viscosity
variant, a box with walls in x only, one with x reservoirs and a periodic box one cell deep.  The GPU
reads the wall-side cell for both sides of a wall face where the oracle spells out the doubled
terms; every comparison is exact.  The fused step refuses walls with FHDX_ERR_UNSUPPORTED.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state, valid

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def x_walls():
    base = P.species_prefix(3, 10, 1)
    p = P.make_params(3, (10, 8, 9), 8e-6, [4.0026, 20.1797, 39.9480], [2.18e-8, 2.58e-8, 3.63e-8])
    P.set_walls(p, [(1, 2), (-1, -1), (-1, -1)], [(2, 1), (-1, -1), (-1, -1)], [(2, 1), (-1, -1), (-1, -1)],
                t_lo=(310.0, 0.0, 0.0), bc_Yk=[0.5, 0.3, 0.2])
    return P.Config("x_walls", p, base.rho0, base.rhobar, base.T_init, base.dt, 1)


def x_reservoirs():
    """x reservoirs (bc_mass 3, no slip, isothermal) with different gas at each end; y walls."""
    base = P.species_prefix(3, 10, 1)
    p = P.make_params(3, (10, 8, 9), 8e-6, [4.0026, 20.1797, 39.9480], [2.18e-8, 2.58e-8, 3.63e-8])
    pa = P.ambient_pressure(p, base.rho0, base.rhobar, base.T_init)
    P.set_walls(p, [(2, 2), (1, 1), (-1, -1)], [(2, 2), (1, 1), (-1, -1)], [(3, 3), (1, 1), (-1, -1)],
                t_lo=(310.0, 0.0, 0.0), t_hi=(295.0, 0.0, 0.0),
                bc_Yk=[[0.5, 0.3, 0.2], list(base.rhobar), [0] * 3, [0] * 3, [0] * 3, [0] * 3],
                p_lo=(1.03 * pa, 0.0, 0.0), p_hi=(pa, 0.0, 0.0))
    return P.Config("x_reservoirs", p, base.rho0, base.rhobar, base.T_init, base.dt, 1)


def one_cell_deep():
    """A periodic three-species box one cell deep in z (the quasi-2-D shape of
    exec/compressible/inputs_test_3d): ghosts are images two periods away."""
    cfg = P.species_prefix(3, 8, 1)
    cfg.params.n_cells[0], cfg.params.n_cells[2] = 9, 1
    return cfg


def visc3_walls():
    cfg = P.wall_deck("slip_wall_adiabatic", 8)
    cfg.params.visc_type = 3
    cfg.params.zeta_ratio = 1.5
    return cfg


CASES = [
    ("slip_wall_adiabatic", lambda: P.wall_deck("slip_wall_adiabatic", 8)),
    ("noslip_reservoir_isothermal", lambda: P.wall_deck("noslip_reservoir_isothermal", 8)),
    ("temp_grad", lambda: P.wall_deck("temp_grad", 16)),
    ("RT_3d", lambda: P.wall_deck("RT_3d", 8)),
    ("giantfluct_3d", lambda: P.wall_deck("giantfluct_3d", 8)),
    ("one_cell_deep_periodic", lambda: one_cell_deep()),
    ("visc_type_3", visc3_walls),
    ("x_walls", x_walls),
    ("x_reservoirs", x_reservoirs),
]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    L.lib()
    return torch, fc, L


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("name,mk", CASES)
def test_set_bc_and_noise_match_oracle(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    g = p.ngc
    say(f"case walls setBC + noise {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=3, amp=0.02, vel=300.0)
    ctx = L.Context(p)

    def scrambled(a):  # valid cells kept, every ghost cell garbage
        b = np.full_like(a, 7.0)
        b[:, g:-g, g:-g, g:-g] = a[:, g:-g, g:-g, g:-g]
        return b

    prim = fc.MultiFab(to_dev(torch, scrambled(st.prim)), (-g,) * 3)
    cu = fc.MultiFab(to_dev(torch, scrambled(st.cu)), (-g,) * 3)
    # rebuild the ghosts: periodic fill, then setBC (timeStep.cpp:207-209)
    fc.FillBoundary(ctx, cu)
    fc.FillBoundary(ctx, prim)
    fc.setBC(ctx, prim, cu)
    torch.cuda.synchronize()
    assert np.array_equal(prim.t.cpu().numpy(), st.prim)
    assert np.array_equal(cu.t.cpu().numpy(), st.cu)
    ns = fc.NoiseSets(p)
    fc.fill_noise(ctx, ns, 7, 3)
    orc.fill_noise(p, st, 7, 3)
    for d in range(3):
        assert np.array_equal(ns.A[d].t.cpu().numpy(), st.sA[d]), d
        assert np.array_equal(ns.B[d].t.cpu().numpy(), st.sB[d]), d
    assert np.array_equal(ns.rA.t.cpu().numpy(), st.rA) and np.array_equal(ns.rB.t.cpu().numpy(), st.rB)
    # faces 0 and n of a wall direction draw different values
    for d in range(3):
        if not p.periodic[d]:
            a = st.sA[d]
            lo, hi = np.take(a, 0, axis=3 - d), np.take(a, -1, axis=3 - d)
            assert not np.array_equal(lo[1:], hi[1:])


@pytest.mark.parametrize("name,mk", CASES)
def test_calculate_flux_matches_oracle(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case walls calculateFlux {name} {tuple(p.n_cells)} adv={p.advection_type} visc={p.visc_type}")
    st = perturbed_state(p, cfg, orc, seed=5, amp=0.02, vel=300.0)
    assert orc.transport(p, st) == 0
    orc.fill_noise(p, st, 5, 2)
    w = 0.41
    for d in range(3):
        st.stoch[d][:] = st.sA[d] + w * st.sB[d]
        st.stoch[d][0] = 0
    st.rancorn[:] = st.rA + w * st.rB
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    for nm in ("cu", "prim", "eta", "zeta", "kappa", "chi", "D"):
        getattr(ds, nm).t.copy_(torch.from_numpy(getattr(st, nm)))
    for d in range(3):
        ds.stochFlux[d].t.copy_(torch.from_numpy(st.stoch[d]))
    ds.rancorn.t.copy_(torch.from_numpy(st.rancorn))
    fc.calculateFlux(ctx, ds.cu, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D, ds.flux,
                     ds.stochFlux, ds.cornx, ds.corny, ds.cornz, ds.visccorn, ds.rancorn, cfg.dt)
    torch.cuda.synchronize()
    assert orc.calculate_flux(p, st, st.cu, cfg.dt) == 0
    for d in range(3):
        got = ds.flux[d].t.cpu().numpy()
        same = np.array_equal(got, st.flux[d])
        if not same:
            bad = np.argwhere(got != st.flux[d])
            say(f"  flux[{d}] differs at {len(bad)} entries, first {bad[:4].tolist()}")
        assert same, d
    for a in range(3):
        for b, arr in enumerate((ds.cornx, ds.corny, ds.cornz)[a]):
            assert np.array_equal(arr.t.cpu().numpy(), st.corn[a][b]), (a, b)
    assert np.array_equal(ds.visccorn.t.cpu().numpy(), st.visccorn)


@pytest.mark.parametrize("name,mk", CASES)
def test_rk3_steps_match_oracle(dev, orc, name, mk):
    """Three reference-shaped steps with Philox noise (setBC after every stage)."""
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case walls RK3step x3 {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=9, amp=0.02, vel=300.0)
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    for step in range(3):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=21, step=step)
    ctx.check()
    assert orc.advance(p, st.cu, st.prim, 3, cfg.dt, seed=21, step0=0) == 0
    got_cu, got_prim = ds.cu.t.cpu().numpy(), ds.prim.t.cpu().numpy()
    say(f"  cu exact={np.array_equal(got_cu, st.cu)} prim exact={np.array_equal(got_prim, st.prim)}")
    assert np.array_equal(got_cu, st.cu)            # ghosts included: setBC after the last stage
    assert np.array_equal(got_prim, st.prim)


def test_set_bc_rejects_bad_shapes(dev):
    torch, fc, L = dev
    p = P.wall_deck("slip_wall_adiabatic", 8).params
    ctx = L.Context(p)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    cu = fc.MultiFab.cell(p, p.nvars)
    with pytest.raises(L.FhdxError) as e:                # prim and cons swapped
        fc.setBC(ctx, cu, prim)
    assert e.value.status == 1
    fc.setBC(ctx, prim, cu)                               # the right shapes pass


@pytest.mark.parametrize("case", ["walls", "one_cell_deep"])
def test_fused_paths_refuse_walls_and_thin_boxes(dev, orc, case):
    torch, fc, L = dev
    cfg = P.wall_deck("slip_wall_adiabatic", 8)
    if case == "one_cell_deep":
        cfg = P.species_prefix(3, 8, 1)
        cfg.params.n_cells[2] = 1
    p = cfg.params
    say(f"case fused refusal {case} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=1, amp=0.01, vel=100.0)
    ctx = L.Context(p)
    cu = fc.MultiFab(to_dev(torch, st.cu), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    with pytest.raises(L.FhdxError) as e:
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, 0)
    assert e.value.status == 2
    with pytest.raises(L.FhdxError) as e:
        fc.rk3_step_injected(ctx, cu, prim, fc.NoiseSets(p), cfg.dt)
    assert e.value.status == 2
    from fhdex_amd import ensemble as E
    ens = E.Ensemble(2, 1 << 30)
    with pytest.raises(L.FhdxError) as e:
        ens.add(p, cfg.dt, 1, st.cu)
    assert e.value.status == 2
    ens.close()


@pytest.mark.parametrize("family", ["slip", "noslip"])
def test_wall_layers_fluctuate_like_the_bulk(dev, family):
    """Argon (two identical species) at 8e-5 cm cells between x walls, periodic y and z,
    primitive interpolation, reference-shaped noisy steps: the variances of rho, j and rho*E
    in the two cell layers touching the walls equal fluctuation theory (rho m/dV,
    rho k_B T/dV, e^2 rho m/dV + rho c_v k_B T^2/dV) as in the bulk.  This is what the sqrt(2)
    and 0 factors of StochFlux are for.  (With conserved interpolation the one-sided weights
    -0.2/0.75/0.5/-0.05 at the face next to a wall lower the normal-momentum variance of the
    wall layer to about 0.85 of theory, a property of the reference's scheme; it is printed,
    not asserted.)"""
    torch, fc, L = dev
    base = P.argon_gas(8)
    for adv in (1, 2):
        p = P.make_params(2, (16, 16, 16), 8e-5, [39.948] * 2, [3.66e-8] * 2, advection_type=adv)
        if family == "slip":
            P.set_walls(p, [(1, 1), (-1, -1), (-1, -1)], [(1, 1), (-1, -1), (-1, -1)], [(1, 1), (-1, -1), (-1, -1)])
        else:
            P.set_walls(p, [(2, 2), (-1, -1), (-1, -1)], [(2, 2), (-1, -1), (-1, -1)], [(1, 1), (-1, -1), (-1, -1)],
                        t_lo=(base.T_init,) * 3, t_hi=(base.T_init,) * 3)
        cfg = P.Config("argon_walls", p, base.rho0, base.rhobar, base.T_init,
                       P.cfl_dt(p, base.rhobar, base.T_init, 0.1), 1)
        from fhdex_amd import ensemble as E
        say(f"case wall-layer equilibrium {family} advection_type {adv} {tuple(p.n_cells)}")
        ctx = L.Context(p)
        ds = fc.DeviceState(p)
        ds.cu.t.copy_(torch.from_numpy(E.uniform_cons(cfg)))
        fc.conservedToPrimitive(ctx, ds.prim, ds.cu)
        fc.FillBoundary(ctx, ds.cu)
        fc.FillBoundary(ctx, ds.prim)
        fc.setBC(ctx, ds.prim, ds.cu)
        g = p.ngc
        step = 0
        for _ in range(1000):
            fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=13, step=step)
            step += 1
        acc = torch.zeros((5, 16), dtype=torch.float64, device="cuda")
        tsum = 0.0
        nsnap = 300
        for _ in range(nsnap):
            for _ in range(10):
                fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=13, step=step)
                step += 1
            v = ds.cu.t[:5, g:-g, g:-g, g:-g]
            m = v.mean(dim=(1, 2, 3), keepdim=True)
            acc += ((v - m) ** 2).mean(dim=(1, 2))
            tsum += float(ds.prim.t[4, g:-g, g:-g, g:-g].mean())
        ctx.check()
        var = (acc / nsnap).cpu().numpy()
        T = tsum / nsnap
        kB, dV, rho = p.k_B, p.dx[0] ** 3, cfg.rho0
        mm = p.molmass[0] * kB / p.Runiv
        cv = p.hcv[0]
        th = np.array([rho * mm / dV] + [rho * kB * T / dV] * 3 +
                      [(cv * T) ** 2 * rho * mm / dV + rho * cv * kB * T * T / dV])
        ratio = var / th[:, None]
        for c, nm in enumerate(["rho", "jx", "jy", "jz", "rhoE"]):
            say(f"  {nm:5s} " + " ".join(f"{x:.3f}" for x in ratio[c]))
        bulk = ratio[:, 3:-3].mean(axis=1)
        assert np.all(np.abs(bulk - 1) < 0.04), bulk
        if adv == 1:
            walls = ratio[:, [0, -1]]
            assert np.all(np.abs(walls - 1) < 0.08), walls


@pytest.mark.parametrize("name", ["slip_wall_adiabatic", "noslip_reservoir_isothermal", "temp_grad", "RT_3d",
                                  "giantfluct_3d"])
def test_wall_decks_at_full_size_match_oracle(dev, orc, name):
    """The wall decks at their own sizes (32^3; 80x10x10; 64x32x32; 256x128 cells, 4 deep)
    for 3 steps (the regression decks' max_step), from InitConsVar's start (uniform; the
    Rayleigh-Taylor layers with gravity; the diffusion barrier), Philox noise: GPU equals the
    oracle bit for bit, total mass is kept, and where every wall is a species wall so is every
    species total."""
    torch, fc, L = dev
    from fhdex_amd import ensemble as E
    cfg = P.wall_deck(name)
    p = cfg.params
    g = p.ngc
    say(f"case wall deck {name} {tuple(p.n_cells)} x {cfg.nsteps}")
    cu0 = E.initial_cons(cfg)
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(cu0))
    fc.conservedToPrimitive(ctx, ds.prim, ds.cu)       # main_driver.cpp:535-542
    fc.FillBoundary(ctx, ds.cu)
    fc.FillBoundary(ctx, ds.prim)
    fc.setBC(ctx, ds.prim, ds.cu)
    for step in range(cfg.nsteps):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=1, step=step)
    ctx.check()
    cu = cu0.copy()
    prim = np.zeros((p.nprimvars,) + cu.shape[1:])
    assert orc.advance(p, cu, prim, cfg.nsteps, cfg.dt, seed=1, step0=0) == 0
    got = ds.cu.t.cpu().numpy()
    say(f"  exact={np.array_equal(got, cu)}")
    assert np.array_equal(got, cu) and np.array_equal(ds.prim.t.cpu().numpy(), prim)
    s0, s1 = valid(p, cu0).sum(axis=(1, 2, 3)), valid(p, got).sum(axis=(1, 2, 3))
    assert abs(s1[0] - s0[0]) <= 1e-14 * s0[0]
    if all(c in (-1, 1) for c in list(p.bc_mass_lo) + list(p.bc_mass_hi)):
        for s in range(p.nspecies):
            assert abs(s1[5 + s] - s0[5 + s]) <= 1e-14 * s0[5 + s]


def test_temp_grad_reaches_steady_conduction(dev):
    """temp_grad without noise on a 20-cell gap (x: no-slip walls at 280.8 K and 508.5 K; y, z:
    slip adiabatic walls), run to steady state (12000 steps of 1e-12 s, ~25 thermal and 15
    acoustic relaxation times).  At rest the heat flux through every x face is the same, T rises
    monotonically, and the wall faces' half-cell gradients place the wall temperatures on the
    profile: the linear extrapolation of the two cells next to each wall to the wall plane
    lands within 0.5 K of t_lo / t_hi (a full-cell gradient there would miss by half a cell's
    temperature step, ~5.7 K)."""
    torch, fc, L = dev
    from fhdex_amd import ensemble as E
    cfg = P.wall_deck("temp_grad", 20)
    p = cfg.params
    p.stoch_stress_form = 0
    g = p.ngc
    nv = p.nvars
    say(f"case temp_grad steady conduction {tuple(p.n_cells)}")
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(E.uniform_cons(cfg)))
    fc.conservedToPrimitive(ctx, ds.prim, ds.cu)
    fc.FillBoundary(ctx, ds.cu)
    fc.FillBoundary(ctx, ds.prim)
    fc.setBC(ctx, ds.prim, ds.cu)
    nsteps = 12000
    for step in range(nsteps):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=1, step=step)
        if step == nsteps - 2:
            before = ds.cu.t.clone()
    ctx.check()
    T = ds.prim.t[4, g:-g, g:-g, g:-g].mean(dim=(0, 1)).cpu().numpy()
    q = -ds.flux[0].t[nv].mean(dim=(0, 1)).cpu().numpy()       # kappa dT/dx on the x faces
    u = ds.prim.t[1:4, g:-g, g:-g, g:-g].abs().max().item()
    t_lo = T[0] - 0.5 * (T[1] - T[0])
    t_hi = T[-1] + 0.5 * (T[-1] - T[-2])
    change = (ds.cu.t - before).abs().max().item() / ds.cu.t.abs().max().item()
    say(f"  T: {T[0]:.3f} .. {T[-1]:.3f} K, extrapolated to the walls {t_lo:.3f} / {t_hi:.3f} K")
    say(f"  q(x faces) {q.min():.6e} .. {q.max():.6e}; max |u| {u:.2e} cm/s; last-step change {change:.2e}")
    assert np.all(np.diff(T) > 0)
    assert q.min() > 0 and (q.max() - q.min()) <= 1e-3 * q.max()
    assert abs(t_lo - 280.8) < 0.5 and abs(t_hi - 508.5) < 0.5
    # at rest up to a residual acoustic ring-down far below the sound speed (3.7e4 cm/s)
    assert u < 5.0 and change < 1e-6


def test_chemistry_between_walls_matches_oracle(dev, orc):
    """The NO2/N2O4 chemical Langevin deck between no-slip isothermal walls in x (periodic y, z):
    three reference-shaped Philox steps, reaction noise included, equal the oracle bit for bit."""
    torch, fc, L = dev
    cfg = P.no2_n2o4(10)
    p = cfg.params
    P.set_walls(p, [(2, 2), (-1, -1), (-1, -1)], [(2, 2), (-1, -1), (-1, -1)], [(1, 1), (-1, -1), (-1, -1)],
                t_lo=(340.0, 0.0, 0.0), t_hi=(360.0, 0.0, 0.0))
    say(f"case chemistry between walls {tuple(p.n_cells)} x 3")
    st = perturbed_state(p, cfg, orc, seed=12, amp=0.02, vel=200.0)
    ctx = L.Context(p)
    ctx.set_chemistry(cfg.chem)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    for step in range(3):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=31, step=step)
    ctx.check()
    assert orc.advance_chem(p, cfg.chem, st.cu, st.prim, 3, cfg.dt, seed=31, step0=0) == 0
    same = np.array_equal(ds.cu.t.cpu().numpy(), st.cu) and np.array_equal(ds.prim.t.cpu().numpy(), st.prim)
    say(f"  exact={same}")
    assert same


def test_closed_walls_keep_totals_at_128_cubed(dev):
    """Argon (four identical species, 8e-6 cm cells, ~3400 atoms of each per cell) at 128^3
    between slip adiabatic species walls on every face, 50 noisy reference-shaped steps at CFL
    0.1: no mass or species crosses a closed wall, so Sigma rho and Sigma rho Y_k stay at their
    start values to rounding (<= 1e-13 of Sigma|.|)."""
    torch, fc, L = dev
    from fhdex_amd import ensemble as E
    cfg = P.wall_deck("slip_wall_adiabatic")
    cfg.params = P.make_params(4, (128, 128, 128), 8e-6, [39.948] * 4, [3.66e-8] * 4)
    P.set_walls(cfg.params, 1, 1, 1)
    cfg.dt = P.cfl_dt(cfg.params, cfg.rhobar, cfg.T_init, 0.1)
    p = cfg.params
    g = p.ngc
    say("case closed walls 128^3 x 50")
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(E.uniform_cons(cfg)))
    fc.conservedToPrimitive(ctx, ds.prim, ds.cu)
    fc.FillBoundary(ctx, ds.cu)
    fc.FillBoundary(ctx, ds.prim)
    fc.setBC(ctx, ds.prim, ds.cu)
    v = lambda: ds.cu.t[:, g:-g, g:-g, g:-g]
    s0 = v().sum(dim=(1, 2, 3)).cpu().numpy()
    for step in range(50):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=3, step=step)
    ctx.check()
    s1 = v().sum(dim=(1, 2, 3)).cpu().numpy()
    a1 = v().abs().sum(dim=(1, 2, 3)).cpu().numpy()
    assert bool(torch.isfinite(ds.cu.t).all())
    for c in [0] + list(range(5, 5 + p.nspecies)):
        drift = abs(s1[c] - s0[c]) / a1[c]
        say(f"  component {c}: drift {drift:.2e} of sum|.|")
        assert drift <= 1e-13, c
