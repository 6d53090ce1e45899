"""Parity of the staggered scheme's HIP path (fhdx_*_stag) with the oracle (oracle_stag.c).

Each test builds a state on the host, runs the same call through the C ABI on the GPU and the
oracle on the CPU from the same inputs and noise key, and requires equal bits: conversion to
primitive variables, the fluxes of one calculateFluxStag, and whole RK3stepStag steps -- for
argon written as two identical species and the four-species gas, the three advection forms,
bulk viscosity and gravity, and the reference's periodic regression deck
(exec/compressible_stag/inputs_regression_periodic_3D) at its own size.  At 64^3 the step is
checked through conservation, and at 32^3 against the equilibrium variances of fluctuation
theory (which pins the noise amplitudes of the stochastic fluxes).
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import face_valid, perturbed_stag_state, stag_uniform_state, valid

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gas(name, shape, **kw):
    tt = kw.pop("transport_type", 1)
    cfg, p = _gas_params(name, shape, **kw)
    p.transport_type = tt
    return cfg, p


def _gas_params(name, shape, **kw):
    if name == "heAr":
        cfg = P.species_prefix(2, 4, 1)
        p = P.make_params(2, shape, 8e-6, [4.0026, 39.948], [2.18e-8, 3.63e-8], **kw)
        cfg.rhobar = [0.3, 0.7]
        return cfg, p
    if name == "argon":
        cfg = P.argon_gas(4)
        p = P.make_params(2, shape, 8e-6, [39.95, 39.95], [3.66e-8, 3.66e-8], **kw)
    elif name == "eight":
        cfg = P.species_prefix(8, 4, 1)
        p = P.make_params(8, shape, 8e-6, P.EIGHT_SPECIES_M, P.EIGHT_SPECIES_D, **kw)
    else:
        cfg = P.four_species(4)
        p = P.make_params(4, shape, 8e-6, list(cfg.params.molmass[:4]), list(cfg.params.diameter[:4]), **kw)
    return cfg, p


def _device(p, host):
    """A StagState holding host's cu, cumom, prim and vel (ghosts included)."""
    import torch
    from fhdex_amd import stag as S
    st = S.StagState(p)
    st.cu.t.copy_(torch.from_numpy(host.cu))
    st.prim.t.copy_(torch.from_numpy(host.prim))
    for d in range(3):
        st.cumom[d].t.copy_(torch.from_numpy(host.mom[d]))
        st.vel[d].t.copy_(torch.from_numpy(host.vel[d]))
    return st


def _same(name, got, want):
    got = got.cpu().numpy() if hasattr(got, "cpu") else got
    if not np.array_equal(got, want):
        diff = np.abs(got - want).reshape(got.shape[0], -1).max(axis=1)
        raise AssertionError(f"{name}: max |diff| per component {diff}")


CASES = [("argon", dict()), ("four", dict()), ("argon", dict(advection_type=0, visc_type=3, zeta_ratio=1.5)),
         ("four", dict(advection_type=1)), ("four", dict(grav=(0.0, -3.0e9, 1.0e9), advection_type=0)),
         ("argon", dict(stoch_stress_form=0)), ("heAr", dict(transport_type=3)), ("eight", dict()),
         ("eight", dict(advection_type=0)), ("heAr", dict(transport_type=2)), ("four", dict(transport_type=2))]
IDS = ["argon", "four", "argon-adv0-visc3", "four-adv1", "four-gravity-adv0", "argon-deterministic",
       "heAr-hcb-transport", "eight", "eight-adv0", "heAr-vw-transport", "four-vw-transport"]


@pytest.mark.parametrize("gas", ["argon", "four", "eight"])
def test_c2p_stag_matches_oracle(orc, gas):
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg, p = _gas(gas, (7, 6, 5))
    say(f"case c2p_stag {gas} {tuple(p.n_cells)}")
    host = perturbed_stag_state(p, cfg, orc, seed=1)
    st = _device(p, host)
    st.prim.t.zero_()
    for d in range(3):
        st.vel[d].t.zero_()
    ctx = L.Context(p)
    S.conservedToPrimitiveStag(ctx, st.prim, st.vel, st.cu, st.cumom)
    ctx.check()
    _same("prim", st.prim.t, host.prim)
    _same("cu", st.cu.t, host.cu)
    for d in range(3):
        _same(f"vel{d}", st.vel[d].t, host.vel[d])
        _same(f"cumom{d}", st.cumom[d].t, host.mom[d])


@pytest.mark.parametrize("gas,kw", CASES, ids=IDS)
def test_flux_stag_matches_oracle(orc, gas, kw):
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg, p = _gas(gas, (7, 6, 5), **kw)
    say(f"case flux_stag {gas} {kw}")
    host = perturbed_stag_state(p, cfg, orc, seed=2, amp=0.03, vel=5e3)
    seed, step, w = 5, 3, 0.37
    host.fill_noise(seed, step)
    assert host.transport() == 0
    host.flux(cfg.dt, w)
    st = _device(p, host)
    ctx = L.Context(p)
    S.calculateTransportCoeffs(ctx, st)
    S.calculateFluxStag(ctx, st, w, cfg.dt, seed, step)
    ctx.check()
    for name in ("eta", "zeta", "kappa", "chi", "D"):
        _same(name, getattr(st, name).t, getattr(host, name))
    for d in range(3):
        _same(f"faceflux{d}", st.faceflux[d].t, host.ff[d])
        _same(f"cenflux{d}", st.cenflux[d].t, host.cf[d])
    for e in range(6):
        _same(f"edgeflux{e}", st.edgeflux[e].t, host.ef[e])


@pytest.mark.parametrize("gas,kw", CASES, ids=IDS)
def test_step_stag_matches_oracle(orc, gas, kw):
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg, p = _gas(gas, (6, 7, 5), **kw)
    say(f"case step_stag {gas} {kw}")
    host = perturbed_stag_state(p, cfg, orc, seed=3, amp=0.03, vel=5e3)
    st = _device(p, host)
    seed, step0, nsteps = 9, 4, 3
    ctx = L.Context(p)
    for s in range(nsteps):
        S.RK3stepStag(ctx, st, cfg.dt, seed, step0 + s)
    ctx.check()
    assert host.advance(nsteps, cfg.dt, seed, step0) == 0
    _same("cu", st.cu.t, host.cu)
    _same("prim", st.prim.t, host.prim)
    for d in range(3):
        _same(f"cumom{d}", st.cumom[d].t, host.mom[d])
        _same(f"vel{d}", st.vel[d].t, host.vel[d])


def test_step_stag_regression_deck_matches_oracle(orc):
    """exec/compressible_stag/inputs_regression_periodic_3D: four species (He, Ne, Ar, Kr at
    Y 1/4), 32^3 cells of 8e-6 cm, rho0 4.83e-4, T 300 K, fixed_dt 1e-12 s, max_step 3, seed 1,
    uniform start (prob_type 1); both the GPU and the oracle run the deck's three steps."""
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg = P.four_species(32)
    p = cfg.params
    say("case step_stag regression deck 32^3 four species, 3 steps")
    host = stag_uniform_state(p, cfg, orc)
    st = _device(p, host)
    ctx = L.Context(p)
    for s in range(3):
        S.RK3stepStag(ctx, st, 1e-12, 1, s)
    ctx.check()
    assert host.advance(3, 1e-12, 1, 0) == 0
    _same("cu", st.cu.t, host.cu)
    _same("prim", st.prim.t, host.prim)
    for d in range(3):
        _same(f"cumom{d}", st.cumom[d].t, host.mom[d])
        _same(f"vel{d}", st.vel[d].t, host.vel[d])


def test_step_stag_conserves_at_64_cubed():
    """Full-size property check: 20 noisy steps of a 64^3 four-species box keep the cell sums of
    rho, rho E, rho Y_k and the face sums of each momentum component to rounding, and the
    duplicate periodic faces equal to face 0."""
    import torch
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg = P.four_species(64)
    p = cfg.params
    say("case step_stag conservation 64^3 four species, 20 steps")
    st = S.StagState(p)
    for c, v in enumerate(E.uniform_values(cfg)):
        st.cu.t[c] = v
    ctx = L.Context(p)
    S.conservedToPrimitiveStag(ctx, st.prim, st.vel, st.cu, st.cumom)
    g = p.ngc
    n = 64

    def sums():
        cu = st.cu.t[:, g:g + n, g:g + n, g:g + n].reshape(p.nvars, -1)
        m = [st.cumom[d].t[0, g:g + n, g:g + n, g:g + n] for d in range(3)]
        return (cu.sum(dim=1).cpu().numpy(), cu.abs().sum(dim=1).cpu().numpy(),
                np.array([float(x.sum()) for x in m]), np.array([float(x.abs().sum()) for x in m]))

    s0, a0, m0, _ = sums()
    for s in range(20):
        S.RK3stepStag(ctx, st, cfg.dt, 7, s)
    ctx.check()
    s1, a1, m1, ma1 = sums()
    keep = [0, 4] + list(range(5, p.nvars))
    assert np.all(np.abs(s1 - s0)[keep] <= 1e-13 * a0[keep]), (s1 - s0) / a0
    assert np.all(np.abs(m1 - m0) <= 1e-13 * ma1), (m1 - m0, ma1)
    assert ma1.min() > 0
    for d in range(3):
        a = st.cumom[d].t[0].movedim(2 - d, 0)
        assert torch.equal(a[g + n], a[g])


def test_stag_equilibrium_variances():
    """A periodic 32^3 argon box at dx = 8e-5 cm started uniform reaches equilibrium under the
    staggered step; over 100 snapshots the variances of cell rho and rho E and of the face
    momenta match fluctuation theory (Var rho = rho m/dV, Var j_d = rho k_B T/dV, Var rho E =
    (15/4) rho (k_B T)^2/(m dV), T the sampled mean) within 1 %.  This pins the amplitudes of
    the stochastic stress (cell and edge noise) and heat flux noise."""
    import torch
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg = P.argon_gas(32)
    p = cfg.params
    for d in range(3):
        p.dx[d] = 8e-5
    cfg.dt = P.cfl_dt(p, cfg.rhobar, cfg.T_init, P.ARGON_CFL)
    say(f"case stag equilibrium argon 32^3 dx 8e-5 dt {cfg.dt:.3e}")
    st = S.StagState(p)
    for c, v in enumerate(E.uniform_values(cfg)):
        st.cu.t[c] = v
    ctx = L.Context(p)
    S.conservedToPrimitiveStag(ctx, st.prim, st.vel, st.cu, st.cumom)
    g, n = p.ngc, 32
    step = 0
    for _ in range(2000):
        S.RK3stepStag(ctx, st, cfg.dt, 2, step)
        step += 1
    acc = torch.zeros(5, dtype=torch.float64, device="cuda")
    tsum = 0.0
    nsnap = 100
    for _ in range(nsnap):
        for _ in range(20):
            S.RK3stepStag(ctx, st, cfg.dt, 2, step)
            step += 1
        v = st.cu.t[[0, 4], g:g + n, g:g + n, g:g + n].reshape(2, -1).var(dim=1)
        m = torch.stack([st.cumom[d].t[0, g:g + n, g:g + n, g:g + n].reshape(-1).var() for d in range(3)])
        acc += torch.cat([v[:1], m, v[1:]])
        tsum += float(st.prim.t[4, g:g + n, g:g + n, g:g + n].mean())
    ctx.check()
    got = (acc / nsnap).cpu().numpy()
    T = tsum / nsnap
    kB, R = p.k_B, p.Runiv
    mol = p.molmass[0] * kB / R
    dV = p.dx[0] * p.dx[1] * p.dx[2]
    rho = cfg.rho0
    want = np.array([rho * mol / dV, rho * kB * T / dV, rho * kB * T / dV, rho * kB * T / dV,
                     3.75 * rho * (kB * T) ** 2 / (mol * dV)])
    ratio = got / want
    for name, gv, w, r in zip(["rho", "jx", "jy", "jz", "rhoE"], got, want, ratio):
        say(f"  Var {name}: {gv:.5e} theory {w:.5e} ratio {r:.4f}")
    assert np.all(np.abs(ratio - 1) < 0.01), ratio


@pytest.mark.parametrize("rtype", [1, 0])
def test_step_stag_chemistry_matches_oracle(orc, rtype):
    """The staggered step with the NO2/N2O4 dimerization of the reference's staggered deck
    (exec/compressible_stag/test_NO2_dimerization_eq): per stage the chemical-Langevin source of
    the stage's prim (reaction noise A + w_s B) enters the cell update; three steps from a
    perturbed state equal the oracle's bit for bit, and the source / ranchem arrays hold the
    last stage's values."""
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    cfg = P.no2_n2o4(4, reaction_type=rtype)
    p = cfg.params
    p.n_cells[:] = (6, 5, 4)
    say(f"case step_stag chemistry reaction_type {rtype}")
    host = perturbed_stag_state(p, cfg, orc, seed=6, amp=0.02, vel=3e3)
    st = _device(p, host)
    ctx = L.Context(p)
    ctx.set_chemistry(cfg.chem)
    for s in range(3):
        S.RK3stepStag(ctx, st, cfg.dt, 4, s)
    ctx.check()
    assert host.advance(3, cfg.dt, 4, 0, chem=cfg.chem) == 0
    _same("cu", st.cu.t, host.cu)
    _same("prim", st.prim.t, host.prim)
    for d in range(3):
        _same(f"cumom{d}", st.cumom[d].t, host.mom[d])
        _same(f"vel{d}", st.vel[d].t, host.vel[d])
    g = p.ngc
    src = st.source.t[5:, g:-g, g:-g, g:-g]
    assert float(src.abs().max()) > 0
    assert float(st.source.t[:5].abs().max()) == 0.0
    if rtype == 1:
        assert float(st.ranchem.t[:, g:-g, g:-g, g:-g].abs().max()) > 0
