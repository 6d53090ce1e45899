"""GPU parity of the parameter branches the default decks do not take, through the C ABI.

Each variant changes one switch of the reference's namelist on a three-species box: primitive
(advection_type 1) or no (0) hyperbolic interpolation (flux.cpp:2016-2372), bulk viscosity
(visc_type 3 with zeta = zeta_ratio * eta, transCoeffs.cpp:84-86), the non-symmetric stress
(visc_type 1), single-component transport (algorithm_type 1: no species fluxes), gravity
(timeStep.cpp:190-201) and deterministic fluxes (stoch_stress_form 0).  One step of the
fused path and of the reference-shaped path, with the same injected noise, against the oracle;
bar 1e-12 relative, exact equality expected.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state, rel_err, valid

pytestmark = pytest.mark.gpu

TOL1 = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


VARIANTS = [
    ("advection_type_1", dict(advection_type=1)),
    ("advection_type_0", dict(advection_type=0)),
    ("visc_type_3", dict(visc_type=3, zeta_ratio=1.5)),
    ("visc_type_1", dict(visc_type=1)),
    ("algorithm_type_1", dict(algorithm_type=1)),
    ("gravity", dict(grav=(2.5e8, -1.0e8, 9.8e7))),
    ("deterministic", dict(stoch_stress_form=0)),
]


def variant(kw):
    base = P.species_prefix(3, 10, 1)
    M = [4.0026, 20.1797, 39.9480]
    d = [2.18e-8, 2.58e-8, 3.63e-8]
    p = P.make_params(3, (10, 9, 8), 8e-6, M, d, **kw)
    return P.Config(base.name, p, base.rho0, base.rhobar, base.T_init, base.dt, 1)


def check(tag, got, want):
    e = rel_err(got, want)
    say(f"  {tag}: max rel err {e.max():.3e} exact={np.array_equal(got, want)}")
    assert e.max() <= TOL1, tag


@pytest.mark.parametrize("name,kw", VARIANTS)
def test_variant_one_step(orc, name, kw):
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    cfg = variant(kw)
    p = cfg.params
    g = p.ngc
    say(f"case variant {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=13, amp=0.02, vel=400.0)
    orc.fill_noise(p, st, 5, 2)
    ctx = L.Context(p)
    # fused path, caller noise
    cu = fc.MultiFab(torch.from_numpy(st.cu.copy()).cuda(), (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns = fc.NoiseSets(p)
    for d in range(3):
        ns.A[d].t.copy_(torch.from_numpy(st.sA[d]))
        ns.B[d].t.copy_(torch.from_numpy(st.sB[d]))
    ns.rA.t.copy_(torch.from_numpy(st.rA))
    ns.rB.t.copy_(torch.from_numpy(st.rB))
    fc.rk3_step_injected(ctx, cu, prim, ns, cfg.dt)
    # reference-shaped path, same noise
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    fc.RK3step(ctx, ds, cfg.dt, noise=ns)
    ctx.check()
    assert orc.rk3_step(p, st, cfg.dt) == 0
    check("fused cu", valid(p, cu.t.cpu().numpy()), valid(p, st.cu))
    check("fused prim", valid(p, prim.t.cpu().numpy()), valid(p, st.prim))
    check("reference-shaped cu", valid(p, ds.cu.t.cpu().numpy()), valid(p, st.cu))
    for d in range(3):
        check(f"reference-shaped flux[{d}]", ds.flux[d].t.cpu().numpy(), st.flux[d])


def test_trace_species_branch(orc):
    """A third species at mass fraction ~2e-14 takes the zeroed-row branch of the species noise
    (flux.cpp:395-401) on every face: calculateFlux and one fused step against the oracle."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    cfg = variant({})
    cfg.rhobar = [0.6, 0.4 - 2e-14, 2e-14]
    p = cfg.params
    g = p.ngc
    say(f"case trace species {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=17, amp=0.02, vel=400.0)
    assert st.prim[6 + 2, g:-g, g:-g, g:-g].max() < 1e-13
    assert orc.transport(p, st) == 0
    orc.fill_noise(p, st, 5, 9)
    w = 0.37
    for d in range(3):
        st.stoch[d][:] = st.sA[d] + w * st.sB[d]
        st.stoch[d][0] = 0
    st.rancorn[:] = st.rA + w * st.rB
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    for name in ("cu", "prim", "eta", "zeta", "kappa", "chi", "D"):
        getattr(ds, name).t.copy_(torch.from_numpy(getattr(st, name)))
    for d in range(3):
        ds.stochFlux[d].t.copy_(torch.from_numpy(st.stoch[d]))
    ds.rancorn.t.copy_(torch.from_numpy(st.rancorn))
    fc.calculateFlux(ctx, ds.cu, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D, ds.flux,
                     ds.stochFlux, ds.cornx, ds.corny, ds.cornz, ds.visccorn, ds.rancorn, cfg.dt)
    torch.cuda.synchronize()
    assert orc.calculate_flux(p, st, st.cu, cfg.dt) == 0
    for d in range(3):
        check(f"flux[{d}]", ds.flux[d].t.cpu().numpy(), st.flux[d])
    # one fused step with the same (injected) noise
    cu = fc.MultiFab(torch.from_numpy(st.cu.copy()).cuda(), (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns = fc.NoiseSets(p)
    for d in range(3):
        ns.A[d].t.copy_(torch.from_numpy(st.sA[d]))
        ns.B[d].t.copy_(torch.from_numpy(st.sB[d]))
    ns.rA.t.copy_(torch.from_numpy(st.rA))
    ns.rB.t.copy_(torch.from_numpy(st.rB))
    fc.rk3_step_injected(ctx, cu, prim, ns, cfg.dt)
    ctx.check()
    assert orc.rk3_step(p, st, cfg.dt) == 0
    check("fused cu", valid(p, cu.t.cpu().numpy()), valid(p, st.cu))


def test_hcb_binary_transport_one_step(orc):
    """transport_type 3 (Hirschfelder-Curtiss-Bird binary forms, compressible_functions.H:
    1357-1478) on a helium-argon mixture: the transport arrays of fhdx_transport, one fused step
    and one reference-shaped step with the same injected noise, against the oracle."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    base = P.species_prefix(2, 10, 1)
    p = P.make_params(2, (10, 9, 8), 8e-6, [4.0026, 39.948], [2.18e-8, 3.63e-8])
    p.transport_type = 3
    cfg = P.Config("heAr_hcb", p, base.rho0, [0.3, 0.7], base.T_init, base.dt, 1)
    g = p.ngc
    say(f"case HCB binary transport {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=19, amp=0.02, vel=400.0)
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    fc.calculateTransportCoeffs(ctx, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D)
    torch.cuda.synchronize()
    assert orc.transport(p, st) == 0
    for name in ("eta", "zeta", "kappa", "chi", "D"):
        check(name, getattr(ds, name).t.cpu().numpy(), getattr(st, name))
    orc.fill_noise(p, st, 5, 2)
    cu = fc.MultiFab(torch.from_numpy(st.cu.copy()).cuda(), (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns = fc.NoiseSets(p)
    for d in range(3):
        ns.A[d].t.copy_(torch.from_numpy(st.sA[d]))
        ns.B[d].t.copy_(torch.from_numpy(st.sB[d]))
    ns.rA.t.copy_(torch.from_numpy(st.rA))
    ns.rB.t.copy_(torch.from_numpy(st.rB))
    fc.rk3_step_injected(ctx, cu, prim, ns, cfg.dt)
    fc.RK3step(ctx, ds, cfg.dt, noise=ns)
    ctx.check()
    assert orc.rk3_step(p, st, cfg.dt) == 0
    check("fused cu", valid(p, cu.t.cpu().numpy()), valid(p, st.cu))
    check("reference-shaped cu", valid(p, ds.cu.t.cpu().numpy()), valid(p, st.cu))


def _vw_config(ns):
    if ns == 2:
        base = P.species_prefix(2, 10, 1)
        p = P.make_params(2, (10, 9, 8), 8e-6, [4.0026, 39.948], [2.18e-8, 3.63e-8])
        rhobar = [0.3, 0.7]
    else:
        base = P.species_prefix(ns, 0, 1, shape=(8, 7, 6))
        p = base.params
        rhobar = base.rhobar
    p.transport_type = 2
    return P.Config(f"vw{ns}", p, base.rho0, rhobar, base.T_init, base.dt, 1)


@pytest.mark.parametrize("ns", [2, 4, 8])
def test_vw_transport_one_step(orc, ns):
    """transport_type 2 (Valk-Waldmann, compressible_functions.H:1047-1150; pivoted LU solves in
    private memory): the transport arrays of fhdx_transport, one fused step and one
    reference-shaped step with the same injected noise, against the oracle."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    cfg = _vw_config(ns)
    p = cfg.params
    g = p.ngc
    say(f"case Valk-Waldmann transport NS={ns} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=23, amp=0.02, vel=400.0)
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    fc.calculateTransportCoeffs(ctx, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D)
    torch.cuda.synchronize()
    assert orc.transport(p, st) == 0
    for name in ("eta", "zeta", "kappa", "chi", "D"):
        check(name, getattr(ds, name).t.cpu().numpy(), getattr(st, name))
    orc.fill_noise(p, st, 5, 2)
    cu = fc.MultiFab(torch.from_numpy(st.cu.copy()).cuda(), (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns_ = fc.NoiseSets(p)
    for d in range(3):
        ns_.A[d].t.copy_(torch.from_numpy(st.sA[d]))
        ns_.B[d].t.copy_(torch.from_numpy(st.sB[d]))
    ns_.rA.t.copy_(torch.from_numpy(st.rA))
    ns_.rB.t.copy_(torch.from_numpy(st.rB))
    fc.rk3_step_injected(ctx, cu, prim, ns_, cfg.dt)
    fc.RK3step(ctx, ds, cfg.dt, noise=ns_)
    ctx.check()
    assert orc.rk3_step(p, st, cfg.dt) == 0
    check("fused cu", valid(p, cu.t.cpu().numpy()), valid(p, st.cu))
    check("reference-shaped cu", valid(p, ds.cu.t.cpu().numpy()), valid(p, st.cu))


def test_vw_singular_matrix_raises():
    """A cell at zero temperature zeroes every collision integral, so the pivoted factorisation
    meets a pivot below 1e-30 (the reference aborts); fhdx_ctx_check raises FHDX_ERR_SINGULAR.
    Every loop of the solves is bounded, and the solve itself is skipped once the flag is set."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    cfg = _vw_config(2)
    p = cfg.params
    say("case Valk-Waldmann singular matrix flag")
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    pr = ds.prim.t
    pr[0] = 4.83e-4
    pr[4] = 300.0
    pr[5] = 1e5
    pr[6] = 0.3
    pr[7] = 0.7
    pr[8] = 0.8
    pr[9] = 0.2
    pr[4, 3, 3, 3] = 0.0
    fc.calculateTransportCoeffs(ctx, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D)
    with pytest.raises(L.FhdxError) as e:
        ctx.check()
    assert e.value.status == 8
    ctx.check()  # the flag is cleared once reported
