"""GPU parity of the HIP path against the CPU oracle, through the C ABI (libfhdx.so).

Bar: the kernels keep the reference's operation order and are built without contraction,
so results are expected to be bit-identical to the oracle.  The asserted tolerance is the
parity target: <= 1e-12 relative (inf-norm per field, max|gpu-cpu|/max|cpu|) after one step
and <= 1e-10 after 100; each test also reports whether the match was exact.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state, rel_err, valid

pytestmark = pytest.mark.gpu

TOL1 = 1e-12     # after one step / one call
TOL100 = 1e-10   # after 100 steps


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    L.lib()
    return torch, fc, L


def to_dev(torch, fc, a, lo):
    return fc.MultiFab(torch.from_numpy(np.ascontiguousarray(a)).cuda(), lo)


def host(m):
    return m.t.cpu().numpy()


def device_state(torch, fc, p, st):
    """DeviceState whose arrays hold copies of the oracle HostState `st`."""
    ds = fc.DeviceState(p)
    g = p.ngc
    for name in ("cu", "cup", "cup2", "prim", "eta", "zeta", "kappa", "chi", "D"):
        getattr(ds, name).t.copy_(torch.from_numpy(getattr(st, name)))
    for d in range(3):
        ds.stochFlux[d].t.copy_(torch.from_numpy(st.stoch[d]))
    ds.rancorn.t.copy_(torch.from_numpy(st.rancorn))
    del g
    return ds


def noise_to_dev(torch, fc, p, st):
    ns = fc.NoiseSets(p)
    for d in range(3):
        ns.A[d].t.copy_(torch.from_numpy(st.sA[d]))
        ns.B[d].t.copy_(torch.from_numpy(st.sB[d]))
    ns.rA.t.copy_(torch.from_numpy(st.rA))
    ns.rB.t.copy_(torch.from_numpy(st.rB))
    return ns


def noise_to_host(st, ns):
    for d in range(3):
        st.sA[d][:] = host(ns.A[d])
        st.sB[d][:] = host(ns.B[d])
    st.rA[:] = host(ns.rA)
    st.rB[:] = host(ns.rB)


def report(tag, got, want):
    e = rel_err(got, want)
    exact = np.array_equal(got, want)
    say(f"  {tag}: max rel err {e.max():.3e} exact={exact}")
    return e


CASES = [("argon", lambda: P.argon_gas(12)), ("species4", lambda: P.four_species(10)),
         ("species5", lambda: P.species_prefix(5, 0, 1, shape=(9, 7, 8))),
         ("species8", lambda: P.species_prefix(8, 0, 1, shape=(8, 6, 7)))]


@pytest.mark.parametrize("name,mk", CASES)
def test_cons_to_prim_and_transport(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case cons_to_prim+transport {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=2, amp=0.03)
    ctx = L.Context(p)
    cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    fc.conservedToPrimitive(ctx, prim, cu)
    torch.cuda.synchronize()
    gp = valid(p, host(prim))
    e = report("prim", gp, valid(p, st.prim))
    assert e.max() <= TOL1
    # transport on the grown box from the oracle's prim (ghosts filled)
    ds = device_state(torch, fc, p, st)
    fc.calculateTransportCoeffs(ctx, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D)
    ctx.check()
    assert orc.transport(p, st) == 0
    for nm in ("eta", "kappa", "chi", "D"):
        e = report(nm, host(getattr(ds, nm)), getattr(st, nm))
        assert e.max() <= TOL1


def test_transport_flags_bad_mass_fraction(dev, orc):
    torch, fc, L = dev
    cfg = P.four_species(8)
    p = cfg.params
    say("case transport error flag")
    st = perturbed_state(p, cfg, orc, seed=1)
    st.prim[6, 4, 4, 4] = 1.0
    ctx = L.Context(p)
    ds = device_state(torch, fc, p, st)
    fc.calculateTransportCoeffs(ctx, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D)
    with pytest.raises(L.FhdxError) as ei:
        ctx.check()
    assert ei.value.status == 4
    ctx.check()   # flag cleared


@pytest.mark.parametrize("name,mk", CASES)
def test_calculate_flux(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case calculateFlux {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=3, amp=0.03)
    assert orc.transport(p, st) == 0
    orc.fill_noise(p, st, 5, 9)
    w = 0.37
    for d in range(3):
        st.stoch[d][:] = st.sA[d] + w * st.sB[d]
        st.stoch[d][0] = 0
    st.rancorn[:] = st.rA + w * st.rB
    ds = device_state(torch, fc, p, st)
    ctx = L.Context(p)
    fc.calculateFlux(ctx, ds.cu, ds.prim, ds.eta, ds.zeta, ds.kappa, ds.chi, ds.D, ds.flux,
                     ds.stochFlux, ds.cornx, ds.corny, ds.cornz, ds.visccorn, ds.rancorn, cfg.dt)
    torch.cuda.synchronize()
    assert orc.calculate_flux(p, st, st.cu, cfg.dt) == 0
    for d in range(3):
        e = report(f"flux[{d}]", host(ds.flux[d]), st.flux[d])
        assert e.max() <= TOL1
    for a, arrs in enumerate((ds.cornx, ds.corny, ds.cornz)):
        for b in range(3):
            e = report(f"corn[{a}][{b}]", host(arrs[b]), st.corn[a][b])
            assert e.max() <= TOL1
    e = report("visccorn", host(ds.visccorn), st.visccorn)
    assert e.max() <= TOL1


def test_fill_noise_matches_oracle(dev, orc):
    torch, fc, L = dev
    cfg = P.four_species(9)
    p = cfg.params
    say("case fill_noise")
    st = orc.HostState(p)
    orc.fill_noise(p, st, 123456789012345, 77, 3)
    ctx = L.Context(p)
    ns = fc.NoiseSets(p)
    fc.fill_noise(ctx, ns, 123456789012345, 77, 3)
    torch.cuda.synchronize()
    # same Philox words and the same IEEE operation sequence of the normal transform: same bits
    for d in range(3):
        for tag, g, h in (("A", ns.A[d], st.sA[d]), ("B", ns.B[d], st.sB[d])):
            e = np.abs(host(g) - h).max()
            say(f"  {tag}[{d}] max abs diff {e:.2e}")
            assert np.array_equal(host(g), h)
    assert np.array_equal(host(ns.rA), st.rA) and np.array_equal(host(ns.rB), st.rB)
    a = host(ns.A[0])
    np.testing.assert_array_equal(a[:, :, :, 0], a[:, :, :, -1])


def _run_ref_step(torch, fc, L, p, cfg, st, injected=True):
    ds = device_state(torch, fc, p, st)
    ns = noise_to_dev(torch, fc, p, st) if injected else None
    ctx = L.Context(p)
    fc.RK3step(ctx, ds, cfg.dt, noise=ns)
    ctx.check()
    return ds


@pytest.mark.parametrize("name,mk", CASES)
def test_rk3_step_reference_shaped_one_step(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case RK3step(ref) injected {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=4)
    orc.fill_noise(p, st, 11, 0)
    ds = _run_ref_step(torch, fc, L, p, cfg, st)
    assert orc.rk3_step(p, st, cfg.dt) == 0
    e = report("cu", valid(p, host(ds.cu)), valid(p, st.cu))
    assert e.max() <= TOL1
    e = report("prim", host(ds.prim), st.prim)
    assert e.max() <= TOL1
    for d in range(3):
        e = report(f"flux[{d}]", host(ds.flux[d]), st.flux[d])
        assert e.max() <= TOL1


@pytest.mark.parametrize("name,mk", CASES)
def test_rk3_step_fused_injected_one_step(dev, orc, name, mk):
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case rk3_step_injected {name} {tuple(p.n_cells)}")
    st = perturbed_state(p, cfg, orc, seed=6)
    orc.fill_noise(p, st, 21, 4)
    ctx = L.Context(p)
    cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns = noise_to_dev(torch, fc, p, st)
    fc.rk3_step_injected(ctx, cu, prim, ns, cfg.dt)
    ctx.check()
    assert orc.rk3_step(p, st, cfg.dt) == 0
    e = report("cu (valid+ghost)", host(cu), st.cu)
    assert e.max() <= TOL1
    e = report("prim (valid+ghost)", host(prim), st.prim)
    assert e.max() <= TOL1


def test_philox_step_equals_injected_fill_noise(dev, orc):
    """The in-kernel noise of rk3_step is exactly fhdx_fill_noise's for the same key."""
    torch, fc, L = dev
    cfg = P.argon_gas(12)
    p = cfg.params
    say("case rk3_step philox vs injected(fill_noise)")
    st = perturbed_state(p, cfg, orc, seed=8)
    ctx = L.Context(p)
    cu1 = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    cu2 = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    pr1 = fc.MultiFab.cell(p, p.nprimvars)
    pr2 = fc.MultiFab.cell(p, p.nprimvars)
    ns = fc.NoiseSets(p)
    for step in range(3):
        fc.rk3_step(ctx, cu1, pr1, cfg.dt, 31, step)
        fc.fill_noise(ctx, ns, 31, step)
        fc.rk3_step_injected(ctx, cu2, pr2, ns, cfg.dt)
    ctx.check()
    assert torch.equal(cu1.t, cu2.t) and torch.equal(pr1.t, pr2.t)


def _multi_step(torch, fc, L, orc, cfg, nsteps, seed):
    p = cfg.params
    st = perturbed_state(p, cfg, orc, seed=seed, amp=0.01, vel=500.0)
    ctx = L.Context(p)
    cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    ns = fc.NoiseSets(p)
    for step in range(nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, step)
        fc.fill_noise(ctx, ns, seed, step)
        noise_to_host(st, ns)
        assert orc.rk3_step(p, st, cfg.dt) == 0
    ctx.check()
    return p, st, host(cu), host(prim)


def test_argon_32_cubed_100_steps(dev, orc):
    """The 100-step parity shape: 32^3 two-identical-species argon, CFL 0.2, Philox noise generated
    in-kernel on the GPU and injected into the oracle step by step."""
    torch, fc, L = dev
    cfg = P.argon_gas(32)
    say("case argon 32^3 x 100 steps")
    p, st, cu, prim = _multi_step(torch, fc, L, orc, cfg, 100, 1)
    e = report("cu after 100", valid(p, cu), valid(p, st.cu))
    assert e.max() <= TOL100
    e = report("prim after 100", valid(p, prim), valid(p, st.prim))
    assert e.max() <= TOL100


def test_four_species_16_cubed_20_steps(dev, orc):
    torch, fc, L = dev
    cfg = P.four_species(16)
    say("case four species 16^3 x 20 steps")
    p, st, cu, prim = _multi_step(torch, fc, L, orc, cfg, 20, 3)
    e = report("cu after 20", valid(p, cu), valid(p, st.cu))
    assert e.max() <= TOL100


@pytest.mark.parametrize("ns", [5, 6, 7, 8])
def test_wide_mixture_8_cubed_10_steps(dev, orc, ns):
    """Five to eight species (the reference's MAX_SPECIES is 8) through the fused path with its
    Philox noise, the oracle stepped with the same noise."""
    torch, fc, L = dev
    cfg = P.species_prefix(ns, 8, 10)
    say(f"case {ns} species 8^3 x 10 steps")
    p, st, cu, prim = _multi_step(torch, fc, L, orc, cfg, 10, 9)
    e = report("cu after 10", valid(p, cu), valid(p, st.cu))
    assert e.max() <= TOL100
    e = report("prim after 10", valid(p, prim), valid(p, st.prim))
    assert e.max() <= TOL100


def test_fused_equals_reference_shaped_path_64_cubed(dev, orc):
    """Both GPU paths, Philox noise from the same key, 3 steps at 64^3: identical bits."""
    torch, fc, L = dev
    cfg = P.argon_gas(64)
    p = cfg.params
    say("case fused vs reference-shaped, argon 64^3 x 3 steps")
    st = perturbed_state(p, cfg, orc, seed=12, amp=0.01, vel=500.0)
    ctx = L.Context(p)
    ds = device_state(torch, fc, p, st)
    cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for step in range(3):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=5, step=step)
        fc.rk3_step(ctx, cu, prim, cfg.dt, 5, step)
    ctx.check()
    assert torch.equal(ds.cu.t, cu.t)
    assert torch.equal(ds.prim.t, prim.t)


@pytest.mark.parametrize("name,mk", [("argon", lambda: P.argon_gas(0, shape=(40, 24, 20))),
                                     ("species8", lambda: P.species_prefix(8, 0, 1, shape=(24, 12, 10)))])
def test_launch_shapes_give_identical_bits(dev, orc, name, mk):
    """Every launch shape the tuner can pick computes the same bits, and those of the oracle
    stepped with the same noise; the tuner leaves cu alone."""
    torch, fc, L = dev
    cfg = mk()
    p = cfg.params
    say(f"case launch shapes, {name} {tuple(p.n_cells)} x 2 steps")
    st = perturbed_state(p, cfg, orc, seed=21, amp=0.01, vel=500.0)
    ctx = L.Context(p)
    ref = None
    shapes = [(tw, rh, kz, inl, 0, 0) for tw in (16, 32, 64) for rh in (1, 3) for kz in (1, 3)
              for inl in (0, 1)] + [(64, 1, 1, 1, 2, 0), (32, 2, 3, 0, 3, 0), (64, 1, 1, 1, 0, 1),
                                    (32, 2, 3, 0, 0, 1), (16, 1, 1, 1, 3, 1)]
    for shape in shapes:
        ctx.set_tune(*shape)
        cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
        prim = fc.MultiFab.cell(p, p.nprimvars)
        for step in range(2):
            fc.rk3_step(ctx, cu, prim, cfg.dt, 9, step)
        got = (host(cu), host(prim))
        if ref is None:
            ref = got
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), shape
    ns = fc.NoiseSets(p)
    for step in range(2):
        fc.fill_noise(ctx, ns, 9, step)
        noise_to_host(st, ns)
        assert orc.rk3_step(p, st, cfg.dt) == 0
    e = report("cu after 2 (every shape)", valid(p, ref[0]), valid(p, st.cu))
    assert e.max() <= TOL1
    cu = to_dev(torch, fc, st.cu, (-p.ngc,) * 3)
    best, trials = ctx.autotune(cu.fab, cfg.dt, trial_steps=1)
    assert np.array_equal(host(cu), st.cu)
    assert len(trials) >= 2 and best.ms_per_step == min(t.ms_per_step for t in trials)
    ctx.check()
