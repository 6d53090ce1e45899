"""Batched ensembles (fhdx_engine_*) on the GPU: every instance must equal its solo run.

The solo run is the single-box fast path (fhdx_rk3_step) on the same state, dt, seed and step
indices; that path is itself bit-identical to the oracle (test_gpu_parity.py).  Bar: identical
bits (the batched kernels run the same per-cell/face/node bodies), asserted as exact equality.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    L.lib()
    return torch, fc, L, E


def solo(dev, cfg, cu0, seed, step0, nsteps):
    torch, fc, L, E = dev
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(np.ascontiguousarray(cu0)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for s in range(nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, step0 + s)
    ctx.check()
    out = cu.t.cpu().numpy(), prim.t.cpu().numpy()
    ctx.close()
    return out


def members(orc):
    """Uneven boxes and species counts 2 to 8, each with a perturbed state."""
    cfgs = [P.argon_gas(0, shape=(12, 8, 16)), P.argon_gas(0, shape=(5, 9, 7)),
            P.species_prefix(3, 8, 10), P.four_species(6), P.argon_gas(16),
            P.species_prefix(5, 0, 10, shape=(6, 7, 5)), P.species_prefix(6, 5, 10),
            P.species_prefix(7, 0, 10, shape=(4, 6, 5)), P.species_prefix(8, 6, 10)]
    return [(cfg, perturbed_state(cfg.params, cfg, orc, seed=40 + q, amp=0.01, vel=500.0).cu)
            for q, cfg in enumerate(cfgs)]


def budget_for(E, cfgs):
    return sum(E.bytes_needed(c.params) for c in cfgs)


@pytest.mark.parametrize("graph", [True, False])
def test_engine_instances_equal_solo_runs(dev, orc, graph):
    torch, fc, L, E = dev
    mem = members(orc)
    say(f"case engine vs solo, {len(mem)} uneven members, graph={graph}")
    ens = E.Ensemble(12, budget_for(E, [c for c, _ in mem]))
    slots = [ens.add(cfg.params, cfg.dt, 100 + q, cu0, step0=3) for q, (cfg, cu0) in enumerate(mem)]
    ens.step(2, graph=graph)
    ens.step(2, graph=graph)
    ens.sync()
    assert not any(ens.errors())
    for q, (cfg, cu0) in enumerate(mem):
        cu, prim = ens.read(slots[q])
        cu_s, prim_s = solo(dev, cfg, cu0, 100 + q, 3, 4)
        say(f"  {cfg.name}: cu exact={np.array_equal(cu, cu_s)} prim exact={np.array_equal(prim, prim_s)}")
        assert np.array_equal(cu, cu_s)
        assert np.array_equal(prim, prim_s)
        assert ens.status(slots[q]) == (7, False)
    ens.close()


def test_engine_alone_equals_in_the_mix(dev, orc):
    """A member that runs alone, and the same member run while others join and leave between
    step calls (membership changes recapture the graph): identical bits."""
    torch, fc, L, E = dev
    say("case engine alone vs in the mix")
    mem = members(orc)
    cfg, cu0 = mem[0]
    alone = E.Ensemble(1, E.bytes_needed(cfg.params))
    s = alone.add(cfg.params, cfg.dt, 7, cu0)
    alone.step(6)
    a_cu, a_prim = alone.read(s)
    alone.close()

    mix = E.Ensemble(8, budget_for(E, [c for c, _ in mem]))
    s = mix.add(cfg.params, cfg.dt, 7, cu0)
    others = [mix.add(c.params, c.dt, 50 + q, u) for q, (c, u) in enumerate(mem[1:3])]
    mix.step(2)
    mix.remove(others[0])
    later = mix.add(mem[3][0].params, mem[3][0].dt, 99, mem[3][1])
    mix.step(1)
    mix.add(mem[4][0].params, mem[4][0].dt, 98, mem[4][1])
    mix.step(3)
    m_cu, m_prim = mix.read(s)
    assert mix.status(s) == (6, False)
    assert mix.status(later) == (4, False)
    assert not any(mix.errors())
    mix.close()
    assert np.array_equal(a_cu, m_cu) and np.array_equal(a_prim, m_prim)


def test_engine_step_limit(dev, orc):
    """An instance admitted with nsteps = 3 idles after its third step."""
    torch, fc, L, E = dev
    say("case engine step limit")
    mem = members(orc)
    (c0, u0), (c1, u1) = mem[0], mem[2]
    ens = E.Ensemble(2, budget_for(E, [c0, c1]))
    a = ens.add(c0.params, c0.dt, 11, u0, step0=0, nsteps=3)
    b = ens.add(c1.params, c1.dt, 12, u1, step0=0, nsteps=0)
    ens.step(5)
    ens.sync()
    assert ens.status(a) == (3, True)
    assert ens.status(b) == (5, False)
    cu, prim = ens.read(a)
    cu_s, prim_s = solo(dev, c0, u0, 11, 0, 3)
    assert np.array_equal(cu, cu_s) and np.array_equal(prim, prim_s)
    cu, _ = ens.read(b)
    cu_s, _ = solo(dev, c1, u1, 12, 0, 5)
    assert np.array_equal(cu, cu_s)
    ens.close()


def test_engine_admission_by_memory_budget(dev, orc):
    torch, fc, L, E = dev
    say("case engine admission")
    cfg = P.argon_gas(8)
    need = E.bytes_needed(cfg.params)
    cu0 = E.uniform_cons(cfg)
    ens = E.Ensemble(4, 2 * need + need // 2)
    s0 = ens.add(cfg.params, cfg.dt, 1, cu0)
    s1 = ens.add(cfg.params, cfg.dt, 2, cu0)
    assert ens.bytes_used == 2 * need
    with pytest.raises(L.FhdxError) as ei:
        ens.add(cfg.params, cfg.dt, 3, cu0)
    assert ei.value.status == 5  # FHDX_ERR_NOMEM
    ens.remove(s0)
    assert ens.bytes_used == need
    s2 = ens.add(cfg.params, cfg.dt, 3, cu0)
    assert s2 == s0
    ens.step(1)
    ens.sync()
    assert ens.status(s1) == (1, False) and ens.status(s2) == (1, False)
    ens.close()


def test_engine_error_flag_is_per_instance(dev, orc):
    """A member with a mass fraction outside (0,1) raises its own flag only."""
    torch, fc, L, E = dev
    say("case engine per-instance error flags")
    cfg = P.argon_gas(8)
    good = E.uniform_cons(cfg)
    bad = good.copy()
    bad[5] *= 2.5   # rho_1 = 1.25 rho: Y_1 > 1 everywhere
    ens = E.Ensemble(3, 3 * E.bytes_needed(cfg.params))
    a = ens.add(cfg.params, cfg.dt, 1, good)
    b = ens.add(cfg.params, cfg.dt, 2, bad)
    ens.step(1)
    flags = ens.errors()
    assert flags[a] == 0 and flags[b] != 0
    assert not any(ens.errors())   # cleared by the read
    ens.close()


def test_engine_rejects_valk_waldmann_transport(dev, orc):
    """transport_type 2 has kernels of its own on the single-box paths; the engine's shared
    kernels carry the default transport only, so admission refuses it with a clear error."""
    torch, fc, L, E = dev
    say("case engine refuses transport_type 2")
    cfg = P.four_species(6)
    cfg.params.transport_type = 2
    cu0 = perturbed_state(cfg.params, cfg, orc, seed=5, amp=0.01, vel=300.0).cu
    ens = E.Ensemble(2, E.bytes_needed(cfg.params))
    with pytest.raises(L.FhdxError) as e:
        ens.add(cfg.params, cfg.dt, 1, cu0)
    assert e.value.status == 2
    assert "Valk-Waldmann" in str(e.value)
    ens.close()
