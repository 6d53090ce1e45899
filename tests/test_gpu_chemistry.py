"""Chemistry on the GPU against the oracle, through the C ABI: the source of
compute_compressible_chemistry_source_CLE (chemistry_functions.cpp:128-193), its reaction noise,
and RK3 steps carrying the +dt*S term (timeStep.cpp:135-167, 244-273, 349-378), on the
NO2/N2O4 deck with the chemical Langevin source.  exp and pow are the same IEEE operation
sequence on both sides, so every comparison is exact.
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


def setup(dev, orc, n=12, seed=4):
    torch, fc, L = dev
    cfg = P.no2_n2o4(n)
    p = cfg.params
    st = perturbed_state(p, cfg, orc, seed=seed, amp=0.02, vel=300.0)
    ctx = L.Context(p)
    ctx.set_chemistry(cfg.chem)
    return cfg, p, st, ctx


def test_source_and_noise_match_oracle(dev, orc):
    torch, fc, L = dev
    cfg, p, st, ctx = setup(dev, orc)
    g = p.ngc
    say(f"case chemistry source + noise {tuple(p.n_cells)}")
    A = fc.MultiFab.cell(p, p.nreaction)
    B = fc.MultiFab.cell(p, p.nreaction)
    fc.fill_chem_noise(ctx, 21, 5, A, B)
    hA = np.zeros(tuple(A.t.shape))
    hB = np.zeros_like(hA)
    orc.fill_chem_noise(p, hA, hB, seed=21, step=5)
    assert np.array_equal(valid(p, A.t.cpu().numpy()), valid(p, hA))
    assert np.array_equal(valid(p, B.t.cpu().numpy()), valid(p, hB))
    ran = 1.0 * hA + 0.7 * hB
    prim = to_dev(torch, fc, st.prim, (-g,) * 3)
    source = fc.MultiFab.cell(p, p.nvars)
    fc.compute_compressible_chemistry_source_CLE(ctx, cfg.dt, prim, source,
                                                 to_dev(torch, fc, ran, (-g,) * 3))
    want = np.zeros(tuple(source.t.shape))
    assert orc.chem_source(p, cfg.chem, st.prim, want, ran, cfg.dt) == 0
    got = source.t.cpu().numpy()
    say(f"  source exact={np.array_equal(got, want)}")
    assert np.array_equal(got, want)
    assert not got[:5].any()


def test_fused_step_with_chemistry_matches_oracle(dev, orc):
    """Three Philox steps: flux and reaction noise from the same key on both sides."""
    torch, fc, L = dev
    cfg, p, st, ctx = setup(dev, orc, seed=6)
    g = p.ngc
    say(f"case fused step with chemistry {tuple(p.n_cells)} x 3")
    cu = to_dev(torch, fc, st.cu, (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for step in range(3):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 17, step)
    ctx.check()
    assert orc.advance_chem(p, cfg.chem, st.cu, st.prim, 3, cfg.dt, seed=17, step0=0) == 0
    for tag, a, b in (("cu", cu, st.cu), ("prim", prim, st.prim)):
        same = np.array_equal(valid(p, a.t.cpu().numpy()), valid(p, b))
        say(f"  {tag} exact={same}")
        assert same, tag


def test_reference_shaped_step_and_launch_shapes_agree(dev, orc):
    """RK3step on the reference's arrays (source MultiFab, ranchem A/B) and the fused step under
    two launch shapes (records from k_cells or from the update kernel) give the same bits."""
    torch, fc, L = dev
    cfg, p, st, ctx = setup(dev, orc, n=10, seed=8)
    g = p.ngc
    say(f"case chemistry: reference-shaped vs fused {tuple(p.n_cells)} x 2")
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(st.cu))
    ds.prim.t.copy_(torch.from_numpy(st.prim))
    outs = []
    for shape in ((64, 1, 1, 1, 0, 0), (32, 2, 1, 0, 0, 1)):
        ctx.set_tune(*shape)
        cu = to_dev(torch, fc, st.cu, (-g,) * 3)
        prim = fc.MultiFab.cell(p, p.nprimvars)
        for step in range(2):
            fc.rk3_step(ctx, cu, prim, cfg.dt, 9, step)
        outs.append((cu.t.cpu().numpy(), prim.t.cpu().numpy()))
    for step in range(2):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=9, step=step)
    ctx.check()
    ref = (valid(p, ds.cu.t.cpu().numpy()), valid(p, ds.prim.t.cpu().numpy()))
    for cu, prim in outs:
        assert np.array_equal(valid(p, cu), ref[0]) and np.array_equal(valid(p, prim), ref[1])


def test_chemistry_rejections(dev, orc):
    torch, fc, L = dev
    cfg = P.no2_n2o4(8)
    p = cfg.params
    g = p.ngc
    st = perturbed_state(p, cfg, orc, seed=1, amp=0.01, vel=100.0)
    ctx = L.Context(p)
    cu = to_dev(torch, fc, st.cu, (-g,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    with pytest.raises(L.FhdxError) as e:            # reactions not set yet
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, 0)
    assert e.value.status == 1
    ctx.set_chemistry(cfg.chem)
    with pytest.raises(L.FhdxError) as e:            # caller noise has no reaction-noise array
        fc.rk3_step_injected(ctx, cu, prim, fc.NoiseSets(p), cfg.dt)
    assert e.value.status == 2
    from fhdex_amd import ensemble as E
    ens = E.Ensemble(2, 1 << 30)
    with pytest.raises(L.FhdxError) as e:            # the ensemble engine has no chemistry
        ens.add(p, cfg.dt, 1, st.cu)
    assert e.value.status == 2
    ens.close()
