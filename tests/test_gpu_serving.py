"""The continuous-batching server (fhdx_server_*) on the GPU.

Every served result must equal the solo run of the same request (the single-box fast path,
fhdx_rk3_step, steps 0..nsteps-1 with the request's seed): identical bits, whatever else was
resident.  A bad request is evicted alone; requests that can never fit are rejected.
"""
import sys
import time

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
    from fhdex_amd import serving as S
    L.lib()
    return torch, fc, L, E, S


def solo(dev, cfg, cu0, seed, nsteps):
    torch, fc, L, E, S = dev
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab(torch.from_numpy(np.ascontiguousarray(cu0)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    for s in range(nsteps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, s)
    ctx.check()
    out = cu.t.cpu().numpy(), prim.t.cpu().numpy()
    ctx.close()
    return out


def mix(orc):
    cfgs = [(P.argon_gas(10), 5), (P.species_prefix(3, 8, 0), 9), (P.four_species(6), 4),
            (P.argon_gas(0, shape=(12, 6, 8)), 7), (P.species_prefix(2, 7, 0), 3),
            (P.argon_gas(8), 6)]
    return [(cfg, n, perturbed_state(cfg.params, cfg, orc, seed=60 + q, amp=0.01, vel=500.0).cu)
            for q, (cfg, n) in enumerate(cfgs)]


@pytest.mark.parametrize("policy,cap", [(0, 0), (1, 3000)])
def test_served_results_equal_solo_runs(dev, orc, policy, cap):
    """Six requests of uneven size, 2-4 species and 3-9 steps, arriving in two waves into a
    server with room for about three at a time and chunks of 2 steps; admission in arrival
    order, or smallest work first with a cap on the cells in flight below one request's size
    (one request is still admitted onto an idle GPU)."""
    torch, fc, L, E, S = dev
    say(f"case serving: 6 requests vs solo runs, policy {policy}, cap {cap}")
    reqs = mix(orc)
    need = sorted(E.bytes_needed(c.params) for c, _, _ in reqs)
    srv = S.Server(4, sum(need[-3:]), chunk_steps=2)
    srv.set_policy(policy, cap)
    outs = {}
    for q, (cfg, n, cu0) in enumerate(reqs):
        cu_out = np.zeros_like(cu0)
        prim_out = np.zeros(E.grown_shape(cfg.params, cfg.params.nprimvars))
        rid = srv.submit(cfg, 500 + q, n, cu_init=cu0, cu_out=cu_out, prim_out=prim_out)
        outs[rid] = (q, cu_out, prim_out)
        if q == 2:
            time.sleep(0.05)
    assert srv.wait(120.0) == 0
    done = srv.poll()
    srv.close()
    assert len(done) == len(reqs)
    for c in done:
        q, cu_out, prim_out = outs[c["id"]]
        cfg, n, cu0 = reqs[q]
        assert c["status"] == 0, c["message"]
        assert c["steps"] == n
        assert c["arrival_ns"] <= c["admit_ns"] <= c["done_ns"]
        cu_s, prim_s = solo(dev, cfg, cu0, 500 + q, n)
        say(f"  {cfg.name} x {n}: exact={np.array_equal(cu_out, cu_s)}")
        assert np.array_equal(cu_out, cu_s)
        assert np.array_equal(prim_out, prim_s)


def test_bad_request_is_evicted_alone(dev, orc):
    torch, fc, L, E, S = dev
    say("case serving: bad request evicted alone")
    cfg = P.argon_gas(8)
    good = perturbed_state(cfg.params, cfg, orc, seed=3, amp=0.01, vel=500.0).cu
    bad = good.copy()
    bad[5] *= 2.5          # Y_1 > 1
    srv = S.Server(4, 4 * E.bytes_needed(cfg.params), chunk_steps=2)
    out_good = np.zeros_like(good)
    ids = {srv.submit(cfg, 1, 6, cu_init=good, cu_out=out_good): "good",
           srv.submit(cfg, 2, 6, cu_init=bad): "bad"}
    assert srv.wait(60.0) == 0
    done = {ids[c["id"]]: c for c in srv.poll()}
    srv.close()
    assert done["bad"]["status"] == 4 and done["bad"]["steps"] <= 2
    assert done["good"]["status"] == 0 and done["good"]["steps"] == 6
    cu_s, _ = solo(dev, cfg, good, 1, 6)
    assert np.array_equal(out_good, cu_s)


def test_requests_that_cannot_run_are_rejected(dev, orc):
    torch, fc, L, E, S = dev
    say("case serving: rejected requests")
    small, big = P.argon_gas(8), P.argon_gas(24)
    srv = S.Server(2, 2 * E.bytes_needed(small.params), chunk_steps=4)
    odd = P.argon_gas(8)
    odd.params.ngc = 1            # stencils need two ghost layers
    walled = P.argon_gas(8)       # walls are served by the reference-shaped step, not batched
    P.set_walls(walled.params, 1, 1, 1)
    ids = {srv.submit(big, 1, 3): "big", srv.submit(odd, 2, 3): "odd",
           srv.submit(walled, 4, 3): "walls", srv.submit(small, 3, 3): "ok"}
    assert srv.wait(60.0) == 0
    done = {ids[c["id"]]: c for c in srv.poll()}
    st = srv.stats()
    srv.close()
    assert done["big"]["status"] == 5 and done["big"]["admit_ns"] == 0
    assert done["odd"]["status"] == 2 and "ngc" in done["odd"]["message"]
    assert done["walls"]["status"] == 2 and "periodic" in done["walls"]["message"]
    assert done["ok"]["status"] == 0 and done["ok"]["steps"] == 3
    assert st["steps"] == 3 and st["cell_updates"] == 3 * 8 ** 3 and st["bytes_used"] == 0
