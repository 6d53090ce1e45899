"""The batched and served workloads at their full size, against solo runs and the oracle.

* Batched ensemble: 512 argon boxes with per-axis edges log-uniform in [16,64] (seed 2024),
  seeds 1000 + i, uniform start, 500 hipGraph-replayed steps -- the benchmark's workload.  No
  error flag; every instance keeps its totals of rho, rho*u, rho*E and rho*Y_k to 5e-14 (flux
  form on a periodic box); 8 sampled instances equal their solo runs (fhdx_rk3_step) bit for bit; the two
  smallest equal the oracle's 500 steps (Philox noise generated on both sides) bit for bit.
* Served stream: the benchmark's 16 requests (edges log-uniform in [16,128], 1-4 species,
  200-2000 steps, stream seed 7) submitted with Poisson arrivals at 50/s to the continuous-
  batching server: all complete, three equal their solo runs and the smallest equals the oracle,
  bit for bit.
"""
import sys

import numpy as np
import pytest

from fhdex_amd import params as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solo(cfg, cu0, seed, nsteps):
    import torch

    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
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


def oracle_run(orc, cfg, cu0, seed, nsteps):
    p = cfg.params
    st = orc.HostState(p)
    st.cu[:] = cu0
    assert orc.advance(p, st.cu, st.prim, nsteps, cfg.dt, None, seed, 0, 0) == 0
    return st.cu, st.prim


def totals_kept(p, cu0, cu, tol=5e-14):
    """Sum of rho, rho*u, rho*E, rho*Y_k over the valid cells, start against end, relative to the
    sum of magnitudes.  Every flux is in divergence form on the periodic box, so the totals move
    only by rounding.  The third RK3 stage multiplies by 2/3 (timeStep.cpp:372), which is not a
    binary fraction: its rounding biases every step by up to ~6e-17 relative, so 500 steps may
    drift by ~3e-14 (the reference's own formula does the same)."""
    g = p.ngc
    v = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    a, b = cu0[v].reshape(p.nvars, -1), cu[v].reshape(p.nvars, -1)
    # rho*u starts at zero: compare its drift with the size of the momenta reached
    scale = np.maximum(np.abs(a).sum(axis=1), np.abs(b).sum(axis=1))
    drift = np.abs(b.sum(axis=1) - a.sum(axis=1)) / scale
    return bool((drift <= tol).all()), drift


def test_ensemble_512_boxes_500_graph_steps(orc):
    from fhdex_amd import ensemble as E

    shapes = E.ensemble_members(512)
    cfgs = [P.argon_gas(0, shape=sh) for sh in shapes]
    cells = sum(a * b * c for a, b, c in shapes)
    say(f"case C4: 512 argon boxes, {cells} cells, 500 graph steps")
    budget = sum(E.bytes_needed(c.params) for c in cfgs)
    ens = E.Ensemble(512, budget)
    starts = [E.uniform_cons(c) for c in cfgs]
    slots = [ens.add(c.params, c.dt, 1000 + q, starts[q]) for q, c in enumerate(cfgs)]
    for _ in range(5):
        ens.step(100, graph=True)
        ens.sync()
        say("  100 steps done")
    flags = ens.errors()
    assert not any(flags), "error flags raised"
    bad = []
    for q, c in enumerate(cfgs):
        cu, _ = ens.read(slots[q])
        ok, drift = totals_kept(c.params, starts[q], cu)
        if not ok:
            bad.append((q, shapes[q], drift))
        assert ens.status(slots[q]) == (500, False)
    assert not bad, f"totals drift: {bad[:3]}"
    say("  every instance keeps its totals")
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(512, size=8, replace=False).tolist())
    for q in sample:
        cu, prim = ens.read(slots[q])
        cu_s, prim_s = solo(cfgs[q], starts[q], 1000 + q, 500)
        same = np.array_equal(cu, cu_s) and np.array_equal(prim, prim_s)
        say(f"  instance {q} {shapes[q]}: equals its solo run: {same}")
        assert same
    order = sorted(range(512), key=lambda q: shapes[q][0] * shapes[q][1] * shapes[q][2])
    for q in order[:2]:
        cu, prim = ens.read(slots[q])
        cu_o, prim_o = oracle_run(orc, cfgs[q], starts[q], 1000 + q, 500)
        same = np.array_equal(cu, cu_o) and np.array_equal(prim, prim_o)
        say(f"  instance {q} {shapes[q]}: equals the oracle after 500 steps: {same}")
        assert same
    ens.close()


@pytest.fixture(scope="module")
def served():
    """The benchmark's served stream, run once: {id: (cfg, nsteps, seed, final cu)} and the
    completions."""
    import time

    from fhdex_amd import ensemble as E
    from fhdex_amd import serving as S

    reqs = S.request_stream(16, 7)
    say(f"case C5: 16 requests, {sum(c.cells * n for c, _, n, _ in reqs)} cell-updates, Poisson 50/s")
    srv = S.Server(256, 160 << 30, chunk_steps=8, graph=True)
    outs = {}
    rng = np.random.default_rng(7)
    t_next = time.perf_counter()
    for (cfg, ns, n, seed), gap in zip(reqs, rng.exponential(1 / 50.0, size=len(reqs))):
        t_next += gap
        time.sleep(max(0.0, t_next - time.perf_counter()))
        out = np.empty(E.grown_shape(cfg.params, cfg.params.nvars))
        outs[srv.submit(cfg, seed, n, cu_out=out)] = (cfg, n, seed, out)
    open_n = srv.wait(600.0)
    done = srv.poll(64)
    srv.close()
    say(f"  served: {len(done)} completions, {open_n} open")
    return outs, done, open_n


def test_served_stream_16_requests(served):
    from fhdex_amd import ensemble as E

    outs, done, open_n = served
    assert open_n == 0
    assert len(done) == 16 and all(c["status"] == 0 for c in done), [c["message"] for c in done]
    assert all(c["steps"] == outs[c["id"]][1] for c in done)
    by_work = sorted(outs, key=lambda i: outs[i][0].cells * outs[i][1])
    for rid in by_work[:3]:
        cfg, n, seed, out = outs[rid]
        cu_s, _ = solo(cfg, E.uniform_cons(cfg), seed, n)
        same = np.array_equal(out, cu_s)
        say(f"  request {rid} {cfg.name} x {n} steps: equals its solo run: {same}")
        assert same


def test_served_request_equals_oracle(served, orc):
    """The smallest request of the stream, served in the mix, against the oracle."""
    from fhdex_amd import ensemble as E

    outs, _, _ = served
    rid = min(outs, key=lambda i: outs[i][0].cells * outs[i][1])
    cfg, n, seed, out = outs[rid]
    cu_o, _ = oracle_run(orc, cfg, E.uniform_cons(cfg), seed, n)
    same = np.array_equal(out, cu_o)
    say(f"  request {rid} {cfg.name} x {n} steps: equals the oracle: {same}")
    assert same
