"""Benchmark of the compressible FHD RK3 step on one MI355X.

Default workload: two-identical-species argon (the reference's single-gas form), 256^3 periodic
cells, FP64, CFL 0.1, in-kernel Philox noise; one "step" is one full RK3 step (three stages)
of the fused HIP path.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def b_alg(nvars: int) -> int:
    """Algorithmic HBM bytes per cell-update: 8 passes over the FP64 state vector per RK3 step
    (stage 1 reads U^n, writes U^1; stages 2 and 3 read U^n and the previous stage, write)."""
    return 64 * nvars


def measured_traffic(n: int):
    """HBM-side bytes per step from the committed rocprofv3 counter summary of this workload
    (profiles/traffic_argon<n>.json): the TCC_EA0 read requests by size (TCC_EA0_RDREQ_32B,
    _64B and the 128-byte remainder) and write requests (TCC_EA0_WRREQ, _64B) times their sizes,
    calibrated against a plain device copy of known size; None if there is no summary."""
    path = os.path.join(ROOT, "profiles", f"traffic_argon{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_step"]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, steps: int, n1: int = 32, steps1: int = 2, n4: int = 64, steps4: int = 4):
    """The CPU oracle (the reference's arithmetic in the reference's operation order, OpenMP over
    z planes), built here with gcc -O3 -march=native -ffp-contract=off for this host, on the host
    cores: all threads on argon n^3 x steps, one thread on argon n1^3 x steps1, and all threads
    on the four-species gas n4^3 x steps4."""
    from fhdex_amd import params as P
    from oracle import ffi as orc

    orc.use_native_build()

    def run(nn, ss, gas=P.argon_gas):
        cfg = gas(nn)
        p = cfg.params
        st = orc.HostState(p)
        orc.init_cons(p, st.cu, cfg.rho0, cfg.rhobar, cfg.T_init)
        t0 = time.perf_counter()
        rc = orc.advance(p, st.cu, st.prim, ss, cfg.dt, None, 1, 0, 0)
        dt = time.perf_counter() - t0
        assert rc == 0
        return nn ** 3 * ss / dt, dt

    threads = orc.num_threads()
    v, t = run(n, steps)
    orc.set_num_threads(1)
    v1, t1 = run(n1, steps1)
    orc.set_num_threads(threads)
    v4, t4 = run(n4, steps4, P.four_species)
    model = cpu_model()
    return {"value": v, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "sample": f"oracle (CPU restatement of the reference step, gcc -O3 -march=native, OpenMP), argon "
                      f"{n}^3 x {steps} RK3 steps on {threads} threads of {model}, {t:.1f} s",
            "value_1thread": v1,
            "sample_1thread": f"argon {n1}^3 x {steps1} RK3 steps on 1 thread, {t1:.1f} s",
            "value_species4": v4,
            "sample_species4": f"four species {n4}^3 x {steps4} RK3 steps on {threads} threads, {t4:.1f} s"}


def config_rate(cfg, steps: int, warmup: int = 2):
    """cell-updates/s of the fused step on one configuration (InitConsVar's start, tuned shape)."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    p = cfg.params
    ctx = L.Context(p)
    if cfg.chem is not None:
        ctx.set_chemistry(cfg.chem)
    cu = fc.MultiFab(torch.from_numpy(E.initial_cons(cfg)).cuda(), (-p.ngc,) * 3)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    best, _ = ctx.autotune(cu.fab, cfg.dt, trial_steps=2)
    for s in range(warmup):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, 1, warmup + s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.check()
    ctx.close()
    cells = p.n_cells[0] * p.n_cells[1] * p.n_cells[2]
    return {"workload": cfg.name, "value": cells * steps / dt, "unit": "cell-updates/s",
            "ms_per_step": dt / steps * 1e3, "steps": steps, "nvars": p.nvars,
            "launch_shape": best.as_dict()}


def wall_rate(cfg, reps: int = 20):
    """cell-updates/s of the reference-shaped step (the path that serves walls) on a wall deck,
    run as the reference runs it: max_step (3) noisy steps from InitConsVar's start with ghosts
    set by setBC, repeated `reps` times with seeds 1.. ; only the steps are timed.  The argon
    decks hold ~50 molecules per species per cell, so now and then a stage leaves a mass
    fraction outside (0,1), where the reference aborts (transCoeffs.cpp:45-46): such runs are
    counted and reported, not raised."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    p = cfg.params
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    cu0 = torch.from_numpy(E.initial_cons(cfg)).cuda()
    total = 0.0
    flagged = 0
    for r in range(reps + 1):
        ds.cu.t.copy_(cu0)
        fc.conservedToPrimitive(ctx, ds.prim, ds.cu)
        fc.FillBoundary(ctx, ds.cu)
        fc.FillBoundary(ctx, ds.prim)
        fc.setBC(ctx, ds.prim, ds.cu)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(cfg.nsteps):
            fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=1 + r, step=s)
        torch.cuda.synchronize()
        if r > 0:  # the first run is warm-up
            total += time.perf_counter() - t0
        try:
            ctx.check()
        except L.FhdxError as e:
            if e.status != 4:  # FHDX_ERR_MASSFRAC: a stage left Y in (0,1), where the reference aborts
                raise
            flagged += 1
    ctx.close()
    steps = reps * cfg.nsteps
    cells = p.n_cells[0] * p.n_cells[1] * p.n_cells[2]
    return {"workload": cfg.name + " (walls, reference-shaped step)", "value": cells * steps / total,
            "unit": "cell-updates/s", "ms_per_step": total / steps * 1e3, "steps": steps, "nvars": p.nvars,
            "runs": reps + 1, "runs_with_massfrac_flag": flagged}


def ref_rate(n: int, steps: int, warmup: int = 1):
    """cell-updates/s of the reference-shaped step (fhdx_rk3_step_ref, the literal RK3step
    drop-in with the reference's own arrays) on periodic argon n^3 from the uniform start,
    Philox noise."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    from fhdex_amd import params as P
    cfg = P.argon_gas(n)
    p = cfg.params
    ctx = L.Context(p)
    ds = fc.DeviceState(p)
    ds.cu.t.copy_(torch.from_numpy(E.uniform_cons(cfg)))
    fc.conservedToPrimitive(ctx, ds.prim, ds.cu)
    fc.FillBoundary(ctx, ds.prim)
    for s in range(warmup):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=1, step=s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        fc.RK3step(ctx, ds, cfg.dt, noise=None, seed=1, step=warmup + s)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.check()
    ctx.close()
    del ds
    return {"workload": f"argon two-species {n}^3 periodic, reference-shaped step (fhdx_rk3_step_ref)",
            "value": n ** 3 * steps / dt, "unit": "cell-updates/s", "ms_per_step": dt / steps * 1e3,
            "steps": steps}


def stag_rate(cfg, steps: int, warmup: int = 2, cpu_steps: int = 0):
    """cell-updates/s of the staggered step (fhdx_rk3_step_stag, RK3stepStag's arithmetic) on a
    periodic configuration from the uniform start, Philox noise, HIP events around the timed
    steps on the launch stream.  With cpu_steps > 0 also the oracle's staggered step on the
    same box for that many steps on the host cores."""
    import torch
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    from fhdex_amd import stag as S
    p = cfg.params
    ctx = L.Context(p)
    st = S.StagState(p)
    for c, v in enumerate(E.uniform_values(cfg)):
        st.cu.t[c] = v
    S.conservedToPrimitiveStag(ctx, st.prim, st.vel, st.cu, st.cumom)
    for s in range(warmup):
        S.RK3stepStag(ctx, st, cfg.dt, 1, s)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for s in range(steps):
        S.RK3stepStag(ctx, st, cfg.dt, 1, warmup + s)
    stop.record()
    torch.cuda.synchronize()
    ctx.check()
    ms = start.elapsed_time(stop) / steps
    cells = p.n_cells[0] * p.n_cells[1] * p.n_cells[2]
    out = {"workload": cfg.name + " (staggered scheme, RK3stepStag)", "value": cells / (ms * 1e-3),
           "unit": "cell-updates/s", "ms_per_step": ms, "steps": steps, "nvars": p.nvars}
    ctx.close()
    del st
    if cpu_steps:
        sys.path.insert(0, ROOT)
        from oracle import ffi as orc
        from tests.helpers import stag_uniform_state
        host = stag_uniform_state(p, cfg, orc)
        t0 = time.perf_counter()
        assert host.advance(cpu_steps, cfg.dt, 1, 0) == 0
        t = time.perf_counter() - t0
        out["cpu_oracle"] = {"value": cells * cpu_steps / t, "unit": "cell-updates/s",
                             "cores": orc.num_threads(), "kind": "port",
                             "sample": f"{cpu_steps} staggered steps of the same box"}
    return out


def measured_flops(n: int):
    """FP64 flops per cell-update from the committed rocprofv3 counter summary of this workload
    (profiles/fp64_argon<n>.json), or None."""
    path = os.path.join(ROOT, "profiles", f"fp64_argon{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["flops_per_cell_update"]


def ensemble_bench(count: int, steps: int, loop_steps: int):
    """512 independent argon boxes of uneven size, one noise seed each: the engine's
    graph-captured batched step against the plain loop (the single-box fast path called for
    each instance in turn).  Both run the same arithmetic; the rate is whole-ensemble
    cell-updates/s with the state resident in HBM."""
    import torch
    from fhdex_amd import compressible as fc
    from fhdex_amd import ensemble as E
    from fhdex_amd import lib as L
    from fhdex_amd import params as P

    shapes = E.ensemble_members(count)
    cfgs = [P.argon_gas(0, shape=sh) for sh in shapes]
    cells = sum(a * b * c for a, b, c in shapes)
    budget = sum(E.bytes_needed(c.params) for c in cfgs)
    ens = E.Ensemble(count, budget)
    for q, cfg in enumerate(cfgs):
        ens.add(cfg.params, cfg.dt, 1000 + q, E.uniform_cons(cfg))
    ens.step(2, graph=True)          # warm-up; captures the graph
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps, graph=True)
    ens.sync()
    t_ens = time.perf_counter() - t0
    bad = sum(1 for f in ens.errors() if f)
    ens.close()
    del ens

    # plain loop: one context per instance, instances stepped one after another, with the
    # single-box path's default launch shape and with records written by the update kernel
    # (fewer launches, which favours small boxes); the faster of the two is the baseline
    def plain_loop(records_in_update):
        ctxs, states = [], []
        for q, cfg in enumerate(cfgs):
            p = cfg.params
            ctx = L.Context(p)
            ctx.set_tune(64, 1, 1, 1, 0, records_in_update)
            cu = fc.MultiFab(torch.from_numpy(E.uniform_cons(cfg)).cuda(), (-p.ngc,) * 3)
            prim = fc.MultiFab.cell(p, p.nprimvars)
            fc.rk3_step(ctx, cu, prim, cfg.dt, 1000 + q, 0)
            ctxs.append(ctx)
            states.append((cu, prim))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q, cfg in enumerate(cfgs):
            cu, prim = states[q]
            for s in range(loop_steps):
                fc.rk3_step(ctxs[q], cu, prim, cfg.dt, 1000 + q, 1 + s)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        for c in ctxs:
            c.close()
        return t

    t_loop0 = plain_loop(0)
    t_loop1 = plain_loop(1)
    t_loop = min(t_loop0, t_loop1)
    r_ens = cells * steps / t_ens
    r_loop = cells * loop_steps / t_loop
    return {"workload": f"{count} argon boxes, per-axis edges log-uniform in [16,64] (seed 2024), "
                        f"seeds 1000+i, uniform start, CFL 0.1 dt",
            "instances": count, "cells": cells, "steps": steps,
            "value": r_ens, "unit": "cell-updates/s", "ms_per_step": t_ens / steps * 1e3,
            "graph": True, "error_flags": bad,
            "plain_loop": {"value": r_loop, "unit": "cell-updates/s", "steps": loop_steps,
                           "ms_per_step": t_loop / loop_steps * 1e3,
                           "ms_per_step_default_shape": t_loop0 / loop_steps * 1e3,
                           "ms_per_step_records_in_update": t_loop1 / loop_steps * 1e3},
            "speedup_vs_plain_loop": r_ens / r_loop,
            "hbm_bytes_held": budget}


def serve_bench(rate: float, count: int, max_instances: int = 256, budget_gb: float = 160.0,
                chunk: int = 8, seed: int = 7, cap_cells: int = 2 << 20, variants=("batched", "smallest", "plain")):
    """The served stream: `count` requests with Poisson arrivals at `rate`/s (edge log-uniform
    in [16,128], 1-4 species, nsteps in [200,2000]), served by the continuous-batching server
    (admission in arrival order, and smallest work first with at most `cap_cells` cells in
    flight) and, for comparison, by the same server with one slot (requests strictly one after
    another).  Latency is arrival to result in host memory; service is admission to result."""
    import numpy as np
    from fhdex_amd import serving as S

    reqs = S.request_stream(count, seed)
    work = sum(c.cells * n for c, _, n, _ in reqs)

    def one(slots, policy=0, cap=0):
        srv = S.Server(slots, int(budget_gb * 2 ** 30), chunk_steps=chunk, graph=True)
        srv.set_policy(policy, cap)
        done, wall, open_n, _ = S.run_stream(srv, reqs, rate, seed=seed)
        srv.close()
        ok = [c for c in done if c["status"] == 0]
        lat = np.array([(c["done_ns"] - c["arrival_ns"]) * 1e-9 for c in ok])
        queue = np.array([(c["admit_ns"] - c["arrival_ns"]) * 1e-9 for c in ok])
        serv = np.array([(c["done_ns"] - c["admit_ns"]) * 1e-9 for c in ok])
        t_first = min(c["arrival_ns"] for c in done)
        t_last = max(c["done_ns"] for c in done)
        span = (t_last - t_first) * 1e-9
        pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None  # noqa: E731
        return {"completed": len(ok), "failed": len(done) - len(ok), "open": open_n,
                "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
                "p50_service_s": pct(serv, 50), "p99_service_s": pct(serv, 99),
                "mean_queue_s": float(queue.mean()) if len(queue) else None,
                "max_queue_s": float(queue.max()) if len(queue) else None,
                "value": work / span, "unit": "cell-updates/s", "span_s": span}

    out = {"workload": f"{count} requests, Poisson {rate:.3g}/s, edge log-uniform [16,128], "
                       f"species 1-4, nsteps 200-2000 (stream seed {seed})",
           "rate_per_s": rate, "requests": count, "cell_updates": work,
           "offered_load": work * rate / count, "slots": max_instances, "budget_GiB": budget_gb,
           "chunk_steps": chunk}
    if "batched" in variants:
        out["batched"] = one(max_instances)
    if "smallest" in variants:
        out["batched_smallest_first"] = dict(one(max_instances, S.Server.SMALLEST_FIRST, cap_cells),
                                             max_active_cells=cap_cells)
    if "plain" in variants:
        out["one_at_a_time"] = one(1)
        if "batched" in out:
            out["speedup_vs_one_at_a_time"] = out["batched"]["value"] / out["one_at_a_time"]["value"]
    return out


def serve_unsaturated(capacity: float, count: int = 64, load: float = 0.6, seed: int = 11):
    """`count` requests of the same mix arriving at the rate whose offered load is `load` of a
    measured capacity (cell-updates/s): below saturation, so a latency is the request's own
    service time, not backlog.  Batched (arrival order) and one at a time."""
    from fhdex_amd import serving as S
    reqs = S.request_stream(count, seed)
    mean_work = sum(c.cells * n for c, _, n, _ in reqs) / count
    rate = load * capacity / mean_work
    r = serve_bench(rate, count, seed=seed, variants=("batched", "plain"))
    r["offered_load_fraction"] = load
    r["capacity_cell_updates_per_s"] = capacity
    return r


def stag_configs(a):
    """The staggered scheme on the reference's periodic regression deck (four species 32^3,
    exec/compressible_stag/inputs_regression_periodic_3D, with the oracle beside it), its
    equilibrium deck size (four species 64^3, inputs_equil_3d), and argon at the headline box."""
    from fhdex_amd import params as P
    cases = {"deck32": lambda: stag_rate(P.four_species(32), 20, cpu_steps=0 if a.no_cpu else 3),
             "equil64": lambda: stag_rate(P.four_species(64), 20),
             "argon": lambda: stag_rate(P.argon_gas(a.n), 10)}
    pick = list(cases) if a.stag_case == "all" else [a.stag_case]
    return [cases[c]() for c in pick]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--ensemble", type=int, default=512, help="ensemble members (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--loop-steps", type=int, default=5)
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ref-steps", type=int, default=3,
                    help="steps of the reference-shaped step at --n (0: skip)")
    ap.add_argument("--ref-only", action="store_true", help="only time the reference-shaped step")
    ap.add_argument("--serve-rates", type=str, default="50",
                    help="comma-separated arrival rates of the served stream ('' to skip)")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-cap", type=int, default=2 << 20,
                    help="cells in flight of the smallest-first served run")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-unsaturated", type=int, default=64,
                    help="requests of a stream at 0.6 of the measured serving capacity (0: skip)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--no-configs", action="store_true", help="skip the 32^3 argon and 128^3 four-species rates")
    ap.add_argument("--no-stag", action="store_true", help="skip the staggered-scheme rates")
    ap.add_argument("--stag-only", action="store_true", help="only time the staggered-scheme step")
    ap.add_argument("--stag-case", type=str, default="all", choices=["all", "deck32", "equil64", "argon"])
    ap.add_argument("--shape", type=str, default="", help="launch shape tile_width,region_tiles,z_levels,noise_inline"
                                                         "[,face_blocks_per_cu[,records_in_update]]")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(boxes would shard by spatial slabs with 2-layer halo exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if not a.no_cpu or not a.no_stag:
        # every CPU leg of this run (the staggered comparison, the CPU baseline) times the
        # host-tuned oracle build, which has to be chosen before the oracle is first loaded
        from oracle import ffi as orc
        orc.use_native_build()

    import torch

    def copy_probe():
        """A plain 1 GiB device-to-device copy, 5 times: read + written bytes per second."""
        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        gbs = 5 * 2 * n * 8 / (time.perf_counter() - t0) / 1e9
        del src, dst
        return gbs

    if a.copy_probe:
        n = (1 << 30) // 8
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "copy_GBs": copy_probe()}), flush=True)
        return

    if a.serve_only:
        caps = []
        for r in [float(v) for v in a.serve_rates.split(",") if v]:
            e = serve_bench(r, a.serve_requests, cap_cells=a.serve_cap)
            caps.append(e["batched"]["value"])
            print(json.dumps({"serving": e}), flush=True)
        if a.serve_unsaturated and caps:
            print(json.dumps({"serving": serve_unsaturated(max(caps), a.serve_unsaturated)}), flush=True)
        return

    if a.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(a.ensemble, a.ensemble_steps, a.loop_steps)}),
              flush=True)
        return

    if a.ref_only:
        print(json.dumps({"reference_shaped": ref_rate(a.n, a.ref_steps, a.warmup)}), flush=True)
        return

    if a.stag_only:
        print(json.dumps({"staggered": stag_configs(a)}), flush=True)
        return

    from fhdex_amd import compressible as fc
    from fhdex_amd import lib as L
    from fhdex_amd import params as P
    from fhdex_amd import timing

    cfg = P.argon_gas(a.n)
    p = cfg.params
    ctx = L.Context(p)
    cu = fc.MultiFab.cell(p, p.nvars)
    prim = fc.MultiFab.cell(p, p.nprimvars)
    # uniform start (InitConsVar prob_type 1), built on the device
    cv = p.hcv[0] * cfg.rhobar[0] + p.hcv[1] * cfg.rhobar[1]
    cu.t[0] = cfg.rho0
    cu.t[4] = cfg.rho0 * cv * cfg.T_init
    cu.t[5] = cfg.rho0 * cfg.rhobar[0]
    cu.t[6] = cfg.rho0 * cfg.rhobar[1]
    seed = 1
    tune = None
    if a.shape:
        ctx.set_tune(*(int(v) for v in a.shape.split(",")))
    elif not a.no_tune:
        best, trials = ctx.autotune(cu.fab, cfg.dt, trial_steps=2)
        tune = {"chosen": best.as_dict(), "trials": [t.as_dict() for t in trials]}
    for s in range(a.warmup):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, s)
    torch.cuda.synchronize()
    ctx.check()
    timing.enable(ctx, True)   # HIP events around every kernel group, on the launch stream
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for s in range(a.steps):
        fc.rk3_step(ctx, cu, prim, cfg.dt, seed, a.warmup + s)
    stop.record()
    torch.cuda.synchronize()
    ctx.check()
    groups = timing.read(ctx)
    timing.enable(ctx, False)
    ms = start.elapsed_time(stop) / a.steps
    cells = a.n ** 3
    rate = cells / (ms * 1e-3)
    nv = p.nvars
    # the unit of algorithmic work is one RK3 step of the whole box: B_alg bytes per cell-update
    step_ms_events = sum(t for _, t in groups.values()) / a.steps
    achieved = b_alg(nv) * cells / (step_ms_events * 1e-3) / 1e9
    dom = max(groups.items(), key=lambda kv: kv[1][1])
    roof = {"bound": "hbm",
            "kernel": "rk3_step: one full RK3 step of the box (every kernel of the step, HIP events on "
                      "the launch stream); algorithmic bytes = B_alg x cells",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.n),
            "avg_launch_ms": step_ms_events,
            "dominant_kernel": {"name": "k_faces" if dom[0] == "faces" else dom[0],
                                "launches_per_step": dom[1][0] / a.steps,
                                "avg_launch_ms": dom[1][1] / max(dom[1][0], 1)},
            "groups_ms_per_step": {k: v[1] / a.steps for k, v in groups.items()},
            "copy_probe_GBs": copy_probe()}
    fl = measured_flops(a.n)
    if fl is not None:
        roof["fp64"] = {"achieved": fl * rate / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": fl * rate / 1e12 / 78.6, "flops_per_cell_update": fl}
    line = {
        "metric": "fp64_cell_updates_per_s",
        "value": rate,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"argon two-species {a.n}^3 periodic, CFL 0.1 dt, Philox noise, "
                               f"fused RK3 step", "cells": cells, "nvars": nv,
                   "b_alg_bytes_per_cell_update": b_alg(nv),
                   "step_hbm_frac_b_alg": rate * b_alg(nv) / 1e9 / HBM_PEAK_GBS},
        "roofline": roof,
        "launch_shape": ctx.get_tune().as_dict() if tune is None else tune,
    }
    ctx.close()
    del cu, prim
    if not a.no_configs:
        vw = P.four_species(64, 10)
        vw.params.transport_type = 2
        vw.name += " Valk-Waldmann transport"
        line["configs"] = [config_rate(P.argon_gas(32), 50), config_rate(P.four_species(128), 10),
                           config_rate(P.species_prefix(8, 64, 10), 10), config_rate(vw, 10),
                           config_rate(P.no2_n2o4(128), 10), config_rate(P.taylor_green(64), 50),
                           wall_rate(P.wall_deck("slip_wall_adiabatic")),
                           wall_rate(P.wall_deck("temp_grad")), wall_rate(P.wall_deck("RT_3d")),
                           wall_rate(P.wall_deck("giantfluct_3d"))]
    if a.ref_steps:
        line["reference_shaped"] = ref_rate(a.n, a.ref_steps)
    if not a.no_stag:
        line["staggered"] = stag_configs(a)
    if a.ensemble:
        line["ensemble"] = ensemble_bench(a.ensemble, a.ensemble_steps, a.loop_steps)
    rates = [float(v) for v in a.serve_rates.split(",") if v]
    if rates:
        line["serving"] = [serve_bench(r, a.serve_requests, cap_cells=a.serve_cap) for r in rates]
        if a.serve_unsaturated:
            cap = max(e["batched"]["value"] for e in line["serving"])
            line["serving"].append(serve_unsaturated(cap, a.serve_unsaturated))
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
