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
    """HBM bytes per step from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this
    workload (profiles/traffic_argon<n>.json), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_argon{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["bytes_per_step"]


def cpu_baseline(n: int, steps: int):
    """The CPU oracle (the reference's arithmetic, OpenMP over z-planes) on the host cores."""
    from fhdex_amd import params as P
    from oracle import ffi as orc
    cfg = P.argon_gas(n)
    p = cfg.params
    st = orc.HostState(p)
    orc.init_cons(p, st.cu, cfg.rho0, cfg.rhobar, cfg.T_init)
    t0 = time.perf_counter()
    rc = orc.advance(p, st.cu, st.prim, steps, cfg.dt, None, 1, 0, 0)
    dt = time.perf_counter() - t0
    assert rc == 0
    return {"value": n ** 3 * steps / dt, "unit": "cell-updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"oracle (CPU restatement of the reference step), argon {n}^3 x {steps} "
                      f"RK3 steps, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if a.copy_probe:
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
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "copy_GBs": gbs}), flush=True)
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
    roof = {"bound": "hbm", "kernel": "rk3_step (all kernel groups of one step, HIP-event timed)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.n),
            "avg_launch_ms": step_ms_events,
            "dominant_group": {"name": dom[0], "launches_per_step": dom[1][0] / a.steps,
                               "ms_per_step": dom[1][1] / a.steps},
            "groups_ms_per_step": {k: v[1] / a.steps for k, v in groups.items()}}
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
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_steps)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
