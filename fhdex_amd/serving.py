"""Continuous-batching request loop (fhdx_server_* C ABI): ensemble runs submitted as they
arrive, admitted by HBM budget and slots, advanced together, results copied to host memory.

The worker loop is native (fhdx_server in libfhdx.so); this module is its ctypes binding plus
the request-stream generator used to measure it.  The reference runs one box per process
(main_driver.cpp:574); there is no counterpart there.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .lib import check, lib
from .ensemble import grown_shape
from .params import MAX_SPECIES, Config, Params


class Request(C.Structure):
    """``fhdx_request`` (include/fhdx.h)."""

    _fields_ = [
        ("params", Params),
        ("dt", C.c_double),
        ("seed", C.c_uint64),
        ("nsteps", C.c_uint64),
        ("cu_init", C.c_void_p),
        ("cu_out", C.c_void_p),
        ("prim_out", C.c_void_p),
        ("rho0", C.c_double),
        ("T0", C.c_double),
        ("Y0", C.c_double * MAX_SPECIES),
    ]


class Completion(C.Structure):
    """``fhdx_completion`` (include/fhdx.h)."""

    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int32),
        ("pad", C.c_int32),
        ("arrival_ns", C.c_int64),
        ("admit_ns", C.c_int64),
        ("done_ns", C.c_int64),
        ("steps", C.c_uint64),
        ("message", C.c_char * 96),
    ]


def _bind():
    L = lib()
    if getattr(L, "_serving_bound", False):
        return L
    vp = C.c_void_p
    L.fhdx_server_create.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_int, C.POINTER(vp)]
    L.fhdx_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
    L.fhdx_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_int, C.POINTER(C.c_int)]
    L.fhdx_server_wait.argtypes = [vp, C.c_double, C.POINTER(C.c_int64)]
    L.fhdx_server_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]
    L.fhdx_server_destroy.argtypes = [vp]
    L.fhdx_server_set_policy.argtypes = [vp, C.c_int, C.c_uint64]
    for n in ("fhdx_server_create", "fhdx_server_submit", "fhdx_server_poll", "fhdx_server_wait",
              "fhdx_server_stats", "fhdx_server_destroy", "fhdx_server_set_policy"):
        getattr(L, n).restype = C.c_int
    L._serving_bound = True
    return L


class Server:
    """Owns an fhdx_server (engine + worker thread)."""

    def __init__(self, max_instances: int, mem_budget_bytes: int, chunk_steps: int = 8,
                 graph: bool = True):
        L = _bind()
        h = C.c_void_p()
        check(L.fhdx_server_create(int(max_instances), int(mem_budget_bytes), int(chunk_steps),
                                   int(bool(graph)), C.byref(h)), "fhdx_server_create")
        self.h = h
        self._keep = {}   # request id -> buffers that must outlive the request

    ARRIVAL, SMALLEST_FIRST = 0, 1

    def set_policy(self, policy: int, max_active_cells: int = 0):
        """fhdx_server_set_policy: admission in arrival order (0) or smallest work first (1),
        with an optional cap on the cells in flight."""
        check(_bind().fhdx_server_set_policy(self.h, int(policy), int(max_active_cells)),
              "fhdx_server_set_policy")

    def close(self):
        if self.h:
            lib().fhdx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cfg: Config, seed: int, nsteps: int, cu_init=None, cu_out=None,
               prim_out=None) -> int:
        """Queue a run of `nsteps` steps of `cfg`.  cu_init: grown start state (numpy), or None
        for the uniform start of cfg; cu_out / prim_out: numpy buffers for the result, or None."""
        r = Request()
        r.params = cfg.params
        r.dt = cfg.dt
        r.seed = int(seed)
        r.nsteps = int(nsteps)
        keep = []
        for name, arr in (("cu_init", cu_init), ("cu_out", cu_out), ("prim_out", prim_out)):
            if arr is not None:
                assert isinstance(arr, np.ndarray) and arr.dtype == np.float64 and arr.flags.c_contiguous
                setattr(r, name, arr.ctypes.data)
                keep.append(arr)
        r.rho0 = cfg.rho0
        r.T0 = cfg.T_init
        for k, y in enumerate(cfg.rhobar):
            r.Y0[k] = y
        rid = C.c_int64(-1)
        check(_bind().fhdx_server_submit(self.h, C.byref(r), C.byref(rid)), "fhdx_server_submit")
        if keep:
            self._keep[rid.value] = keep
        return rid.value

    def poll(self, max_n: int = 256):
        buf = (Completion * max_n)()
        n = C.c_int(0)
        check(_bind().fhdx_server_poll(self.h, buf, max_n, C.byref(n)), "fhdx_server_poll")
        out = []
        for c in buf[:n.value]:
            self._keep.pop(c.id, None)
            out.append({"id": c.id, "status": c.status, "arrival_ns": c.arrival_ns,
                        "admit_ns": c.admit_ns, "done_ns": c.done_ns, "steps": c.steps,
                        "message": c.message.decode(errors="replace")})
        return out

    def wait(self, timeout_s: float = 600.0) -> int:
        """Waits for every submitted request; returns how many are still open."""
        n = C.c_int64(0)
        _bind().fhdx_server_wait(self.h, float(timeout_s), C.byref(n))
        return n.value

    def stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(_bind().fhdx_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c)),
              "fhdx_server_stats")
        return {"steps": a.value, "cell_updates": b.value, "bytes_used": c.value}


def request_stream(count: int, seed: int, edge=(16, 128), species=(1, 4), nsteps=(200, 2000)):
    """The served request mix: cubic edge log-uniform in [16,128], species count uniform in
    1..4 (1 is the argon gas as two identical species; 2..4 are prefixes of the four-species
    list with equal mass fractions), nsteps uniform in [200,2000], one seed each."""
    from .params import species_prefix
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        e = int(np.rint(np.exp(rng.uniform(np.log(edge[0]), np.log(edge[1])))))
        ns = int(rng.integers(species[0], species[1] + 1))
        n = int(rng.integers(nsteps[0], nsteps[1] + 1))
        out.append((species_prefix(ns, e, n), ns, n, 7000 + q))
    return out


def run_stream(server: Server, reqs, rate: float, seed: int = 1, timeout_s: float = 900.0,
               results: bool = True):
    """Submit `reqs` with Poisson arrivals at `rate` per second (exponential gaps from `seed`),
    wait for all of them and return the completions plus the wall time.  results: each request
    gets a host buffer for its final conserved state, so a completion time is the time its
    result reached host memory."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    t0 = time.perf_counter()
    t_next = t0
    ids = {}
    for (cfg, ns, n, s), gap in zip(reqs, gaps):
        t_next += gap
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        out = np.empty(grown_shape(cfg.params, cfg.params.nvars)) if results else None
        ids[server.submit(cfg, s, n, cu_out=out)] = (cfg, ns, n)
    open_n = server.wait(timeout_s)
    wall = time.perf_counter() - t0
    done = server.poll(len(reqs) + 16)
    return done, wall, open_n, ids
