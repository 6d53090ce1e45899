"""Batched ensembles over the fhdx_engine_* C ABI: many independent periodic boxes, each with
its own size, species count, time step and noise seed, advanced by shared launches.

An instance's result is bit-identical to `compressible.rk3_step` on the same state, dt, seed
and step index.  The reference runs one box per process (main_driver.cpp:574); this module has
no counterpart there.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import check, lib
from .params import Config, Fab, Params


def grown_shape(p: Params, ncomp: int):
    g = p.ngc
    n0, n1, n2 = p.n_cells
    return (ncomp, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g)


def uniform_values(cfg: Config):
    """Per-component values of the uniform start (InitConsVar prob_type 1,
    compressible_functions.cpp:163-196): rho0, zero momentum, rho0 * sum_k Y_k (e0_k + cv_k T),
    rho0 * Y_k."""
    p = cfg.params
    ns = p.nspecies
    ie = 0.0
    for k in range(ns):
        ie += cfg.rhobar[k] * (p.e0[k] + p.hcv[k] * cfg.T_init)
    return [cfg.rho0, 0.0, 0.0, 0.0, cfg.rho0 * ie] + [cfg.rho0 * cfg.rhobar[k] for k in range(ns)]


def uniform_cons(cfg: Config) -> np.ndarray:
    """The uniform start on the box grown by ngc, as a host array."""
    p = cfg.params
    cu = np.zeros(grown_shape(p, p.nvars), dtype=np.float64)
    for c, v in enumerate(uniform_values(cfg)):
        cu[c] = v
    return cu


def bytes_needed(p: Params) -> int:
    return int(lib().fhdx_engine_bytes_needed(C.byref(p)))


class Ensemble:
    """Owns an fhdx_engine: slots, an HBM budget and one stream."""

    def __init__(self, max_instances: int, mem_budget_bytes: int):
        h = C.c_void_p()
        check(lib().fhdx_engine_create(int(max_instances), int(mem_budget_bytes), C.byref(h)),
              "fhdx_engine_create")
        self.h = h
        self.max_instances = int(max_instances)
        self.params = {}

    def close(self):
        if self.h:
            lib().fhdx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, p: Params, dt: float, seed: int, cu_init, step0: int = 0, nsteps: int = 0) -> int:
        """Admit an instance.  cu_init: grown conserved state, a C-contiguous float64 numpy array
        or device tensor shaped (nvars, nz+2g, ny+2g, nx+2g)."""
        if isinstance(cu_init, np.ndarray):
            a = np.ascontiguousarray(cu_init, dtype=np.float64)
            assert a.shape == grown_shape(p, p.nvars)
            ptr = a.ctypes.data
        else:
            assert cu_init.is_contiguous() and tuple(cu_init.shape) == grown_shape(p, p.nvars)
            ptr = cu_init.data_ptr()
        slot = C.c_int(-1)
        check(lib().fhdx_engine_add(self.h, C.byref(p), float(dt), int(seed), int(step0),
                                    int(nsteps), C.c_void_p(ptr), C.byref(slot)), "fhdx_engine_add")
        self.params[slot.value] = p.copy()
        return slot.value

    def add_uniform(self, cfg: Config, seed: int, step0: int = 0, nsteps: int = 0) -> int:
        """Admit an instance with the uniform start of `cfg` (filled on the device)."""
        p = cfg.params
        vals = uniform_values(cfg)
        arr = (C.c_double * len(vals))(*vals)
        slot = C.c_int(-1)
        check(lib().fhdx_engine_add_uniform(self.h, C.byref(p), float(cfg.dt), int(seed), int(step0),
                                            int(nsteps), arr, C.byref(slot)), "fhdx_engine_add_uniform")
        self.params[slot.value] = p.copy()
        return slot.value

    def remove(self, slot: int):
        check(lib().fhdx_engine_remove(self.h, int(slot)), "fhdx_engine_remove")
        self.params.pop(slot, None)

    def step(self, nsteps: int, graph: bool = True):
        check(lib().fhdx_engine_step(self.h, int(nsteps), int(bool(graph))), "fhdx_engine_step")

    def sync(self):
        check(lib().fhdx_engine_sync(self.h), "fhdx_engine_sync")

    def status(self, slot: int):
        step = C.c_uint64()
        done = C.c_int()
        check(lib().fhdx_engine_state(self.h, int(slot), None, None, C.byref(step), C.byref(done)),
              "fhdx_engine_state")
        return step.value, bool(done.value)

    def device_fabs(self, slot: int):
        cu, prim = Fab(), Fab()
        check(lib().fhdx_engine_state(self.h, int(slot), C.byref(cu), C.byref(prim), None, None),
              "fhdx_engine_state")
        return cu, prim

    def read(self, slot: int):
        """(cu, prim) of a slot as host arrays, after the queued steps."""
        p = self.params[slot]
        cu = np.empty(grown_shape(p, p.nvars), dtype=np.float64)
        prim = np.empty(grown_shape(p, p.nprimvars), dtype=np.float64)
        check(lib().fhdx_engine_read(self.h, int(slot), C.c_void_p(cu.ctypes.data),
                                     C.c_void_p(prim.ctypes.data)), "fhdx_engine_read")
        return cu, prim

    def errors(self):
        flags = (C.c_int * self.max_instances)()
        check(lib().fhdx_engine_errors(self.h, flags), "fhdx_engine_errors")
        return list(flags)

    @property
    def bytes_used(self) -> int:
        return int(lib().fhdx_engine_bytes_used(self.h))
