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


def ensemble_members(count: int, seed: int = 2024, lo: int = 16, hi: int = 64):
    """Box shapes of the batched-ensemble workload: per-axis edge round(exp(U(ln lo, ln hi)))
    drawn independently per axis from a fixed seed (512 argon boxes, seeds 1000 + i, in the
    benchmark)."""
    rng = np.random.default_rng(seed)
    e = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))).astype(int)
    return [tuple(int(v) for v in row) for row in e]


def uniform_values(cfg: Config):
    """Per-component values of the uniform start (InitConsVar prob_type 1,
    compressible_functions.cpp:163-196): rho0, zero momentum, rho0 * e, rho0 * Y_k, with the
    internal energy e = e0mix + T cvmix of GetEnergy (compressible_functions.H:1498-1511):
    cvmix = sum_k Y_k cv_k and e0mix = sum_k Y_k e0_k summed separately, in that order."""
    p = cfg.params
    ns = p.nspecies
    cvmix = 0.0
    e0mix = 0.0
    for k in range(ns):
        cvmix += cfg.rhobar[k] * p.hcv[k]
        e0mix += cfg.rhobar[k] * p.e0[k]
    ie = e0mix + cfg.T_init * cvmix
    return [cfg.rho0, 0.0, 0.0, 0.0, cfg.rho0 * ie] + [cfg.rho0 * cfg.rhobar[k] for k in range(ns)]


def uniform_cons(cfg: Config) -> np.ndarray:
    """The uniform start on the box grown by ngc, as a host array."""
    p = cfg.params
    cu = np.zeros(grown_shape(p, p.nvars), dtype=np.float64)
    for c, v in enumerate(uniform_values(cfg)):
        cu[c] = v
    return cu


def initial_cons(cfg: Config) -> np.ndarray:
    """InitConsVar's conserved start (compressible_functions.cpp:103-413) on the grown box, as a
    host array, for cfg.prob_type
      1 uniform;
      2 Rayleigh-Taylor layers (:172-214): four species, mass fractions 0.4/0.4/0.1/0.1 above
        the mid-plane in z and 0.1/0.1/0.4/0.4 below, density rho0 exp(g_z z / (R_mix T)) at
        hydrostatic equilibrium;
      3 diffusion barrier (:216-227): mass fractions linear in y between the y walls'
        compositions;
      4 Taylor-Green vortex (:228-262): two species;
      100 sinusoidal density in y at the uniform energy density (:292-302), ghosts filled.
    The domain is [0, n dx].  For types 2, 3 and 4 the ghost cells keep the uniform values:
    fill them with fill_boundary / setBC, as the reference does after InitConsVar
    (main_driver.cpp:532-542)."""
    cu = uniform_cons(cfg)
    if cfg.prob_type == 1:
        return cu
    p = cfg.params
    ns = p.nspecies
    g = p.ngc
    n0, n1, n2 = p.n_cells
    v = (slice(None), slice(g, g + n2), slice(g, g + n1), slice(g, g + n0))
    rho = np.full((n2, n1, n0), cfg.rho0)
    if cfg.prob_type == 100:
        # sinusoidal density in y at the uniform energy density (:292-302), mass fractions kept;
        # the ghost cells are the periodic images
        y = ((np.arange(n1) + 0.5) * p.dx[1])[None, :, None]
        Ly = p.dx[1] * n1
        Yk = [cu[5 + l][v[1:]] / cu[0][v[1:]] for l in range(ns)]
        r = cfg.rho0 + 0.1 * cfg.rho0 * np.sin(2. * np.arccos(-1.) * y / Ly)
        cu[v][0] = r
        for l in range(ns):
            cu[v][5 + l] = cu[v][0] * Yk[l]
        for ax, n in ((3, n0), (2, n1), (1, n2)):           # periodic images, x, then y, then z
            src = (np.arange(-g, n + g) % n) + g
            cu = np.take(cu, src, axis=ax)
        return np.ascontiguousarray(cu)
    if cfg.prob_type == 4:
        # Taylor-Green vortex (:228-262): positions from the lower corner, all scaled by the
        # x extent; rho from the pressure at T_init; rho E = p / (5/3 - 1) + kinetic energy
        if ns != 2:
            raise ValueError("prob_type 4 (Taylor-Green) is a two-species start")
        pi = np.arccos(-1.0)
        Lf = n0 * p.dx[0]
        avogadro = p.Runiv / p.k_B
        velscale, pscale = 9237.0, 884.147e3
        rhoscale = p.molmass[0] / avogadro * pscale / (p.k_B * cfg.T_init)
        x = ((np.arange(n0) + 0.5) * p.dx[0])[None, None, :]
        y = ((np.arange(n1) + 0.5) * p.dx[1])[None, :, None]
        z = ((np.arange(n2) + 0.5) * p.dx[2])[:, None, None]
        pres = pscale + rhoscale * velscale * velscale * np.cos(2. * pi * x / Lf) * np.cos(4. * pi * y / Lf) \
            * (np.cos(4. * pi * z / Lf) + 2.)
        r = (p.molmass[0] / avogadro) * pres / (p.k_B * cfg.T_init)
        m1 = velscale * r * np.sin(2. * pi * x / Lf) * np.cos(2. * pi * y / Lf) * np.cos(2. * pi * z / Lf)
        m2 = -velscale * r * np.cos(2. * pi * x / Lf) * np.sin(2. * pi * y / Lf) * np.cos(2. * pi * z / Lf)
        cu[v][0] = r
        cu[v][1] = m1
        cu[v][2] = m2
        cu[v][3] = 0.0
        cu[v][4] = pres / (5. / 3. - 1.) + 0.5 * (m1 * m1 + m2 * m2) / r
        cu[v][5] = 0.5 * r
        cu[v][6] = 0.5 * r
        return cu
    if cfg.prob_type == 2:
        if ns != 4:
            raise ValueError("prob_type 2 (Rayleigh-Taylor) is a four-species start")
        posz = ((np.arange(n2) + 0.5) * p.dx[2])[:, None, None]
        top = posz - n2 * p.dx[2] / 2.0 >= 0.0
        Y = np.empty((ns, n2, 1, 1))
        for l, (hi, lo) in enumerate(((0.4, 0.1), (0.4, 0.1), (0.1, 0.4), (0.1, 0.4))):
            Y[l] = np.where(top, hi, lo)
        molmix = 1.0 / sum(Y[l] / p.molmass[l] for l in range(ns))
        pamb = rho * (p.Runiv / molmix) * cfg.T_init
        rgasmix = p.Runiv / molmix
        alpha = p.grav[2] / (rgasmix * cfg.T_init)
        rho = pamb * np.exp(alpha * posz) / (rgasmix * cfg.T_init)
    elif cfg.prob_type == 3:
        posy = ((np.arange(n1) + 0.5) * p.dx[1])[None, :, None]
        Ly = n1 * p.dx[1]
        Y = np.empty((ns, 1, n1, 1))
        for l in range(ns):
            Y[l] = (p.bc_Yk[3][l] - p.bc_Yk[2][l]) / Ly * posy + p.bc_Yk[2][l]
    else:
        raise ValueError(f"prob_type {cfg.prob_type} is not built (types 1 to 4 and 100 are)")
    cv = sum(Y[l] * p.hcv[l] for l in range(ns))
    e0 = sum(Y[l] * p.e0[l] for l in range(ns))
    cu[v][0] = rho
    cu[v][1:4] = 0.0
    cu[v][4] = rho * (e0 + cfg.T_init * cv)
    for l in range(ns):
        cu[v][5 + l] = rho * Y[l]
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
