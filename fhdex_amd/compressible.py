"""Host-side mirror of src_compressible's interface over the C ABI.

Function names and argument meaning follow the reference (compressible_functions.H:23-55):
``conservedToPrimitive``, ``calculateTransportCoeffs``, ``calculateFlux``, ``RK3step``, plus
the fast fused step ``rk3_step``.  Arrays are MultiFab-shaped float64 device tensors laid out
(ncomp, z, y, x) with the reference's ghost widths; `DeviceState` allocates the whole argument
list of RK3step.  torch is used only for device memory and streams.
"""
from __future__ import annotations

import ctypes as C

from .lib import Context, check, fab_of_tensor, lib
from .params import Fab, Params


class MultiFab:
    """One FAB: a device tensor plus its index box.  The constructors take the valid box as
    (lo, hi) global cell indices, inclusive; by default the whole domain (a single-box MultiFab)."""

    def __init__(self, t, lo):
        self.t = t
        self.lo = tuple(lo)

    @property
    def fab(self) -> Fab:
        return fab_of_tensor(self.t, self.lo)

    @staticmethod
    def _box(p: Params, box):
        if box is None:
            return (0, 0, 0), tuple(n - 1 for n in p.n_cells)
        return tuple(box[0]), tuple(box[1])

    @staticmethod
    def cell(p: Params, ncomp: int, device="cuda", box=None):
        import torch
        lo, hi = MultiFab._box(p, box)
        g = p.ngc
        n0, n1, n2 = (hi[d] - lo[d] + 1 for d in range(3))
        t = torch.zeros((ncomp, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g), dtype=torch.float64, device=device)
        return MultiFab(t, tuple(l - g for l in lo))

    @staticmethod
    def face(p: Params, d: int, ncomp: int, device="cuda", box=None):
        import torch
        lo, hi = MultiFab._box(p, box)
        n0, n1, n2 = (hi[e] - lo[e] + 1 for e in range(3))
        t = torch.zeros((ncomp, n2 + (d == 2), n1 + (d == 1), n0 + (d == 0)), dtype=torch.float64,
                        device=device)
        return MultiFab(t, lo)

    @staticmethod
    def node(p: Params, ncomp: int = 1, device="cuda", box=None):
        import torch
        lo, hi = MultiFab._box(p, box)
        n0, n1, n2 = (hi[e] - lo[e] + 1 for e in range(3))
        t = torch.zeros((ncomp, n2 + 1, n1 + 1, n0 + 1), dtype=torch.float64, device=device)
        return MultiFab(t, lo)


def _fa(mfs):
    return (Fab * 3)(*[m.fab for m in mfs])


class DeviceState:
    """Every MultiFab of RK3step's argument list (main_driver.cpp:160-297) on the device, for one
    box (the whole domain by default); with chemistry also source (nvars) and ranchem
    (nreaction), as main_driver.cpp:241-245 defines them."""

    def __init__(self, p: Params, device="cuda", box=None):
        self.p = p
        self.box = box
        nv, NS = p.nvars, p.nspecies
        cell = lambda nc: MultiFab.cell(p, nc, device, box)  # noqa: E731
        node = lambda: MultiFab.node(p, 1, device, box)  # noqa: E731
        self.cu, self.cup, self.cup2 = cell(nv), cell(nv), cell(nv)
        self.prim = cell(p.nprimvars)
        self.eta, self.zeta, self.kappa = cell(1), cell(1), cell(1)
        self.chi, self.D = cell(NS), cell(NS * NS)
        self.flux = [MultiFab.face(p, d, nv + 4, device, box) for d in range(3)]
        self.stochFlux = [MultiFab.face(p, d, nv, device, box) for d in range(3)]
        self.cornx = [node() for _ in range(3)]
        self.corny = [node() for _ in range(3)]
        self.cornz = [node() for _ in range(3)]
        self.visccorn = node()
        self.rancorn = node()
        self.source = cell(nv) if p.nreaction > 0 else None
        self.ranchem = cell(p.nreaction) if p.nreaction > 0 else None


class NoiseSets:
    """The A/B noise of one step (stochFlux_A/B, rancorn_A/B of timeStep.cpp:38-53) for one box;
    with chemistry also ranchem_A/B (timeStep.cpp:64-72), as cell FABs with ngc ghosts."""

    def __init__(self, p: Params, device="cuda", box=None):
        nv = p.nvars
        self.A = [MultiFab.face(p, d, nv, device, box) for d in range(3)]
        self.B = [MultiFab.face(p, d, nv, device, box) for d in range(3)]
        self.rA = MultiFab.node(p, 1, device, box)
        self.rB = MultiFab.node(p, 1, device, box)
        self.chemA = MultiFab.cell(p, p.nreaction, device, box) if p.nreaction > 0 else None
        self.chemB = MultiFab.cell(p, p.nreaction, device, box) if p.nreaction > 0 else None


def conservedToPrimitive(ctx: Context, prim: MultiFab, cons: MultiFab, stream=None):
    check(lib().fhdx_cons_to_prim(ctx.h, prim.fab, cons.fab, stream), "conservedToPrimitive")


def FillBoundary(ctx: Context, mf: MultiFab, stream=None):
    check(lib().fhdx_fill_boundary(ctx.h, mf.fab, stream), "FillBoundary")


def setBC(ctx: Context, prim: MultiFab, cons: MultiFab, stream=None):
    """setBC (boundary.cpp:172-1209): the ghost cells outside walls, after FillBoundary."""
    check(lib().fhdx_set_bc(ctx.h, prim.fab, cons.fab, stream), "setBC")


def calculateTransportCoeffs(ctx: Context, prim, eta, zeta, kappa, chi, D, stream=None):
    check(lib().fhdx_transport(ctx.h, prim.fab, eta.fab, zeta.fab, kappa.fab, chi.fab, D.fab, stream),
          "calculateTransportCoeffs")


def calculateFlux(ctx: Context, cons, prim, eta, zeta, kappa, chi, D, flux, stochFlux, cornx, corny,
                  cornz, visccorn, rancorn, dt, stream=None):
    check(lib().fhdx_calculate_flux(ctx.h, cons.fab, prim.fab, eta.fab, zeta.fab, kappa.fab, chi.fab,
                                    D.fab, _fa(flux), _fa(stochFlux), _fa(cornx), _fa(corny),
                                    _fa(cornz), visccorn.fab, rancorn.fab, dt, stream),
          "calculateFlux")


def fill_noise(ctx: Context, noise: NoiseSets, seed: int, step: int, instance: int = 0, stream=None):
    check(lib().fhdx_fill_noise(ctx.h, seed, step, instance, _fa(noise.A), _fa(noise.B),
                                noise.rA.fab, noise.rB.fab, stream), "fill_noise")


def RK3step(ctx: Context, st: DeviceState, dt: float, noise: NoiseSets | None = None, seed: int = 1,
            step: int = 0, stream=None):
    """RK3step (timeStep.cpp:11-427) on the reference's arrays; noise from `noise` if given,
    else Philox (seed, step)."""
    if noise is not None:
        a, b = _fa(noise.A), _fa(noise.B)
        ra, rb = noise.rA.fab, noise.rB.fab
        args = (a, b, C.pointer(ra), C.pointer(rb))
    else:
        args = (None, None, None, None)
    check(lib().fhdx_rk3_step_ref(ctx.h, st.cu.fab, st.cup.fab, st.cup2.fab, st.prim.fab, st.eta.fab,
                                  st.zeta.fab, st.kappa.fab, st.chi.fab, st.D.fab, _fa(st.flux),
                                  _fa(st.stochFlux), _fa(st.cornx), _fa(st.corny), _fa(st.cornz),
                                  st.visccorn.fab, st.rancorn.fab, *args, dt, seed, step, stream),
          "RK3step")


def RK3stage(ctx: Context, stage: int, st: DeviceState, noise: NoiseSets, dt: float, stream=None):
    """Stage `stage` (0, 1, 2) of RK3step on one FAB (fhdx_rk3_stage_ref): no ghost cell is
    filled; between stages the caller exchanges the ghosts of the stage output and prim across
    the FABs and applies setBC, as timeStep.cpp:204-209 does."""
    z = Fab()
    chem = ctx.p.nreaction > 0
    check(lib().fhdx_rk3_stage_ref(ctx.h, int(stage), st.cu.fab, st.cup.fab, st.cup2.fab, st.prim.fab,
                                   st.eta.fab, st.zeta.fab, st.kappa.fab, st.chi.fab, st.D.fab,
                                   _fa(st.flux), _fa(st.stochFlux), _fa(st.cornx), _fa(st.corny),
                                   _fa(st.cornz), st.visccorn.fab, st.rancorn.fab, _fa(noise.A),
                                   _fa(noise.B), noise.rA.fab, noise.rB.fab,
                                   st.source.fab if chem else z, st.ranchem.fab if chem else z,
                                   noise.chemA.fab if chem else z, noise.chemB.fab if chem else z,
                                   float(dt), stream), "RK3stage")


def rk3_step(ctx: Context, cu: MultiFab, prim: MultiFab, dt: float, seed: int, step: int, stream=None):
    """Fast fused RK3 step with in-kernel Philox noise."""
    check(lib().fhdx_rk3_step(ctx.h, cu.fab, prim.fab, dt, seed, step, stream), "rk3_step")


def rk3_step_injected(ctx: Context, cu: MultiFab, prim: MultiFab, noise: NoiseSets, dt: float,
                      stream=None):
    check(lib().fhdx_rk3_step_injected(ctx.h, cu.fab, prim.fab, _fa(noise.A), _fa(noise.B),
                                       noise.rA.fab, noise.rB.fab, dt, stream), "rk3_step_injected")


class StatsState:
    """The statistics MultiFabs of main_driver.cpp:205-238 on the device (zero-initialised):
    cuMeans, cuVars (nvars), primMeans (nprimvars), primVars (nprimvars+5), spatialCross (8),
    miscStats (10), plus miscVals[20] on the host."""

    def __init__(self, p: Params, device="cuda"):
        import numpy as np
        self.consMean = MultiFab.cell(p, p.nvars, device)
        self.consVar = MultiFab.cell(p, p.nvars, device)
        self.primMean = MultiFab.cell(p, p.nprimvars, device)
        self.primVar = MultiFab.cell(p, p.nprimvars + 5, device)
        self.spatialCross = MultiFab.cell(p, 8, device)
        self.miscStats = MultiFab.cell(p, 10, device)
        self.miscVals = np.zeros(20)


def evaluateStats(ctx: Context, cons: MultiFab, prim: MultiFab, st: StatsState, steps: int,
                  cross_cell: int = 0, stream=None):
    """evaluateStats (stats.cpp:5-363): updates the running statistics in `st`; steps is the
    running sample count (statsCount, starting at 1)."""
    check(lib().fhdx_evaluate_stats(ctx.h, cons.fab, st.consMean.fab, st.consVar.fab, prim.fab,
                                    st.primMean.fab, st.primVar.fab, st.spatialCross.fab,
                                    st.miscStats.fab,
                                    st.miscVals.ctypes.data_as(C.POINTER(C.c_double)), int(steps),
                                    int(cross_cell), stream), "evaluateStats")


# ---- on-disk formats (writePlotFile.cpp, checkpoint.cpp) -------------------------------------

def _fields(cu: MultiFab, prim: MultiFab, st: StatsState, eta: MultiFab, kappa: MultiFab):
    from .lib import Fields
    mfs = (cu, st.consMean, st.consVar, prim, st.primMean, st.primVar, st.spatialCross,
           st.miscStats, eta, kappa)
    host = {not m.t.is_cuda for m in mfs}
    if len(host) != 1:
        raise ValueError("the fields must all be device arrays or all host arrays")
    f = Fields(*[m.fab for m in mfs])
    return f, (1 if host.pop() else 0)


def WritePlotFile(p: Params, step: int, time: float, prob_lo, cu: MultiFab, prim: MultiFab,
                  st: StatsState, eta: MultiFab, kappa: MultiFab, plot_base_name: str = "plt",
                  plot_means: int = 1, plot_vars: int = 1, stream=None) -> str:
    """WritePlotFile (writePlotFile.cpp:7-190): an AMReX plotfile named
    <plot_base_name><step, 9 digits>; returns that directory name."""
    f, host = _fields(cu, prim, st, eta, kappa)
    lo = (C.c_double * 3)(*prob_lo)
    check(lib().fhdx_write_plotfile(C.byref(p), plot_base_name.encode(), int(step), float(time), lo,
                                    C.byref(f), int(plot_means), int(plot_vars), host, stream),
          "WritePlotFile")
    return f"{plot_base_name}{int(step):09d}"


def WriteCheckPoint(p: Params, step: int, time: float, statsCount: int, seed: int, cu: MultiFab,
                    prim: MultiFab, st: StatsState, eta: MultiFab, kappa: MultiFab,
                    chk_base_name: str = "chk", stream=None) -> str:
    """WriteCheckPoint (checkpoint.cpp:22-133) plus the Philox seed; returns the directory."""
    f, host = _fields(cu, prim, st, eta, kappa)
    check(lib().fhdx_write_checkpoint(C.byref(p), chk_base_name.encode(), int(step), float(time),
                                      int(statsCount), int(seed), C.byref(f), host, stream),
          "WriteCheckPoint")
    return f"{chk_base_name}{int(step):09d}"


def ReadCheckPoint(p: Params, restart: int, cu: MultiFab, prim: MultiFab, st: StatsState,
                   eta: MultiFab, kappa: MultiFab, chk_base_name: str = "chk", reset_stats: int = 0,
                   stream=None):
    """ReadCheckPoint (checkpoint.cpp:135-255) into existing arrays.  Returns (step, time,
    statsCount, seed) with step = the checkpoint's step + 1, as the reference does; seed is None
    for a checkpoint that records none."""
    f, host = _fields(cu, prim, st, eta, kappa)
    step, sc, has = C.c_int(0), C.c_int(0), C.c_int(0)
    t = C.c_double(0.0)
    seed = C.c_uint64(0)
    check(lib().fhdx_read_checkpoint(C.byref(p), chk_base_name.encode(), int(restart),
                                     int(reset_stats), C.byref(step), C.byref(t), C.byref(sc),
                                     C.byref(seed), C.byref(has), C.byref(f), host, stream),
          "ReadCheckPoint")
    return step.value, t.value, sc.value, (seed.value if has.value else None)


# ---- chemistry (chemistry_functions.cpp:128-193) ------------------------------------------------

def compute_compressible_chemistry_source_CLE(ctx: Context, dt: float, prim: MultiFab,
                                              source: MultiFab, ranchem: MultiFab, stream=None):
    """The species source (components 5..) of the valid cells from prim and the weighted
    reaction noise ranchem; the context's reactions come from Context.set_chemistry."""
    check(lib().fhdx_chemistry_source(ctx.h, prim.fab, source.fab, ranchem.fab, float(dt), stream),
          "compute_compressible_chemistry_source_CLE")


def fill_chem_noise(ctx: Context, seed: int, step: int, A: MultiFab, B: MultiFab, instance: int = 0,
                    stream=None):
    """The reaction noise sets of key (seed, step, instance), as the steps draw them."""
    check(lib().fhdx_fill_chem_noise(ctx.h, seed, step, instance, A.fab, B.fab, stream),
          "fill_chem_noise")
