"""Host-side mirror of src_compressible_stag's interface over the C ABI.

Names and argument meaning follow the reference (compressible_functions_stag.H:113-143,
conservedPrimitiveConversionsStag.cpp:5-6): ``conservedToPrimitiveStag``, ``calculateFluxStag``
and ``RK3stepStag``.  ``StagState`` allocates RK3stepStag's argument list as the staggered
main_driver defines it (main_driver.cpp:616-631, 1023-1050): momenta and velocities on faces
with ngc ghosts, the fluxes on faces, edges and cells.  Arrays are float64 device tensors laid
out (ncomp, z, y, x); torch is only the allocator.
"""
from __future__ import annotations

import ctypes as C

from .compressible import MultiFab
from .lib import Context, check, lib
from .params import Fab, Params

# edge types xy, xz, yz: their two nodal axes
EDGE_AXES = ((0, 1), (0, 2), (1, 2))
# edgeflux_x[0] xy, [1] xz; edgeflux_y[0] xy, [1] yz; edgeflux_z[0] xz, [1] yz
EDGE_FLUX_TYPES = (0, 1, 0, 2, 1, 2)


def face_mf(p: Params, d: int, ncomp: int, g: int, device="cuda") -> MultiFab:
    """Face MultiFab nodal in d with g ghost layers (convert(ba, nodal_flag_dir[d]), ngc)."""
    import torch
    n0, n1, n2 = p.n_cells
    t = torch.zeros((ncomp, n2 + 2 * g + (d == 2), n1 + 2 * g + (d == 1), n0 + 2 * g + (d == 0)),
                    dtype=torch.float64, device=device)
    return MultiFab(t, (-g, -g, -g))


def edge_mf(p: Params, e: int, device="cuda") -> MultiFab:
    """Edge MultiFab of type e (nodal along both of its axes), no ghosts."""
    import torch
    a, b = EDGE_AXES[e]
    n0, n1, n2 = p.n_cells
    t = torch.zeros((1, n2 + (2 in (a, b)), n1 + (1 in (a, b)), n0 + (0 in (a, b))), dtype=torch.float64,
                    device=device)
    return MultiFab(t, (0, 0, 0))


def cell_mf(p: Params, ncomp: int, g: int, device="cuda") -> MultiFab:
    import torch
    n0, n1, n2 = p.n_cells
    t = torch.zeros((ncomp, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g), dtype=torch.float64, device=device)
    return MultiFab(t, (-g, -g, -g))


class StagState:
    """Every MultiFab of RK3stepStag's argument list on the device (one periodic box): cu, prim,
    cumom[3], vel[3], eta, zeta, kappa, chi, D, faceflux[3], edgeflux[6] (x[0], x[1], y[0],
    y[1], z[0], z[1]) and cenflux[3]."""

    def __init__(self, p: Params, device="cuda"):
        self.p = p
        g, nv, NS = p.ngc, p.nvars, p.nspecies
        self.cu = cell_mf(p, nv, g, device)
        self.prim = cell_mf(p, p.nprimvars, g, device)
        self.cumom = [face_mf(p, d, 1, g, device) for d in range(3)]
        self.vel = [face_mf(p, d, 1, g, device) for d in range(3)]
        self.eta, self.zeta, self.kappa = (cell_mf(p, 1, g, device) for _ in range(3))
        self.chi = cell_mf(p, NS, g, device)
        self.D = cell_mf(p, NS * NS, g, device)
        self.faceflux = [face_mf(p, d, nv + 4, 0, device) for d in range(3)]
        self.edgeflux = [edge_mf(p, e, device) for e in EDGE_FLUX_TYPES]
        self.cenflux = [cell_mf(p, 1, 1, device) for _ in range(3)]
        # chemistry (main_driver.cpp defines source with nvars and ranchem with nreaction comps)
        self.source = cell_mf(p, nv, g, device) if p.nreaction > 0 else None
        self.ranchem = cell_mf(p, p.nreaction, g, device) if p.nreaction > 0 else None


def _f3(mfs):
    return (Fab * 3)(*[m.fab for m in mfs])


def _f6(mfs):
    return (Fab * 6)(*[m.fab for m in mfs])


def conservedToPrimitiveStag(ctx: Context, prim: MultiFab, vel, cons: MultiFab, cumom, stream=None):
    """conservedToPrimitiveStag with the periodic ghost fills around it: face velocities, prim
    and cons' cell-centred momentum components, all ghosts current afterwards."""
    check(lib().fhdx_cons_to_prim_stag(ctx.h, prim.fab, _f3(vel), cons.fab, _f3(cumom), stream),
          "conservedToPrimitiveStag")


def calculateFluxStag(ctx: Context, st: StagState, weight: float, dt: float, seed: int = 1, step: int = 0,
                      stream=None):
    """calculateFluxStag of st's state with the noise A + weight B of the (seed, step) sets; the
    transport coefficients in st are inputs."""
    check(lib().fhdx_calculate_flux_stag(ctx.h, st.cu.fab, _f3(st.cumom), st.prim.fab, _f3(st.vel),
                                         st.eta.fab, st.zeta.fab, st.kappa.fab, st.chi.fab, st.D.fab,
                                         _f3(st.faceflux), _f6(st.edgeflux), _f3(st.cenflux),
                                         float(weight), float(dt), int(seed), int(step), stream),
          "calculateFluxStag")


def RK3stepStag(ctx: Context, st: StagState, dt: float, seed: int = 1, step: int = 0, stream=None):
    """RK3stepStag (timeStepStag.cpp:13-1245): cu and cumom advanced in place, prim, vel, the
    transport coefficients and the fluxes left as the reference leaves them; with chemistry
    (Context.set_chemistry) also source and ranchem."""
    z = Fab()
    src = st.source.fab if st.source is not None else z
    ran = st.ranchem.fab if st.ranchem is not None else z
    check(lib().fhdx_rk3_step_stag(ctx.h, st.cu.fab, _f3(st.cumom), st.prim.fab, _f3(st.vel), st.eta.fab,
                                   st.zeta.fab, st.kappa.fab, st.chi.fab, st.D.fab, _f3(st.faceflux),
                                   _f6(st.edgeflux), _f3(st.cenflux), src, ran, float(dt), int(seed), int(step),
                                   stream), "RK3stepStag")


def calculateTransportCoeffs(ctx: Context, st: StagState, stream=None):
    """calculateTransportCoeffs (transCoeffs.cpp:7-90) of st.prim on the grown box."""
    check(lib().fhdx_transport(ctx.h, st.prim.fab, st.eta.fab, st.zeta.fab, st.kappa.fab, st.chi.fab,
                               st.D.fab, stream), "calculateTransportCoeffs")
