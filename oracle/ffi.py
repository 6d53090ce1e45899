"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg as the checker.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fhdex_amd.params import Chemistry, Fab, Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def use_native_build():
    """Build liborc_native.so (gcc -march=native for this host, same operation order) and load
    it instead of liborc.so.  For timing the CPU baseline on the machine that runs it; must be
    called before the oracle is first used in the process (a second call is a no-op)."""
    global LIB
    native = os.path.join(HERE, "liborc_native.so")
    if LIB == native:
        return
    if _lib is not None:
        raise RuntimeError("the oracle is already loaded in this process")
    subprocess.check_call(["make", "-s", "-C", HERE, "-B", "native"])
    LIB = native


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER(Params)
        FA = Fab * 3
        L.orc_cons_to_prim.argtypes = [P, Fab, Fab]
        L.orc_transport.argtypes = [P, Fab, Fab, Fab, Fab, Fab, Fab]
        L.orc_fill_boundary.argtypes = [P, Fab]
        L.orc_set_bc.argtypes = [P, Fab, Fab]
        L.orc_calculate_flux.argtypes = [P, Fab, Fab, Fab, Fab, Fab, Fab, Fab, FA, FA, FA, FA,
                                         FA, Fab, Fab, C.c_double]
        L.orc_rk3_step.argtypes = [P, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab, FA, FA, FA,
                                   FA, FA, Fab, Fab, FA, FA, Fab, Fab, C.c_double]
        L.orc_advance.argtypes = [P, Fab, Fab, C.c_int, C.c_double, C.c_void_p, C.c_uint64,
                                  C.c_uint64, C.c_uint32]
        L.orc_noise_doubles_per_step.argtypes = [P]
        L.orc_noise_doubles_per_step.restype = C.c_long
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]
        L.orc_normal_pair.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.orc_fill_noise.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint32, FA, FA, Fab, Fab]
        L.orc_init_cons.argtypes = [P, Fab, C.c_int, C.c_double, C.POINTER(C.c_double),
                                    C.c_double]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_set_num_threads.restype = None
        L.orc_evaluate_stats.argtypes = [P, Fab, Fab, Fab, Fab, Fab, Fab, Fab, Fab,
                                         C.POINTER(C.c_double), C.c_int, C.c_int]
        CH = C.POINTER(Chemistry)
        L.orc_chem_exp.argtypes = [C.c_double]
        L.orc_chem_exp.restype = C.c_double
        L.orc_chem_pow.argtypes = [C.c_double, C.c_double]
        L.orc_chem_pow.restype = C.c_double
        L.orc_chem_source.argtypes = [P, CH, Fab, Fab, Fab, C.c_double]
        L.orc_fill_chem_noise.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint32, Fab, Fab]
        L.orc_advance_chem.argtypes = [P, CH, Fab, Fab, C.c_int, C.c_double, C.c_uint64,
                                       C.c_uint64, C.c_uint32]
        _bind_stag(L)
        _lib = L
    return _lib


# ---- host FABs ---------------------------------------------------------------------------

def fab_of(arr: np.ndarray, lo) -> Fab:
    """Describe a C-contiguous float64 array of shape (ncomp, nz, ny, nx) whose first element
    is index `lo`."""
    assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"] and arr.ndim == 4
    f = Fab()
    f.p = arr.ctypes.data
    nc, nz, ny, nx = arr.shape
    for d, n in enumerate((nx, ny, nz)):
        f.lo[d] = lo[d]
        f.hi[d] = lo[d] + n - 1
    f.ncomp = nc
    return f


class HostState:
    """All arrays of the reference's RK3step argument list, on the host (numpy)."""

    def __init__(self, p: Params):
        self.p = p
        n0, n1, n2 = p.n_cells
        g = p.ngc
        NS = p.nspecies
        nv = p.nvars

        def cell(nc):
            return np.zeros((nc, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g))

        def face(d, nc):
            return np.zeros((nc, n2 + (d == 2), n1 + (d == 1), n0 + (d == 0)))

        def node(nc=1):
            return np.zeros((nc, n2 + 1, n1 + 1, n0 + 1))

        self.cu, self.cup, self.cup2 = cell(nv), cell(nv), cell(nv)
        self.prim = cell(p.nprimvars)
        self.eta, self.zeta, self.kappa = cell(1), cell(1), cell(1)
        self.chi, self.D = cell(NS), cell(NS * NS)
        self.flux = [face(d, nv + 4) for d in range(3)]
        self.stoch = [face(d, nv) for d in range(3)]
        self.corn = [[node() for _ in range(3)] for _ in range(3)]
        self.visccorn, self.rancorn = node(), node()
        self.sA = [face(d, nv) for d in range(3)]
        self.sB = [face(d, nv) for d in range(3)]
        self.rA, self.rB = node(), node()
        self.glo = (-g, -g, -g)

    def cf(self, a):
        return fab_of(a, self.glo)

    @staticmethod
    def zf(a):
        return fab_of(a, (0, 0, 0))

    def fa(self, arrs):
        return (Fab * 3)(*[self.zf(a) for a in arrs])


def init_cons(p: Params, cu: np.ndarray, rho0, rhobar, T_init, prob_type=1):
    g = p.ngc
    rb = (C.c_double * 8)(*(list(rhobar) + [0.0] * (8 - len(rhobar))))
    return lib().orc_init_cons(C.byref(p), fab_of(cu, (-g, -g, -g)), prob_type, rho0, rb,
                               T_init)


def cons_to_prim(p, prim, cons):
    g = p.ngc
    return lib().orc_cons_to_prim(C.byref(p), fab_of(prim, (-g,) * 3), fab_of(cons, (-g,) * 3))


def fill_boundary(p, a):
    g = p.ngc
    return lib().orc_fill_boundary(C.byref(p), fab_of(a, (-g,) * 3))


def set_bc(p, prim, cons):
    """setBC on host cell arrays with ngc ghosts."""
    g = p.ngc
    return lib().orc_set_bc(C.byref(p), fab_of(prim, (-g,) * 3), fab_of(cons, (-g,) * 3))


def transport(p, st: HostState):
    return lib().orc_transport(C.byref(p), st.cf(st.prim), st.cf(st.eta), st.cf(st.zeta),
                               st.cf(st.kappa), st.cf(st.chi), st.cf(st.D))


def calculate_flux(p, st: HostState, cons, dt):
    return lib().orc_calculate_flux(
        C.byref(p), st.cf(cons), st.cf(st.prim), st.cf(st.eta), st.cf(st.zeta),
        st.cf(st.kappa), st.cf(st.chi), st.cf(st.D), st.fa(st.flux), st.fa(st.stoch),
        st.fa(st.corn[0]), st.fa(st.corn[1]), st.fa(st.corn[2]), st.zf(st.visccorn),
        st.zf(st.rancorn), dt)


def rk3_step(p, st: HostState, dt):
    """One RK3step with the noise sets currently held in st.sA/sB/rA/rB."""
    return lib().orc_rk3_step(
        C.byref(p), st.cf(st.cu), st.cf(st.cup), st.cf(st.cup2), st.cf(st.prim),
        st.cf(st.eta), st.cf(st.zeta), st.cf(st.kappa), st.cf(st.chi), st.cf(st.D),
        st.fa(st.flux), st.fa(st.stoch), st.fa(st.corn[0]), st.fa(st.corn[1]),
        st.fa(st.corn[2]), st.zf(st.visccorn), st.zf(st.rancorn), st.fa(st.sA), st.fa(st.sB),
        st.zf(st.rA), st.zf(st.rB), dt)


def fill_noise(p, st: HostState, seed, step, instance=0):
    return lib().orc_fill_noise(C.byref(p), seed, step, instance, st.fa(st.sA), st.fa(st.sB),
                                st.zf(st.rA), st.zf(st.rB))


def advance(p, cu, prim, nsteps, dt, noise=None, seed=1, step0=0, instance=0):
    g = p.ngc
    ptr = None if noise is None else noise.ctypes.data
    return lib().orc_advance(C.byref(p), fab_of(cu, (-g,) * 3), fab_of(prim, (-g,) * 3),
                             nsteps, dt, ptr, seed, step0, instance)


def chem_exp(x):
    return lib().orc_chem_exp(float(x))


def chem_pow(x, y):
    return lib().orc_chem_pow(float(x), float(y))


def chem_source(p, chem, prim, source, ranchem, dt):
    """compute_compressible_chemistry_source_CLE on host cell arrays (ngc ghosts)."""
    g = p.ngc
    return lib().orc_chem_source(C.byref(p), C.byref(chem), fab_of(prim, (-g,) * 3),
                                 fab_of(source, (-g,) * 3), fab_of(ranchem, (-g,) * 3), dt)


def fill_chem_noise(p, A, B, seed, step, instance=0):
    g = p.ngc
    return lib().orc_fill_chem_noise(C.byref(p), seed, step, instance, fab_of(A, (-g,) * 3),
                                     fab_of(B, (-g,) * 3))


def advance_chem(p, chem, cu, prim, nsteps, dt, seed=1, step0=0, instance=0):
    """orc_advance with the chemistry source and Philox noise (flux and reaction noise)."""
    g = p.ngc
    return lib().orc_advance_chem(C.byref(p), C.byref(chem), fab_of(cu, (-g,) * 3),
                                  fab_of(prim, (-g,) * 3), nsteps, dt, seed, step0, instance)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def normal_pair(words):
    w = (C.c_uint32 * 4)(*words)
    a, b = C.c_double(), C.c_double()
    lib().orc_normal_pair(w, C.byref(a), C.byref(b))
    return a.value, b.value


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


class StatsState:
    """The statistics MultiFabs of main_driver.cpp:205-238 (host, zero-initialised)."""

    def __init__(self, p: Params):
        g = p.ngc
        n0, n1, n2 = p.n_cells
        sh = lambda nc: np.zeros((nc, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g))  # noqa: E731
        self.p = p
        self.consMean, self.consVar = sh(p.nvars), sh(p.nvars)
        self.primMean, self.primVar = sh(p.nprimvars), sh(p.nprimvars + 5)
        self.spatialCross, self.miscStats = sh(8), sh(10)
        self.miscVals = np.zeros(20)
        self.lo = (-g, -g, -g)


def evaluate_stats(p, cons, prim, ss: StatsState, steps, cross_cell):
    """evaluateStats (stats.cpp:5-363) on host arrays; updates ss in place."""
    lo = ss.lo
    rc = lib().orc_evaluate_stats(C.byref(p), fab_of(cons, lo), fab_of(ss.consMean, lo), fab_of(ss.consVar, lo),
                                  fab_of(prim, lo), fab_of(ss.primMean, lo), fab_of(ss.primVar, lo),
                                  fab_of(ss.spatialCross, lo), fab_of(ss.miscStats, lo),
                                  ss.miscVals.ctypes.data_as(C.POINTER(C.c_double)), int(steps),
                                  int(cross_cell))
    assert rc == 0, rc


# ---- staggered scheme (oracle_stag.c) ------------------------------------------------------

EDGE_AXES = ((0, 1), (0, 2), (1, 2))       # edge types xy, xz, yz
EDGE_FLUX_TYPES = (0, 1, 0, 2, 1, 2)        # edgeflux_x[0..1], _y[0..1], _z[0..1]


class StagHost:
    """Host arrays of RK3stepStag's argument list (timeStepStag.cpp:13-26) and its noise sets:
    cu, prim, transport coefficients (cells, ngc ghosts), cumom/vel (faces nodal in d, ngc
    ghosts), faceflux (faces, no ghosts), the six edge fluxes (edges, no ghosts), cenflux
    (cells, one ghost); noise: face[d] (nvars comps), edge[e] (xy, xz, yz), cen[d] (cells,
    one ghost).  Arrays are (ncomp, z, y, x), C order."""

    def __init__(self, p: Params):
        self.p = p
        g, NS, nv = p.ngc, p.nspecies, p.nvars
        self.g = g
        self.cu, self.prim = self.cell(nv), self.cell(p.nprimvars)
        self.eta, self.zeta, self.kappa = self.cell(1), self.cell(1), self.cell(1)
        self.chi, self.D = self.cell(NS), self.cell(NS * NS)
        self.mom = [self.face(d, 1) for d in range(3)]
        self.vel = [self.face(d, 1) for d in range(3)]
        self.ff = [self.face(d, nv + 4, 0) for d in range(3)]
        self.ef = [self.edge(e) for e in EDGE_FLUX_TYPES]
        self.cf = [self.cell(1, 1) for _ in range(3)]
        self.fA = [self.face(d, nv, 0) for d in range(3)]
        self.fB = [self.face(d, nv, 0) for d in range(3)]
        self.eA = [self.edge(e) for e in range(3)]
        self.eB = [self.edge(e) for e in range(3)]
        self.cA = [self.cell(1, 1) for _ in range(3)]
        self.cB = [self.cell(1, 1) for _ in range(3)]

    def cell(self, nc, g=None):
        g = self.g if g is None else g
        n0, n1, n2 = self.p.n_cells
        return np.zeros((nc, n2 + 2 * g, n1 + 2 * g, n0 + 2 * g))

    def face(self, d, nc, g=None):
        g = self.g if g is None else g
        n0, n1, n2 = self.p.n_cells
        return np.zeros((nc, n2 + 2 * g + (d == 2), n1 + 2 * g + (d == 1), n0 + 2 * g + (d == 0)))

    def edge(self, e, nc=1):
        a, b = EDGE_AXES[e]
        n0, n1, n2 = self.p.n_cells
        return np.zeros((nc, n2 + (2 in (a, b)), n1 + (1 in (a, b)), n0 + (0 in (a, b))))

    @staticmethod
    def ghosts_of(arr, p, nodal=-1):
        """Ghost width of a cell (nodal -1) or face array from its shape."""
        n0 = p.n_cells[0]
        return (arr.shape[3] - n0 - (nodal == 0)) // 2

    def F(self, arr, nodal=-1):
        g = self.ghosts_of(arr, self.p, nodal)
        return fab_of(arr, (-g, -g, -g))

    def faces(self, arrs):
        return (Fab * 3)(*[self.F(a, d) for d, a in enumerate(arrs)])

    def cells(self, arrs):
        return (Fab * 3)(*[self.F(a) for a in arrs])

    def edges(self, arrs):
        return (Fab * len(arrs))(*[fab_of(a, (0, 0, 0)) for a in arrs])

    def fill_noise(self, seed, step, instance=0):
        rc = lib().orc_fill_noise_stag(C.byref(self.p), seed, step, instance, self.faces(self.fA),
                                       self.faces(self.fB), self.edges(self.eA), self.edges(self.eB),
                                       self.cells(self.cA), self.cells(self.cB))
        assert rc == 0, rc

    def c2p(self):
        rc = lib().orc_c2p_stag(C.byref(self.p), self.F(self.prim), self.faces(self.vel), self.F(self.cu),
                                self.faces(self.mom))
        assert rc == 0, rc

    def transport(self):
        return lib().orc_transport(C.byref(self.p), self.F(self.prim), self.F(self.eta), self.F(self.zeta),
                                   self.F(self.kappa), self.F(self.chi), self.F(self.D))

    def flux(self, dt, w):
        """calculateFluxStag of the current state with noise A + w B of the held sets."""
        nv = self.p.nvars
        sf = [self.fA[d] + 0.0 for d in range(3)]
        for d in range(3):
            sf[d][4:nv] = 1.0 * self.fA[d][4:nv] + w * self.fB[d][4:nv]
        se = [1.0 * self.eA[e] + w * self.eB[e] for e in range(3)]
        sc = [1.0 * self.cA[d] + w * self.cB[d] for d in range(3)]
        rc = lib().orc_flux_stag(C.byref(self.p), self.F(self.cu), self.faces(self.mom), self.F(self.prim),
                                 self.faces(self.vel), self.F(self.eta), self.F(self.zeta), self.F(self.kappa),
                                 self.F(self.chi), self.F(self.D), self.faces(self.ff), self.edges(self.ef),
                                 self.cells(self.cf), self.faces(sf), self.edges(se), self.cells(sc), dt)
        assert rc == 0, rc

    def step(self, dt):
        """One RK3stepStag with the held noise sets."""
        return lib().orc_rk3_step_stag(
            C.byref(self.p), self.F(self.cu), self.faces(self.mom), self.F(self.prim), self.faces(self.vel),
            self.F(self.eta), self.F(self.zeta), self.F(self.kappa), self.F(self.chi), self.F(self.D),
            self.faces(self.ff), self.edges(self.ef), self.cells(self.cf), self.faces(self.fA),
            self.faces(self.fB), self.edges(self.eA), self.edges(self.eB), self.cells(self.cA),
            self.cells(self.cB), dt)

    def advance(self, nsteps, dt, seed=1, step0=0, instance=0, chem=None):
        """nsteps of RK3stepStag with Philox noise (seed, step0 + s, instance); with `chem` (a
        Chemistry) the reaction source of every stage too."""
        if chem is not None:
            return lib().orc_advance_stag_chem(C.byref(self.p), C.byref(chem), self.F(self.cu), self.faces(self.mom),
                                               self.F(self.prim), self.faces(self.vel), nsteps, dt, seed, step0,
                                               instance)
        return lib().orc_advance_stag(C.byref(self.p), self.F(self.cu), self.faces(self.mom), self.F(self.prim),
                                      self.faces(self.vel), nsteps, dt, seed, step0, instance)


def _bind_stag(L):
    P = C.POINTER(Params)
    FA = Fab * 3
    E6 = Fab * 6
    L.orc_c2p_stag.argtypes = [P, Fab, FA, Fab, FA]
    L.orc_flux_stag.argtypes = [P, Fab, FA, Fab, FA, Fab, Fab, Fab, Fab, Fab, FA, E6, FA, FA, FA, FA,
                                C.c_double]
    L.orc_fill_noise_stag.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint32, FA, FA, FA, FA, FA, FA]
    L.orc_rk3_step_stag.argtypes = [P, Fab, FA, Fab, FA, Fab, Fab, Fab, Fab, Fab, FA, E6, FA, FA, FA, FA,
                                    FA, FA, FA, C.c_double]
    L.orc_advance_stag.argtypes = [P, Fab, FA, Fab, FA, C.c_int, C.c_double, C.c_uint64, C.c_uint64,
                                   C.c_uint32]
    L.orc_advance_stag_chem.argtypes = [P, C.POINTER(Chemistry), Fab, FA, Fab, FA, C.c_int, C.c_double,
                                        C.c_uint64, C.c_uint64, C.c_uint32]
