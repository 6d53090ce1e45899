"""Shared state builders for the tests (host numpy arrays; the oracle is the checker)."""
from __future__ import annotations

import numpy as np

from fhdex_amd.params import Params


def perturbed_state(p: Params, cfg, orc, seed=0, amp=1e-2, vel=2e3):
    """A conserved state near the uniform start of `cfg`: density, velocity, temperature and
    composition perturbed cell by cell (all mass fractions stay in (0,1)); ghost cells are
    periodic images, or set by setBC outside walls."""
    rng = np.random.default_rng(seed)
    st = orc.HostState(p)
    orc.init_cons(p, st.cu, cfg.rho0, cfg.rhobar, cfg.T_init)
    g = p.ngc
    NS = p.nspecies
    n0, n1, n2 = p.n_cells
    sl = (slice(g, g + n2), slice(g, g + n1), slice(g, g + n0))
    shape = (n2, n1, n0)
    rho = cfg.rho0 * (1 + amp * rng.standard_normal(shape))
    Y = np.array(cfg.rhobar)[:, None, None, None] * (1 + amp * rng.standard_normal((NS,) + shape))
    Y /= Y.sum(axis=0)
    u = vel * rng.standard_normal((3,) + shape)
    T = cfg.T_init * (1 + amp * rng.standard_normal(shape))
    cv = sum(Y[n] * p.hcv[n] for n in range(NS))
    e = cv * T + sum(Y[n] * p.e0[n] for n in range(NS))
    st.cu[0][sl] = rho
    for d in range(3):
        st.cu[1 + d][sl] = rho * u[d]
    st.cu[4][sl] = rho * (e + 0.5 * (u ** 2).sum(axis=0))
    for n in range(NS):
        st.cu[5 + n][sl] = rho * Y[n]
    orc.fill_boundary(p, st.cu)
    orc.cons_to_prim(p, st.prim, st.cu)
    orc.fill_boundary(p, st.prim)
    if p.has_walls:
        orc.set_bc(p, st.prim, st.cu)
    return st


def valid(p: Params, a):
    g = p.ngc
    n0, n1, n2 = p.n_cells
    return a[:, g:g + n2, g:g + n1, g:g + n0]


def rel_err(a, b):
    """Per-component max|a-b| / max|b| (inf-norm relative, the parity metric)."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    den = np.abs(b).max(axis=1)
    den[den == 0] = 1.0
    return np.abs(a - b).max(axis=1) / den


def perturbed_stag_state(p: Params, cfg, orc, seed=0, amp=1e-2, vel=2e3):
    """A staggered-scheme state (oracle StagHost) near the uniform start of `cfg`: cell density,
    temperature and composition perturbed cell by cell (mass fractions stay in (0,1)), face
    momenta rho_face * u with u ~ vel N(0,1), the duplicate periodic face n equal to face 0; the
    energy puts the cell at its perturbed temperature with the staggered kinetic energy
    (conservedPrimitiveConversionsStag.cpp:99-106).  Ghosts and prim/vel from orc_c2p_stag."""
    rng = np.random.default_rng(seed)
    st = orc.StagHost(p)
    g = p.ngc
    NS = p.nspecies
    n0, n1, n2 = p.n_cells
    shape = (n2, n1, n0)
    sl = (slice(g, g + n2), slice(g, g + n1), slice(g, g + n0))
    rho = cfg.rho0 * (1 + amp * rng.standard_normal(shape))
    Y = np.array(cfg.rhobar)[:, None, None, None] * (1 + amp * rng.standard_normal((NS,) + shape))
    Y /= Y.sum(axis=0)
    T = cfg.T_init * (1 + amp * rng.standard_normal(shape))
    mom = []
    for d in range(3):
        rl = np.roll(rho, 1, axis=2 - d)          # the lower neighbour along d (periodic)
        m = 0.5 * (rho + rl) * vel * rng.standard_normal(shape)
        full = np.concatenate([m, m.take([0], axis=2 - d)], axis=2 - d)   # face n = face 0
        a = st.mom[d]
        a[0, g:g + n2 + (d == 2), g:g + n1 + (d == 1), g:g + n0 + (d == 0)] = full
        mom.append(m)
    kin = sum((m + np.roll(m, -1, axis=2 - d)) ** 2 for d, m in enumerate(mom)) * (0.125 / rho)
    cv = sum(Y[n] * p.hcv[n] for n in range(NS))
    e0 = sum(Y[n] * p.e0[n] for n in range(NS))
    st.cu[0][sl] = rho
    st.cu[4][sl] = rho * (e0 + T * cv) + kin
    for n in range(NS):
        st.cu[5 + n][sl] = rho * Y[n]
    st.c2p()
    return st


def stag_uniform_state(p: Params, cfg, orc):
    """The staggered uniform start (InitConsVarStag prob_type 1, compressible_functions_stag.cpp:
    36-56): rho0, rho0 e(T_init), rho0 Y_k in every cell, zero face momenta."""
    from fhdex_amd.ensemble import uniform_values
    st = orc.StagHost(p)
    for c, v in enumerate(uniform_values(cfg)):
        st.cu[c] = v
    st.c2p()
    return st


def face_valid(p: Params, a, d, g=None):
    """The n distinct faces 0..n-1 of a face array nodal in d with g ghost layers."""
    g = p.ngc if g is None else g
    n0, n1, n2 = p.n_cells
    return a[:, g:g + n2, g:g + n1, g:g + n0]
