"""CPU tests that pin the staggered-scheme oracle (oracle/oracle_stag.c).

The reference ships no golden outputs for RK3stepStag either, so the restatement is pinned by
invariants (conservation of mass, face momentum, energy and species with the noise on; a
uniform gas at rest stays put; duplicate periodic faces stay equal), closed forms (the
conserved-to-primitive conversion of a constructed state, shear-wave and sound-wave rates with
the staggered grid's wavenumbers), cyclic symmetry of the fluxes under relabelling of the axes,
and the reference's z-face quirks showing up exactly there.
"""
import math

import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import face_valid, perturbed_stag_state, rel_err, stag_uniform_state, valid


def _four(shape, **kw):
    cfg = P.four_species(4)
    p = P.make_params(4, shape, 8e-6, list(cfg.params.molmass[:4]), list(cfg.params.diameter[:4]), **kw)
    return cfg, p


def _argon(shape, **kw):
    cfg = P.argon_gas(4)
    p = P.make_params(2, shape, 8e-6, [39.95, 39.95], [3.66e-8, 3.66e-8], **kw)
    return cfg, p


def _eight(shape, **kw):
    cfg = P.species_prefix(8, 4, 1)
    p = P.make_params(8, shape, 8e-6, P.EIGHT_SPECIES_M, P.EIGHT_SPECIES_D, **kw)
    return cfg, p


GASES = {"argon": _argon, "four": _four, "eight": _eight}


@pytest.mark.parametrize("gas", ["argon", "four", "eight"])
def test_stag_step_conserves_mass_momentum_energy_species(orc, gas):
    """Flux form on a periodic box: sum rho, rho E, rho Y_k over cells and sum of each momentum
    component over the n distinct faces of its direction are unchanged by three noisy steps to
    rounding; the duplicate face n keeps equal to face 0 bit for bit (what OverrideSync and
    FillBoundary leave in the reference)."""
    cfg, p = GASES[gas]((6, 5, 4))
    st = perturbed_stag_state(p, cfg, orc, seed=3)
    tot0 = valid(p, st.cu).sum(axis=(1, 2, 3))
    abs0 = np.abs(valid(p, st.cu)).sum(axis=(1, 2, 3))
    m0 = [face_valid(p, st.mom[d], d).sum() for d in range(3)]
    ma = [np.abs(face_valid(p, st.mom[d], d)).sum() for d in range(3)]
    for s in range(3):
        st.fill_noise(7, s)
        assert st.step(cfg.dt) == 0
    tot1 = valid(p, st.cu).sum(axis=(1, 2, 3))
    keep = [0, 4] + list(range(5, p.nvars))
    assert np.all(np.abs(tot1 - tot0)[keep] <= 2e-14 * abs0[keep])
    g = p.ngc
    for d in range(3):
        assert abs(face_valid(p, st.mom[d], d).sum() - m0[d]) <= 2e-14 * ma[d]
        a = np.moveaxis(st.mom[d][0], 2 - d, 0)
        n = p.n_cells[d]
        assert np.array_equal(a[g + n], a[g])
    # the state moved (the noise is on)
    assert np.abs(valid(p, st.cu)[0] - cfg.rho0).max() > 0


def test_stag_uniform_gas_at_rest_stays_put_without_noise(orc):
    cfg, p = _four((5, 4, 6), stoch_stress_form=0)
    st = stag_uniform_state(p, cfg, orc)
    cu0 = valid(p, st.cu).copy()
    for s in range(3):
        st.fill_noise(1, s)
        assert st.step(cfg.dt) == 0
    assert np.array_equal(valid(p, st.cu), cu0)
    for d in range(3):
        assert not np.any(st.mom[d])


def test_stag_cons_to_prim_recovers_construction(orc):
    """prim from a state built from known rho, T, Y and face velocities: face velocity =
    2 m / (rho_i + rho_{i-1}), cell velocity the mean of its two faces, T from
    rho E - (1/8 rho) sum_d (m_lo + m_hi)^2, p = rho R T / M_mix, X_k = Y_k M_mix / M_k."""
    cfg, p = _four((6, 5, 4))
    rng = np.random.default_rng(5)
    st = orc.StagHost(p)
    g, NS = p.ngc, p.nspecies
    n0, n1, n2 = p.n_cells
    sh = (n2, n1, n0)
    rho = cfg.rho0 * (1 + 0.05 * rng.standard_normal(sh))
    T = 300 * (1 + 0.05 * rng.standard_normal(sh))
    Y = rng.uniform(0.1, 1.0, (NS,) + sh)
    Y /= Y.sum(axis=0)
    u = [3e3 * rng.standard_normal(sh) for _ in range(3)]
    sl = (slice(g, g + n2), slice(g, g + n1), slice(g, g + n0))
    m = []
    for d in range(3):
        rl = np.roll(rho, 1, axis=2 - d)
        md = u[d] * 0.5 * (rho + rl)
        m.append(md)
        full = np.concatenate([md, md.take([0], axis=2 - d)], axis=2 - d)
        st.mom[d][0, g:g + n2 + (d == 2), g:g + n1 + (d == 1), g:g + n0 + (d == 0)] = full
    ucell = [0.5 * (m[d] + np.roll(m[d], -1, axis=2 - d)) / rho for d in range(3)]
    cv = sum(Y[n] * p.hcv[n] for n in range(NS))
    st.cu[0][sl] = rho
    st.cu[4][sl] = rho * cv * T + 0.5 * rho * sum(x * x for x in ucell)
    for n in range(NS):
        st.cu[5 + n][sl] = rho * Y[n]
    st.c2p()
    pr = valid(p, st.prim)
    assert np.abs(pr[4] / T - 1).max() < 1e-12
    for d in range(3):
        vd = face_valid(p, st.vel[d], d)[0]
        assert np.abs(vd - u[d]).max() <= 1e-12 * np.abs(u[d]).max()
        uc = 0.5 * (u[d] * 0.5 * (rho + np.roll(rho, 1, axis=2 - d)) / rho
                    + np.roll(u[d] * 0.5 * (rho + np.roll(rho, 1, axis=2 - d)), -1, axis=2 - d) / rho)
        assert np.abs(valid(p, st.cu)[1 + d] - rho * uc).max() <= 1e-12 * np.abs(rho * uc).max()
    Mmix = 1.0 / sum(Y[n] / p.molmass[n] for n in range(NS))
    assert np.abs(pr[5] / (rho * P.R_UNIV / Mmix * T) - 1).max() < 1e-12
    for n in range(NS):
        assert np.abs(pr[6 + NS + n] - Y[n] * Mmix / p.molmass[n]).max() < 1e-14
    # ghosts are periodic images: the lo ghost layer in x is the hi valid layer
    assert np.array_equal(st.prim[:, :, :, g - 1], st.prim[:, :, :, g + n0 - 1])
    assert np.array_equal(st.vel[0][:, :, :, g - 1], st.vel[0][:, :, :, g + n0 - 1])


def _run(orc, p, cfg, st, nsteps, probe):
    out = [probe(st)]
    for s in range(nsteps):
        assert st.step(cfg.dt) == 0
        out.append(probe(st))
    return np.array(out)


def test_stag_shear_wave_decays_at_kinematic_viscosity(orc):
    """x momentum m = rho A sin(k y) on the x faces decays as exp(-nu k_eff^2 t), nu = eta/rho:
    the staggered stress tau_xy = eta (du/dy + dv/dx) on xy edges and its difference back onto
    the faces form the 3-point Laplacian, k_eff^2 = (2 - 2 cos(k dy)) / dy^2."""
    n = 32
    cfg, p = _argon((4, n, 4), stoch_stress_form=0)
    st = stag_uniform_state(p, cfg, orc)
    g = p.ngc
    dy = p.dx[1]
    k = 2 * math.pi / (n * dy)
    y = (np.arange(n) + 0.5) * dy
    s = np.sin(k * y)
    A = 100.0
    st.mom[0][0, :, g:g + n, :] = cfg.rho0 * A * s[None, :, None]
    st.cu[4, g:-g, g:-g, g:-g] += 0.5 * cfg.rho0 * (A * s[None, :, None]) ** 2
    st.c2p()

    def probe(st):
        u = st.vel[0][0, g, g:g + n, g]
        return 2 * (u * s).mean()

    amp = _run(orc, p, cfg, st, 60, probe)
    t = cfg.dt * np.arange(len(amp))
    rate = -np.polyfit(t, np.log(amp), 1)[0]
    eta = 2.0812e-4
    keff2 = (2 - 2 * math.cos(k * dy)) / dy ** 2
    assert rate == pytest.approx(eta / cfg.rho0 * keff2, rel=5e-3)


@pytest.mark.parametrize("adv,visc,zr", [(2, 2, -1.0), (1, 2, -1.0), (0, 2, -1.0), (2, 3, 1.5)])
def test_stag_sound_wave_speed_and_damping(orc, adv, visc, zr):
    """A standing adiabatic sound wave on the staggered grid: the pressure gradient (cell to
    face) and the mass-flux divergence (face to cell) are both 2-point differences, so the
    linearised Navier-Stokes-Fourier system carries k1 = 2 sin(k h / 2) / h and k1^2 for the
    viscous and conductive terms.  Frequency and damping are pinned for the three advection
    forms (which agree at linear order) and for bulk viscosity (visc_type 3)."""
    from scipy.optimize import curve_fit
    n = 32
    cfg, p = _argon((4, n, 4), stoch_stress_form=0, advection_type=adv, visc_type=visc, zeta_ratio=zr)
    g = p.ngc
    dy = p.dx[1]
    k = 2 * math.pi / (n * dy)
    y = (np.arange(n) + 0.5) * dy
    s = np.sin(k * y)[None, :, None]
    eps = 1e-4
    gam = 5.0 / 3.0
    cv = p.hcv[0]
    st = orc.StagHost(p)
    rho = cfg.rho0 * (1 + eps * s)
    T = cfg.T_init * (1 + (gam - 1) * eps * s)
    st.cu[0, g:-g, g:-g, g:-g] = rho
    st.cu[4, g:-g, g:-g, g:-g] = rho * cv * T
    st.cu[5, g:-g, g:-g, g:-g] = 0.5 * rho
    st.cu[6, g:-g, g:-g, g:-g] = 0.5 * rho
    st.c2p()

    def probe(st):
        r = st.cu[0, g, g:-g, g] / cfg.rho0 - 1
        return 2 * (r * s[0, :, 0]).mean()

    nsteps = 240
    a = _run(orc, p, cfg, st, nsteps, probe)
    t = cfg.dt * np.arange(nsteps + 1)
    h = dy
    k1 = 2 * math.sin(k * h / 2) / h
    k2sq = k1 * k1
    eta, kappa = 2.0812e-4, 1624.3
    zeta = zr * eta if visc == 3 else 0.0
    rho0, T0 = cfg.rho0, cfg.T_init
    Rm = P.R_UNIV / 39.95
    p0 = rho0 * Rm * T0
    M = np.array([[0, -1j * k1 * rho0, 0],
                  [-1j * k1 * Rm * T0 / rho0, -(4 / 3 * eta + zeta) / rho0 * k2sq, -1j * k1 * Rm],
                  [0, -1j * k1 * p0 / (rho0 * cv), -kappa * k2sq / (rho0 * cv)]])
    ev = np.linalg.eigvals(M)
    sound = ev[np.argmax(np.abs(ev.imag))]
    G0, w0 = -sound.real, abs(sound.imag)
    heat = ev[np.argmin(np.abs(ev.imag))].real

    def model(t, A, G, w, ph, B):
        return A * np.exp(-G * t) * np.cos(w * t + ph) + B * np.exp(heat * t)

    popt, _ = curve_fit(model, t, a, p0=[eps, G0, w0, 0.0, 0.0])
    assert popt[2] == pytest.approx(w0, rel=1e-3)
    assert popt[1] == pytest.approx(G0, rel=1e-2)


# ---- cyclic relabelling of the axes ---------------------------------------------------------

def _rot(a):
    """new x = old z, new y = old x, new z = old y, on (comp, z, y, x) arrays."""
    return np.ascontiguousarray(np.transpose(a, (0, 2, 3, 1)))


# old direction -> new direction under the relabelling: x -> y, y -> z, z -> x
NEW_DIR = {0: 1, 1: 2, 2: 0}
# old edge flux index -> new (x_v, x_w, y_u, y_w, z_u, z_v)
NEW_EDGE_FLUX = {0: 3, 1: 2, 2: 5, 3: 4, 4: 0, 5: 1}
# old edge type (xy, xz, yz) -> new
NEW_EDGE = {0: 2, 1: 0, 2: 1}


def _rotated_host(orc, p, st):
    st2 = orc.StagHost(p)
    st2.cu[:] = _rot(st.cu)
    for d in range(3):
        st2.mom[NEW_DIR[d]][:] = _rot(st.mom[d])
        st2.fA[NEW_DIR[d]][:] = _rot(st.fA[d])
        st2.fB[NEW_DIR[d]][:] = _rot(st.fB[d])
        st2.cA[NEW_DIR[d]][:] = _rot(st.cA[d])
        st2.cB[NEW_DIR[d]][:] = _rot(st.cB[d])
    for e in range(3):
        st2.eA[NEW_EDGE[e]][:] = _rot(st.eA[e])
        st2.eB[NEW_EDGE[e]][:] = _rot(st.eB[e])
    st2.c2p()
    return st2


@pytest.mark.parametrize("gas,noise", [("argon", False), ("argon", True), ("four", False), ("four", True)])
def test_stag_fluxes_are_cyclic_symmetric(orc, gas, noise):
    """Relabelling the axes cyclically (and the noise with them) maps every face, edge and
    cell-centred flux onto the corresponding one to rounding, except the z-face Dufour term of
    the diffusive flux, where the reference adds chi(i,j,k) twice (fluxStag.cpp:1398), and the
    energy flux that includes it.  Catches swapped indices, wrong transverse pairs, a wrong
    edge or noise array and sign slips."""
    cfg, p = (_argon if gas == "argon" else _four)((6, 6, 6), stoch_stress_form=1 if noise else 0)
    st = perturbed_stag_state(p, cfg, orc, seed=21, amp=0.05)
    st.fill_noise(9, 0)
    w = 0.7
    assert st.transport() == 0
    st.flux(cfg.dt, w)
    st2 = _rotated_host(orc, p, st)
    assert st2.transport() == 0
    st2.flux(cfg.dt, w)
    nv = p.nvars
    quirky = gas == "four"
    for d in range(3):
        nd = NEW_DIR[d]
        expect = _rot(st.ff[d])
        got = st2.ff[nd]
        diff = np.abs(got - expect).reshape(nv + 4, -1).max(axis=1)
        scale = np.abs(expect).reshape(nv + 4, -1).max(axis=1)
        scale[nv + 3] = max(scale[nv + 3], 1e-2 * scale[nv])
        for c in range(nv + 4):
            if quirky and (d == 2 or nd == 2) and c in (4, nv + 3):
                continue
            assert diff[c] <= 1e-11 * max(scale[c], 1e-300), (d, c, diff[c], scale[c])
        ce = _rot(st.cf[d])
        assert np.abs(st2.cf[nd] - ce).max() <= 1e-11 * np.abs(ce).max()
    for e in range(6):
        ee = _rot(st.ef[e])
        assert np.abs(st2.ef[NEW_EDGE_FLUX[e]] - ee).max() <= 1e-11 * np.abs(ee).max()
    if quirky:
        # old y faces become new z faces: their Dufour term differs by the chi(k)+chi(k) read
        expect = _rot(st.ff[1])
        err = rel_err(st2.ff[2], expect)
        assert err[nv + 3] > 1e-9


def test_stag_advection_type_0_z_face_density_quirk(orc):
    """advection_type 0 divides the z faces' lower kinetic energy by the density of cell
    (i, j-1, k) (fluxStag.cpp:1754): scaling the density of one cell layer in y moves the z-face
    energy flux of the next layer up, which the x faces, reading their own cells, do not."""
    cfg, p = _argon((6, 6, 6), stoch_stress_form=0, advection_type=0)
    st = perturbed_stag_state(p, cfg, orc, seed=8, amp=0.05, vel=3e4)
    assert st.transport() == 0
    st.flux(cfg.dt, 0.0)
    fz = st.ff[2][4].copy()
    # change only the density used by the kinetic-energy factor: the cons array of layer j = 2
    g = p.ngc
    st.cu[0, :, g + 2, :] *= 1.001
    st.flux(cfg.dt, 0.0)
    dz = np.abs(st.ff[2][4] - fz).max(axis=(0, 2))     # per y layer j of the z faces
    assert dz[3] > 0            # z faces at j = 3 read rho(j-1 = 2) through the quirk


# ---- closed forms of individual flux terms --------------------------------------------------

@pytest.mark.parametrize("adv", [0, 1, 2])
def test_stag_uniform_flow_fluxes(orc, adv):
    """A uniform four-species gas moving at (U, V, W): every gradient vanishes, so the fluxes are
    the Euler ones -- mass rho U_d, energy (rho E + p) U_d for each advection form (type 0 through
    sum_k Y_k h_k + |u|^2/2, fluxStag.cpp:1538-1567), species rho Y_k U_d, cell-centred momentum
    rho U_d^2 + p and edge momentum rho U_a U_b."""
    cfg, p = _four((4, 5, 6), stoch_stress_form=0, advection_type=adv)
    U = (2.0e3, -3.0e3, 1.5e3)
    st = orc.StagHost(p)
    NS, nv = p.nspecies, p.nvars
    rho, T = cfg.rho0, cfg.T_init
    Y = [0.1, 0.2, 0.3, 0.4]
    cv = sum(Y[n] * p.hcv[n] for n in range(NS))
    E = rho * cv * T + 0.5 * rho * sum(u * u for u in U)
    st.cu[0] = rho
    st.cu[4] = E
    for n in range(NS):
        st.cu[5 + n] = rho * Y[n]
    for d in range(3):
        st.mom[d][:] = rho * U[d]
    st.c2p()
    Mmix = 1.0 / sum(Y[n] / p.molmass[n] for n in range(NS))
    pres = rho * P.R_UNIV / Mmix * T
    assert st.transport() == 0
    st.flux(cfg.dt, 0.0)
    for d in range(3):
        f = st.ff[d]
        assert np.abs(f[0] / (rho * U[d]) - 1).max() < 1e-13
        assert np.abs(f[4] / ((E + pres) * U[d]) - 1).max() < 1e-12
        for n in range(NS):
            assert np.abs(f[5 + n] / (rho * Y[n] * U[d]) - 1).max() < 1e-13
        assert np.abs(st.cf[d] / (rho * U[d] ** 2 + pres) - 1).max() < 1e-13
    from oracle.ffi import EDGE_AXES, EDGE_FLUX_TYPES
    for e, t in enumerate(EDGE_FLUX_TYPES):
        a, b = EDGE_AXES[t]
        assert np.abs(st.ef[e] / (rho * U[a] * U[b]) - 1).max() < 1e-13, e


def _shear_host(orc, p, cfg, A, axis_u, axis_k, n):
    """Uniform argon with velocity component axis_u varying as A sin(k x_axis_k) (cell
    positions), uniform T: face momenta rho u, energy rho cv T + the staggered kinetic energy."""
    st = orc.StagHost(p)
    g = p.ngc
    h = p.dx[axis_k]
    k = 2 * math.pi / (n * h)
    if axis_u == axis_k:       # faces of the varying axis sit half a cell below the cell centres
        x = np.arange(-g, n + g + 1) * h
    else:
        x = (np.arange(-g, n + g) + 0.5) * h
    u = A * np.sin(k * x)
    shape = [1, 1, 1]
    shape[2 - axis_k] = len(u)
    st.mom[axis_u][0] = cfg.rho0 * u.reshape(shape)
    st.cu[0] = cfg.rho0
    for n_ in range(2):
        st.cu[5 + n_] = 0.5 * cfg.rho0
    m = st.mom[axis_u][0][g:-g or None, g:-g or None, g:-g or None]
    st.cu[4] = cfg.rho0 * p.hcv[0] * cfg.T_init
    st.c2p()
    # put the kinetic energy in so that T stays uniform
    mm = st.mom[axis_u][0]
    sl = [slice(g, -g)] * 3
    lo = mm[g:g + p.n_cells[2], g:g + p.n_cells[1], g:g + p.n_cells[0]]
    up_idx = [slice(g, g + p.n_cells[2]), slice(g, g + p.n_cells[1]), slice(g, g + p.n_cells[0])]
    up_idx[2 - axis_u] = slice(g + 1, g + 1 + p.n_cells[axis_u])
    up = mm[tuple(up_idx)]
    st.cu[4][tuple(sl)] += 0.125 / cfg.rho0 * (lo + up) ** 2
    st.c2p()
    return st, k, h


def test_stag_viscous_heating_is_stress_times_face_velocity(orc):
    """Energy flux of the viscous work, term by term, for flows where the stresses are known:
    a shear flow u_x(y) gives on y faces -tau_xy (u_j + u_{j-1})/2 with tau_xy = eta (u_j -
    u_{j-1})/dy (deterministic), plus the stochastic part sqrt(2 k_B eta T / (dV dt)) W_j (u_j +
    u_{j-1})/2 for an edge noise W_j; a compression u_x(x) gives on x faces -u_i (tau_xx(i-1) +
    tau_xx(i))/2 with tau_xx = 4/3 eta du/dx, and with a cell noise Z_i the stochastic part
    u_i (t(i-1) + t(i))/2, t = (fac1 + fac2) Z (trace = Z).  Pins which velocities and stresses
    each face pairs (the cyclic test cannot: a pairing mistake made in every direction keeps the
    symmetry)."""
    n = 16
    A = 50.0
    kB = P.K_B
    # shear: u_x(y) on x faces, look at y faces
    cfg, p = _argon((4, n, 4), stoch_stress_form=1)
    st, k, h = _shear_host(orc, p, cfg, A, 0, 1, n)
    g, nv = p.ngc, p.nvars
    T = cfg.T_init
    assert np.abs(valid(p, st.prim)[4] / T - 1).max() < 1e-12
    for arrs in (st.fA, st.fB, st.eA, st.eB, st.cA, st.cB):
        for a in arrs:
            a[:] = 0
    W = np.cos(3 * k * np.arange(n + 1) * h) + 0.25
    st.eA[0][0] = W[None, :, None]                          # xy edges, one value per edge row j
    assert st.transport() == 0
    eta = st.eta[0, g, g, g]                                # uniform: the GIO value at T
    assert np.ptp(st.eta) <= 1e-15 * eta
    st.flux(cfg.dt, 0.0)
    u = st.vel[0][0, g, g - 1:g + n, g]                     # rows j = -1 .. n-1
    uj, ujm = u[1:], u[:-1]
    tau = eta * (uj - ujm) / h
    dt = cfg.dt
    vol = p.dx[0] * p.dx[1] * p.dx[2]
    fac = math.sqrt(2.0 * kB * eta * T / vol / dt)
    expect = fac * W[:n] * 0.5 * (uj + ujm) - tau * 0.5 * (uj + ujm)
    got = st.ff[1][nv + 2, 0, :n, 0]
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()
    # compression: u_x(x) on x faces, look at x faces
    cfg, p = _argon((n, 4, 4), stoch_stress_form=1)
    st, k, h = _shear_host(orc, p, cfg, A, 0, 0, n)
    for arrs in (st.fA, st.fB, st.eA, st.eB, st.cA, st.cB):
        for a in arrs:
            a[:] = 0
    Z = np.sin(2 * k * (np.arange(-1, n + 1) + 0.5) * h) + 0.5
    st.cA[0][0] = Z[None, None, :]                          # cells -1 .. n (one ghost layer)
    assert st.transport() == 0
    eta = st.eta[0, g, g, g]
    st.flux(cfg.dt, 0.0)
    v = st.vel[0][0, g, g, g - 1:g + n + 2]                 # faces -1 .. n+1
    ui = v[1:n + 2]                                         # faces 0 .. n
    txx = 4.0 / 3.0 * eta * (v[1:] - v[:-1]) / h            # cells -1 .. n
    fac1 = math.sqrt(2.0 * kB * eta * T / vol / dt)
    ts = (fac1 - fac1 / 3.0) * Z                            # cells -1 .. n
    expect = 0.5 * ui * (ts[:-1] + ts[1:]) - 0.5 * ui * (txx[:-1] + txx[1:])
    got = st.ff[0][nv + 1, 0, 0, :]
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()


@pytest.mark.parametrize("alg", [2, 1])
def test_stag_label_diffusion_of_identical_species(orc, alg):
    """Two identical species at rest with a sinusoidal mass fraction: only the labels move, at
    D_12 k_eff^2 (hard-sphere binary diffusivity, k_eff^2 = (2 - 2 cos kh)/h^2) with species
    fluxes (algorithm_type 2) and not at all without (algorithm_type 1)."""
    n = 32
    cfg, p = _argon((4, n, 4), stoch_stress_form=0, algorithm_type=alg)
    g = p.ngc
    h = p.dx[1]
    k = 2 * math.pi / (n * h)
    s = np.sin(k * (np.arange(n) + 0.5) * h)[None, :, None]
    eps = 0.1
    st = stag_uniform_state(p, cfg, orc)
    st.cu[5, g:-g, g:-g, g:-g] = cfg.rho0 * (0.5 + eps * s)
    st.cu[6, g:-g, g:-g, g:-g] = cfg.rho0 * (0.5 - eps * s)
    st.c2p()

    def probe(st):
        y = st.cu[5, g, g:-g, g] / cfg.rho0 - 0.5
        return 2 * (y * s[0, :, 0]).mean()

    amp = _run(orc, p, cfg, st, 40, probe)
    if alg == 1:
        assert np.abs(amp - amp[0]).max() <= 1e-14
        return
    t = cfg.dt * np.arange(len(amp))
    rate = -np.polyfit(t, np.log(amp), 1)[0]
    kB, T, d = P.K_B, cfg.T_init, 3.66e-8
    m = 39.95 * P.K_B / P.R_UNIV
    pres = cfg.rho0 / m * kB * T
    D12 = 3 / 16 * math.sqrt(2 * math.pi * (kB * T) ** 3 / (m / 2)) / (pres * math.pi * d * d)
    assert rate == pytest.approx(D12 * (2 - 2 * math.cos(k * h)) / h ** 2, rel=5e-3)


def test_stag_trace_species_gets_no_species_noise(orc):
    """A species at mass fraction <= 1e-12 on both sides of a face has its row and column of the
    face's noise matrix zeroed (fluxStag.cpp:412-419): its stochastic flux is exactly zero."""
    base = P.species_prefix(3, 6, 1)
    Y = [0.6, 0.4 - 2e-14, 2e-14]

    def fluxes(stoch):
        p = P.make_params(3, (6, 6, 6), 8e-6, [4.0026, 20.1797, 39.9480], [2.18e-8, 2.58e-8, 3.63e-8],
                          stoch_stress_form=stoch)
        from fhdex_amd.ensemble import uniform_values
        cfg = P.Config("trace", p, base.rho0, Y, base.T_init, base.dt, 1)
        st = stag_uniform_state(p, cfg, orc)
        st.fill_noise(5, 9)
        assert st.transport() == 0
        st.flux(base.dt, 0.37)
        return [f.copy() for f in st.ff]

    noisy, quiet = fluxes(1), fluxes(0)
    for d in range(3):
        assert np.array_equal(noisy[d][7], quiet[d][7]), d
        for s in (5, 6):
            assert np.abs(noisy[d][s] - quiet[d][s]).max() > 0, (d, s)


# ---- chemistry ------------------------------------------------------------------------------

def _no2(shape, **kw):
    cfg = P.no2_n2o4(4, **kw)
    p = cfg.params
    p.n_cells[:] = shape
    return cfg, p


def test_stag_chemistry_source_enters_as_dt_times_source(orc):
    """A uniform NO2/N2O4 gas off equilibrium with the stochastic fluxes off and reaction_type 0
    (no reaction noise): every flux vanishes, so one staggered step changes rho Y_k by the RK3
    integral of the reaction source, dt S_k up to the source's own change within the step (S from
    the reference's rate law at the start, oracle/oracle.c::orc_chem_source; the composition and,
    through e0, the temperature move by ~1e-4 of the rate over 1e-12 s), and keeps rho, rho E and
    the face momenta (zero)."""
    cfg, p = _no2((4, 4, 4), reaction_type=0)
    p.stoch_stress_form = 0
    Y = [0.9, 0.1]                           # far from the deck's 0.724 / 0.276 balance
    cfg.rhobar = Y
    st = stag_uniform_state(p, cfg, orc)
    cu0 = valid(p, st.cu).copy()
    # the source of the start
    S = np.zeros((p.nvars,) + st.cu.shape[1:])
    ran = np.zeros((p.nreaction,) + st.cu.shape[1:])
    g = p.ngc
    assert orc.chem_source(p, cfg.chem, st.prim, S, ran, cfg.dt) == 0
    assert st.advance(1, cfg.dt, 1, 0, chem=cfg.chem) == 0
    d = valid(p, st.cu) - cu0
    want = cfg.dt * valid(p, S)
    for k in range(2):
        assert np.abs(d[5 + k] / want[5 + k] - 1).max() < 1e-3, k
    assert np.abs(want[5]).min() > 0
    assert np.abs(d[0]).max() <= 1e-15 * cfg.rho0
    assert np.abs(d[4]).max() <= 1e-13 * abs(cu0[4]).max()
    for dd in range(3):
        assert not np.any(st.mom[dd])


def test_stag_chemistry_conserves_mass_with_noise(orc):
    """With fluctuating hydrodynamics and the chemical Langevin noise on (reaction_type 1), three
    steps keep sum rho, sum rho E and the face momentum sums to rounding (M(N2O4) = 2 M(NO2):
    each reaction conserves mass), while the species totals move."""
    cfg, p = _no2((6, 5, 4), reaction_type=1)
    st = perturbed_stag_state(p, cfg, orc, seed=4, amp=0.02, vel=3e3)
    tot0 = valid(p, st.cu).sum(axis=(1, 2, 3))
    abs0 = np.abs(valid(p, st.cu)).sum(axis=(1, 2, 3))
    m0 = [face_valid(p, st.mom[d], d).sum() for d in range(3)]
    ma = [np.abs(face_valid(p, st.mom[d], d)).sum() for d in range(3)]
    assert st.advance(3, cfg.dt, 5, 0, chem=cfg.chem) == 0
    tot1 = valid(p, st.cu).sum(axis=(1, 2, 3))
    assert abs(tot1[0] - tot0[0]) <= 2e-14 * abs0[0]
    assert abs(tot1[4] - tot0[4]) <= 2e-14 * abs0[4]
    assert abs(tot1[5] - tot0[5]) > 1e-10 * abs0[5]
    for d in range(3):
        assert abs(face_valid(p, st.mom[d], d).sum() - m0[d]) <= 2e-14 * ma[d]
