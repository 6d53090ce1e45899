"""CPU tests that pin the oracle's mixture transport (GIO + D_GIO1) for species of distinct mass.

With identical species every mass-ratio term of the Giovangigli matrices collapses, so the
single-gas closed forms in test_oracle.py cannot catch a swapped or mistyped ratio.  Here the
oracle's eta, kappa, chi and D for rigid-sphere mixtures of DISTINCT masses are compared with
the textbook Chapman-Enskog first approximations, written in forms that share no algebra with
GIO's linear solves:

* binary mixtures: the closed forms of Hirschfelder, Curtiss & Bird, "Molecular Theory of Gases
  and Liquids" (1954), sec. 8.2 -- viscosity (1 + Z_eta)/(X_eta + Y_eta) (eq. 8.2-22), thermal
  conductivity (1 + Z_lambda)/(X_lambda + Y_lambda) with the U coefficients (eqs. 8.2-36/37),
  the thermal diffusion ratio k_T from S^(1), S^(2) (eq. 8.2-50) and the binary diffusion
  coefficient of rigid spheres (eq. 8.2-44).  The reference carries the same binary forms,
  multiplied by rigid-sphere higher-approximation factors 1.016, 1.025 and 1.299, as its
  transport_type 3 (src_compressible/compressible_functions.H:1355-1478); the first
  approximation drops those factors, and that is what GIO computes
  (compressible_functions.H:1153-1352);
* any number of species: the determinant (matrix) forms of the same section, with the
  interaction viscosities [eta_ij] and conductivities [lambda_ij] of reduced mass 2 m_i m_j/
  (m_i + m_j), eta = x^T H^-1 x and lambda = 4 x^T (-L)^-1 x;
* splitting a species into two identical labelled species changes no transport coefficient;
* a trace species leaves the host gas's own closed forms.

Collision integrals of rigid spheres: Omega* = A* = B* = C* = 1.
"""
import math

import numpy as np
import pytest

from fhdex_amd import params as P

KB, RU = P.K_B, P.R_UNIV
HE_AR = ((4.0026, 39.948), (2.18e-8, 3.63e-8))
NE_KR = ((20.1797, 83.8), (2.58e-8, 4.16e-8))
AR_HE = ((39.948, 4.0026), (3.63e-8, 2.18e-8))  # heavy species first: every ratio inverted
# the four-species gas of exec/compressible/inputs_profiling_gpu:82-83
C3 = ((4.0026, 20.1797, 39.948, 83.8), (2.18e-8, 2.58e-8, 3.63e-8, 4.16e-8))
BINARY_Y = [(0.5, 0.5), (0.2, 0.8), (0.9, 0.1), (0.01, 0.99)]
T0, RHO0 = 300.0, 4.83e-4


def uniform_transport(orc, M, d, Y, T=T0, rho=RHO0, transport_type=1):
    """The oracle's eta, kappa, chi, D (D(l*NS+k) = rho Diff_kl Y_k) and X, p for a uniform state."""
    NS = len(M)
    p = P.make_params(NS, (4, 4, 4), 8e-6, M, d)
    p.transport_type = transport_type
    st = orc.HostState(p)
    Y = np.asarray(Y, float)
    Mv = np.asarray(M, float)
    X = (Y / Mv) / np.sum(Y / Mv)
    pres = rho * RU * T * np.sum(Y / Mv)
    st.prim[0], st.prim[4], st.prim[5] = rho, T, pres
    for n in range(NS):
        st.prim[6 + n] = Y[n]
        st.prim[6 + NS + n] = X[n]
    assert orc.transport(p, st) == 0
    c = (slice(None), 2, 2, 2)
    return st.eta[c][0], st.kappa[c][0], st.chi[c].copy(), st.D[c].copy(), X, pres


def pure_eta(m, d, T):
    return 5.0 / (16.0 * d * d) * math.sqrt(m * KB * T / math.pi)


def pure_lambda(m, d, T):
    return 75.0 / 64.0 * KB / (d * d) * math.sqrt(KB * T / (math.pi * m))


def hcb_binary(M, d, X, T):
    """HCB sec. 8.2 closed forms for a binary rigid-sphere mixture: eta, lambda and k_T of
    species 1 (first approximations)."""
    m1, m2 = (Mi * KB / RU for Mi in M)
    x1, x2 = X
    m12 = 2.0 * m1 * m2 / (m1 + m2)
    d12 = 0.5 * (d[0] + d[1])
    e1, e2, e12 = pure_eta(m1, d[0], T), pure_eta(m2, d[1], T), pure_eta(m12, d12, T)
    r = (m1 + m2) ** 2 / (4.0 * m1 * m2)
    Xe = x1 * x1 / e1 + 2.0 * x1 * x2 / e12 + x2 * x2 / e2
    Ye = 0.6 * (x1 * x1 / e1 * (m1 / m2) + 2.0 * x1 * x2 * r * e12 / (e1 * e2) + x2 * x2 / e2 * (m2 / m1))
    Ze = 0.6 * (x1 * x1 * m1 / m2 + 2.0 * x1 * x2 * (r * (e12 / e1 + e12 / e2) - 1.0) + x2 * x2 * m2 / m1)
    eta = (1.0 + Ze) / (Xe + Ye)
    l1, l2, l12 = pure_lambda(m1, d[0], T), pure_lambda(m2, d[1], T), pure_lambda(m12, d12, T)
    q = (m1 - m2) ** 2 / (m1 * m2)
    U1 = 4.0 / 15.0 - 17.0 / 60.0 * m1 / m2 + 0.5 * q
    U2 = 4.0 / 15.0 - 17.0 / 60.0 * m2 / m1 + 0.5 * q
    UY = 4.0 / 15.0 * r * l12 * l12 / (l1 * l2) - 17.0 / 60.0 + 13.0 / 32.0 * q
    UZ = 4.0 / 15.0 * (r * (l12 / l1 + l12 / l2) - 1.0) - 17.0 / 60.0
    Xl = x1 * x1 / l1 + 2.0 * x1 * x2 / l12 + x2 * x2 / l2
    Yl = x1 * x1 / l1 * U1 + 2.0 * x1 * x2 / l12 * UY + x2 * x2 / l2 * U2
    Zl = x1 * x1 * U1 + 2.0 * x1 * x2 * UZ + x2 * x2 * U2
    lam = (1.0 + Zl) / (Xl + Yl)
    S1 = (m1 + m2) / (2.0 * m2) * l12 / l1 - 15.0 / 8.0 * (m2 - m1) / m1 - 1.0
    S2 = (m1 + m2) / (2.0 * m1) * l12 / l2 - 15.0 / 8.0 * (m1 - m2) / m2 - 1.0
    kT1 = x1 * x2 / (6.0 * l12) * (x1 * S1 - x2 * S2) / (Xl + Yl)
    return eta, lam, kT1


def hcb_matrices(M, d, X, T):
    """HCB sec. 8.2 determinant forms of the mixture viscosity and conductivity (rigid spheres)."""
    m = np.asarray(M, float) * KB / RU
    dd = np.asarray(d, float)
    NS = len(m)
    eta_ij = np.empty((NS, NS))
    lam_ij = np.empty((NS, NS))
    for i in range(NS):
        for j in range(NS):
            mij = 2.0 * m[i] * m[j] / (m[i] + m[j])
            dij = 0.5 * (dd[i] + dd[j])
            eta_ij[i, j] = pure_eta(mij, dij, T)
            lam_ij[i, j] = pure_lambda(mij, dij, T)
    H = np.empty((NS, NS))
    L = np.empty((NS, NS))
    for i in range(NS):
        H[i, i] = X[i] ** 2 / eta_ij[i, i]
        L[i, i] = -4.0 * X[i] ** 2 / lam_ij[i, i]
        for k in range(NS):
            if k == i:
                continue
            mm = m[i] * m[k] / (m[i] + m[k]) ** 2
            H[i, i] += 2.0 * X[i] * X[k] / eta_ij[i, k] * mm * (5.0 / 3.0 + m[k] / m[i])
            L[i, i] -= (2.0 * X[i] * X[k] * (7.5 * m[i] ** 2 + 6.25 * m[k] ** 2 - 3.0 * m[k] ** 2 +
                                               4.0 * m[i] * m[k]) / ((m[i] + m[k]) ** 2 * lam_ij[i, k]))
            H[i, k] = -2.0 * X[i] * X[k] / eta_ij[i, k] * mm * (5.0 / 3.0 - 1.0)
            L[i, k] = 2.0 * X[i] * X[k] * m[i] * m[k] * (55.0 / 4.0 - 3.0 - 4.0) / ((m[i] + m[k]) ** 2 * lam_ij[i, k])
    return X @ np.linalg.solve(H, X), 4.0 * X @ np.linalg.solve(-L, X)


def binary_D12(M, d, T, pres):
    """First-approximation binary diffusion coefficient of rigid spheres (HCB eq. 8.2-44)."""
    m1, m2 = (Mi * KB / RU for Mi in M)
    mu = m1 * m2 / (m1 + m2)
    s = 0.5 * (d[0] + d[1])
    return 3.0 / 16.0 * math.sqrt(2.0 * math.pi * (KB * T) ** 3 / mu) / (pres * math.pi * s * s)


PAIRS = [HE_AR, NE_KR, AR_HE]


@pytest.mark.parametrize("gas", PAIRS, ids=["He-Ar", "Ne-Kr", "Ar-He"])
@pytest.mark.parametrize("Y", BINARY_Y)
def test_binary_viscosity_and_conductivity_match_hcb(orc, gas, Y):
    M, d = gas
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    he, hl, _ = hcb_binary(M, d, X, T0)
    assert eta == pytest.approx(he, rel=1e-13)
    assert kap == pytest.approx(hl, rel=1e-13)
    # the mixture differs from either pure gas by far more than the tolerance
    m = [Mi * KB / RU for Mi in M]
    assert min(abs(eta / pure_eta(m[n], d[n], T0) - 1) for n in range(2)) > 1e-4


@pytest.mark.parametrize("gas", PAIRS, ids=["He-Ar", "Ne-Kr", "Ar-He"])
@pytest.mark.parametrize("Y", BINARY_Y)
def test_binary_thermal_diffusion_ratio_matches_hcb(orc, gas, Y):
    """The rescaled ratios of GIO are chi_1 = -k_T/X_1, chi_2 = k_T/X_2 (k_T of species 2 in
    HCB's labelling, i.e. -k_T of species 1); the heavy species gathers at the cold side."""
    M, d = gas
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    _, _, kT1 = hcb_binary(M, d, X, T0)
    kT2 = -kT1
    assert X[1] * chi[1] == pytest.approx(kT2, rel=1e-12)
    assert X[0] * chi[0] == pytest.approx(-kT2, rel=1e-12)
    heavy = int(np.argmax(M))
    assert chi[heavy] > 0 > chi[1 - heavy]


@pytest.mark.parametrize("gas", PAIRS, ids=["He-Ar", "Ne-Kr", "Ar-He"])
@pytest.mark.parametrize("Y", BINARY_Y)
def test_binary_fick_coefficient(orc, gas, Y):
    """With d_2 = -d_1 the species flux is F_1 = -(D(0,0) - D(1*2+0)) d_1, and Fick's law in mole
    fractions for a binary mixture is F_1 = -rho D_12 M_1 M_2 / Mbar^2 grad X_1."""
    M, d = gas
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    Mbar = float(np.dot(X, M))
    assert D[0] - D[2] == pytest.approx(RHO0 * binary_D12(M, d, T0, pres) * M[0] * M[1] / Mbar ** 2, rel=1e-12)
    assert D[3] - D[1] == pytest.approx(D[0] - D[2], rel=1e-12)


@pytest.mark.parametrize("Y", [(0.25, 0.25, 0.25, 0.25), (0.1, 0.2, 0.3, 0.4), (0.7, 0.1, 0.1, 0.1),
                               (0.01, 0.01, 0.01, 0.97)])
def test_four_species_match_hcb_determinants(orc, Y):
    M, d = C3
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    he, hl = hcb_matrices(M, d, X, T0)
    assert eta == pytest.approx(he, rel=1e-13)
    assert kap == pytest.approx(hl, rel=1e-13)
    # sum_i X_i chi_i = 0: the ratios are k_T,i / X_i and the k_T,i sum to zero
    assert abs(np.dot(X, chi)) <= 1e-13 * np.abs(chi).max()


@pytest.mark.parametrize("Y", [(0.5, 0.5), (0.2, 0.8)])
def test_split_species_leave_transport_unchanged(orc, Y):
    """He-Ar written as four species (two labels of each gas) gives the binary eta, kappa and
    chi, and the diffusive flux of each gas, summed over its labels, is the binary one.  The
    flux agrees to the truncation of D_GIO1's three-term series (jmax = 3,
    compressible_functions.H:663, 745-765), which is exact for two species and leaves ~4e-4 for four; a wrong mass ratio or
    index moves it by O(1)."""
    M, d = HE_AR
    eb, kb, cb, Db, Xb, _ = uniform_transport(orc, M, d, Y)
    f = (0.3, 0.7, 0.6, 0.4)
    Y4 = (Y[0] * f[0], Y[0] * f[1], Y[1] * f[2], Y[1] * f[3])
    M4 = (M[0], M[0], M[1], M[1])
    d4 = (d[0], d[0], d[1], d[1])
    e4, k4, c4, D4, X4, _ = uniform_transport(orc, M4, d4, Y4)
    assert e4 == pytest.approx(eb, rel=1e-13)
    assert k4 == pytest.approx(kb, rel=1e-13)
    np.testing.assert_allclose(c4, cb[[0, 0, 1, 1]], rtol=1e-12)
    # driving forces with the labels in fixed proportion: d_l = (X_l / X_gas) d_gas, d_Ar = -d_He
    dgas = np.array([1.0, -1.0])
    gas_of = [0, 0, 1, 1]
    d_l = np.array([X4[l] / Xb[gas_of[l]] * dgas[gas_of[l]] for l in range(4)])
    D4m = D4.reshape(4, 4)  # [l, k]
    F4 = -(D4m * d_l[:, None]).sum(axis=0)
    Fb = -(Db.reshape(2, 2) * dgas[:, None]).sum(axis=0)
    np.testing.assert_allclose([F4[0] + F4[1], F4[2] + F4[3]], Fb, rtol=2e-3)


@pytest.mark.parametrize("host", range(4))
def test_trace_species_leave_host_closed_forms(orc, host):
    """A host gas with the other three species at Y = 1e-9 each has its own eta and kappa to
    O(1e-7) (the light species' mole fractions are up to 21 times their mass fractions), and the trace species' effect on eta is linear in their amount."""
    M, d = C3
    m = [Mi * KB / RU for Mi in M]
    out = []
    for tr in (1e-9, 2e-9):
        Y = [tr] * 4
        Y[host] = 1.0 - 3.0 * tr
        eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
        out.append((eta, kap))
        assert eta == pytest.approx(pure_eta(m[host], d[host], T0), rel=1e-7)
        assert kap == pytest.approx(pure_lambda(m[host], d[host], T0), rel=1e-6)
    e0 = pure_eta(m[host], d[host], T0)
    # doubling the trace amount doubles the departure (first order in the trace mole fractions)
    assert (out[1][0] - e0) == pytest.approx(2.0 * (out[0][0] - e0), rel=1e-3)


@pytest.mark.parametrize("Y,tol", [((0.25, 0.25, 0.25, 0.25), 3e-3), ((0.1, 0.2, 0.3, 0.4), 3e-3),
                                   ((0.7, 0.1, 0.1, 0.1), 3e-3), ((1e-6, 1e-6, 1e-6, 1 - 3e-6), 1e-10),
                                   ((1e-4, 1e-4, 0.5, 0.4998), 1e-7)])
def test_four_species_diffusion_solves_stefan_maxwell(orc, Y, tol):
    """The species velocities V_k = -sum_l Diff_kl d_l of D_GIO1 solve the Stefan-Maxwell
    relations d_i = sum_j X_i X_j (V_j - V_i) / D_ij with sum_k Y_k V_k = 0 (the binary D_ij of
    HCB eq. 8.2-44), to the truncation of the three-term series: ~2e-3 for comparable mole
    fractions, and far less where trace species make the series converge at once.  The
    mixture-averaged preconditioner inside the series changes only that truncation, so its
    exact form is pinned by the restatement alone (the series is exact for two species)."""
    M, d = C3
    NS = 4
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    Diff = D.reshape(NS, NS).T / (RHO0 * np.asarray(Y)[:, None])  # [k, l]
    A = np.zeros((NS, NS))
    for i in range(NS):
        for j in range(NS):
            if i != j:
                a = X[i] * X[j] / binary_D12((M[i], M[j]), (d[i], d[j]), T0, pres)
                A[i, j] += a
                A[i, i] -= a
    K = np.vstack([A, np.asarray(Y)])
    rng = np.random.default_rng(7)
    for _ in range(8):
        dv = rng.standard_normal(NS)
        dv -= dv.mean()
        V = np.linalg.lstsq(K, np.concatenate([dv, [0.0]]), rcond=None)[0]
        assert np.abs(-Diff @ dv - V).max() <= tol * np.abs(V).max()


# ---- transport_type 3: the reference's Hirschfelder-Curtiss-Bird binary forms ------------------

AR_AR = ((39.95, 39.95), (3.66e-8, 3.66e-8))


@pytest.mark.parametrize("gas", PAIRS + [AR_AR], ids=["He-Ar", "Ne-Kr", "Ar-He", "Ar-Ar"])
@pytest.mark.parametrize("Y", BINARY_Y)
def test_hcb_binary_transport_matches_textbook_times_corrections(orc, gas, Y):
    """IdealMixtureTransportHCBBin (compressible_functions.H:1357-1478) against HCB sec. 8.2
    written independently above: viscosity 1.016 (1 + Z_eta)/(X_eta + Y_eta), conductivity
    1.025 (1 + Z_lambda)/(X_lambda + Y_lambda), thermal diffusion X_2 chi_2 = -X_1 chi_1 =
    1.299 k_T(2); the Fick coefficient is 1.133 rho D_12 M_1 M_2 / Mbar^2 times 2 mu / mbar,
    because the reference divides the interaction viscosity by rho where HCB's (6/5) eta_12 /
    (n m_12) has n m_12 = n 2 mu (the two agree at equal mass fractions and for one gas)."""
    M, d = gas
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y, transport_type=3)
    he, hl, kT1 = hcb_binary(M, d, X, T0)
    assert eta == pytest.approx(1.016 * he, rel=1e-13)
    assert kap == pytest.approx(1.025 * hl, rel=1e-13)
    kT2 = -kT1
    if M[0] == M[1]:
        assert chi[0] == 0.0 and chi[1] == 0.0
    else:
        assert X[1] * chi[1] == pytest.approx(1.299 * kT2, rel=1e-12)
        assert X[0] * chi[0] == pytest.approx(-1.299 * kT2, rel=1e-12)
    m1, m2 = (Mi * KB / RU for Mi in M)
    mu = m1 * m2 / (m1 + m2)
    Mbar = float(np.dot(X, M))
    mbar = Mbar * KB / RU
    fick = RHO0 * binary_D12(M, d, T0, pres) * M[0] * M[1] / Mbar ** 2
    assert D[0] - D[2] == pytest.approx(1.133 * (2.0 * mu / mbar) * fick, rel=1e-12)
    assert D[3] - D[1] == pytest.approx(D[0] - D[2], rel=1e-12)
    assert D[0] + D[1] == pytest.approx(0.0, abs=1e-15 * abs(D[0]))
    assert D[2] + D[3] == pytest.approx(0.0, abs=1e-15 * abs(D[3]))


def test_hcb_and_gio_agree_up_to_the_correction_factors(orc):
    """For rigid spheres both are the first Chapman-Enskog approximation: eta, kappa and chi of
    transport_type 3 are those of the Giovangigli solve (type 1) times 1.016, 1.025 and 1.299."""
    for M, d in (HE_AR, NE_KR):
        for Y in BINARY_Y:
            e1, k1, c1, D1, X, _ = uniform_transport(orc, M, d, Y)
            e3, k3, c3, D3, _, _ = uniform_transport(orc, M, d, Y, transport_type=3)
            assert e3 == pytest.approx(1.016 * e1, rel=1e-13)
            assert k3 == pytest.approx(1.025 * k1, rel=1e-13)
            assert np.allclose(c3, 1.299 * c1, rtol=1e-12, atol=0)


# ---- five to eight species (the reference's MAX_SPECIES is 8) ----------------------------------

def _wide_Y(NS, kind):
    if kind == "equal":
        return [1.0 / NS] * NS
    if kind == "dirichlet":
        return list(np.random.default_rng(NS).dirichlet(np.ones(NS)))
    Y = [1e-6] * NS  # one host gas with the others in trace
    Y[2] = 1.0 - (NS - 1) * 1e-6
    return Y


@pytest.mark.parametrize("NS", [5, 6, 7, 8])
@pytest.mark.parametrize("kind", ["equal", "dirichlet", "trace"])
def test_many_species_match_hcb_determinants(orc, NS, kind):
    """Up to eight distinct species (params.EIGHT_SPECIES_*): GIO's eta and kappa equal the HCB
    determinant forms, and sum_i X_i chi_i = 0."""
    M, d = P.EIGHT_SPECIES_M[:NS], P.EIGHT_SPECIES_D[:NS]
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, _wide_Y(NS, kind))
    he, hl = hcb_matrices(M, d, X, T0)
    assert eta == pytest.approx(he, rel=1e-13)
    assert kap == pytest.approx(hl, rel=1e-13)
    assert abs(np.dot(X, chi)) <= 1e-13 * np.abs(chi).max()


@pytest.mark.parametrize("NS", [5, 8])
@pytest.mark.parametrize("kind", ["equal", "dirichlet"])
def test_many_species_diffusion_solves_stefan_maxwell(orc, NS, kind):
    """D_GIO1's species velocities solve the Stefan-Maxwell relations for five and eight
    species, to the truncation of its three-term series (2e-3 here)."""
    M, d = P.EIGHT_SPECIES_M[:NS], P.EIGHT_SPECIES_D[:NS]
    Y = _wide_Y(NS, kind)
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y)
    Diff = D.reshape(NS, NS).T / (RHO0 * np.asarray(Y)[:, None])
    A = np.zeros((NS, NS))
    for i in range(NS):
        for j in range(NS):
            if i != j:
                a = X[i] * X[j] / binary_D12((M[i], M[j]), (d[i], d[j]), T0, pres)
                A[i, j] += a
                A[i, i] -= a
    K = np.vstack([A, np.asarray(Y)])
    rng = np.random.default_rng(11)
    for _ in range(8):
        dv = rng.standard_normal(NS)
        dv -= dv.mean()
        V = np.linalg.lstsq(K, np.concatenate([dv, [0.0]]), rcond=None)[0]
        assert np.abs(-Diff @ dv - V).max() <= 2e-3 * np.abs(V).max()


def test_eight_labels_of_two_gases_leave_transport_unchanged(orc):
    """He-Ar written as eight species (four labels of each gas) gives the binary eta, kappa and
    the binary chi on every label."""
    M, d = HE_AR
    Y = (0.2, 0.8)
    eb, kb, cb, Db, Xb, _ = uniform_transport(orc, M, d, Y)
    f = (0.1, 0.2, 0.3, 0.4)
    Y8 = tuple(Y[0] * x for x in f) + tuple(Y[1] * x for x in f)
    e8, k8, c8, D8, X8, _ = uniform_transport(orc, (M[0],) * 4 + (M[1],) * 4, (d[0],) * 4 + (d[1],) * 4, Y8)
    assert e8 == pytest.approx(eb, rel=1e-13)
    assert k8 == pytest.approx(kb, rel=1e-13)
    np.testing.assert_allclose(c8, cb[[0] * 4 + [1] * 4], rtol=1e-12)


# ---- transport_type 2: the reference's Valk-Waldmann forms (compressible_functions.H:1047-1150) ---
#
# ViscLin solves the first Sonine approximation of the viscosity (HCB 7.4-56, 62, 63), the same
# approximation as the HCB determinant form, so eta agrees to rounding for any mixture.  ThermalDiff
# (Valk) gives X_1 chi_1 = k_T of HCB eq. 8.2-50 for a binary gas.  LambdaLin's conductivity
# (two Sonine terms with the thermal-diffusion correction of HCB 7.4-65) reduces to the pure-gas
# value for identical species and stays within 3e-3 of HCB's first approximation [lambda]_1 for
# distinct masses (2.4e-3 at most here, always below); a sign or coefficient slip in its Q
# matrix, its rhs or the correction moves it by 6e-3 to O(1).  Relabelling the species permutes
# every output.  D is D_GIO's series on the same binary diffusion
# coefficients (HCB eq. 8.2-44), exact for two species.

VW_PAIRS = PAIRS + [AR_AR]
VW_IDS = ["He-Ar", "Ne-Kr", "Ar-He", "Ar-Ar"]


@pytest.mark.parametrize("gas", VW_PAIRS, ids=VW_IDS)
@pytest.mark.parametrize("Y", BINARY_Y)
def test_vw_binary_matches_hcb(orc, gas, Y):
    M, d = gas
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y, transport_type=2)
    he, hl, hk = hcb_binary(M, d, X, T0)
    assert eta == pytest.approx(he, rel=1e-13)
    assert kap == pytest.approx(hl, rel=3e-3)
    if M[0] == M[1]:
        assert kap == pytest.approx(hl, rel=1e-13)
        assert np.abs(chi).max() <= 1e-14
    else:
        assert X[0] * chi[0] == pytest.approx(hk, rel=1e-12)
    assert abs(np.dot(X, chi)) <= 1e-13 * max(np.abs(chi).max(), 1.0)
    # Fick coefficient of the binary series: rho D12 M1 M2 / Mbar^2 (the series is exact for NS = 2)
    Dm = D.reshape(2, 2)  # [l, k]
    Mbar = 1.0 / (Y[0] / M[0] + Y[1] / M[1])
    want = RHO0 * binary_D12(M, d, T0, pres) * M[0] * M[1] / Mbar ** 2
    assert Dm[0, 0] - Dm[1, 0] == pytest.approx(want, rel=1e-10)


@pytest.mark.parametrize("NS,kind", [(4, "equal"), (4, "dirichlet"), (8, "equal"), (8, "dirichlet"),
                                     (8, "trace")])
def test_vw_many_species(orc, NS, kind):
    """Four and eight distinct species: eta equals the HCB determinant form, sum_i X_i chi_i = 0, and
    the diffusion coefficients solve the Stefan-Maxwell relations to the series truncation."""
    M, d = P.EIGHT_SPECIES_M[:NS], P.EIGHT_SPECIES_D[:NS]
    Y = _wide_Y(NS, kind)
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y, transport_type=2)
    he, hl = hcb_matrices(M, d, X, T0)
    assert eta == pytest.approx(he, rel=1e-13)
    assert kap == pytest.approx(hl, rel=3e-3)
    assert abs(np.dot(X, chi)) <= 1e-13 * np.abs(chi).max()
    Diff = D.reshape(NS, NS).T / (RHO0 * np.asarray(Y)[:, None])
    A = np.zeros((NS, NS))
    for i in range(NS):
        for j in range(NS):
            if i != j:
                a = X[i] * X[j] / binary_D12((M[i], M[j]), (d[i], d[j]), T0, pres)
                A[i, j] += a
                A[i, i] -= a
    K = np.vstack([A, np.asarray(Y)])
    rng = np.random.default_rng(5)
    for _ in range(6):
        dv = rng.standard_normal(NS)
        dv -= dv.mean()
        V = np.linalg.lstsq(K, np.concatenate([dv, [0.0]]), rcond=None)[0]
        assert np.abs(-Diff @ dv - V).max() <= 3e-3 * np.abs(V).max()


@pytest.mark.parametrize("host", range(4))
def test_vw_trace_species_leave_host_closed_forms(orc, host):
    """A host gas with three trace species (Y = 1e-9): eta and kappa are the host's pure-gas
    values to O(1e-7)."""
    M, d = C3
    m = [Mi * KB / RU for Mi in M]
    Y = [1e-9] * 4
    Y[host] = 1.0 - 3e-9
    eta, kap, chi, D, X, pres = uniform_transport(orc, M, d, Y, transport_type=2)
    assert eta == pytest.approx(pure_eta(m[host], d[host], T0), rel=1e-7)
    assert kap == pytest.approx(pure_lambda(m[host], d[host], T0), rel=1e-6)


def test_vw_singular_matrix_is_reported(orc):
    """A zero temperature makes every collision integral zero: the pivoted factorisation meets a
    pivot below 1e-30, where the reference aborts ("singular matrix in decomp"); the oracle
    reports FHDX_ERR_SINGULAR."""
    M, d = HE_AR
    p = P.make_params(2, (4, 4, 4), 8e-6, M, d)
    p.transport_type = 2
    st = orc.HostState(p)
    st.prim[0], st.prim[4], st.prim[5] = RHO0, 0.0, 1e5
    st.prim[6], st.prim[7] = 0.5, 0.5
    st.prim[8], st.prim[9] = 0.9, 0.1
    assert orc.transport(p, st) == 8


def test_vw_relabelling_permutes_every_output(orc):
    """The four-species gas listed in reverse with a Dirichlet composition: eta, kappa equal, chi and
    D permuted (to 1e-12; the pivoted factorisation sees the rows in another order)."""
    M, d = C3
    Y = list(np.random.default_rng(3).dirichlet(np.ones(4)))
    perm = [3, 1, 0, 2]
    e1, k1, c1, D1, X1, _ = uniform_transport(orc, M, d, Y, transport_type=2)
    e2, k2, c2, D2, X2, _ = uniform_transport(orc, [M[q] for q in perm], [d[q] for q in perm],
                                              [Y[q] for q in perm], transport_type=2)
    assert e2 == pytest.approx(e1, rel=1e-12)
    assert k2 == pytest.approx(k1, rel=1e-12)
    np.testing.assert_allclose(c2, c1[perm], rtol=1e-11, atol=1e-14 * np.abs(c1).max())
    Dm1, Dm2 = D1.reshape(4, 4), D2.reshape(4, 4)
    np.testing.assert_allclose(Dm2, Dm1[np.ix_(perm, perm)], rtol=1e-10, atol=1e-12 * np.abs(Dm1).max())
