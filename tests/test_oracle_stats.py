"""The oracle's evaluateStats (stats.cpp:5-363) against what the statistics must be.

Running means must be the arithmetic means of the samples; with two samples x1, x2 the running
variance recursion gives (x2 - x1)^2 / 8; the primitive means are conservedToPrimitive of the
mean state; fields constant in y and z have yz-slice averages equal to the field; with one
sample every covariance-type entry (instant minus mean products) vanishes.
"""
import numpy as np
import pytest

from fhdex_amd import params as P
from tests.helpers import perturbed_state, valid


def samples(orc, cfg, n, seed=0):
    p = cfg.params
    out = []
    for q in range(n):
        st = perturbed_state(p, cfg, orc, seed=seed + q, amp=0.01, vel=300.0)
        out.append((st.cu.copy(), st.prim.copy()))
    return out


@pytest.fixture
def cfg():
    return P.argon_gas(0, shape=(8, 6, 5))


def test_means_are_sample_means_and_prim_of_mean(orc, cfg):
    p = cfg.params
    xs = samples(orc, cfg, 3)
    ss = orc.StatsState(p)
    for s, (cu, prim) in enumerate(xs, start=1):
        orc.evaluate_stats(p, cu, prim, ss, s, 2)
    want = np.mean([valid(p, cu) for cu, _ in xs], axis=0)
    got = valid(p, ss.consMean)
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max(axis=None)
    # primMean = conservedToPrimitive of the mean state (rho, u, T, p)
    mean_state = np.zeros_like(xs[0][0])
    mean_state[...] = ss.consMean
    pr = np.zeros((p.nprimvars,) + mean_state.shape[1:])
    orc.cons_to_prim(p, pr, mean_state)
    for c in range(6):
        a, b = valid(p, ss.primMean)[c], valid(p, pr)[c]
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), c


def test_two_sample_variance_closed_form(orc, cfg):
    p = cfg.params
    (x1, p1), (x2, p2) = samples(orc, cfg, 2, seed=5)
    ss = orc.StatsState(p)
    orc.evaluate_stats(p, x1, p1, ss, 1, 0)
    assert np.all(valid(p, ss.consVar)[:5] == 0.0)
    orc.evaluate_stats(p, x2, p2, ss, 2, 0)
    want = (valid(p, x2)[:5] - valid(p, x1)[:5]) ** 2 / 8.0
    got = valid(p, ss.consVar)[:5]
    for c in range(5):
        assert np.abs(got[c] - want[c]).max() <= 1e-12 * np.abs(want[c]).max(), c
    # primVar[0] is consVar[0]; the temperature variance uses the instantaneous prim T
    assert np.array_equal(valid(p, ss.primVar)[0], got[0])
    dT = valid(p, p2)[4] - valid(p, ss.primMean)[4]
    assert np.allclose(valid(p, ss.primVar)[4], dT * dT / 2.0, rtol=1e-12, atol=0)


def test_yz_constant_fields_and_one_sample_correlations(orc, cfg):
    """A state varying only in x: slice averages are the profile itself; after one sample the
    mean equals the instant, so spatialCross[3] (<T* T> - <T*><T>) vanishes and
    spatialCross[2] = T(x*) T(x)."""
    p = cfg.params
    st = orc.HostState(p)
    orc.init_cons(p, st.cu, cfg.rho0, cfg.rhobar, cfg.T_init)
    g = p.ngc
    n0 = p.n_cells[0]
    x = np.arange(n0)
    prof = 1.0 + 0.05 * np.sin(2 * np.pi * x / n0)
    st.cu[:, :, :, g:g + n0] *= prof[None, None, None, :]
    orc.fill_boundary(p, st.cu)
    orc.cons_to_prim(p, st.prim, st.cu)
    ss = orc.StatsState(p)
    xs = 3
    orc.evaluate_stats(p, st.cu, st.prim, ss, 1, xs)
    T = valid(p, st.prim)[4][0, 0, :]
    sc = valid(p, ss.spatialCross)
    assert np.allclose(sc[1][0, 0, :], T, rtol=1e-13, atol=0)           # yz-average of mean T
    assert np.allclose(sc[0], T[xs], rtol=1e-13, atol=0)                  # mean T at x*
    assert np.allclose(sc[2][0, 0, :], T[xs] * T, rtol=1e-13, atol=0)
    assert np.abs(sc[3]).max() <= 1e-12 * T.max() ** 2
    rho = valid(p, st.cu)[0][0, 0, :]
    assert np.isclose(ss.miscVals[3], rho[xs], rtol=1e-13)               # mean rho at x*
    assert np.isclose(ss.miscVals[4], rho[xs], rtol=1e-13)               # instant rho at x*


def test_rejects_bad_arguments(orc, cfg):
    p = cfg.params
    (cu, prim), = samples(orc, cfg, 1)
    ss = orc.StatsState(p)
    with pytest.raises(AssertionError):
        orc.evaluate_stats(p, cu, prim, ss, 0, 0)
    with pytest.raises(AssertionError):
        orc.evaluate_stats(p, cu, prim, ss, 1, p.n_cells[0])
