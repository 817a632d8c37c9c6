"""The reference's run statistics (block intervals, per-component enthalpy of adsorption, Widom chemical potential):
the oracle pinned by closed forms and known distributions, and the library's host implementation (C ABI, no device
work) against the oracle."""
import numpy as np
import pytest
from scipy import stats

import raspa3_amd as ra
from oracle import statistics as so

KT300 = 300.0 * so.KELVIN_TO_ENERGY


def test_t_deviates_are_student_t_quantiles():
    for df in range(1, 21):
        exact = stats.t.ppf(0.975, df)
        # the reference's table is printed to 4 significant digits
        assert abs(so.T95[df] - exact) <= 5e-4 * exact + 1e-12, (df, so.T95[df], exact)


def test_block_interval_covers_95_percent():
    rng = np.random.default_rng(3)
    hits = 0
    trials = 4000
    for _ in range(trials):
        b = rng.normal(1.7, 0.3, 5)
        hits += abs(b.mean() - 1.7) <= so.block_error(b, b.mean())
    assert abs(hits / trials - 0.95) < 0.012


def test_half_filled_blocks_are_dropped():
    mean, ci = so.block_average([10.0, 10.2, 9.8, 1.0], [10.0, 10.0, 10.0, 4.0])
    assert mean == pytest.approx(31.0 / 34.0)
    assert ci == pytest.approx(so.block_error([1.0, 1.02, 0.98], mean))


def linear_model_blocks(a, b, c, nb=5, per=4000, seed=0, noise=0.0):
    rng = np.random.default_rng(seed)
    w = np.full(nb, float(per))
    su, sn, sun, snn = np.zeros(nb), np.zeros((nb, 2)), np.zeros((nb, 2)), np.zeros((nb, 2, 2))
    for k in range(nb):
        n1 = rng.poisson(6.0, per).astype(float)
        n2 = rng.poisson(2.0, per) + (rng.random(per) < 0.3) * n1 // 3
        n = np.stack([n1, n2], 1)
        u = a * n1 + b * n2 + c + noise * rng.normal(size=per)
        su[k], sn[k], sun[k] = u.sum(), n.sum(0), (u[:, None] * n).sum(0)
        snn[k] = n.T @ n
    return w, su, sn, sun, snn


def test_enthalpy_is_regression_coefficient_minus_kT():
    a, b = -3100.0, -1250.0
    w, su, sn, sun, snn = linear_model_blocks(a, b, 17.0)
    mean, ci = so.enthalpy_of_adsorption(w, su, sn, sun, snn, 300.0)
    assert np.allclose(mean, [a - KT300, b - KT300], rtol=1e-9)
    assert np.all(ci < 1e-6 * abs(a))


def test_enthalpy_one_component_reduces_to_scalar_formula():
    rng = np.random.default_rng(5)
    nb = 5
    w = rng.uniform(50, 60, nb)
    n = rng.uniform(2, 4, nb) * w
    nn = (n / w) ** 2 * w + rng.uniform(0.5, 1.5, nb) * w
    u = rng.uniform(-900, -800, nb) * w
    un = u / w * n / w * w + rng.uniform(-300, -200, nb) * w
    mean, _ = so.enthalpy_of_adsorption(w, u, n[:, None], un[:, None], nn[:, None, None], 250.0)
    scalar = [((un[k] / w[k]) - (u[k] / w[k]) * (n[k] / w[k])) / ((nn[k] / w[k]) - (n[k] / w[k]) ** 2)
              - 250.0 * so.KELVIN_TO_ENERGY for k in range(nb)]
    assert mean[0] == pytest.approx(np.mean(scalar), rel=1e-12)


def test_widom_closed_forms():
    w = np.full(5, 100.0)
    st = so.widom_statistics(w, 1.0 * w, 0.002 * w, 300.0)
    assert st["mu_excess"] == 0.0 and st["mu_excess_ci"] == 0.0
    # with W = 1 the fugacity is the ideal-gas pressure rho kT
    assert st["fugacity"] == pytest.approx(0.002 * KT300, rel=1e-12)
    st = so.widom_statistics(w, 0.25 * w, 0.002 * w, 300.0)
    assert st["mu_excess"] == pytest.approx(-KT300 * np.log(0.25), rel=1e-12)


def random_terms(seed, nc=2, nb=5):
    rng = np.random.default_rng(seed)
    w = rng.uniform(80, 120, nb)
    n = rng.uniform(1, 8, (nb, nc))
    su = rng.uniform(-2e3, -1e3, nb) * w
    sn = n * w[:, None]
    sun = (su / w)[:, None] * n * w[:, None] * rng.uniform(0.97, 1.03, (nb, nc))
    snn = np.einsum("bi,bj->bij", n, n) * w[:, None, None]
    snn += np.eye(nc)[None] * rng.uniform(0.5, 2.0, (nb, 1, 1)) * w[:, None, None]
    return w, su, sn, sun, snn


def test_library_statistics_match_oracle():
    rng = np.random.default_rng(9)
    b = rng.normal(0.3, 0.05, 5)
    assert ra.block_error(b, b.mean()) == pytest.approx(so.block_error(b, b.mean()), rel=1e-13)
    sums, weights = rng.uniform(1, 2, 5) * 40, np.array([40.0, 41, 39, 40, 12])
    assert np.allclose(ra.block_average(sums, weights), so.block_average(sums, weights), rtol=1e-13)
    for nc in (1, 2, 3):
        terms = random_terms(nc, nc)
        m, c = ra.enthalpy_of_adsorption(*terms, 300.0)
        mo, co = so.enthalpy_of_adsorption(*terms, 300.0)
        assert np.allclose(m, mo, rtol=1e-11) and np.allclose(c, co, rtol=1e-9, atol=1e-9)
    w = rng.uniform(90, 110, 5)
    sw, sr = rng.uniform(0.1, 0.3, 5) * w, rng.uniform(1e-3, 2e-3, 5) * w
    lib_st, or_st = ra.widom_statistics(w, sw, sr, 310.0), so.widom_statistics(w, sw, sr, 310.0)
    for k in or_st:
        assert lib_st[k] == pytest.approx(or_st[k], rel=1e-11, abs=1e-14), k


def test_library_mc_statistics_of_a_state():
    w, su, sn, sun, snn = linear_model_blocks(-2000.0, -900.0, 3.0, per=500, seed=2, noise=40.0)
    state = dict(block_count=w, block_energy=su, block_n=sn, block_energy_n=sun, block_nn=snn)
    st = ra.mc_statistics(state, 2, 300.0)
    lo = [so.block_average(sn[:, c], w) for c in range(2)]
    assert np.allclose(st["loading"], [m for m, _ in lo], rtol=1e-13)
    mo, co = so.enthalpy_of_adsorption(w, su, sn, sun, snn, 300.0)
    assert np.allclose(st["enthalpy"], mo, rtol=1e-11) and np.allclose(st["enthalpy_ci"], co, rtol=1e-8)
    # the noisy linear model: the interval covers the exact coefficients
    assert np.all(np.abs(st["enthalpy"] - (np.array([-2000.0, -900.0]) - KT300)) < 3 * st["enthalpy_ci"] + 1e-9)
