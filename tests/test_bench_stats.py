"""The statistics behind the bench's DMC energies (bench._series_mean): the integrated
autocorrelation time and the error of the mean of a correlated series, checked on AR(1)
series whose values are known in closed form, and the rule that leaves out steps whose
energy spikes."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def ar1(phi, n, seed, mean=0.4, sigma=1e-3):
    """x_t = phi x_{t-1} + e_t, stationary from the first step"""
    rng = np.random.default_rng(seed)
    e = rng.normal(0.0, sigma * np.sqrt(1.0 - phi * phi), n)
    x = np.empty(n)
    x[0] = rng.normal(0.0, sigma)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return mean + x


def test_tau_int_of_ar1():
    """tau_int = (1 + phi) / (2 (1 - phi)) for AR(1) (1/2 for independent steps)"""
    import bench
    for phi in (0.0, 0.5, 0.9, 0.97):
        want = 0.5 * (1.0 + phi) / (1.0 - phi)
        taus = [bench._series_mean(ar1(phi, 40000, s))[2] for s in range(4)]
        assert abs(np.mean(taus) - want) <= 0.15 * want + 0.05, (phi, np.mean(taus), want)


def test_error_of_the_mean_covers_the_spread():
    """the stated error against the spread of the means of independent series"""
    import bench
    phi, n = 0.9, 8000
    res = [bench._series_mean(ar1(phi, n, 100 + s)) for s in range(40)]
    means = np.array([r[0] for r in res])
    errs = np.array([r[1] for r in res])
    spread = means.std(ddof=1)
    assert 0.75 < np.mean(errs) / spread < 1.3, (np.mean(errs), spread)
    # the naive error of independent steps would be sqrt(2 tau_int) = 4.4 times too small
    naive = np.mean([ar1(phi, n, 100 + s).std(ddof=1) / np.sqrt(n) for s in range(5)])
    assert np.mean(errs) > 3.0 * naive


def test_spikes_are_left_out():
    """steps more than 10 robust standard deviations from the median are left out and
    counted; the others keep their mean"""
    import bench
    x = ar1(0.5, 5000, 7)
    clean = bench._series_mean(x)
    y = x.copy()
    y[[100, 2000, 4000]] = [-50.0, 3.0, np.nan]
    m, se, tau, nout = bench._series_mean(y)
    assert nout == 3
    keep = np.ones(len(x), bool)
    keep[[100, 2000, 4000]] = False
    assert abs(m - x[keep].mean()) < 1e-15
    assert abs(m - clean[0]) < 3 * clean[1]
