"""The statistics behind the bench's DMC energies (bench._series_mean, bench._replica_stats):
the integrated autocorrelation time and the error of the mean of a correlated series, checked
on AR(1) series whose values are known in closed form, the rule that leaves out steps whose
energy spikes, and the error of an energy from independent replicas."""
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


def test_replica_errors_closed_form():
    """the mean of the replica means, the spread error std / sqrt(R), the combined
    autocorrelation error sqrt(sum se^2) / R, and the larger of the two quoted"""
    import bench
    r = bench._replica_stats([(1.0, 1e-3, 2.0, 0), (2.0, 1e-3, 4.0, 1), (3.0, 1e-3, 2.0, 0),
                              (4.0, 1e-3, 4.0, 2)])
    assert r["e"] == 2.5
    assert abs(r["err_replicas"] - np.std([1.0, 2.0, 3.0, 4.0], ddof=1) / 2.0) < 1e-15
    assert abs(r["err_autocorr"] - 5e-4) < 1e-15  # sqrt(4 * 1e-6) / 4
    assert r["err"] == r["err_replicas"]
    assert r["tau_int"] == 3.0 and r["outlier_steps"] == 3
    r = bench._replica_stats([(0.4, 0.1, 1.0, 0)] * 4)
    assert r["err_replicas"] == 0.0 and abs(r["err"] - 0.05) < 1e-15


def test_replica_error_covers_a_slow_mode():
    """replicas whose energy carries an offset constant over the run (a mode slower than
    the run, which no autocorrelation window sees): the autocorrelation errors miss it, the
    replica spread does not, and the quoted error covers the spread of the combined means"""
    import bench
    rng = np.random.default_rng(11)
    slow, nrep, sets = 3e-4, 4, 60
    means, errs, errs_ac = [], [], []
    for k in range(sets):
        reps = [bench._series_mean(ar1(0.5, 2000, 1000 * k + j) + rng.normal(0.0, slow))
                for j in range(nrep)]
        r = bench._replica_stats(reps)
        means.append(r["e"])
        errs.append(r["err"])
        errs_ac.append(r["err_autocorr"])
    spread = np.std(means, ddof=1)
    assert np.mean(errs_ac) < 0.5 * spread  # the autocorrelation errors alone fall short
    assert 0.75 < np.mean(errs) / spread < 1.4, (np.mean(errs), spread)


def test_quadratic_extrapolation():
    """the least-squares quadratic returns E(0) of an exact quadratic with chi^2 = 0, through
    three points the closed form (8 E3 - 6 E2 + E1) / 3 with the same combination of the
    errors, and a chi^2 of the expected size for noisy data"""
    import bench
    taus = [1e-3, 5e-4, 2.5e-4, 1.25e-4]
    es = [0.4 - 50.0 * t + 2e4 * t * t for t in taus]
    r = bench._quadratic_extrapolation(taus, es, [1e-4] * 4)
    assert abs(r["e"] - 0.4) < 1e-12 and r["chi2"] < 1e-12 and r["dof"] == 1
    # three points at tau, tau/2, tau/4: E0 = (8 E3 - 6 E2 + E1) / 3, the error the same
    # combination of the errors
    s = [1e-4, 2e-4, 3e-4]
    r3 = bench._quadratic_extrapolation(taus[1:], es[1:], s)
    assert abs(r3["e"] - 0.4) < 1e-12
    want = np.sqrt((8 * s[2]) ** 2 + (6 * s[1]) ** 2 + s[0] ** 2) / 3
    assert abs(r3["err"] - want) < 1e-12 * want + 1e-15
    rng = np.random.default_rng(3)
    chi = [bench._quadratic_extrapolation(taus, [e + rng.normal(0, 1e-4) for e in es],
                                          [1e-4] * 4)["chi2"] for _ in range(400)]
    assert 0.8 < np.mean(chi) < 1.2  # chi^2 with one degree of freedom
