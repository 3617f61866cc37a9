"""Invariants that pin the oracle where the reference has no counterpart.

* The single-particle-move sweep with Sherman-Morrison updates (no reference trajectory)
  is pinned by re-inversion: refreshing the inverse before every move (k = 1) must give the
  same sweep as pure Sherman-Morrison updates (k = npart).
* Gradient and Laplacian of log psi against central finite differences (the reference's own
  chkder check, wavefunction.f90:289-349).
* Single-particle ratios against full recomputation (chkone, wavefunction.f90:351-391).
* The Lagrange interpolation is exact for cubics; the LCG equals big-integer arithmetic.
"""
import numpy as np
import pytest


def test_lcg_closed_form(oracle_mod):
    s = 12345
    for _ in range(100):
        ns, u = oracle_mod.ran1(s)
        assert ns == (s * 44485709377909 + 96309754297) % (1 << 48)
        assert u == ((ns | 1) / float(1 << 48))
        s = ns


@pytest.mark.parametrize("n,norb", [(14, 3), (38, 4)])
def test_hpsi_finite_differences(oracle_mod, n, norb):
    # the tabulated u, u' and u'' are interpolated separately, so finite differences agree
    # only to the interpolation error; a fine table makes that negligible
    s = oracle_mod.System(npart=n, norb=norb, ntab=16384)
    rng = np.random.default_rng(n)
    x = s.el * (rng.random((n, 3)) - 0.5)
    o = s.hpsi(x)
    h = 1e-5
    lap = 0.0
    psi0 = o["psil"]
    for i in range(n):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            op, om = s.hpsi(xp), s.hpsi(xm)
            g = (op["psil"] - om["psil"]) / (2 * h)
            assert abs(g - o["dpsi"][i, d]) < 1e-5 * (1 + abs(g))
            lap += (op["psil"] - 2 * psi0 + om["psil"]) / h**2
    # laplacian(psi)/psi = lap(log psi) + |grad log psi|^2
    d2 = lap + np.sum(o["dpsi"] ** 2)
    assert abs(d2 - o["d2psi"]) < 1e-3 * (1 + abs(o["d2psi"]))


def test_ratio_full_recompute(oracle_mod):
    s = oracle_mod.System(npart=14, norb=3)
    rng = np.random.default_rng(3)
    x = s.el * (rng.random((14, 3)) - 0.5)
    base = s.hpsi(x)
    for ip in range(14):
        xn = x[ip] + rng.normal(0, 0.4, 3)
        x2 = x.copy()
        x2[ip] = xn
        new = s.hpsi(x2)
        ref = new["is_"] * base["is_"] * np.exp(new["psil"] - base["psil"])
        assert s.ratio1(x, ip, xn) == pytest.approx(ref, rel=1e-10)


@pytest.mark.parametrize("n,norb", [(14, 3), (38, 4)])
def test_sm_sweep_matches_reinversion(oracle_mod, n, norb):
    s = oracle_mod.System(npart=n, norb=norb, dt=0.002)
    x, irn = s.init_walkers(11, 24)
    dpsi, sc = s.init_hpsi(x)
    a = s.sm_step1(x, dpsi, sc, irn, krefresh=1)
    b = s.sm_step1(x, dpsi, sc, irn, krefresh=n)
    np.testing.assert_array_equal(a[5], b[5])       # accepted-move counts
    np.testing.assert_array_equal(a[3], b[3])       # RNG states
    np.testing.assert_allclose(a[0], b[0], atol=1e-10)
    np.testing.assert_allclose(a[2][:, 1:4], b[2][:, 1:4], rtol=1e-9)
    assert a[5].sum() > 0.5 * n * len(x)            # most moves accepted at this dt


def test_sm_sweep_draws_five_per_move(oracle_mod):
    s = oracle_mod.System(npart=14, norb=3)
    x, irn = s.init_walkers(5, 4)
    dpsi, sc = s.init_hpsi(x)
    out = s.sm_step1(x, dpsi, sc, irn, krefresh=14)
    for w in range(4):
        st = int(irn[w])
        for _ in range(5 * 14):
            st, _u = oracle_mod.ran1(st)
        assert st == out[3][w]
