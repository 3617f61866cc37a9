"""The oracle (oracle/dmc_oracle.c) against golden fixtures written by the reference itself.

Fixtures come from tests/golden/make_golden.py, which runs the reference's own Fortran
modules (compiled by `make -C oracle ref`).  Integer quantities (k vectors, RNG states,
populations, branching multisets) must match exactly; floating-point ones to 1e-10 relative.
"""
import numpy as np
import pytest

from conftest import golden

SYS = {14: 3, 38: 4, 66: 5}


@pytest.mark.parametrize("n", [14, 38, 66])
def test_setup_tables(oracle_mod, n):
    g = golden(f"setup_N{n}.npz")
    s = oracle_mod.System(npart=n, norb=int(g["norb"]))
    assert s.el == pytest.approx(float(g["el"]), rel=1e-15)
    assert s.eunit == pytest.approx(float(g["eunit"]), rel=1e-15)
    assert s.etrial == pytest.approx(float(g["etrial"]), rel=1e-15)
    ak, ak2, numsh = s.kvec()
    assert ak.shape[0] == int(g["nk"])
    np.testing.assert_array_equal(numsh, g["numsh"])
    np.testing.assert_allclose(ak, g["ak"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(ak2, g["ak2"], rtol=1e-14, atol=1e-14)
    r = g["r"]
    rad = np.array([s.radial(x) for x in r])
    np.testing.assert_allclose(rad, g["radial"], rtol=1e-12, atol=1e-12)
    be = np.array([s.beta(x) for x in r])
    np.testing.assert_allclose(be, g["beta"], rtol=1e-12, atol=1e-12)
    db = np.array([s.derbeta(x) for x in r[1:]])   # derbeta divides by r
    np.testing.assert_allclose(db, g["derbeta"][1:], rtol=1e-9, atol=1e-12)
    vp = np.array([s.vpot(x) for x in r])
    np.testing.assert_allclose(vp, g["vpot"], rtol=1e-13, atol=0)


def test_rng_streams(oracle_mod):
    g = golden("setup_N14.npz")
    for j, seed in enumerate(g["seeds"]):
        s1 = s2 = int(seed)
        for i in range(g["ran1_state"].shape[1]):
            s1, u1 = oracle_mod.ran1(s1)
            s2, u2 = oracle_mod.ran2(s2)
            assert s1 == g["ran1_state"][j, i] and u1 == g["ran1_u"][j, i]
            assert s2 == g["ran2_state"][j, i] and u2 == g["ran2_u"][j, i]
        st, gg = oracle_mod.gaussian(int(seed) & ((1 << 48) - 1), 7)
        assert st == g["gauss7_state"][j]
        np.testing.assert_allclose(gg, g["gauss7"][j], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("tag,n,norb", [("N14", 14, 3), ("N38", 38, 4), ("N66", 66, 5),
                                        ("N66_norb8", 66, 8)])
def test_hpsi(oracle_mod, tag, n, norb):
    g = golden(f"hpsi_{tag}.npz")
    s = oracle_mod.System(npart=n, norb=norb)
    for k in range(len(g["x"])):
        x = g["x"][k]
        u, du, d2u, v = s.getphi(x)
        scale = abs(g["d2u"][k]) + abs(g["v"][k]) + 1.0
        assert abs(u - g["u"][k]) <= 1e-10 * (abs(g["u"][k]) + 1)
        assert abs(d2u - g["d2u"][k]) <= 1e-10 * scale
        assert abs(v - g["v"][k]) <= 1e-10 * scale
        np.testing.assert_allclose(du, g["du"][k], rtol=1e-10, atol=1e-10)
        o = s.hpsi(x)
        assert o["is_"] == int(g["is"][k])
        assert abs(o["psil"] - g["psil"][k]) <= 1e-10 * abs(g["psil"][k])
        el_ref = -0.5 * g["d2psi"][k] + g["v"][k]
        el = -0.5 * o["d2psi"] + o["v"]
        assert abs(el - el_ref) <= 1e-10 * (abs(el_ref) + abs(g["v"][k]) + 1)
        gs = np.abs(g["dpsi"][k]).max() + 1
        np.testing.assert_allclose(o["dpsi"], g["dpsi"][k], rtol=0, atol=1e-10 * gs)


@pytest.mark.parametrize("tag,n,norb", [("N14", 14, 3), ("N66", 66, 5)])
def test_single_particle_ratio(oracle_mod, tag, n, norb):
    g = golden(f"ratio_{tag}.npz")
    s = oracle_mod.System(npart=n, norb=norb)
    for k in range(len(g["x"])):
        for m in range(g["ip"].shape[1]):
            r = s.ratio1(g["x"][k], g["ip"][k, m] - 1, g["xnew"][k, m])
            ref = g["ratio"][k, m]
            assert abs(r - ref) <= 1e-10 * abs(ref) + 1e-300


@pytest.mark.parametrize("tag,n,norb", [("N14", 14, 3), ("N66_norb8", 66, 8)])
def test_getder(oracle_mod, tag, n, norb):
    g = golden(f"der_{tag}.npz")
    s = oracle_mod.System(npart=n, norb=norb)
    for k in range(len(g["x"])):
        d = s.getder(g["x"][k])
        np.testing.assert_allclose(d, g["dlogpsi"][k], rtol=1e-9,
                                   atol=1e-9 * np.abs(g["dlogpsi"][k]).max())


def _sorted(irn, *arrs):
    o = np.argsort(irn, kind="stable")
    return (irn[o],) + tuple(a[o] for a in arrs)


def _run_reference_trajectory(oracle_mod, g, nthreads=0):
    n = int(g["npart"])
    s = oracle_mod.System(npart=n, norb=int(g["norb"]), nwalk=int(g["nwalk"]),
                          dt=float(g["dt"]))
    x, irn = s.init_walkers(int(g["seed"]), int(g["nwalk"]))
    dpsi, sc = s.init_hpsi(x)
    return s, x, dpsi, sc, irn


@pytest.mark.parametrize("tag", ["N14_short", "N38_short", "N66_short"])
def test_dmc_trajectory_multisets(oracle_mod, tag):
    """Walker multisets after step1 + branch + population control, keyed by RNG state."""
    g = golden(f"dmc_{tag}.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    dumps = set(int(d) for d in g["dumps"])
    nsteps = len(g["pop"])
    nstep = int(g["nstep"])

    def check(step):
        ri, rs, rx = _sorted(g[f"d{step}_irn"], g[f"d{step}_sc"], g[f"d{step}_x"])
        oi, os_, ox = _sorted(irn, sc, x)
        np.testing.assert_array_equal(oi, ri)
        np.testing.assert_array_equal(os_[:, 0], rs[:, 0])          # weights (+-1)
        np.testing.assert_array_equal(os_[:, 4], rs[:, 4])          # signs
        np.testing.assert_allclose(os_[:, 1], rs[:, 1], rtol=1e-10)  # psil
        np.testing.assert_allclose(ox, rx, rtol=0, atol=1e-10)

    if 0 in dumps:
        check(0)
    est = np.zeros(42)
    for j in range(1, nsteps + 1):
        x, dpsi, sc, irn, nt, est = s.dmc_step(x, dpsi, sc, irn, opcomp=(j % nstep == 0),
                                                est=est)
        assert nt == tuple(int(v) for v in g["pop"][j - 1])
        if j in dumps:
            check(j)


def test_dmc_cfg1_population_trace(oracle_mod):
    """N=14, 1000 walkers, 500 steps: walker counts after branching and after population
    control must match the reference bit for bit at every step, and the final (irn, weight)
    multiset must be identical; block energies agree to 1e-8 relative."""
    g = golden("dmc_N14_cfg1.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    nstep = int(g["nstep"])
    est = np.zeros(42)
    blocks = []
    for j in range(1, len(g["pop"]) + 1):
        x, dpsi, sc, irn, nt, est = s.dmc_step(x, dpsi, sc, irn, opcomp=(j % nstep == 0),
                                                est=est)
        assert nt == tuple(int(v) for v in g["pop"][j - 1]), f"step {j}"
        if j % nstep == 0:
            e = est.reshape(14, 3)
            blocks.append(e[:, 0] / e[:, 2])
            est = np.zeros(42)
    np.testing.assert_array_equal(np.sort(irn), np.sort(g["d500_irn"]))
    ref = g["est"][:, :, 0]
    got = np.array(blocks)
    for i in (0, 1, 2, 5):  # total, kinetic, potential, growth energy
        np.testing.assert_allclose(got[:, i], ref[:, i], rtol=1e-8)


@pytest.mark.parametrize("tag", ["vmc_N14", "vmc_N66"])
def test_vmc_trajectory(oracle_mod, tag):
    """Variational runs (step1 with idmc = 0): walkers after every dumped step and the six
    block estimators against the reference."""
    g = golden(f"dmc_{tag}.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    dumps = set(int(d) for d in g["dumps"])
    nstep = int(g["nstep"])
    est = np.zeros(42)
    for j in range(1, len(g["pop"]) + 1):
        x, dpsi, sc, irn, est = s.vmc_step1(x, dpsi, sc, irn, est=est)
        assert len(x) == int(g["pop"][j - 1][1])
        if j in dumps:
            ri, rs, rx = _sorted(g[f"d{j}_irn"], g[f"d{j}_sc"], g[f"d{j}_x"])
            oi, os_, ox = _sorted(irn, sc, x)
            np.testing.assert_array_equal(oi, ri)
            np.testing.assert_array_equal(os_[:, 4], rs[:, 4])
            np.testing.assert_allclose(os_[:, 1], rs[:, 1], rtol=1e-10)
            np.testing.assert_allclose(ox, rx, rtol=0, atol=1e-10)
        if j % nstep == 0:
            e = est.reshape(14, 3)
            ref = g["est"][j // nstep - 1]
            for i in range(6):
                np.testing.assert_allclose(e[i, 0] / e[i, 2], ref[i, 0], rtol=1e-9, atol=1e-30)
            est = np.zeros(42)


@pytest.mark.parametrize("tag,n,norb", [("N14", 14, 3), ("N66", 66, 5)])
def test_optimizer_update(oracle_mod, tag, n, norb):
    """calder + updatepar on configurations: mean energy and new parameters."""
    g = golden(f"opt_{tag}.npz")
    s = oracle_mod.System(npart=n, norb=norb)
    ene, par = oracle_mod.calder_updatepar(s, g["x"], float(g["delta"]), float(g["eps"]))
    assert ene == pytest.approx(float(g["ene"]), rel=1e-10)
    np.testing.assert_allclose(par, g["par"], rtol=1e-7, atol=1e-12)


def test_optimisation_run(oracle_mod):
    """A whole optimisation iteration (idmc = -1): VMC blocks, calder over the walkers after
    each block, updatepar: the reference's mean energy and new parameters."""
    g = golden("dmc_optrun_N14.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    confs = []
    for b in range(int(g["nav"])):
        for j in range(int(g["nstep"])):
            x, dpsi, sc, irn, _ = s.vmc_step1(x, dpsi, sc, irn)
        confs.append(x.copy())
    ene, par = oracle_mod.calder_updatepar(s, np.concatenate(confs), 0.7, 1e-4)
    assert ene == pytest.approx(float(g["opt_ene"]), rel=1e-10)
    np.testing.assert_allclose(par, g["opt_par"], rtol=1e-9)


@pytest.mark.parametrize("tag", ["mdmc_N14_short", "mdmc_N66_short"])
def test_metropolis_dmc_trajectory(oracle_mod, tag):
    """Metropolis DMC (step1 with idmc = 2, step.f90:161-226): walker multisets after every
    dumped step, the population trace, and the 13 block estimators of dmcbcs.f90:204-218
    against the reference's run."""
    g = golden(f"dmc_{tag}.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    dumps = set(int(d) for d in g["dumps"])
    nstep = int(g["nstep"])

    def check(step):
        ri, rs, rx = _sorted(g[f"d{step}_irn"], g[f"d{step}_sc"], g[f"d{step}_x"])
        oi, os_, ox = _sorted(irn, sc, x)
        np.testing.assert_array_equal(oi, ri)
        np.testing.assert_array_equal(os_[:, 0], rs[:, 0])
        np.testing.assert_array_equal(os_[:, 4], rs[:, 4])
        np.testing.assert_allclose(os_[:, 1], rs[:, 1], rtol=1e-10)
        np.testing.assert_allclose(ox, rx, rtol=0, atol=1e-10)

    if 0 in dumps:
        check(0)
    est = np.zeros(42)
    blocks = []
    for j in range(1, len(g["pop"]) + 1):
        x, dpsi, sc, irn, nt, est = s.dmc_step(x, dpsi, sc, irn, opcomp=(j % nstep == 0),
                                                est=est, metropolis=True)
        assert nt == tuple(int(v) for v in g["pop"][j - 1]), f"step {j}"
        if j in dumps:
            check(j)
        if j % nstep == 0:
            e = est.reshape(14, 3)
            blocks.append(e[:13, 0] / e[:13, 2])
            est = np.zeros(42)
    np.testing.assert_allclose(np.array(blocks), g["est"][:, :13, 0], rtol=1e-8, atol=1e-300)


def test_metropolis_dmc_stack_overflow(oracle_mod):
    """At dt = 0.002 the reference's Metropolis DMC overflows its stacks (capacity 10 nwalk,
    stack.f90:88-91) in step 6; the oracle's population trace matches up to there and its
    sixth step overflows the same capacity."""
    g = golden("mdmc_N14_overflow.npz")
    s, x, dpsi, sc, irn = _run_reference_trajectory(oracle_mod, g)
    cap = int(g["capacity"])
    for j, ref in enumerate(g["pop"], 1):
        x, dpsi, sc, irn, nt, _ = s.dmc_step(x, dpsi, sc, irn, cap=cap, metropolis=True)
        assert nt == tuple(int(v) for v in ref), f"step {j}"
    with pytest.raises(RuntimeError):
        s.dmc_step(x, dpsi, sc, irn, cap=cap, metropolis=True)
