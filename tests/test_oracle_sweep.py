"""The single-particle-move sweep of the oracle, pinned move by move against hpsi.

The reference has no single-particle DMC, so the oracle's sweep (dmco_sm_step1) is checked
against quantities that are pinned to the reference elsewhere: hpsi (ln|psi|, its sign and
grad ln psi; wavefunction.f90:161-201, golden fixtures at 1e-10) and the ran1/gaussian
streams (ran.f90, bit-exact fixtures).  For every move of every walker the sweep's record
(dmco_sm_step1_trace) must agree with an evaluation from scratch:

* the drift is sigma^2 grad_a ln psi(x) of a fresh hpsi, cut as step.f90:66-67 cuts it;
* the displacement is the drift plus sigma times the walker's next three gaussians, the
  acceptance draw is the fifth ran1 draw of the move;
* the ratio det * Jastrow equals psi(x')/psi(x) from two fresh hpsi (sign included), and the
  fixed-node test refuses exactly the moves that change the sign of psi;
* the log of the Metropolis ratio is the detailed-balance expression of step.f90:80-93
  restricted to the mover,
      2 (ln|psi(x')| - ln|psi(x)|) + ln G(x' -> x) - ln G(x -> x'),
      G(x -> x') = exp(-|x' - x - D(x)|^2 / (2 sigma^2)),  D = cut(sigma^2 grad_a ln psi),
  with D(x') from a fresh hpsi at the proposed configuration.

Planted mistakes in the oracle (a sign slip in the down-spin determinant gradient, a factor
slip in the Green's-function term, a dropped factor 2 on the Jastrow part of the acceptance)
must each fail these checks.  A statistical check closes the loop: with the weights off, the
sweep is a Metropolis chain of |psi|^2, so its averages (ln|psi|, the local energy, the pair
distance) must agree with the reference's own long variational run (step1 with idmc = 0,
all-particle moves, at N = 2: the fixture vmclong_N2 written from the reference by
tests/golden/make_golden.py).
"""
import numpy as np
import pytest

from conftest import golden

TOL = dict(pos=0.0, drift=1e-10, cut=1e-10, gauss=1e-14, draw=0.0, prop=1e-13, ratio=1e-10,
           node=0, logp=1e-9, decision=0, grad_new=1e-10, final=0.0)


def _wrap(x, el):
    return x - el * np.rint(x / el)


def sweep_errors(oracle_mod, s_sweep, s_ref, x, irn, krefresh):
    """largest deviation of each check over every move of one sweep of every walker (the
    sweep from s_sweep, the fresh evaluations from s_ref)"""
    nw, n = x.shape[0], s_ref.npart
    m = s_ref.params["mass"]
    dcut = s_ref.params["dcut"]
    dpsi, sc = s_ref.init_hpsi(x)
    xo, _, _, _, _, nacc, tr = s_sweep.sm_step1_trace(x, dpsi, sc, irn, krefresh)
    el, sig = s_ref.el, s_ref.sigma
    sig2 = sig * sig
    xs = _wrap(x, el)
    before = np.empty((nw, n, n, 3))
    after = np.empty((nw, n, n, 3))
    for ip in range(n):
        before[:, ip] = xs
        after[:, ip] = xs
        after[:, ip, ip] = tr[:, ip, 13:16]
        acc = tr[:, ip, 20] == 1.0
        xs[acc, ip] = tr[acc, ip, 13:16]
    db, sb = s_ref.init_hpsi(before.reshape(-1, n, 3))
    da, sa = s_ref.init_hpsi(after.reshape(-1, n, 3))
    db, da = db.reshape(nw, n, n, 3), da.reshape(nw, n, n, 3)
    sb, sa = sb.reshape(nw, n, 6), sa.reshape(nw, n, 6)
    err = {k: 0.0 for k in TOL}

    def worse(k, v):
        err[k] = max(err[k], float(v))

    def cut(d):
        return d if dcut < 0 else np.clip(d, -sig * dcut, sig * dcut)

    def rel(a, b):
        return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))

    for w in range(nw):
        st = int(irn[w])
        for ip in range(n):
            t = tr[w, ip]
            worse("pos", np.max(np.abs(t[1:4] - before[w, ip, ip])))
            g_old = db[w, ip, ip] * np.sqrt(m)  # grad_ip ln psi (hpsi's dpsi carries 1/sqrt m)
            worse("drift", rel(t[4:7], g_old))
            d_old = cut(sig2 * g_old)
            worse("cut", rel(t[7:10], d_old))
            st, g = oracle_mod.gaussian(st, 3)
            st, u = oracle_mod.ran1(st)
            worse("gauss", np.max(np.abs(t[10:13] - t[7:10] - sig * g)))
            worse("draw", abs(t[18] - u))
            worse("prop", np.max(np.abs(t[13:16] - _wrap(before[w, ip, ip] + t[10:13], el))))
            psi_b, is_b = sb[w, ip, 1], sb[w, ip, 4]
            psi_a, is_a = sa[w, ip, 1], sa[w, ip, 4]
            ratio = is_a * is_b * np.exp(psi_a - psi_b)
            worse("ratio", abs(t[16] * t[17] - ratio) / abs(ratio))
            refused = np.isnan(t[19])
            worse("node", int(refused != (is_a != is_b)))
            if refused:
                worse("decision", int(t[20] != 0.0))
                continue
            g_new = da[w, ip, ip] * np.sqrt(m)
            worse("grad_new", rel(t[21:24], g_new))
            d_new = cut(sig2 * g_new)
            delta = t[10:13]
            logp = 2.0 * (psi_a - psi_b) + (np.sum((delta - d_old) ** 2) -
                                            np.sum((delta + d_new) ** 2)) / (2.0 * sig2)
            worse("logp", abs(t[19] - logp) / max(1.0, abs(logp)))
            p = np.exp(min(0.0, logp))
            if abs(u - p) > 1e-8:
                worse("decision", int(t[20] != float(u < p)))
    worse("final", np.max(np.abs(xo - xs)))
    assert nacc.sum() == int(np.sum(tr[:, :, 20]))
    return err, tr


CASES = [  # npart, norb, dt, dcut, krefresh
    (14, 3, 0.02, -1.0, 14),
    (14, 3, 0.05, 2.0, 5),
    (66, 5, 0.001, 2.0, 66),
    (66, 5, 0.01, -1.0, 16),
]


def _configs(s, nc, seed):
    rng = np.random.default_rng(seed)
    x = s.el * (rng.random((nc, s.npart, 3)) - 0.5)
    irn = rng.integers(1, 1 << 48, nc, dtype=np.int64)
    return x, irn


@pytest.mark.parametrize("n,norb,dt,dcut,k", CASES)
def test_sweep_matches_fresh_hpsi_move_by_move(oracle_mod, n, norb, dt, dcut, k):
    s = oracle_mod.System(npart=n, norb=norb, dt=dt, dcut=dcut)
    x, irn = _configs(s, 64, 1000 + n + k)
    err, tr = sweep_errors(oracle_mod, s, s, x, irn, k)
    bad = {key: (v, TOL[key]) for key, v in err.items() if v > TOL[key]}
    assert not bad, bad
    # the checks are not vacuous: both spins move, moves are refused by the node test or the
    # Metropolis test, and with a cut some drifts are cut
    acc = tr[:, :, 20]
    assert 0.5 < acc.mean() < 1.0
    assert np.isnan(tr[:, :, 19]).sum() > 0 or (acc == 0).sum() > 0
    if dcut >= 0:
        assert np.any(np.abs(tr[:, :, 7:10]) == s.sigma * dcut)


@pytest.mark.parametrize("plant,key", [(1, "drift"), (2, "logp"), (3, "logp")])
def test_planted_mistakes_fail_the_checks(oracle_mod, plant, key):
    """each planted mistake (DMCO_PLANT in dmc_oracle.c) is caught by the move-by-move
    checks, at N = 14 and N = 66"""
    for n, norb, dt, dcut, k in (CASES[1], CASES[2]):
        s = oracle_mod.System(npart=n, norb=norb, dt=dt, dcut=dcut)
        p = oracle_mod.System(npart=n, norb=norb, dt=dt, dcut=dcut,
                              lib_=oracle_mod.planted(plant))
        x, irn = _configs(s, 8, 2000 + n)
        err, _ = sweep_errors(oracle_mod, p, s, x, irn, k)
        assert err[key] > 1e3 * TOL[key], (plant, n, err)


def chain_statistics(oracle_mod, lib_=None, nwalk=512, neq=500, nsweep=2500, dcut=2.0):
    """The single-particle sweep with the weights off (a Metropolis chain of |psi|^2, fn off so
    that it may cross nodes as the variational step does) at N = 2, dt = 0.05: per-walker
    time averages of ln|psi|, of the local energy clipped to [-50, 50] (a bounded function of
    the configuration, so its expectation under |psi|^2 is finite and shared) and of the
    pair distance; returns {name: (mean, error of the mean)} over the independent chains"""
    s = oracle_mod.System(npart=2, norb=3, dt=0.05, dcut=dcut, fn=0, nwalk=nwalk, lib_=lib_)
    x, irn = s.init_walkers(11, nwalk)
    dpsi, sc = s.init_hpsi(x)
    st = {"psil": [], "elocal": [], "r12": []}
    for it in range(neq + nsweep):
        x, dpsi, sc, irn, _, _ = s.sm_step1(x, dpsi, sc, irn, krefresh=2)
        if it >= neq and it % 10 == 0:
            d = x[:, 0] - x[:, 1]
            d -= s.el * np.rint(d / s.el)
            st["psil"].append(sc[:, 1].copy())
            st["elocal"].append(np.clip(-s.params["hbar"] * sc[:, 2] + sc[:, 3], -50.0, 50.0))
            st["r12"].append(np.sqrt(np.sum(d * d, axis=1)))
    out = {}
    for k, v in st.items():
        pw = np.array(v).mean(axis=0)
        out[k] = (pw.mean(), pw.std(ddof=1) / np.sqrt(len(pw)))
    return out


def reference_statistics():
    """the same averages over the reference's variational run (fixture vmclong_N2: 512
    chains, steps 500-3000 every 50, dt = 0.05, drift cut at sigma)"""
    ref = golden("vmclong_N2.npz")
    out = {}
    for k in ("psil", "elocal", "r12"):
        a = ref[k].astype(np.float64)
        if k == "elocal":
            a = np.clip(a, -50.0, 50.0)
        pw = a.mean(axis=0)
        out[k] = (pw.mean(), pw.std(ddof=1) / np.sqrt(len(pw)))
    return out


def _z(a, b):
    return {k: (a[k][0] - b[k][0]) / np.hypot(a[k][1], b[k][1]) for k in a}


def test_sweep_samples_psi_squared(oracle_mod):
    """With the weights off the single-particle sweep is a Metropolis chain of |psi|^2; its
    averages must agree within 2 sigma with the reference's own variational run of the same
    deck (all-particle Metropolis moves, step.f90:78-126, another proposal: drift cut at 2
    sigma here, at sigma there).  A direct quadrature of |psi|^2 over the relative
    coordinate (64M uniform points weighted by |psi|^2) gives ln|psi| = -0.359(5),
    clipped E_L = 0.895(12), r_12 = 0.4715(9), in line with both."""
    z = _z(chain_statistics(oracle_mod), reference_statistics())
    assert all(abs(v) < 2.0 for v in z.values()), z


@pytest.mark.parametrize("plant", [2, 3])
def test_planted_acceptance_mistakes_bias_the_chain(oracle_mod, plant):
    """A wrong acceptance ratio samples another distribution.  (A wrong drift that the
    proposal and the acceptance share, planted mistake 1, still samples |psi|^2: the
    move-by-move drift check is the one that catches it.)"""
    z = _z(chain_statistics(oracle_mod, lib_=oracle_mod.planted(plant), nsweep=1000),
           reference_statistics())
    assert max(abs(v) for v in z.values()) > 5.0, z
