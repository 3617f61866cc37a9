"""Throughput of the DMC walker-propagation path on one MI355X.

Headline workload: the N=66 unitary Fermi gas (33 up + 33 down, cosh potential with
mu = 40, 5 k-shells, 1024-point tables), 32768 walkers from seed 7, tau = 0.001, each step a
sweep of 66 single-particle moves with Sherman-Morrison inverse updates and an FP64 LU
refresh every 66 moves, then FP64 local energy, weights, branching and population control.
The mixed-precision propagator (FP32 moves, FP64 refresh/energies/weights/sign test) is the
reported value; the plain FP64 propagator runs beside it for the ratio, and the per-step
local-energy error of the mixed path against FP64 is measured on the same configurations.

    python bench.py [--steps K] [--warmup W] [--workload mixed|fp64|allparticle|batch|observables]

Prints one JSON line.  A "step" is one DMC step of the whole resident population; value is
walker-steps per second (sum of the populations of the timed steps / wall time), with the
walkers resident in HBM before the timed region starts.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed FP32; one FP32 FMA per lane and
#                           clock without packing gives 78.6)
N, NORB, NWALK, DT, SEED = 66, 5, 32768, 0.001, 7
DCUT_SM = 2.0


def flops_per_walker_step(mode, n=N // 2, k=N, kq=17, part="step"):
    """Algorithmic flop counts per walker-step (C = 2K - 1 plane-wave columns).

    all-particle: F_AP = 10 n^2 C + 2 n^3 + 16 n^2 + 70 n^2 + 90 n^2 (622,908 at N=66);
    single-particle sweep: F_SM(k) = N(8nC + 4n^2 + 164n) [the moves]
                                     + (N/k)(2n^3 + 2n^2 C) [refresh]
                                     + N(2nC + 2n) + 16 n^2 [local energy]
    (1,528,956 at N=66, k=66); part="moves" gives the first term alone (1,219,680), the
    work of the move kernel."""
    c = 2 * kq - 1
    if mode == "allparticle":
        return 10 * n * n * c + 2 * n ** 3 + 16 * n * n + 70 * n * n + 90 * n * n
    nn = 2 * n
    moves = nn * (8 * n * c + 4 * n * n + 164 * n)
    if part == "moves":
        return moves
    return (moves + (nn / k) * (2 * n ** 3 + 2 * n * n * c) + nn * (2 * n * c + 2 * n)
            + 16 * n * n)


def make_engine(mode, nwalk=NWALK, krefresh=N, dt=DT, walkers=None, capacity=None, dcut=None,
                **kw):
    """the headline system with the mode's propagator; walkers = (x, irn) to load, else the
    reference's initial walkers from seed 7"""
    from dmcbs_project_amd import Engine, DMC_FP64, DMC_MIXED, DMC_ALLPARTICLE, DMC_SINGLEPARTICLE
    prec = DMC_MIXED if mode == "mixed" else DMC_FP64
    move = DMC_ALLPARTICLE if mode == "allparticle" else DMC_SINGLEPARTICLE
    # single-particle moves cut the drift at 2 sigma per component (the reference's dcut,
    # step.f90:32-37); the all-particle replay keeps the reference deck's dcut = -1
    if dcut is None:
        dcut = -1.0 if mode == "allparticle" else DCUT_SM
    e = Engine(npart=N, norb=NORB, nwalk=nwalk, dt=dt, capacity=capacity or 2 * nwalk,
               precision=prec, movetype=move, krefresh=krefresh, dcut=dcut, **kw)
    if walkers is None:
        e.init_walkers(SEED)
    else:
        e.push_walkers(*walkers)
    return e


def timed_run(mode, steps, warmup):
    e = make_engine(mode)
    for _ in range(warmup):
        e.step()
    e.sync()
    e.kernel_timing(True)
    e.kernel_times(reset=True)
    pops = []  # populations entering each timed step (the walkers its kernel processes)
    nw = e.numstack()
    t0 = time.perf_counter()
    for _ in range(steps):
        pops.append(nw)
        nw = e.step()[1]
    e.sync()
    wall = time.perf_counter() - t0
    kms, launches = e.kernel_times(reset=True)
    e.kernel_timing(False)
    ws = float(np.sum(pops))
    res = dict(mode=mode, walker_steps=ws, wall=wall, value=ws / wall,
               ms_per_step=1e3 * wall / steps, kernel_ms=kms / max(launches, 1),
               launches=launches, pop_mean=float(np.mean(pops)))
    acc = e.acceptance()
    if acc[1]:
        res["acceptance"] = acc[0] / acc[1]
    e.close()
    return res


def _wrap(x, el):
    return x - el * np.rint(x / el)


def _pct(v):
    v = np.asarray(v, np.float64)
    if v.size == 0:
        return None
    return dict(median=float(np.median(v)), p90=float(np.percentile(v, 90)),
                p99=float(np.percentile(v, 99)), max=float(v.max()),
                frac_below_1e6=float(np.mean(v <= 1e-6)), n=int(v.size))


def _replay(x0, rec, el):
    """configurations before each move and after the sweep, from the starting positions and
    the recorded proposals and decisions"""
    xs = _wrap(np.array(x0, np.float64), el)
    before = []
    for m in range(rec.shape[1]):
        before.append(xs.copy())
        acc = rec[:, m, 2] == 1.0
        xs[acc, m] = rec[acc, m, 4:7]
    return before, xs


def precision_study(nconf=4096, ks=(N, 33, 16), nequil=200):
    """Error of the mixed-precision sweep against FP64 on the same configurations.

    nconf walkers from seed 7 are equilibrated by nequil FP64 steps.  From that population
    the mixed and the FP64 sweeps each make one step with the same random streams, every
    move recorded (dmc_move_trace).  Reported per refresh interval k:
    * ratio: for every mixed move, its psi ratio (FP32 pair values and FP32 inverse carried
      through the Sherman-Morrison updates, FP64 sum) against the FP64 move1 (dmc_ratio1)
      on the same configuration and proposal;
    * decisions: the moves whose accept/reject differs from the FP64 sweep's, and the walkers
      with any such flip (their trajectories part);
    * E_L: the FP64 local energy after the sweep, mixed against FP64, over the walkers
      whose decisions all agree (their positions differ only by the FP32 rounding of the
      displacements) and over all walkers (trajectory divergence included).
    fixed_config: the FP64 hpsi that gives the mixed path's E_L against the oracle on the
    same configurations (the E_L error with the configuration held fixed)."""
    from dmcbs_project_amd import DMC_FP64, DMC_MIXED
    eq = make_engine("fp64", nwalk=nconf, capacity=8 * nconf)
    for _ in range(nequil):
        eq.step()
    w0 = eq.walkers()
    x0, irn0 = w0["x"], w0["irn"]
    el = eq.el
    out = dict(configurations=int(len(irn0)), equilibration=f"{nequil} FP64 steps")
    try:
        import oracle
        s = oracle.System(npart=N, norb=NORB, nwalk=NWALK, dt=DT, dcut=DCUT_SM)
        sel = np.arange(0, len(irn0), max(1, len(irn0) // 64))[:64]
        h = eq.hpsi(x0[sel])
        eg = -0.5 * h["d2psi"] + h["v"]
        eo = np.array([(lambda o: -0.5 * o["d2psi"] + o["v"])(s.hpsi(x0[i])) for i in sel])
        out["fixed_config_elocal_rel_err_max"] = float(np.max(np.abs(eg - eo) / np.abs(eo)))
    except Exception as ex:  # the oracle is only the checker here
        out["fixed_config_elocal_rel_err_max"] = str(ex)
    for k in ks:
        rec, final, el_ = {}, {}, {}
        for prec in (DMC_FP64, DMC_MIXED):
            e = make_engine("mixed" if prec == DMC_MIXED else "fp64", nwalk=len(irn0),
                            capacity=8 * len(irn0), krefresh=k, walkers=(x0, irn0),
                            popup=1e9, popdn=0.0)
            e.move_trace(True)
            e.step1()
            rec[prec] = e.read_move_trace()
            e.branch(1.0)
            e.close()
        before, final32 = _replay(x0, rec[DMC_MIXED], el)
        _, final64 = _replay(x0, rec[DMC_FP64], el)
        r = rec[DMC_MIXED]
        ratio32 = r[:, :, 0] * np.exp(r[:, :, 1])
        ratio64 = np.stack([eq.move1(before[m], m, r[:, m, 4:7]) for m in range(N)], axis=1)
        rerr = np.abs(ratio32 - ratio64) / np.abs(ratio64)
        flips = rec[DMC_MIXED][:, :, 2] != rec[DMC_FP64][:, :, 2]
        same = ~flips.any(axis=1)
        for tag, xf in (("mixed", final32), ("fp64", final64)):
            h = eq.hpsi(xf)
            el_[tag] = -0.5 * h["d2psi"] + h["v"]
        elerr = np.abs(el_["mixed"] - el_["fp64"]) / np.abs(el_["fp64"])
        out[str(k)] = dict(
            ratio_rel_err=_pct(rerr), node_sign_flips=int(np.sum(np.sign(ratio32) !=
                                                                 np.sign(ratio64))),
            moves=int(rerr.size), moves_flipped=int(flips.sum()),
            walkers_diverged_frac=float(1.0 - same.mean()),
            elocal_rel_err_same_path=_pct(elerr[same]),
            elocal_rel_err_all=_pct(elerr))
    eq.close()
    return out


def _energy_series(e, nequil, nmeas):
    """the population's energy estimator E_L / (N E_FG) (estimator 1, step.f90:137-156) of
    every step after nequil steps"""
    for _ in range(nequil):
        e.step()
    e.estimators(reset=True)
    x = np.empty(nmeas)
    for i in range(nmeas):
        e.step(opcomp=True)
        est = e.estimators(reset=True)
        x[i] = est[0, 0] / est[0, 2]
    return x


def _series_mean(x):
    """mean and standard error of a series of per-step energies.

    Steps whose energy lies more than 10 robust standard deviations (1.4826 times the median
    absolute deviation) from the median are left out: in such a step one walker whose local
    energy spikes (a pair in the narrow potential well, or next to a node) carries an uncapped
    estimator weight exp(-(E_L - E_T) dt) (step.f90:137-156) and decides the step alone.
    Successive steps of one population are correlated, so the error is sigma sqrt(2 tau_int
    / n) with the integrated autocorrelation time tau_int summed up to the first lag
    M >= 5 tau_int (the automatic window of Madras and Sokal, J. Stat. Phys. 50 (1988) 109).
    Returns (mean, error, tau_int in steps, steps left out)."""
    fin = x[np.isfinite(x)]  # a step whose estimator weights sum to zero gives 0/0
    med = float(np.median(fin))
    mad = 1.4826 * float(np.median(np.abs(fin - med)))
    keep = np.isfinite(x) & (np.abs(x - med) <= 10.0 * max(mad, 1e-12))
    y = x[keep]
    n = len(y)
    d = y - y.mean()
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf = acf / acf[0] if acf[0] > 0 else np.zeros(n)
    tau = 0.5
    for m in range(1, n):
        tau += float(acf[m])
        if m >= 5.0 * tau:
            break
    se = float(y.std(ddof=1) * np.sqrt(2.0 * tau / n))
    return float(y.mean()), se, tau, int((~keep).sum())


def _replica_stats(reps):
    """one energy from R independent replicas, each a list (mean, error, tau_int, steps left
    out) from _series_mean: the mean of the replica means (equal populations), and two
    errors, the spread of the replica means, std / sqrt(R), which needs no model of the
    correlation in time, and the replicas' autocorrelation errors combined, sqrt(sum se^2) / R.
    The quoted error is the larger of the two."""
    m = np.array([r[0] for r in reps], float)
    se = np.array([r[1] for r in reps], float)
    nrep = len(reps)
    err_rep = float(m.std(ddof=1) / np.sqrt(nrep)) if nrep > 1 else float("nan")
    err_ac = float(np.sqrt(np.sum(se ** 2)) / nrep)
    err = max(err_rep, err_ac) if np.isfinite(err_rep) else err_ac
    return dict(e=float(m.mean()), err=err, err_replicas=err_rep, err_autocorr=err_ac,
                replica_means=[float(v) for v in m],
                tau_int=float(np.mean([r[2] for r in reps])),
                outlier_steps=int(sum(r[3] for r in reps)))


def _quadratic_extrapolation(taus, es, errs):
    """E(0) of the weighted least-squares fit E(tau) = E0 + a tau + b tau^2, its error from
    the fit's covariance, and chi^2 with its degrees of freedom"""
    t = np.asarray(taus, float)
    y = np.asarray(es, float)
    w = 1.0 / np.asarray(errs, float) ** 2
    A = np.stack([np.ones_like(t), t, t * t], axis=1)
    cov = np.linalg.inv(A.T @ (A * w[:, None]))
    coef = cov @ (A.T @ (w * y))
    chi2 = float(np.sum(w * (A @ coef - y) ** 2))
    return dict(e=float(coef[0]), err=float(np.sqrt(cov[0, 0])), chi2=chi2, dof=len(t) - 3)


def energy_study(nequil=6000, nsm=4000, nfp64=1000, nap=3000, nrep=4):
    """DMC energies of the headline system from an equilibrated population.

    One population (32768 walkers) is equilibrated by nequil mixed-precision steps at tau
    from the reference's initial walkers (copies of one configuration; their energy settles
    after about 4.5 units of imaginary time), then split into nrep groups, each run as an
    independent population (its own trial energy and population control) of 32768 / nrep
    walkers for every propagator below, so that each energy has nrep independent replicas
    (_replica_stats):
    * the mixed and the FP64 single-particle sweeps at tau (same walkers and random streams;
      the FP64 run is nfp64 steps, compared replica by replica with the mixed run's first
      nfp64 steps), and the mixed sweep at tau/2 (nsm steps each, tau/2 after 1000 steps of
      re-equilibration);
      The single-particle energy at tau = 0 comes from the line through tau and tau/2.
    * the reference's own algorithm, the all-particle step (the device replay that reproduces
      the reference's trajectories), at tau, tau/2, tau/4 and tau/8: nap f steps after 1000 f
      steps (one unit of imaginary time) of re-equilibration (f = 1, 2, 4, 8).  Its time-step
      error is large and not linear at these steps: E(0) comes from the quadratic through the
      three smallest steps, (8 E(tau/8) - 6 E(tau/4) + E(tau/2)) / 3, and, as a check of that
      model, from the weighted least-squares quadratic through all four (with its chi^2).
    Every propagator cuts the drift at 2 sigma (the reference's dcut, step.f90:32-37): with
    the deck's dcut = -1 the all-particle step from an equilibrated population lets walkers
    next to a node overshoot into regions of very negative local energy whose uncapped
    weights swamp the estimator (E = -132 +- 132 in one such run).  Each replica's energy
    comes from its per-step energies (_series_mean)."""
    e = make_engine("mixed")
    for _ in range(nequil):
        e.step()
    w = e.walkers()
    e.close()
    per = len(w["irn"]) // nrep
    groups = [(w["x"][r * per:(r + 1) * per], w["irn"][r * per:(r + 1) * per])
              for r in range(nrep)]
    series, runs = {}, {}
    for label, mode, dt, neq, nmeas in (("mixed", "mixed", DT, 0, nsm),
                                         ("fp64", "fp64", DT, 0, nfp64),
                                         ("mixed_half_tau", "mixed", DT / 2, 1000, nsm),
                                         ("allparticle", "allparticle", DT, 1000, nap),
                                         ("allparticle_half_tau", "allparticle", DT / 2, 2000,
                                          2 * nap),
                                         ("allparticle_quarter_tau", "allparticle", DT / 4, 4000,
                                          4 * nap),
                                         ("allparticle_eighth_tau", "allparticle", DT / 8, 8000,
                                          8 * nap)):
        series[label] = []
        for r, g in enumerate(groups):
            e = make_engine(mode, nwalk=per, dt=dt, walkers=g, dcut=DCUT_SM)
            series[label].append(_energy_series(e, neq, nmeas))
            e.close()
            print(f"bench: energy {label} replica {r}: {_series_mean(series[label][-1])[0]:.5f}",
                  file=sys.stderr, flush=True)
        runs[label] = dict(_replica_stats([_series_mean(x) for x in series[label]]),
                           steps=nmeas)
        print(f"bench: energy {label}: {runs[label]}", file=sys.stderr, flush=True)

    def diff(a, b):
        d = a[0] - b[0]
        sig = float(np.hypot(a[1], b[1]))
        return dict(diff=d, sigma=sig, diff_over_sigma=d / sig if sig > 0 else None)

    def ev(label):
        return runs[label]["e"], runs[label]["err"]

    # mixed against FP64 replica by replica: the two start from the same walkers and streams
    pairs = [_series_mean(a[:nfp64])[0] - _series_mean(b)[0]
             for a, b in zip(series["mixed"], series["fp64"])]
    dpair = float(np.mean(pairs))
    spair = float(np.std(pairs, ddof=1) / np.sqrt(nrep))
    e2, e4, e8 = (ev("allparticle_half_tau"), ev("allparticle_quarter_tau"),
                  ev("allparticle_eighth_tau"))
    ap0 = ((8 * e8[0] - 6 * e4[0] + e2[0]) / 3,
           float(np.sqrt((8 * e8[1]) ** 2 + (6 * e4[1]) ** 2 + e2[1] ** 2) / 3))
    ap0_lin = (2 * e8[0] - e4[0], float(np.hypot(2 * e8[1], e4[1])))
    ap_labels = ("allparticle", "allparticle_half_tau", "allparticle_quarter_tau",
                 "allparticle_eighth_tau")
    lsq = _quadratic_extrapolation([DT, DT / 2, DT / 4, DT / 8],
                                   [ev(k)[0] for k in ap_labels], [ev(k)[1] for k in ap_labels])
    sm = ev("mixed_half_tau")
    # the single-particle sweep's own time-step error, linear at these steps
    sp0 = (2 * sm[0] - ev("mixed")[0], float(np.hypot(2 * sm[1], ev("mixed")[1])))
    return dict(unit="E_L / (N E_FG)", tau=DT,
                start=f"{nequil} mixed steps from the reference's initial walkers, then "
                      f"{nrep} independent populations of {per} walkers",
                error=f"{nrep} replicas per energy: the larger of the spread of the replica "
                      "means and their autocorrelation errors (per-step energies, integrated "
                      "autocorrelation time, automatic window; steps beyond 10 robust sigma "
                      "left out)",
                runs=runs,
                mixed_vs_fp64=dict(diff=dpair, sigma=spair,
                                   diff_over_sigma=dpair / spair if spair > 0 else None,
                                   steps=nfp64, paired="replica by replica"),
                singleparticle_tau_vs_half_tau=diff(ev("mixed"), sm),
                allparticle_extrapolated=dict(e=ap0[0], err=ap0[1],
                                              fit="quadratic through tau/2, tau/4, tau/8"),
                allparticle_extrapolated_lsq=dict(lsq, fit="weighted least-squares quadratic "
                                                           "through tau .. tau/8"),
                allparticle_extrapolated_linear=dict(e=ap0_lin[0], err=ap0_lin[1],
                                                     fit="tau/4, tau/8"),
                singleparticle_extrapolated=dict(e=sp0[0], err=sp0[1], fit="tau, tau/2"),
                singleparticle_vs_allparticle_extrapolated=diff(sm, ap0),
                singleparticle_extrapolated_vs_allparticle_extrapolated=diff(sp0, ap0),
                singleparticle_vs_allparticle_extrapolated_lsq=diff(sm, (lsq["e"], lsq["err"])))


BATCH_SYSTEMS = [(14, 3, 8192), (38, 4, 16384), (66, 5, 24576)]


def batch_run(steps, warmup, tune=True):
    """N=14, 38 and 66 resident together (8k/16k/24k walkers, mixed precision, single-
    particle moves, k = N), stepped by the batch graph; the same systems stepped one after
    another on their own for comparison.  Walker-steps count the populations entering every
    step (read back after each step)."""
    from dmcbs_project_amd import Engine, Batch, DMC_MIXED, DMC_SINGLEPARTICLE

    def make():
        out = []
        for n, norb, nw in BATCH_SYSTEMS:
            e = Engine(npart=n, norb=norb, nwalk=nw, dt=DT, capacity=2 * nw,
                       precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=n,
                       dcut=DCUT_SM)
            e.init_walkers(SEED + n)
            out.append(e)
        return out

    res = {}
    es = make()
    b = Batch(es)
    if tune:
        choice, ms = b.tune(reps=3)
        res["wpg"] = choice
        # ms per launch at 1, 2, 4, 8 walkers per workgroup and per-system launches with the
        # tables read from global memory (0) or staged in LDS (-1)
        res["tune_ms"] = {k: [round(float(v), 4) for v in row]
                          for k, row in zip(("moves", "refresh", "end"), ms)}
    for _ in range(warmup):
        b.step()
    pops = b.sync()
    ws = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ws += sum(pops)
        b.step()
        pops = b.sync()
    wall = time.perf_counter() - t0
    res.update(value=ws / wall, ms_per_step=1e3 * wall / steps, populations=pops)
    b.close()
    for e in es:
        e.close()
    es = make()
    for _ in range(warmup):
        for e in es:
            e.step()
    ws = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in es:
            ws += e.numstack()
            e.step()
    for e in es:
        e.sync()
    wall = time.perf_counter() - t0
    res["separate_value"] = ws / wall
    res["batch_over_separate"] = res["value"] / res["separate_value"]
    for e in es:
        e.close()
    return res


def observables_run(steps, warmup, ntab=100, ntabq=10, ntabk=20):
    """g(r) with density profile (addgofr, ntab bins) and S(q) (addsofq, ntabq shells) over
    the resident headline population after 3 mixed-precision steps: one "step" is one block's
    operator pass over every walker (dmcbcs.f90:465-471), host sums included."""
    e = make_engine("mixed")
    for _ in range(3):
        e.step()
    gacc, sacc = e.addgofr(ntab), e.addsofq(ntabq)
    for _ in range(warmup):
        e.addgofr(ntab, gacc)
        e.addsofq(ntabq, sacc)
    e.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.addgofr(ntab, gacc)
        e.addsofq(ntabq, sacc)
    e.sync()
    dt = time.perf_counter() - t0
    nw = e.numstack()
    # n(k) (addenk, 20 shells of |m|^2 <= 20): 660 single-particle ratios per walker
    kacc = e.addenk(ntabk)
    e.sync()
    t1 = time.perf_counter()
    for _ in range(max(1, steps // 5)):
        e.addenk(ntabk, kacc)
    e.sync()
    dtk = (time.perf_counter() - t1) / max(1, steps // 5)
    return dict(value=steps * nw / dt, ms_per_step=1e3 * dt / steps, population=nw,
                nq=len(e.sofq_kvec(ntabq)[1]), nofk_ms=1e3 * dtk,
                nofk_ratios_per_s=nw * N * 10 / dtk)


def traffic_from_profiles(workload):
    """HBM bytes per launch of the dominant kernel from the rocprofv3 --pmc summaries in
    profiles/ (FETCH_SIZE + WRITE_SIZE, each scaled by its reading for a copy of known
    size), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def copy_calibration(nbytes=1 << 30):
    """A plain device copy of known size (for scaling the FETCH_SIZE/WRITE_SIZE counters)."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b = a.clone()
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "sum": float(b[:8].sum())}))


def _parse_ref_dmc(raw, nav, nstep, nwalk):
    """walker-steps of every block from the driver's dmc output (per step the populations
    after branch and after population control, per block 14 x 2 estimator values)"""
    pos, prev, out = 8, nwalk, []
    for _ in range(nav):
        pops = np.frombuffer(raw, "<i8", 2 * nstep, pos).reshape(nstep, 2)[:, 1]
        pos += 16 * nstep + 8 * 28
        entering = np.concatenate([[prev], pops[:-1]])
        out.append(int(entering.sum()))
        prev = int(pops[-1])
    return out


def cpu_reference_baseline(nwalk=64, nav=11, nstep=10):
    """The reference's own FP64 CPU path (oracle/_ref/ref_driver: the reference's modules
    compiled with amdflang -O2, all-particle DMC, nompi) on the host cores: one single-
    threaded process per core, each nwalk walkers for nav blocks of nstep steps.  The rate of
    each process is its walker-steps over its block times (the per-block CPU timer of
    dmcbcs.f90:434,450,487, printed by the driver), the first block left out; start-up,
    table set-up and the initial hpsi are outside every block.  value = the sum of the
    processes' rates (they run at once)."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    from oracle import refdeck
    cores = min(16, len(os.sched_getaffinity(0)))
    deck, _ = refdeck.make_deck(npart=N, norb=NORB, nwalk=nwalk, nav=nav, nstep=nstep, dt=DT,
                                seed=SEED)
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
    rates, times = [], []
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        np.array([0], "<i8").tofile(fin)
        procs = []
        for i in range(cores):
            d = os.path.join(td, f"p{i}")
            os.mkdir(d)
            procs.append(subprocess.Popen([drv, "dmc", fin, os.path.join(d, "o.bin")],
                                          stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                                          cwd=d, env=env))
        for p in procs:
            p.stdin.write(deck.encode())
            p.stdin.close()
        for i, p in enumerate(procs):
            txt = p.stdout.read().decode(errors="replace")
            p.wait(timeout=900)
            bt = [float(l.split("=")[1].split()[0]) for l in txt.splitlines()
                  if l.startswith("Time for block")]
            with open(os.path.join(td, f"p{i}", "o.bin"), "rb") as fh:
                ws = _parse_ref_dmc(fh.read(), nav, nstep, nwalk)
            rates.append(sum(ws[1:]) / sum(bt[1:]))
            times.append(sum(bt[1:]))
    return dict(value=float(sum(rates)), unit="walker-steps/s", cores=cores, kind="reference",
                per_core=float(np.mean(rates)),
                sample=f"reference FP64 all-particle DMC at N=66 (its Fortran modules, "
                       f"amdflang -O2, nompi, OpenBLAS 1 thread): {cores} processes x {nwalk} "
                       f"walkers x {nav} blocks of {nstep} steps, rate from the per-block CPU "
                       f"timer over blocks 2-{nav} ({np.mean(times):.1f} s per process)")


def cpu_reference_observables(cores=None):
    """The reference's own addgofr + addsofq (ntabg = 100, ntabq = 10) and addenk (ntab = 20)
    through oracle/_ref/ref_driver's obs and enk modes on the host cores, one single-threaded
    process per core on a bounded sample of uniform N=66 configurations (walkers/s, set-up
    included)."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    from oracle import refdeck
    cores = cores or min(16, len(os.sched_getaffinity(0)))
    deck, _ = refdeck.make_deck(npart=N, norb=NORB)
    el = (N / 0.0337737278) ** (1.0 / 3.0)
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
    out = {}
    for mode, head, nc in (("obs", [100, 10], 3000), ("enk", [20], 48)):
        rng = np.random.default_rng(SEED)
        parts = [np.array(head + [nc], "<i8").tobytes()]
        for k in range(nc):
            parts.append((el * (rng.random((N, 3)) - 0.5)).astype("<f8").tobytes())
            if mode == "enk":
                parts.append(np.array([1 + 7919 * k], "<i8").tobytes())
            parts.append(np.array([1.0], "<f8").tobytes())
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "in.bin")
            with open(fin, "wb") as fh:
                fh.write(b"".join(parts))
            dirs = [os.path.join(td, f"p{i}") for i in range(cores)]
            for d in dirs:
                os.mkdir(d)
            t0 = time.perf_counter()
            procs = [subprocess.Popen([drv, mode, fin, os.path.join(d, "o.bin")],
                                      stdin=subprocess.PIPE, stdout=subprocess.DEVNULL, cwd=d,
                                      env=env) for d in dirs]
            for p in procs:
                p.stdin.write(deck.encode())
                p.stdin.close()
            for p in procs:
                p.wait(timeout=600)
            wall = time.perf_counter() - t0
        out[mode] = dict(value=cores * nc / wall, unit="walkers/s", cores=cores,
                         kind="reference",
                         sample=f"reference {'addgofr + addsofq' if mode == 'obs' else 'addenk'}"
                                f" (amdflang -O2): {cores} processes x {nc} configurations "
                                f"at N={N}")
    return out


def cpu_port_baseline():
    """The oracle's FP64 single-particle-move sweep (the workload's algorithm) on all host
    cores (OpenMP), bounded sample."""
    import oracle
    cores = min(16, len(os.sched_getaffinity(0)))
    s = oracle.System(npart=N, norb=NORB, nwalk=NWALK, dt=DT, dcut=DCUT_SM)
    nw = 8 * cores
    x, irn = s.init_walkers(SEED, nw)
    dpsi, sc = s.init_hpsi(x)
    t0 = time.perf_counter()
    s.sm_step1(x, dpsi, sc, irn, N, nthreads=cores)
    wall = time.perf_counter() - t0
    return dict(value=nw / wall, unit="walker-steps/s", cores=cores, kind="port",
                sample=f"oracle FP64 single-particle sweep: {nw} walkers x 1 step at N=66")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)  # the headline configuration's length
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed",
                    choices=["mixed", "fp64", "allparticle", "batch", "observables"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the DMC energy study (about a minute of GPU time)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy three times (counter calibration) and exit")
    ap.add_argument("--quick", action="store_true",
                    help="only the named workload (no FP64 leg, no E_L comparison, no CPU)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (the walkers "
              "would shard across GPUs by population, each GPU stepping its own share; only "
              "the one-GPU share is built and measured)", file=sys.stderr)
        sys.exit(2)
    print(f"bench: workload={a.workload} N={N} walkers={NWALK} steps={a.steps} "
          f"warmup={a.warmup}", file=sys.stderr, flush=True)
    if a.workload == "batch":
        r = batch_run(a.steps, a.warmup)
        print(json.dumps({
            "metric": "DMC walker-steps/sec, N=14/38/66 resident together (8k/16k/24k "
                      "walkers), one graph per step", "value": r["value"],
            "unit": "walker-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic: reference initial "
            "walkers per system", "config": {"workload": "batch N14/N38/N66 8k/16k/24k "
                                                     "mixed, single-particle moves, k=N",
                                             "dt": DT, "dcut": DCUT_SM},
            "batch": r}), flush=True)
        return
    if a.workload == "observables":
        r = observables_run(a.steps, a.warmup)
        cpu = None if (a.no_cpu or a.quick) else cpu_reference_observables()
        print(json.dumps({
            "metric": "walkers/sec through g(r) + density (100 bins) and S(q) (10 shells)",
            "value": r["value"], "unit": "walkers/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic: reference initial walkers after 3 DMC steps",
            "config": {"workload": f"observables N={N} {r['population']} walkers, "
                                   f"ntabg=100, ntabq=10 ({r['nq']} q vectors)"},
            "cpu_baseline": cpu["obs"] if cpu else None,
            "nofk": {"ms_per_pass": r["nofk_ms"], "ntab": 20,
                     "walkers_per_s": r["population"] / (1e-3 * r["nofk_ms"]),
                     "ratios_per_s": r["nofk_ratios_per_s"],
                     "cpu_baseline": cpu["enk"] if cpu else None}}), flush=True)
        return
    main_run = timed_run(a.workload, a.steps, a.warmup)
    print(f"bench: {a.workload}: {main_run}", file=sys.stderr, flush=True)
    extra = {}
    if a.quick:
        a.no_cpu = True
    if a.workload == "mixed" and not a.quick:
        fp = timed_run("fp64", a.steps, a.warmup)
        print(f"bench: fp64: {fp}", file=sys.stderr, flush=True)
        extra["fp64_gpu_value"] = fp["value"]
        extra["mixed_over_fp64"] = main_run["value"] / fp["value"]
        # the studies beside the headline: one that fails is reported in its own field
        # (with the error), and the line still carries the measured headline
        def study(key, fn):
            try:
                extra[key] = fn()
            except Exception as ex:  # noqa: BLE001 - reported, not hidden
                extra[key] = {"failed": f"{type(ex).__name__}: {ex}"}
            print(f"bench: {key}: {extra[key]}", file=sys.stderr, flush=True)

        study("precision", precision_study)
        if not a.no_energy:
            study("dmc_energy", energy_study)
        study("batch", lambda: batch_run(2 * a.steps, 3))  # 2000 steps by default
    sm = a.workload != "allparticle"
    f = flops_per_walker_step(a.workload, part="moves" if sm else "step")
    achieved = f * main_run["pop_mean"] / (main_run["kernel_ms"] * 1e-3) / 1e12
    kname = {"mixed": "k_sm_moves32x2<33>", "fp64": "k_sm_moves64<33,1>",
             "allparticle": "k_step<33>"}[a.workload]
    peak = PEAK_FP32_TFLOPS if a.workload == "mixed" else PEAK_FP64_TFLOPS
    roof = dict(bound="valu", achieved=achieved, peak=peak, unit="TFLOP/s",
                frac=achieved / peak, traffic=traffic_from_profiles(a.workload),
                kernel=kname, kernel_ms=main_run["kernel_ms"], flops_per_launch=f *
                main_run["pop_mean"],
                frac_of_unpacked_issue=achieved / PEAK_FP64_TFLOPS,
                note=("vector-ALU bound, no MFMA (the per-walker matrices are 33 x 33); "
                      + ("peak = MI355X's published FP32 vector rate, which needs packed "
                         "FP32; this build emits no packed FP32, so one FP32 FMA per lane "
                         "and clock (78.6 TF/s) is its issue ceiling: "
                         "frac_of_unpacked_issue" if a.workload == "mixed" else
                         "peak = MI355X's FP64 vector rate")))
    cpu = None
    if not a.no_cpu:
        cpu = cpu_reference_baseline()
        try:
            extra["cpu_port"] = cpu_port_baseline()
        except Exception as ex:  # the oracle is a reported baseline only
            extra["cpu_port"] = str(ex)
    wl = {"mixed": "N66 unitary gas, single-particle moves + SM, mixed FP32/FP64, k=66",
          "fp64": "N66 unitary gas, single-particle moves + SM, FP64, k=66",
          "allparticle": "N66 unitary gas, all-particle moves (reference replay), FP64"}
    line = {
        "metric": "DMC walker-steps/sec at N=66, 32k walkers: mixed vs FP64 GPU; E_L rel err",
        "value": main_run["value"], "unit": "walker-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": main_run["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if a.workload == "mixed" else "f64",
        "data": "synthetic: reference initial walkers (one uniform configuration permuted, "
                "seed 7)",
        "config": {"workload": wl[a.workload], "npart": N, "norb": NORB, "walkers": NWALK,
                   "dt": DT, "krefresh": N, "ntab": 1024, "amu": 40.0,
                   "dcut": -1.0 if a.workload == "allparticle" else DCUT_SM},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    if "acceptance" in main_run:
        line["acceptance"] = main_run["acceptance"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
