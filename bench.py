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

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
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


def make_engine(mode, nwalk=NWALK, krefresh=N):
    from dmcbs_project_amd import Engine, DMC_FP64, DMC_MIXED, DMC_ALLPARTICLE, DMC_SINGLEPARTICLE
    prec = DMC_MIXED if mode == "mixed" else DMC_FP64
    move = DMC_ALLPARTICLE if mode == "allparticle" else DMC_SINGLEPARTICLE
    # single-particle moves cut the drift at 2 sigma per component (the reference's dcut,
    # step.f90:32-37); the all-particle replay keeps the reference deck's dcut = -1
    dcut = -1.0 if mode == "allparticle" else DCUT_SM
    e = Engine(npart=N, norb=NORB, nwalk=nwalk, dt=DT, capacity=2 * nwalk, precision=prec,
               movetype=move, krefresh=krefresh, dcut=dcut)
    e.init_walkers(SEED)
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


def elocal_error(nconf=4096, krefresh=N):
    """Per-step local-energy error of the mixed-precision sweep against the FP64 sweep.

    4096 configurations are equilibrated by three FP64 steps from seed 7; both propagators
    then make one step from these same walkers and random streams (population control
    off).  Walkers are matched by their RNG state, which identifies the source walker and
    the copy; for each match the relative difference of the stored local energies
    (E_L = -hbar d2psi + v) is reported, with the fraction of walkers whose accepted-move
    pattern led to the same local energy within 1e-6."""
    from dmcbs_project_amd import Engine, DMC_FP64, DMC_MIXED, DMC_SINGLEPARTICLE
    kw = dict(npart=N, norb=NORB, nwalk=nconf, dt=DT, capacity=8 * nconf, krefresh=krefresh,
              movetype=DMC_SINGLEPARTICLE, dcut=DCUT_SM)
    eq = Engine(precision=DMC_FP64, **kw)
    eq.init_walkers(SEED)
    for _ in range(3):
        eq.step()
    w0 = eq.walkers()
    eq.close()
    out = {}
    for prec in (DMC_FP64, DMC_MIXED):
        e = Engine(precision=prec, popup=1e9, popdn=0.0, **kw)
        e.push_walkers(w0["x"], w0["irn"])
        e.step()
        w = e.walkers()
        out[prec] = (w["irn"], -0.5 * w["d2psi"] + w["v"])
        e.close()
    common, i64, i32 = np.intersect1d(out[DMC_FP64][0], out[DMC_MIXED][0], return_indices=True)
    e64 = out[DMC_FP64][1][i64]
    e32 = out[DMC_MIXED][1][i32]
    rel = np.abs(e32 - e64) / np.maximum(np.abs(e64), 1e-300)
    # the same difference against the population's mean |E_L| (a walker whose E_L passes
    # near zero makes the per-walker ratio large without a large error)
    rels = np.abs(e32 - e64) / np.mean(np.abs(e64))
    return dict(krefresh=krefresh, median=float(np.median(rel)),
                p90=float(np.percentile(rel, 90)), p99=float(np.percentile(rel, 99)),
                max=float(rel.max()), frac_below_1e6=float(np.mean(rel <= 1e-6)),
                frac_below_1e6_of_mean=float(np.mean(rels <= 1e-6)),
                matched=int(len(common)), walkers=int(len(w0["irn"])))


def energy_agreement(nequil=200, nblocks=20, blocklen=20):
    """DMC energy of the mixed and the FP64 propagators on the headline workload: both start
    from the same walkers (seed 7) and random streams, equilibrate nequil steps, then sample
    the energy estimator (E_L / (N E_FG), weighted, estimator.f90:72-82) on every step of
    nblocks blocks of blocklen steps.  Block means give each mean and its standard error;
    the two runs are independent Markov chains after a few steps, so their difference is
    compared with the combined error."""
    out = {}
    for mode in ("mixed", "fp64"):
        e = make_engine(mode)
        for _ in range(nequil):
            e.step()
        e.estimators(reset=True)
        blocks = []
        for _ in range(nblocks):
            for _ in range(blocklen):
                e.step(opcomp=True)
            est = e.estimators(reset=True)
            blocks.append(est[0, 0] / est[0, 2])
        b = np.array(blocks)
        out[mode] = (float(b.mean()), float(b.std(ddof=1) / np.sqrt(nblocks)))
        e.close()
    (m32, s32), (m64, s64) = out["mixed"], out["fp64"]
    sig = float(np.hypot(s32, s64))
    return dict(e_mixed=m32, err_mixed=s32, e_fp64=m64, err_fp64=s64, diff=m32 - m64,
                diff_over_sigma=(m32 - m64) / sig if sig > 0 else None,
                steps=f"{nequil} equilibration + {nblocks} x {blocklen} sampled",
                unit="E_L / (N E_FG)")


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
        # ms per launch at 1, 2, 4, 8 walkers per workgroup and per-system launches (0)
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


def cpu_reference_baseline(seconds=20.0):
    """The reference's own FP64 CPU path (oracle/_ref/ref_driver, all-particle DMC, nompi)
    on the host cores: one single-threaded process per core on a bounded sample."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    from oracle import refdeck
    cores = min(16, len(os.sched_getaffinity(0)))
    nwalk, nstep = 64, 4
    deck, _ = refdeck.make_deck(npart=N, norb=NORB, nwalk=nwalk, nav=1, nstep=nstep, dt=DT,
                                seed=SEED)
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        np.array([0], "<i8").tofile(fin)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([drv, "dmc", fin, os.path.join(td, f"o{i}.bin")],
                                  stdin=subprocess.PIPE, stdout=subprocess.DEVNULL, cwd=td,
                                  env=env) for i in range(cores)]
        for p in procs:
            p.stdin.write(deck.encode())
            p.stdin.close()
        ws = 0
        for i, p in enumerate(procs):
            p.wait(timeout=600)
            raw = np.fromfile(os.path.join(td, f"o{i}.bin"), "<i8")
            pops = raw[1:1 + 2 * nstep].reshape(nstep, 2)[:, 1]
            ws += int(pops.sum())
        wall = time.perf_counter() - t0
    return dict(value=ws / wall, unit="walker-steps/s", cores=cores, kind="reference",
                sample=f"reference FP64 all-particle DMC (nompi, amdflang -O2, OpenBLAS 1 "
                       f"thread): {cores} processes x {nwalk} walkers x {nstep} steps at N=66")


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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed",
                    choices=["mixed", "fp64", "allparticle", "batch", "observables"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
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
        extra["elocal_rel_err"] = elocal_error()
        extra["elocal_rel_err_k"] = {str(k): elocal_error(krefresh=k) for k in (16, 33)}
        extra["dmc_energy"] = energy_agreement()
        print(f"bench: energy: {extra['dmc_energy']}", file=sys.stderr, flush=True)
        br = batch_run(min(a.steps, 30), 3)
        print(f"bench: batch: {br}", file=sys.stderr, flush=True)
        extra["batch"] = br
    sm = a.workload != "allparticle"
    f = flops_per_walker_step(a.workload, part="moves" if sm else "step")
    achieved = f * main_run["pop_mean"] / (main_run["kernel_ms"] * 1e-3) / 1e12
    kname = {"mixed": "k_sm_moves32<33,1>", "fp64": "k_sm_moves64<33,1>",
             "allparticle": "k_step<33>"}[a.workload]
    roof = dict(bound="mfma", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                frac=achieved / PEAK_FP64_TFLOPS, traffic=traffic_from_profiles(a.workload),
                kernel=kname, kernel_ms=main_run["kernel_ms"], flops_per_launch=f *
                main_run["pop_mean"],
                note="vector-ALU peak: no MFMA is used; 78.6 TF/s is MI355X's FP64 peak and "
                     "the non-packed FP32 VALU rate (packed FP32 is not emitted)")
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
