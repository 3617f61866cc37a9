"""Write the golden fixtures under tests/golden/ from the reference itself.

Runs oracle/_ref/ref_driver (the reference's Fortran modules compiled by
`make -C oracle ref`) on decks written by refdeck.py and stores inputs and outputs as .npz.
Only this script touches the reference build; the tests read the .npz files.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py --vmc      # the variational and optimiser ones
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import refdeck  # noqa: E402
import refrun  # noqa: E402
from refrun import pack, Reader  # noqa: E402

SYSTEMS = {14: 3, 38: 4, 66: 5}


def box(n, rho=0.0337737278):
    return (n / rho) ** (1.0 / 3.0)


def setup_fixture(n, norb):
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    el = box(n)
    rr = np.concatenate([[0.0], np.linspace(1e-4, 0.5 * el, 97), [0.5 * el - 1e-9],
                         np.random.default_rng(n).uniform(0.0, 0.5 * el, 30)])
    seeds = np.array([1, 7, 11, 17, 281474976710655, 123456789012345], dtype=np.int64)
    ns = 64
    data, _ = refrun.run("setup", deck, pack([norb], [len(rr)], rr, [len(seeds)], seeds, [ns]))
    r = Reader(data)
    el_, eunit, etrial = r.f(3)
    nk = r.i()
    numsh = r.i(norb)
    ak = r.f(3 * nk).reshape(nk, 3)
    ak2 = r.f(nk)
    rad = np.array([r.f(9) for _ in rr])
    s1 = np.zeros((len(seeds), ns), np.int64)
    u1 = np.zeros((len(seeds), ns))
    s2 = np.zeros((len(seeds), ns), np.int64)
    u2 = np.zeros((len(seeds), ns))
    for j in range(len(seeds)):
        for i in range(ns):
            s1[j, i] = r.i()
            u1[j, i] = r.f()
            s2[j, i] = r.i()
            u2[j, i] = r.f()
    g7 = np.zeros((len(seeds), 7))
    gstate = np.zeros(len(seeds), np.int64)
    for j in range(len(seeds)):
        g7[j] = r.f(7)
        gstate[j] = r.i()
    assert r.done()
    np.savez(os.path.join(HERE, f"setup_N{n}.npz"), npart=n, norb=norb, el=el_, eunit=eunit,
             etrial=etrial, nk=nk, numsh=numsh, ak=ak, ak2=ak2, r=rr, radial=rad[:, 0:3],
             beta=rad[:, 3:6], derbeta=rad[:, 6:8], vpot=rad[:, 8], seeds=seeds,
             ran1_state=s1, ran1_u=u1, ran2_state=s2, ran2_u=u2, gauss7=g7,
             gauss7_state=gstate, deck=deck)
    return el_


def random_configs(n, nc, seed):
    el = box(n)
    rng = np.random.default_rng(seed)
    return el * (rng.random((nc, n, 3)) - 0.5)


def hpsi_fixture(n, norb, configs, tag):
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    nc = len(configs)
    data, _ = refrun.run("hpsi", deck, pack([nc], configs))
    r = Reader(data)
    out = {k: [] for k in ("u", "d2u", "v", "phil", "d2phi", "isdet", "du", "dphi", "psil",
                           "d2psi", "is", "psigl", "d2psig", "dpsi", "dpsig")}
    for _ in range(nc):
        u, d2u, v, phil, d2phi, isd = r.f(6)
        du = r.f(3 * n).reshape(n, 3)
        dphi = r.f(3 * n).reshape(n, 3)
        psil, d2psi, v2, is_, psigl, d2psig = r.f(6)
        dpsi = r.f(3 * n).reshape(n, 3)
        dpsig = r.f(3 * n).reshape(n, 3)
        for k, val in (("u", u), ("d2u", d2u), ("v", v), ("phil", phil), ("d2phi", d2phi),
                       ("isdet", isd), ("du", du), ("dphi", dphi), ("psil", psil),
                       ("d2psi", d2psi), ("is", is_), ("psigl", psigl), ("d2psig", d2psig),
                       ("dpsi", dpsi), ("dpsig", dpsig)):
            out[k].append(val)
    assert r.done()
    np.savez(os.path.join(HERE, f"hpsi_{tag}.npz"), npart=n, norb=norb, x=configs, deck=deck,
             **{k: np.array(v) for k, v in out.items()})


def ratio_fixture(n, norb, configs, nm, seed, tag):
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    rng = np.random.default_rng(seed)
    nc = len(configs)
    ip = rng.integers(1, n + 1, size=(nc, nm))
    step = rng.normal(0.0, 0.3, size=(nc, nm, 3))
    xnew = configs[np.arange(nc)[:, None], ip - 1] + step
    parts = [pack([nc], [nm])]
    for k in range(nc):
        parts.append(pack(configs[k]))
        for m in range(nm):
            parts.append(pack([ip[k, m]], xnew[k, m]))
    data, _ = refrun.run("ratio", deck, b"".join(parts))
    r = Reader(data)
    rat = r.f(nc * nm).reshape(nc, nm)
    assert r.done()
    np.savez(os.path.join(HERE, f"ratio_{tag}.npz"), npart=n, norb=norb, x=configs, ip=ip,
             xnew=xnew, ratio=rat, deck=deck)


def der_fixture(n, norb, configs, tag):
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    nc = len(configs)
    data, _ = refrun.run("der", deck, pack([nc], configs))
    r = Reader(data)
    npar = r.i()
    par = r.f(npar)
    dp = r.f(nc * npar).reshape(nc, npar)
    assert r.done()
    np.savez(os.path.join(HERE, f"der_{tag}.npz"), npart=n, norb=norb, x=configs, par=par,
             dlogpsi=dp, deck=deck)


def opt_fixture(n, norb, configs, tag, delta=0.7, eps=1.0e-4):
    """calder over configurations, updatepar, then the new parameters (do_opt)."""
    deck, p = refdeck.make_deck(npart=n, norb=norb, idmc=-1, delta=delta, eps=eps)
    nc = len(configs)
    data, _ = refrun.run("opt", deck, pack([nc], configs))
    r = Reader(data)
    ene = r.f()
    npar = r.i()
    par = r.f(npar)
    assert r.done()
    np.savez(os.path.join(HERE, f"opt_{tag}.npz"), npart=n, norb=norb, x=configs,
             vorb=np.asarray(refdeck.VORB10[:norb]), ene=ene, par=par, delta=delta, eps=eps,
             deck=deck)


def read_dump(r, n):
    nw = r.i()
    irn = np.zeros(nw, np.int64)
    sc = np.zeros((nw, 6))
    x = np.zeros((nw, n, 3))
    dpsi = np.zeros((nw, n, 3))
    for k in range(nw):
        irn[k] = r.i()
        sc[k] = r.f(6)
        x[k] = r.f(3 * n).reshape(n, 3)
        dpsi[k] = r.f(3 * n).reshape(n, 3)
    return irn, sc, x, dpsi


def dmc_fixture(tag, dumps, keep_x=True, init=None, **deckargs):
    deck, p = refdeck.make_deck(**deckargs)
    n = p["npart"]
    parts = [pack([len(dumps)], np.asarray(dumps, np.int64))]
    mode = "dmc"
    if init is not None:
        mode = "dmcfile"
        x0, irn0 = init
        parts.append(pack([len(x0)]))
        for k in range(len(x0)):
            parts.append(pack(x0[k], [irn0[k]]))
    data, _ = refrun.run(mode, deck, b"".join(parts))
    r = Reader(data)
    nsteps = r.i()
    nstep = p["nstep"]
    out = {}
    pop = np.zeros((nsteps, 2), np.int64)
    est = []
    if 0 in dumps:
        out["d0"] = read_dump(r, n)
    for jj in range(1, nsteps + 1):
        pop[jj - 1] = r.i(2)
        if jj in dumps:
            out[f"d{jj}"] = read_dump(r, n)
        if jj % nstep == 0:
            est.append(r.f(28).reshape(14, 2))
    opt = None
    if p["idmc"] in (-1, -2):
        ene = r.f()
        opt = (ene, r.f(r.i()))
    assert r.done()
    arrs = dict(npart=n, norb=p["norb"], pop=pop, est=np.array(est), deck=deck,
                dumps=np.asarray(dumps, np.int64), nwalk=p["nwalk"], nav=p["nav"],
                nstep=nstep, dt=p["dt"], seed=p["seed"])
    for key, (irn, sc, x, dpsi) in out.items():
        arrs[f"{key}_irn"] = irn
        arrs[f"{key}_sc"] = sc
        if keep_x:
            arrs[f"{key}_x"] = x
            arrs[f"{key}_dpsi"] = dpsi
    if init is not None:
        arrs["init_x"], arrs["init_irn"] = init
    if opt is not None:
        arrs["opt_ene"], arrs["opt_par"] = opt
        arrs["idmc"] = p["idmc"]
    np.savez_compressed(os.path.join(HERE, f"dmc_{tag}.npz"), **arrs)
    return out


def main_vmc():
    """Variational runs (step1 with idmc = 0) and the optimiser (calder, updatepar)."""
    dmc_fixture("vmc_N14", [0, 1, 2, 10, 20], idmc=0, npart=14, norb=3, nwalk=128, nav=2,
                nstep=10, dt=0.002, seed=11)
    dmc_fixture("vmc_N66", [0, 1, 3], idmc=0, npart=66, norb=5, nwalk=24, nav=1, nstep=3,
                dt=0.001, seed=11)
    # a whole optimisation iteration (VMC blocks with calder after each, then updatepar)
    dmc_fixture("optrun_N14", [], idmc=-1, npart=14, norb=3, nwalk=64, nav=3, nstep=5,
                dt=0.002, seed=11)
    opt_fixture(14, 3, random_configs(14, 64, 6014), "N14")
    opt_fixture(66, 5, random_configs(66, 16, 6066), "N66")
    vmc_long_fixture()


def _lcg_jump(s, k):
    """ran1's state after k draws from state s (ran.f90:8-27: s <- s * A + C mod 2^48)"""
    for _ in range(k):
        s = (s * 44485709377909 + 96309754297) % (1 << 48)
    return s


def vmc_long_fixture(n=2, nwalk=512, dt=0.05, dcut=1.0, first=500, last=3000, every=50,
                     seed=11):
    """A long variational run (step1 with idmc = 0: all-particle Metropolis moves that sample
    |psi|^2 exactly), N = 2 (psi = phi(r_1 - r_2) J(r_12): one relative coordinate, fast
    mixing).  Kept: ln|psi|, the local energy and the minimum-image pair distance of every
    walker every 50 steps from step 500 to 3000 (51 samples of 512 independent chains, in
    the walker order of step 500), against which a single-particle Metropolis chain of the
    same |psi|^2 is compared.  The drift is cut at sigma (dcut = 1): without the cut the
    drift sigma^2/r near the pair's coalescence overshoots, the chain rarely enters or
    leaves that region, and 3000 steps at this dt do not converge there (ln|psi| 5 sigma
    below a direct quadrature of |psi|^2; at dt = 0.01 or with the cut they agree)."""
    deckargs = dict(idmc=0, npart=n, norb=3, nwalk=nwalk, nav=last // 20, nstep=20, dt=dt,
                    dcut=dcut, seed=seed)
    deck, p = refdeck.make_deck(**deckargs)
    el_box = box(n)
    dumps = np.arange(first, last + 1, every, dtype=np.int64)
    data, _ = refrun.run("dmc", deck, pack([len(dumps)], dumps))
    r = Reader(data)
    nsteps = r.i()
    psil, el, r12, key = [], [], [], None
    # a variational step draws 3N gaussians (3N ran1 draws, N even) and one acceptance draw
    jump = every * (3 * n + 1)
    for jj in range(1, nsteps + 1):
        r.i(2)
        if jj in dumps:
            irn, sc, x, _ = read_dump(r, n)
            if key is None:
                order = np.arange(len(irn))
                key = [int(v) for v in irn]
            else:
                # the stack's pop order changes from step to step; the walker of slot i is
                # the one whose RNG state is slot i's state jumped ahead by the draws between
                key = [_lcg_jump(v, jump) for v in key]
                pos = {int(v): k for k, v in enumerate(irn)}
                order = np.array([pos[v] for v in key])
            d = x[order, 0] - x[order, 1]
            d -= el_box * np.rint(d / el_box)
            psil.append(sc[order, 1])
            el.append(-p["hbar"] * sc[order, 2] + sc[order, 3])
            r12.append(np.sqrt(np.sum(d * d, axis=1)))
        if jj % p["nstep"] == 0:
            r.f(28)
    assert r.done()
    np.savez_compressed(os.path.join(HERE, f"vmclong_N{n}.npz"), npart=n, norb=p["norb"],
                        nwalk=p["nwalk"], dt=p["dt"], dcut=p["dcut"], seed=p["seed"],
                        steps=dumps, psil=np.array(psil, np.float32),
                        elocal=np.array(el, np.float32), r12=np.array(r12, np.float32), deck=deck)


def main_mdmc():
    """Metropolis DMC (step1 with idmc = 2, step.f90:161-226): short trajectories with walker
    dumps at N = 14 and 66.  With this deck the branch grows the population step by step
    (walkers whose proposals keep being rejected are weighted exp(-(E_old - E_T) dt) again
    and again); at dt = 0.002 the reference stops with a stack overflow in its sixth step,
    whose population trace up to the stop is kept as well."""
    dmc_fixture("mdmc_N14_short", [0, 1, 2, 5, 10, 20], idmc=2, npart=14, norb=3, nwalk=200,
                nav=2, nstep=10, dt=0.0005, seed=1)
    dmc_fixture("mdmc_N66_short", [0, 1, 2, 4], idmc=2, npart=66, norb=5, nwalk=32, nav=2,
                nstep=2, dt=0.001, seed=7)
    deck, p = refdeck.make_deck(idmc=2, npart=14, norb=3, nwalk=200, nav=2, nstep=10,
                                dt=0.002, seed=1)
    data, out = refrun.run("dmc", deck, pack([0], np.zeros(0, np.int64)))
    assert "stack overflow" in out
    r = Reader(data)
    nsteps = r.i()
    pops = []
    while not r.done():
        pops.append(r.i(2))
    np.savez(os.path.join(HERE, "mdmc_N14_overflow.npz"), npart=14, norb=3, nwalk=200,
             dt=0.002, seed=1, nsteps=nsteps, pop=np.array(pops), capacity=2000, deck=deck)


def _table(text):
    """rows of numbers of one output table (comment and blank lines skipped)"""
    rows = []
    for line in text.splitlines():
        t = line.strip()
        if not t or t.startswith("#"):
            continue
        rows.append([float(v.replace("D", "E")) for v in t.split()])
    return np.array(rows)


def obs_fixture(n, norb, configs, weights, ntg, ntq, tag):
    """g(r), density profile and S(q) of weighted configurations through the reference's
    gofr.f90/sofq.f90 (setup, add per configuration, update, write), as the operators of
    operators.f90 apply them at the end of a block"""
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    nc = len(configs)
    body = [pack([ntg], [ntq], [nc])]
    for x, w in zip(configs, weights):
        body.append(pack(x, [w]))
    data, _, texts = refrun.run("obs", deck, b"".join(body),
                                files=("gofr.out", "rho.out", "sofq.out"))
    r = Reader(data)
    assert int(r.f()) == nc
    el, kf = r.f(), r.f()
    np.savez(os.path.join(HERE, f"obs_{tag}.npz"), npart=n, norb=norb, x=configs, w=weights,
             el=el, kf=kf,
             ntg=ntg, ntq=ntq, gofr=_table(texts["gofr.out"]), rho=_table(texts["rho.out"]),
             sofq=_table(texts["sofq.out"]), deck=deck)


def enk_fixture(n, norb, configs, irns, weights, ntk, tag):
    """n(k) of weighted configurations with their RNG states through the reference's
    momentum3d.f90 (setupenk, addenk per configuration, updateenk, writeenk)"""
    deck, p = refdeck.make_deck(npart=n, norb=norb)
    nc = len(configs)
    body = [pack([ntk], [nc])]
    for x, irn, w in zip(configs, irns, weights):
        body.append(pack(x, np.array([irn], np.int64), [w]))
    data, _, texts = refrun.run("enk", deck, b"".join(body), files=("nofk.out",))
    r = Reader(data)
    assert int(r.f()) == nc
    el, kf = r.f(), r.f()
    np.savez(os.path.join(HERE, f"enk_{tag}.npz"), npart=n, norb=norb, x=configs, irn=irns,
             w=weights, ntk=ntk, el=el, kf=kf, nofk=_table(texts["nofk.out"]), deck=deck)


def main_obs():
    # uniform configurations with weights of both signs (is * wt0 = -1 walkers included)
    for n, norb, nc, ntg, ntq, seed in ((14, 3, 24, 40, 6, 7014), (66, 5, 12, 60, 8, 7066)):
        rng = np.random.default_rng(seed)
        w = rng.uniform(0.5, 1.5, nc) * np.where(rng.random(nc) < 0.2, -1.0, 1.0)
        obs_fixture(n, norb, random_configs(n, nc, seed), w, ntg, ntq, f"N{n}")


def main_enk():
    # n(k): 48-bit RNG states and weights of both signs
    for n, norb, nc, ntk, seed in ((14, 3, 6, 20, 8014), (66, 5, 3, 12, 8066)):
        rng = np.random.default_rng(seed)
        w = rng.uniform(0.5, 1.5, nc) * np.where(rng.random(nc) < 0.3, -1.0, 1.0)
        irn = rng.integers(1, 1 << 48, nc, dtype=np.int64)
        enk_fixture(n, norb, random_configs(n, nc, seed), irn, w, ntk, f"N{n}")


def main():
    main_vmc()
    for n, norb in SYSTEMS.items():
        setup_fixture(n, norb)
    # hpsi on uniform configurations (all three systems) and with 8 shells at N=66
    hpsi_fixture(14, 3, random_configs(14, 32, 1014), "N14")
    hpsi_fixture(38, 4, random_configs(38, 12, 1038), "N38")
    hpsi_fixture(66, 5, random_configs(66, 12, 1066), "N66")
    hpsi_fixture(66, 8, random_configs(66, 4, 2066), "N66_norb8")
    # single-particle ratios (bcsone * getphione through move1)
    ratio_fixture(14, 3, random_configs(14, 8, 3014), 6, 4014, "N14")
    ratio_fixture(66, 5, random_configs(66, 4, 3066), 6, 4066, "N66")
    # d log psi / d alpha_s (getderbcs)
    der_fixture(14, 3, random_configs(14, 8, 5014), "N14")
    der_fixture(66, 8, random_configs(66, 4, 5066), "N66_norb8")
    # short DMC trajectories with walker dumps (multiset parity of branching)
    dmc_fixture("N14_short", [0, 1, 2, 5, 10, 20], npart=14, norb=3, nwalk=200, nav=2,
                nstep=10, dt=0.002, seed=1)
    dmc_fixture("N38_short", [0, 1, 3], npart=38, norb=4, nwalk=48, nav=1, nstep=3,
                dt=0.001, seed=5)
    dmc_fixture("N66_short", [0, 1, 2, 4], npart=66, norb=5, nwalk=32, nav=2, nstep=2,
                dt=0.001, seed=7)
    # the N=14 production configuration: 1000 walkers, 500 steps, population trace,
    # block estimators and the final (irn, weight) multiset
    dmc_fixture("N14_cfg1", [500], keep_x=False, npart=14, norb=3, nwalk=1000, nav=5,
                nstep=100, dt=0.002, seed=1)


if __name__ == "__main__":
    if "--vmc" in sys.argv:
        main_vmc()
    elif "--vmclong" in sys.argv:
        vmc_long_fixture()
    elif "--mdmc" in sys.argv:
        main_mdmc()
    elif "--obs" in sys.argv:
        main_obs()
    elif "--enk" in sys.argv:
        main_enk()
    else:
        main()
