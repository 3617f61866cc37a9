"""Host-side mirror of the reference's module interfaces for the walker-propagation path.

* ``Engine.setstep`` / ``Engine.step1`` / ``Engine.branch`` follow step.f90 (setstep,
  step1 with idmc = 1, branch and the population control of dmcbcs.f90:441-445);
* ``Engine.push_walkers`` / ``Engine.walkers`` / ``Engine.numstack`` follow stack.f90;
* ``Engine.hpsi`` / ``Engine.move1`` / ``Engine.getderpsi`` / ``Engine.setdetparam``
  follow wavefunction.f90.

All arithmetic runs on the device through libdmc_gpu.so.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lib as L

# s-wave pairing amplitudes of the production deck (dmcbcs.dat:29-38)
VORB10 = (0.00197519728006648144, 0.00249764752028349595, 0.00194383232328550247,
          0.00081339556934447224, 0.00033078733782252270, 0.00019008071412562492,
          0.00019983385861822818, 0.00016699298789053479, 0.00016360289711864207,
          0.00012098992317746450)


@dataclass
class Params:
    """Deck values (dmcbcs.f90:43-66 and the module decks) plus device options."""
    npart: int = 14
    norb: int = 3
    ntab: int = 1024
    fn: bool = True
    units: bool = True
    nwalk: int = 1000
    capacity: int = 0
    precision: int = L.DMC_FP64
    movetype: int = L.DMC_ALLPARTICLE
    krefresh: int = 0
    rho: float = 0.0337737278
    hbar: float = 0.5
    mass: float = 1.0
    dt: float = 0.002
    etrial: float = 0.40
    dcut: float = -1.0
    wtcut: float = 10.0
    popup: float = 1.1
    popdn: float = 0.9
    bbeta: float = 0.5
    dbeta: float = 5.0
    bfact: float = 1.0
    v0: float = 1.0
    amu: float = 40.0
    cjas1: float = 1.0
    cjas2: float = 0.0
    djas: float = 1.15
    v0op: float = 0.0
    v0eq: float = 0.0
    vorb: tuple = field(default=None)

    def to_c(self):
        p = L.dmc_params()
        for k in ("npart", "norb", "ntab", "nwalk", "capacity", "precision", "movetype",
                  "krefresh"):
            setattr(p, k, int(getattr(self, k)))
        p.fn = int(bool(self.fn))
        p.units = int(bool(self.units))
        for k in ("rho", "hbar", "mass", "dt", "etrial", "dcut", "wtcut", "popup", "popdn",
                  "bbeta", "dbeta", "bfact", "v0", "amu", "cjas1", "cjas2", "djas", "v0op",
                  "v0eq"):
            setattr(p, k, float(getattr(self, k)))
        vo = self.vorb if self.vorb is not None else VORB10[: self.norb]
        for i, v in enumerate(vo):
            p.vorb[i] = float(v)
        return p


class Engine:
    """One resident walker population of one system on the device."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else Params(**kw)
        self.lib = L.load()
        self._p = self.params.to_c()
        h = C.c_void_p()
        L.check(self.lib.dmc_create(C.byref(self._p), C.byref(h)))
        self.h = h
        info = np.zeros(8)
        L.check(self.lib.dmc_info(self.h, info))
        self.el, self.eunit, self.etrial, self.nfac = info[:4]
        self.nk, self.ntab0, self.sigma, self.capacity = (int(info[4]), int(info[5]), info[6],
                                                          int(info[7]))
        self.npart = self.params.npart

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # --- wavefunction.f90 ------------------------------------------------------------
    def hpsi(self, x):
        """hpsi on a batch of configurations x[nw][npart][3]."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.npart, 3)
        nw = len(x)
        psil, d2psi, v = np.zeros(nw), np.zeros(nw), np.zeros(nw)
        is_ = np.zeros(nw, np.int32)
        dpsi = np.zeros_like(x)
        L.check(self.lib.dmc_eval(self.h, nw, x.reshape(-1), psil, is_, dpsi.reshape(-1),
                                  d2psi, v))
        return dict(psil=psil, is_=is_, dpsi=dpsi, d2psi=d2psi, v=v)

    def move1(self, x, ip, xnew):
        """psi(x with particle ip moved to xnew)/psi(x), batched; ip is 0-based."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.npart, 3)
        nw = len(x)
        ip = np.ascontiguousarray(np.broadcast_to(ip, (nw,)), np.int32)
        xn = np.ascontiguousarray(xnew, np.float64).reshape(nw, 3)
        out = np.zeros(nw)
        L.check(self.lib.dmc_ratio1(self.h, nw, x.reshape(-1), ip, xn.reshape(-1), out))
        return out

    def getderpsi(self, x, with_elocal=False):
        """d log psi / d(alpha_1..alpha_norb, b, d) and optionally d E_L / d theta."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, self.npart, 3)
        nw = len(x)
        npar = self.params.norb + 2
        dl = np.zeros((nw, npar))
        de = np.zeros((nw, npar)) if with_elocal else None
        L.check(self.lib.dmc_adjoint(self.h, nw, x.reshape(-1), dl.reshape(-1),
                                     de.ctypes.data if de is not None else None))
        return (dl, de) if with_elocal else dl

    # --- optimizer.f90 -------------------------------------------------------------------
    def calder(self):
        """calder (optimizer.f90:90-112) for every resident walker, on the device."""
        L.check(self.lib.dmc_opt_accumulate(self.h))

    def updatepar(self, delta, eps, optbeta=False):
        """updatepar (optimizer.f90:114-169) from the accumulated sums: f = -2<O E> +
        2<O><E>, S = <O O^T> - <O><O>^T + eps I, theta += delta S^-1 f; the new parameters
        are set (setdetparam) and the sums cleared.  Returns (<E>, new parameters)."""
        P = self.params.norb + 2
        sums = np.zeros(3 + 2 * P + P * P)
        L.check(self.lib.dmc_opt_sums(self.h, sums, 1))
        n = sums[0]
        if n <= 0:
            raise L.DMCError(1, "no calder samples")
        k = P if optbeta else self.params.norb
        ovec = sums[3:3 + P][:k] / n
        ovecw = sums[3 + P:3 + 2 * P][:k] / n
        pmat = sums[3 + 2 * P:].reshape(P, P)[:k, :k] / n
        esum = sums[1] / n
        fvec = -2.0 * ovecw + 2.0 * ovec * esum
        pmat = pmat - np.outer(ovec, ovec) + eps * np.eye(k)
        dpar = np.linalg.solve(pmat, fvec)
        par = np.concatenate([self.vorb(), [self.params.bbeta, self.params.dbeta]])[:k]
        par = par + delta * dpar
        vorb = par[:self.params.norb].copy()
        if optbeta:
            self.setdetparam(vorb, par[-2], par[-1])
        else:
            self.setdetparam(vorb)
        return float(esum), par

    def vorb(self):
        """The current alpha_1..alpha_norb (getdetparam, wavefunction.f90:246-258)."""
        vo = self.params.vorb if self.params.vorb is not None else VORB10[: self.params.norb]
        return np.array(vo[: self.params.norb], np.float64)

    def setdetparam(self, vorb, bbeta=None, dbeta=None):
        vorb = np.ascontiguousarray(vorb, np.float64)
        self.params.vorb = tuple(float(v) for v in vorb)
        L.check(self.lib.dmc_set_params(self.h, vorb,
                                        self.params.bbeta if bbeta is None else float(bbeta),
                                        self.params.dbeta if dbeta is None else float(dbeta)))
        if bbeta is not None:
            self.params.bbeta = float(bbeta)
        if dbeta is not None:
            self.params.dbeta = float(dbeta)

    # --- stack.f90 ---------------------------------------------------------------------
    def init_walkers(self, seed, nwalk=None):
        L.check(self.lib.dmc_init_walkers(self.h, int(seed),
                                          int(self.params.nwalk if nwalk is None else nwalk)))

    def push_walkers(self, x, irn):
        x = np.ascontiguousarray(x, np.float64)
        irn = np.ascontiguousarray(irn, np.int64)
        L.check(self.lib.dmc_load_walkers(self.h, len(irn), x.reshape(-1), irn))

    def numstack(self):
        n = C.c_int32()
        L.check(self.lib.dmc_population(self.h, C.byref(n)))
        return n.value

    def walkers(self):
        n = self.numstack()
        nw = C.c_int32()
        x = np.zeros((n, self.npart, 3))
        dpsi = np.zeros_like(x)
        sc = np.zeros((n, 6))
        irn = np.zeros(n, np.int64)
        L.check(self.lib.dmc_read_walkers(self.h, n, C.byref(nw), x.ctypes.data,
                                          dpsi.ctypes.data, sc.ctypes.data, irn.ctypes.data))
        return dict(x=x, dpsi=dpsi, sc=sc, irn=irn, weight=sc[:, 0], psil=sc[:, 1],
                    d2psi=sc[:, 2], v=sc[:, 3], is_=sc[:, 4], wt0=sc[:, 5])

    # --- step.f90 ------------------------------------------------------------------------
    def step(self, opcomp=False):
        """step1 + branch + population control; returns (nt after branch, nt after control)."""
        nt = np.zeros(2, np.int32)
        L.check(self.lib.dmc_step(self.h, int(bool(opcomp)), nt))
        return int(nt[0]), int(nt[1])

    def step1(self, opcomp=False):
        """step1 (step.f90:43-285) on the resident population; dmc_branch must follow."""
        L.check(self.lib.dmc_propagate(self.h, int(bool(opcomp))))

    def branch(self, factor=1.0):
        """branch (step.f90:287-315); returns the new population."""
        nt = C.c_int32()
        L.check(self.lib.dmc_branch(self.h, float(factor), C.byref(nt)))
        return nt.value

    def run(self, nsteps):
        nt = C.c_int32()
        L.check(self.lib.dmc_run(self.h, int(nsteps), C.byref(nt)))
        return nt.value

    def estimators(self, reset=False):
        est = np.zeros(42)
        L.check(self.lib.dmc_estimators(self.h, est.ctypes.data, int(bool(reset))))
        return est.reshape(14, 3)

    def wtisto(self):
        h = np.zeros(500, np.int64)
        L.check(self.lib.dmc_wtisto(self.h, h))
        return h

    def acceptance(self):
        a, p = C.c_int64(), C.c_int64()
        L.check(self.lib.dmc_acceptance(self.h, C.byref(a), C.byref(p)))
        return a.value, p.value

    # --- operators.f90: g(r), density profile, S(q) --------------------------------------
    def addgofr(self, ntab, acc=None):
        """addgofr (gofr.f90:52-98) of every walker of the resident population, added into
        acc = (gnow, gerr [ntab+1][3], dnow, derr [2 ntab+1][6], wsblk [1]); returns acc."""
        if acc is None:
            acc = (np.zeros((ntab + 1, 3)), np.zeros((ntab + 1, 3)),
                   np.zeros((2 * ntab + 1, 6)), np.zeros((2 * ntab + 1, 6)), np.zeros(1))
        L.check(self.lib.dmc_gofr(self.h, int(ntab), *(a.reshape(-1) for a in acc)))
        return acc

    def sofq_kvec(self, ntabq):
        """setupk as setupsofq calls it (sofq.f90:26): (numsh [ntabq], ak [nqtot][3])"""
        nq = C.c_int32()
        L.check(self.lib.dmc_sofq_kvec(self.h, int(ntabq), 0, C.byref(nq), None, None))
        numsh = np.zeros(ntabq, np.int32)
        ak = np.zeros((nq.value, 3))
        L.check(self.lib.dmc_sofq_kvec(self.h, int(ntabq), nq.value, C.byref(nq),
                                       numsh.ctypes.data, ak.ctypes.data))
        return numsh, ak

    def addsofq(self, ntabq, acc=None):
        """addsofq (sofq.f90:40-65) of every walker of the resident population, added into
        acc = (snow, serr [nqtot][3], wsblk [1]); returns acc."""
        if acc is None:
            nq = C.c_int32()
            L.check(self.lib.dmc_sofq_kvec(self.h, int(ntabq), 0, C.byref(nq), None, None))
            acc = (np.zeros((nq.value, 3)), np.zeros((nq.value, 3)), np.zeros(1))
        L.check(self.lib.dmc_sofq(self.h, int(ntabq), *(a.reshape(-1) for a in acc)))
        return acc

    def nofk_shells(self, ntab):
        """setupenk (momentum3d.f90:13-79): (k of each shell [nsh], vectors per shell [nsh])"""
        n = C.c_int32()
        L.check(self.lib.dmc_nofk_shells(self.h, int(ntab), 0, C.byref(n), None, None))
        kmag, nk = np.zeros(n.value), np.zeros(n.value, np.int32)
        L.check(self.lib.dmc_nofk_shells(self.h, int(ntab), n.value, C.byref(n),
                                         kmag.ctypes.data, nk.ctypes.data))
        return kmag, nk

    def addenk(self, ntab, acc=None):
        """addenk (momentum3d.f90:89-120) of every walker of the resident population, added
        into acc = (enknow [nsh], enkerr [nsh], wsblk [1]); returns acc."""
        if acc is None:
            n = len(self.nofk_shells(ntab)[1])
            acc = (np.zeros(n), np.zeros(n), np.zeros(1))
        L.check(self.lib.dmc_nofk(self.h, int(ntab), *acc))
        return acc

    def move_trace(self, enable=True):
        """record every single-particle move of the following steps (dmc_move_trace)"""
        L.check(self.lib.dmc_move_trace(self.h, int(bool(enable))))

    def read_move_trace(self):
        """records of the last propagation [nw][npart][12] (dmc_read_move_trace): ratio,
        Jastrow log-ratio, accepted, draw, proposed position (3), drift (3), log Metropolis
        ratio"""
        nw = C.c_int32()
        L.check(self.lib.dmc_read_move_trace(self.h, 0, C.byref(nw), None))
        rec = np.zeros((nw.value, self.npart, 12))
        L.check(self.lib.dmc_read_move_trace(self.h, nw.value, C.byref(nw), rec.ctypes.data))
        return rec

    def kernel_timing(self, enable=True):
        L.check(self.lib.dmc_kernel_timing(self.h, int(bool(enable))))

    def kernel_times(self, reset=False):
        ms, n = C.c_double(), C.c_int64()
        L.check(self.lib.dmc_kernel_times(self.h, C.byref(ms), C.byref(n), int(bool(reset))))
        return ms.value, n.value

    def sync(self):
        L.check(self.lib.dmc_sync(self.h))


class Batch:
    """Several engines (different particle numbers) stepped together by the device: one
    captured graph per step with one persistent launch per phase over all systems
    (dmc_batch_*).  Steps are asynchronous; sync() returns the populations."""

    def __init__(self, engines):
        self.engines = list(engines)
        self.lib = L.load()
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        L.check(self.lib.dmc_batch_create(arr, len(self.engines), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.dmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, opcomp=False):
        L.check(self.lib.dmc_batch_step(self.h, int(bool(opcomp))))

    def run(self, nsteps):
        L.check(self.lib.dmc_batch_run(self.h, int(nsteps)))

    def sync(self):
        nt = np.zeros(len(self.engines), np.int32)
        L.check(self.lib.dmc_batch_sync(self.h, nt))
        return [int(v) for v in nt]

    def set_wpg(self, moves=1, refresh=1, end=1):
        L.check(self.lib.dmc_batch_set_wpg(self.h, int(moves), int(refresh), int(end)))

    def tune(self, reps=3):
        """Launch setting of every phase from timed launches; returns the choice (walkers
        per workgroup, 0 = one launch per system, -1 = one launch per system of the kernels
        that stage their tables in LDS) and the ms per launch
        [phase][setting = 1, 2, 4, 8, 0, -1] (-1 ms: not run)."""
        w = np.zeros(3, np.int32)
        ms = np.zeros(18)
        L.check(self.lib.dmc_batch_tune(self.h, int(reps), w, ms))
        return dict(zip(("moves", "refresh", "end"), (int(v) for v in w))), ms.reshape(3, 6)

