"""ctypes binding of the oracle (the CPU restatement in dmc_oracle.c).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this package, and only as the checker or the CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libdmc_oracle.so")
MAXORB = 16

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("npart", C.c_int32), ("norb", C.c_int32), ("ntab", C.c_int32),
                ("fn", C.c_int32), ("units", C.c_int32), ("nwalk", C.c_int32)] + [
        (k, C.c_double) for k in ("rho", "hbar", "mass", "dt", "etrial", "dcut", "wtcut",
                                  "popup", "popdn", "bbeta", "dbeta", "bfact", "v0", "amu",
                                  "cjas1", "cjas2", "djas", "v0op", "v0eq")] + [
        ("vorb", C.c_double * MAXORB)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_planted = {}


def planted(k):
    """The oracle built with planted mistake k (DMCO_PLANT in dmc_oracle.c), for the tests
    that show a check catches that mistake"""
    if k not in _planted:
        subprocess.run(["make", "-s", "-C", HERE, f"_build/libdmc_oracle_plant{k}.so"], check=True)
        _planted[k] = _bind(C.CDLL(os.path.join(HERE, "_build", f"libdmc_oracle_plant{k}.so")))
    return _planted[k]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


def _bind(L):
    """argument and result types of the oracle's functions"""
    L.dmco_create.restype = C.c_void_p
    L.dmco_create.argtypes = [C.POINTER(Params)]
    L.dmco_destroy.argtypes = [C.c_void_p]
    L.dmco_info.argtypes = [C.c_void_p, _f64p]
    L.dmco_kvec.argtypes = [C.c_void_p, _f64p, _f64p, _i32p]
    L.dmco_kvec.restype = C.c_int
    L.dmco_tables.argtypes = [C.c_void_p] + [_f64p] * 7
    for f in ("dmco_radial", "dmco_beta"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_double, _f64p]
    L.dmco_derbeta.argtypes = [C.c_void_p, C.c_double, _f64p]
    L.dmco_vpot.argtypes = [C.c_void_p, C.c_double]
    L.dmco_vpot.restype = C.c_double
    L.dmco_getphi.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.dmco_bcs.argtypes = [C.c_void_p, _f64p, _f64p, _i32p, _f64p, _f64p, _f64p]
    L.dmco_hpsi.argtypes = [C.c_void_p, _f64p, _f64p, _f64p]
    L.dmco_ratio1.argtypes = [C.c_void_p, _f64p, C.c_int32, _f64p]
    L.dmco_ratio1.restype = C.c_double
    L.dmco_getder.argtypes = [C.c_void_p, _f64p, C.c_int32, _f64p]
    L.dmco_set_params.argtypes = [C.c_void_p, _f64p, C.c_double, C.c_double]
    L.dmco_ran1.argtypes = [C.POINTER(C.c_int64)]
    L.dmco_ran1.restype = C.c_double
    L.dmco_ran2.argtypes = [C.POINTER(C.c_int64)]
    L.dmco_ran2.restype = C.c_double
    L.dmco_gaussian.argtypes = [C.POINTER(C.c_int64), C.c_int32, _f64p]
    L.dmco_init_walkers.argtypes = [C.c_void_p, C.c_int64, C.c_int32, _f64p, _i64p]
    L.dmco_init_hpsi.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p]
    L.dmco_step1.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                             C.c_int32, _f64p, _f64p, _f64p, _f64p, _i64p, _f64p,
                             C.c_int32]
    L.dmco_vmc_step1.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                                 _f64p, _f64p, _f64p, _f64p, _i64p, C.c_int32]
    L.dmco_branch.argtypes = [C.c_int32, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                              C.c_double, C.c_int32, _f64p, _f64p, _f64p, _i64p]
    L.dmco_branch.restype = C.c_int
    L.dmco_dmc_step.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                                C.c_int32, _f64p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                                _i32p, C.c_int32]
    L.dmco_dmc_step.restype = C.c_int
    L.dmco_mdmc_step.argtypes = L.dmco_dmc_step.argtypes
    L.dmco_mdmc_step.restype = C.c_int
    L.dmco_sm_step1.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                                C.c_int32, _f64p, _f64p, _f64p, _i64p, _f64p, _i32p,
                                C.c_int32]
    L.dmco_sm_step1_trace.argtypes = [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _i64p,
                                      C.c_int32, _f64p, _f64p, _f64p, _i64p, _f64p, _i32p,
                                      _f64p, C.c_int32]
    L.dmco_kshells.argtypes = [C.c_double, C.c_int32, _i32p, _f64p, C.c_int32]
    L.dmco_kshells.restype = C.c_int
    L.dmco_addgofr.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_int32, _f64p,
                               C.c_double, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.dmco_addsofq.argtypes = [C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p, C.c_double,
                               _f64p, _f64p, _f64p]
    L.dmco_enk_shells.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _i32p,
                                  C.POINTER(C.c_int32)]
    L.dmco_enk_shells.restype = C.c_int
    L.dmco_addenk.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p, _i32p, _f64p, C.c_int64,
                              C.c_double, _f64p, _f64p, _f64p]
    return L


def enk_shells(ntab, maxk=8192):
    """setupenk (momentum3d.f90:13-79): (ik [ntot][3], indk [nsh], nk [nsh])"""
    ik = np.zeros(3 * maxk, np.int32)
    indk = np.zeros(maxk, np.int32)
    nk = np.zeros(maxk, np.int32)
    ntot = C.c_int32()
    nsh = lib().dmco_enk_shells(int(ntab), int(maxk), ik, indk, nk, C.byref(ntot))
    if nsh < 0:
        raise ValueError("too many n(k) vectors")
    return ik[: 3 * ntot.value].reshape(-1, 3).copy(), indk[:nsh].copy(), nk[:nsh].copy()


def kshells(el, nsh, maxk=4096):
    """setupk (kshell.f90:6-156): (numsh[nsh], ak[nk][3]) of the first nsh shells"""
    numsh = np.zeros(nsh, np.int32)
    ak = np.zeros(3 * maxk)
    nk = lib().dmco_kshells(float(el), int(nsh), numsh, ak, int(maxk))
    if nk < 0:
        raise ValueError(f"setupk failed ({nk})")
    return numsh, ak[: 3 * nk].reshape(nk, 3).copy()


def addgofr(configs, weights, el, ntab, nup=None):
    """addgofr (gofr.f90:52-98) over configurations [nc][npart][3] with their weights:
    the accumulators gnow, gerr [ntab+1][3], dnow, derr [2 ntab+1][6] and wsblk"""
    configs = np.ascontiguousarray(configs, np.float64)
    npart = configs.shape[1]
    nup = npart // 2 if nup is None else nup
    g, ge = np.zeros((ntab + 1) * 3), np.zeros((ntab + 1) * 3)
    d, de = np.zeros((2 * ntab + 1) * 6), np.zeros((2 * ntab + 1) * 6)
    ws = np.zeros(1)
    for x, w in zip(configs, weights):
        lib().dmco_addgofr(npart, nup, float(el), int(ntab), np.ascontiguousarray(x).ravel(),
                           float(w), g, ge, d, de, ws)
    return (g.reshape(-1, 3), ge.reshape(-1, 3), d.reshape(-1, 6), de.reshape(-1, 6),
            float(ws[0]))


def addsofq(configs, weights, ak, nup=None):
    """addsofq (sofq.f90:40-65) over configurations with their weights: snow, serr
    [nq][3] and wsblk"""
    configs = np.ascontiguousarray(configs, np.float64)
    npart = configs.shape[1]
    nup = npart // 2 if nup is None else nup
    ak = np.ascontiguousarray(ak, np.float64)
    nq = ak.shape[0]
    sn, se, ws = np.zeros(3 * nq), np.zeros(3 * nq), np.zeros(1)
    for x, w in zip(configs, weights):
        lib().dmco_addsofq(npart, nup, nq, ak.ravel(), np.ascontiguousarray(x).ravel(),
                           float(w), sn, se, ws)
    return sn.reshape(nq, 3), se.reshape(nq, 3), float(ws[0])


DEFAULTS = dict(npart=14, norb=3, ntab=1024, fn=1, units=1, nwalk=1000, rho=0.0337737278,
                hbar=0.5, mass=1.0, dt=0.002, etrial=0.40, dcut=-1.0, wtcut=10.0, popup=1.1,
                popdn=0.9, bbeta=0.5, dbeta=5.0, bfact=1.0, v0=1.0, amu=40.0, cjas1=1.0,
                cjas2=0.0, djas=1.15, v0op=0.0, v0eq=0.0)

# s-wave pairing amplitudes of the production deck (dmcbcs.dat:29-38)
VORB10 = [0.00197519728006648144, 0.00249764752028349595, 0.00194383232328550247,
          0.00081339556934447224, 0.00033078733782252270, 0.00019008071412562492,
          0.00019983385861822818, 0.00016699298789053479, 0.00016360289711864207,
          0.00012098992317746450]


class System:
    """One system (deck) of the oracle."""

    def __init__(self, vorb=None, lib_=None, **kw):
        p = dict(DEFAULTS)
        p.update(kw)
        self.params = p
        P = Params()
        for k, v in p.items():
            setattr(P, k, v)
        vo = list(vorb) if vorb is not None else VORB10[: p["norb"]]
        for i, v in enumerate(vo):
            P.vorb[i] = v
        self.vorb = np.array(vo, dtype=np.float64)
        self.L = lib_ if lib_ is not None else lib()
        self.h = self.L.dmco_create(C.byref(P))
        if not self.h:
            raise RuntimeError("dmco_create failed")
        info = np.zeros(8)
        self.L.dmco_info(self.h, info)
        self.el, self.eunit, self.etrial, self.nfac = info[:4]
        self.nk, self.ntab0, self.sigma, self.scale = int(info[4]), int(info[5]), info[6], info[7]
        self.npart = p["npart"]
        self.n = self.npart // 2

    def __del__(self):
        if getattr(self, "h", None):
            self.L.dmco_destroy(self.h)
            self.h = None

    def kvec(self):
        ak = np.zeros(3 * 512)
        ak2 = np.zeros(512)
        numsh = np.zeros(MAXORB, np.int32)
        nk = self.L.dmco_kvec(self.h, ak, ak2, numsh)
        return ak[: 3 * nk].reshape(nk, 3), ak2[:nk], numsh[: self.params["norb"]]

    def tables(self):
        nt = self.params["ntab"] + 1
        t = [np.zeros(nt) for _ in range(7)]
        self.L.dmco_tables(self.h, *t)
        return dict(zip(("v", "u", "du", "d2u", "b", "db", "d2b"), t))

    def radial(self, r):
        o = np.zeros(3)
        self.L.dmco_radial(self.h, float(r), o)
        return o

    def beta(self, r):
        o = np.zeros(3)
        self.L.dmco_beta(self.h, float(r), o)
        return o

    def derbeta(self, r):
        o = np.zeros(2)
        self.L.dmco_derbeta(self.h, float(r), o)
        return o

    def vpot(self, r):
        return self.L.dmco_vpot(self.h, float(r))

    def getphi(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        u, d2u, v = np.zeros(1), np.zeros(1), np.zeros(1)
        du = np.zeros(3 * self.npart)
        self.L.dmco_getphi(self.h, x, u, du, d2u, v)
        return u[0], du.reshape(-1, 3), d2u[0], v[0]

    def bcs(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        phil, d2phi = np.zeros(1), np.zeros(1)
        is_ = np.zeros(1, np.int32)
        dphi = np.zeros(3 * self.npart)
        sinv = np.zeros(self.n * self.n)
        self.L.dmco_bcs(self.h, x, phil, is_, dphi, d2phi, sinv)
        return phil[0], int(is_[0]), dphi.reshape(-1, 3), d2phi[0], sinv.reshape(self.n, self.n).T

    def hpsi(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        o = np.zeros(4)
        dpsi = np.zeros(3 * self.npart)
        self.L.dmco_hpsi(self.h, x, o, dpsi)
        return dict(psil=o[0], d2psi=o[1], v=o[2], is_=int(o[3]), dpsi=dpsi.reshape(-1, 3))

    def ratio1(self, x, ip, xnew):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        return self.L.dmco_ratio1(self.h, x, int(ip), np.ascontiguousarray(xnew, np.float64))

    def getder(self, x, optbeta=False):
        """d log psi / d (alpha_1..alpha_norb[, b, d]) (getderbcs)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        o = np.zeros(self.params["norb"] + (2 if optbeta else 0))
        self.L.dmco_getder(self.h, x, int(bool(optbeta)), o)
        return o

    def set_params(self, vorb, bbeta, dbeta):
        self.vorb = np.ascontiguousarray(vorb, np.float64)
        self.L.dmco_set_params(self.h, self.vorb, float(bbeta), float(dbeta))

    def addenk(self, configs, irns, weights, ntab):
        """addenk (momentum3d.f90:89-120) over configurations with their RNG states and
        weights: (enknow [nsh], enkerr [nsh], wsblk)"""
        ik, indk, nk = enk_shells(ntab)
        nsh = len(nk)
        en, ee, ws = np.zeros(nsh), np.zeros(nsh), np.zeros(1)
        for x, irn, w in zip(configs, irns, weights):
            self.L.dmco_addenk(self.h, nsh, ik.ravel(), indk, nk,
                               np.ascontiguousarray(x, np.float64).ravel(), int(irn), float(w),
                               en, ee, ws)
        return en, ee, float(ws[0])

    def init_walkers(self, seed, nwalk):
        x = np.zeros(nwalk * 3 * self.npart)
        irn = np.zeros(nwalk, np.int64)
        self.L.dmco_init_walkers(self.h, int(seed), int(nwalk), x, irn)
        return x.reshape(nwalk, self.npart, 3), irn

    def init_hpsi(self, x):
        nw = len(x)
        x = np.ascontiguousarray(x, np.float64)
        dpsi = np.zeros_like(x)
        sc = np.zeros((nw, 6))
        self.L.dmco_init_hpsi(self.h, nw, x.reshape(-1), dpsi.reshape(-1), sc.reshape(-1))
        return dpsi, sc

    def step1(self, x, dpsi, sc, irn, opcomp=False, est=None, nthreads=0):
        nw = len(x)
        est = np.zeros(42) if est is None else est
        xo, do = np.zeros_like(x), np.zeros_like(dpsi)
        so, ro, wt = np.zeros_like(sc), np.zeros_like(irn), np.zeros(nw)
        self.L.dmco_step1(self.h, nw, np.ascontiguousarray(x).reshape(-1),
                          np.ascontiguousarray(dpsi).reshape(-1), np.ascontiguousarray(sc).reshape(-1),
                          np.ascontiguousarray(irn), int(opcomp), est, xo.reshape(-1),
                          do.reshape(-1), so.reshape(-1), ro, wt, nthreads)
        return xo, do, so, ro, wt, est

    def vmc_step1(self, x, dpsi, sc, irn, est=None, nthreads=0):
        """step1 with idmc = 0 (Metropolis); returns x, dpsi, sc, irn, est."""
        nw = len(x)
        est = np.zeros(42) if est is None else est
        xo, do = np.zeros_like(x), np.zeros_like(dpsi)
        so, ro = np.zeros_like(sc), np.zeros_like(irn)
        self.L.dmco_vmc_step1(self.h, nw, np.ascontiguousarray(x).reshape(-1),
                              np.ascontiguousarray(dpsi).reshape(-1),
                              np.ascontiguousarray(sc).reshape(-1), np.ascontiguousarray(irn),
                              est, xo.reshape(-1), do.reshape(-1), so.reshape(-1), ro, nthreads)
        return xo, do, so, ro, est

    def dmc_step(self, x, dpsi, sc, irn, opcomp=False, est=None, cap=None, nthreads=0,
                 metropolis=False):
        """step1 + branch + population control; metropolis: step1 with idmc = 2"""
        nw = len(x)
        cap = cap or max(4 * nw, 64)
        est = np.zeros(42) if est is None else est
        xo = np.zeros((cap, self.npart, 3))
        do = np.zeros_like(xo)
        so = np.zeros((cap, 6))
        ro = np.zeros(cap, np.int64)
        nt = np.zeros(2, np.int32)
        fn = self.L.dmco_mdmc_step if metropolis else self.L.dmco_dmc_step
        rc = fn(self.h, nw, np.ascontiguousarray(x).reshape(-1),
                                  np.ascontiguousarray(dpsi).reshape(-1),
                                  np.ascontiguousarray(sc).reshape(-1), np.ascontiguousarray(irn),
                                  int(opcomp), est, cap, xo.reshape(-1), do.reshape(-1),
                                  so.reshape(-1), ro, nt, nthreads)
        if rc:
            raise RuntimeError("population overflow")
        m = int(nt[1])
        return xo[:m], do[:m], so[:m], ro[:m], (int(nt[0]), int(nt[1])), est

    def sm_step1(self, x, dpsi, sc, irn, krefresh, nthreads=0):
        nw = len(x)
        xo, do = np.zeros_like(x), np.zeros_like(dpsi)
        so, ro, wt = np.zeros_like(sc), np.zeros_like(irn), np.zeros(nw)
        nacc = np.zeros(nw, np.int32)
        self.L.dmco_sm_step1(self.h, nw, np.ascontiguousarray(x).reshape(-1),
                             np.ascontiguousarray(dpsi).reshape(-1),
                             np.ascontiguousarray(sc).reshape(-1), np.ascontiguousarray(irn),
                             int(krefresh), xo.reshape(-1), do.reshape(-1), so.reshape(-1), ro,
                             wt, nacc, nthreads)
        return xo, do, so, ro, wt, nacc

    def sm_step1_trace(self, x, dpsi, sc, irn, krefresh, nthreads=0):
        """sm_step1 plus the record of every move, [nw][npart][TRACE] (dmc_oracle.h)"""
        nw = len(x)
        xo, do = np.zeros_like(x), np.zeros_like(dpsi)
        so, ro, wt = np.zeros_like(sc), np.zeros_like(irn), np.zeros(nw)
        nacc = np.zeros(nw, np.int32)
        tr = np.zeros(nw * self.npart * SM_TRACE)
        self.L.dmco_sm_step1_trace(self.h, nw, np.ascontiguousarray(x).reshape(-1),
                                   np.ascontiguousarray(dpsi).reshape(-1),
                                   np.ascontiguousarray(sc).reshape(-1),
                                   np.ascontiguousarray(irn), int(krefresh), xo.reshape(-1),
                                   do.reshape(-1), so.reshape(-1), ro, wt, nacc, tr, nthreads)
        return xo, do, so, ro, wt, nacc, tr.reshape(nw, self.npart, SM_TRACE)


SM_TRACE = 24  # DMCO_SM_TRACE


def calder_updatepar(system, configs, delta, eps, optbeta=False):
    """One optimisation update of the determinant parameters (optimizer.f90:90-169): calder
    on every configuration (E_L per particle in units of eunit, d log psi from getderbcs),
    then updatepar: f = -2<O E> + 2<O><E>, S = <O O> - <O><O> + eps, new = old + delta S^-1 f.
    Returns (<E>, new parameters)."""
    p = system.params
    hbar, npart = p["hbar"], system.npart
    par = np.concatenate([system.vorb, [p["bbeta"], p["dbeta"]]]) if optbeta else system.vorb.copy()
    npar = len(par)
    ovec, ovecw = np.zeros(npar), np.zeros(npar)
    pmat = np.zeros((npar, npar))
    esum = esum2 = 0.0
    for x in configs:
        o = system.hpsi(x)
        ene = (-hbar * o["d2psi"] + o["v"]) / system.eunit / npart
        d = system.getder(x, optbeta=optbeta)
        ovec += d
        esum += ene
        esum2 += ene * ene
        ovecw += d * ene
        pmat += np.outer(d, d)
    n = len(configs)
    ovec, ovecw, pmat, esum = ovec / n, ovecw / n, pmat / n, esum / n
    fvec = -2.0 * ovecw + 2.0 * ovec * esum
    pmat = pmat - np.outer(ovec, ovec) + eps * np.eye(npar)
    dpar = np.linalg.inv(pmat) @ fvec
    return esum, par + delta * dpar


def ran1(s):
    st = C.c_int64(int(s))
    u = lib().dmco_ran1(C.byref(st))
    return st.value, u


def ran2(s):
    st = C.c_int64(int(s))
    u = lib().dmco_ran2(C.byref(st))
    return st.value, u


def gaussian(s, n):
    st = C.c_int64(int(s))
    g = np.zeros(n)
    lib().dmco_gaussian(C.byref(st), int(n), g)
    return st.value, g
