"""Block normalisation and output tables of the observables g(r), density profile and S(q).

The sums are accumulated on the device (``Engine.addgofr`` and ``Engine.addsofq``, the
reference's addgofr gofr.f90:52-98 and addsofq sofq.f90:40-65 over the whole population).
What follows is the host arithmetic of the reference modules:

* ``updategofr`` / ``gofr_rows`` / ``rho_rows`` follow gofr.f90:100-177;
* ``updatesofq`` / ``sofq_rows`` follow sofq.f90:67-119;
* ``updateenk`` / ``nofk_rows`` follow momentum3d.f90:149-190 (n(k), ``Engine.addenk``);
* ``Gofr``, ``Sofq`` and ``Nofk`` keep the module state (setup, zero, add, update, write) as
  operators.f90 drives it once per block (dmcbcs.f90:465-478).

Tables are written in the reference's text format (``e15.7`` columns, appended).
"""
import math

import numpy as np


def fortran_e(v, w=15, d=7):
    """Fortran ``Ew.d`` edit descriptor: 0.ddddddd E+xx, right-justified in w columns."""
    v = float(v)
    if v == 0.0 or not math.isfinite(v):
        body = "0." + "0" * d + "E+00" if v == 0.0 else str(v)
        return body.rjust(w)
    s = f"{abs(v):.{d - 1}E}"
    mant, ex = s.split("E")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    sign = "-" if v < 0 else ""
    exp = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) < 100 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return (sign + "0." + digits + exp).rjust(w)


def _write(filename, header, rows, totalw, tau):
    with open(filename, "a") as fh:
        fh.write(header + "\n")
        fh.write("# total weight = " + fortran_e(totalw) + "\n")
        fh.write("# tau = " + fortran_e(tau) + "\n")
        for r in rows:
            fh.write("".join(fortran_e(v) for v in r) + "\n")
        fh.write(" \n \n")


# ---------------------------------------------------------------------------------- g(r)
def updategofr(gnow, gerr, dnow, derr, wsblk):
    """gofr.f90:100-134: block averages and errors; returns (g, ge, d, de, totalw)."""
    norm = 1.0 / wsblk
    g = np.asarray(gnow) * norm
    e2 = np.asarray(gerr) * norm
    ge = np.sqrt(np.abs(e2 - g ** 2) / wsblk)
    d = np.asarray(dnow) * norm
    de2 = np.asarray(derr) * norm
    de = np.sqrt(np.abs(de2 - d ** 2) / wsblk)
    return g, ge, d, de, wsblk


def gofr_rows(g, ge, npart, el, ntab):
    """gofr.f90:148-158 (ndim = 3): rows r, then g and its error for up-down, up-up,
    down-down, each divided by rho * npart * (shell volume)."""
    pi = 4.0 * math.atan(1.0)
    dr = 0.5 * el / ntab
    rho = npart / el ** 3
    rows = np.zeros((ntab, 7))
    for i in range(ntab):
        r0, r1 = i * dr, (i + 1) * dr
        vol = 4.0 * pi * (r1 ** 3 - r0 ** 3) / 3.0
        facl = 1.0 / (rho * npart * vol)
        rows[i, 0] = (i + 0.5) * dr
        for c in range(3):
            rows[i, 1 + 2 * c] = g[i, c] * facl
            rows[i, 2 + 2 * c] = ge[i, c] * facl
    return rows


def rho_rows(d, de, npart, el, ntab):
    """gofr.f90:166-173: rows x, then the x, y, z profiles of spin up and spin down with
    their errors (columns x-up, x-down, y-up, y-down, z-up, z-down), divided by npart."""
    dr = 0.5 * el / ntab
    facl = 1.0 / npart
    rows = np.zeros((2 * ntab + 1, 13))
    for b in range(2 * ntab + 1):
        rows[b, 0] = (b - ntab) * dr
        for k, c in enumerate((0, 3, 1, 4, 2, 5)):
            rows[b, 1 + 2 * k] = d[b, c] * facl
            rows[b, 2 + 2 * k] = de[b, c] * facl
    return rows


class Gofr:
    """module gofr (gofr.f90) over a device population: setup, zero, add, update, write."""

    def __init__(self, engine, ntab):
        self.engine, self.ntab = engine, int(ntab)
        self.npart, self.el = engine.npart, engine.el
        self.zero()

    def zero(self):
        n = self.ntab
        self.acc = (np.zeros((n + 1, 3)), np.zeros((n + 1, 3)), np.zeros((2 * n + 1, 6)),
                    np.zeros((2 * n + 1, 6)), np.zeros(1))
        self.result = None

    def add(self):
        """addgofr over every walker of the engine's population."""
        self.engine.addgofr(self.ntab, self.acc)

    def update(self):
        gnow, gerr, dnow, derr, ws = self.acc
        self.result = updategofr(gnow, gerr, dnow, derr, float(ws[0]))
        self.zero_sums()
        return self.result

    def zero_sums(self):
        for a in self.acc:
            a[...] = 0.0

    def write(self, filename="gofr.out", tau=0.0, rhofile="rho.out"):
        g, ge, d, de, totalw = self.result
        _write(filename, "# r*kF, g(r), err", gofr_rows(g, ge, self.npart, self.el, self.ntab),
               totalw, tau)
        _write(rhofile, "# r, xrhoup, err, xrhodn, err, yrhoup, err, yrhodn, err, zrhoup, err, "
               "zrhodn, err", rho_rows(d, de, self.npart, self.el, self.ntab), totalw, tau)


# ---------------------------------------------------------------------------------- S(q)
def updatesofq(snow, serr, wsblk):
    """sofq.f90:67-89: returns (s, se, totalw)."""
    norm = 1.0 / wsblk
    s = np.asarray(snow) * norm
    e2 = np.asarray(serr) * norm
    return s, np.sqrt(np.abs(e2 - s ** 2) / wsblk), wsblk


def sofq_rows(s, se, numsh, ak, npart, kf):
    """sofq.f90:102-115: one row per shell, q/kF then the shell averages of S_upup, S_dndn,
    S_updn with their errors, divided by npart."""
    rows = np.zeros((len(numsh), 7))
    k = 0
    for iq, n in enumerate(numsh):
        ave, err = np.zeros(3), np.zeros(3)
        for _ in range(int(n)):
            ave += s[k]
            err += se[k]
            k += 1
        ave = ave / n / npart
        err = err / n / npart
        q = math.sqrt(float(np.sum(ak[k - 1] ** 2)))
        rows[iq, 0] = q / kf
        rows[iq, 1::2] = ave
        rows[iq, 2::2] = err
    return rows


def kfermi(npart, el, units=True):
    """k_F of dmcbcs.f90:148-150 for ndim = 3 (1 without units)."""
    if not units:
        return 1.0
    rho = npart / el ** 3
    return (3.0 * math.pi ** 2 * rho) ** (1.0 / 3.0)


class Sofq:
    """module sofq (sofq.f90) over a device population: setup, zero, add, update, write."""

    def __init__(self, engine, ntabq):
        self.engine, self.ntabq = engine, int(ntabq)
        self.npart = engine.npart
        self.numsh, self.ak = engine.sofq_kvec(self.ntabq)
        self.kf = kfermi(engine.npart, engine.el, engine.params.units)
        self.zero()

    def zero(self):
        nq = len(self.ak)
        self.acc = (np.zeros((nq, 3)), np.zeros((nq, 3)), np.zeros(1))
        self.result = None

    def add(self):
        """addsofq over every walker of the engine's population."""
        self.engine.addsofq(self.ntabq, self.acc)

    def update(self):
        snow, serr, ws = self.acc
        self.result = updatesofq(snow, serr, float(ws[0]))
        for a in self.acc:
            a[...] = 0.0
        return self.result

    def write(self, filename="sofq.out", tau=0.0):
        s, se, totalw = self.result
        _write(filename, "# q/kF, s(q), err",
               sofq_rows(s, se, self.numsh, self.ak, self.npart, self.kf), totalw, tau)


# ---------------------------------------------------------------------------------- n(k)
def updateenk(enknow, enkerr, wsblk):
    """momentum3d.f90:149-171: returns (enk, enke, totalw)."""
    norm = 1.0 / wsblk
    e = np.asarray(enknow) * norm
    e2 = np.asarray(enkerr) * norm
    return e, np.sqrt(np.abs(e2 - e ** 2) / wsblk), wsblk


def nofk_rows(enk, enke, kmag):
    """momentum3d.f90:184-187: rows k (the magnitude of the shell's first vector; the header
    says k/kF but the reference writes k), n(k), err."""
    return np.stack([np.asarray(kmag), np.asarray(enk), np.asarray(enke)], axis=1)


class Nofk:
    """module pdistribution (momentum3d.f90) over a device population: setup, zero, add,
    update, write."""

    def __init__(self, engine, ntab):
        self.engine, self.ntab = engine, int(ntab)
        self.kmag, self.nk = engine.nofk_shells(self.ntab)
        self.zero()

    def zero(self):
        n = len(self.nk)
        self.acc = (np.zeros(n), np.zeros(n), np.zeros(1))
        self.result = None

    def add(self):
        """addenk over every walker of the engine's population."""
        self.engine.addenk(self.ntab, self.acc)

    def update(self):
        enknow, enkerr, ws = self.acc
        self.result = updateenk(enknow, enkerr, float(ws[0]))
        for a in self.acc:
            a[...] = 0.0
        return self.result

    def write(self, filename="nofk.out", tau=0.0):
        enk, enke, totalw = self.result
        _write(filename, "# k/kF, n(k), err", nofk_rows(enk, enke, self.kmag), totalw, tau)
