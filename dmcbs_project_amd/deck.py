"""The reference's input deck and walker file on the host side.

* ``read_deck`` reads the positional stdin deck in the order of the reference's reads:
  the run header (dmcbcs.f90:43-66), setpsi (wavefunction.f90:36-37 and the BCS deck of
  psislaterbcs.f90:47-68 and betafun.f90:19-21), setphi (coshpot.f90:35-41), the four
  operator table sizes (operators.f90:33-41) and, for optimisation runs, setupopt
  (optimizer.f90:23-27).  It returns the engine's ``Params`` and the run settings; deck
  options the device path does not support raise ``ValueError``, as the Fortran host stops
  on them.
* ``read_walkers`` / ``write_walkers`` read and write the walker file fort.9
  (dmcbcs.f90:253-293, 548-567): n and tau as (i10,f15.8), then for every walker x as
  (6e15.7), weight, psil, psigl as (e15.7,2e25.15) and the RNG state as (i20).
"""
import numpy as np

from . import lib as L
from .api import Params
from .observables import fortran_e


def _lines(text):
    for line in text.splitlines():
        yield line


def _real(tok):
    return float(tok.replace("d", "e").replace("D", "E"))


def _logical(tok):
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise ValueError(f"not a logical: {tok!r}")


class _Reader:
    """List-directed reads of one record per line: the first tokens of the line are the
    values, the rest of the line is ignored."""

    def __init__(self, text):
        self.it = _lines(text)

    def tokens(self, n):
        line = next(self.it)
        tok = line.split()
        if len(tok) < n:
            raise ValueError(f"deck line {line!r}: {n} values expected")
        return tok[:n]

    def int(self):
        return int(self.tokens(1)[0])

    def real(self):
        return _real(self.tokens(1)[0])

    def logical(self):
        return _logical(self.tokens(1)[0])

    def word(self):
        return next(self.it)[:8].strip()


def read_deck(text, **device):
    """The reference's deck -> (Params, run settings).  device: Params fields the deck does
    not hold (precision, movetype, krefresh, capacity); variational decks (idmc = 0, -1) run
    the FP64 Metropolis step."""
    r = _Reader(text)
    run = {}
    run["seed"] = r.int()
    run["idmc"] = r.int()
    run["isite"] = r.logical()
    nwalk = r.int()
    run["neq"] = r.int()
    run["nav"] = r.int()
    run["nstep"] = r.int()
    dt = r.real()
    etrial = r.real()
    units = r.logical()
    dcut = r.real()
    fn = r.logical()
    popup = r.real()
    popdn = r.real()
    run["dobal"] = r.real()
    wtcut = r.real()
    ntab = r.int()
    hbar = r.real()
    ndim = r.int()
    npart = r.int()
    n1, m1 = r.tokens(2)
    n2, m2 = r.tokens(2)
    nbin, mass = (int(n1), int(n2)), (_real(m1), _real(m2))
    if ndim != 3:
        raise ValueError("three dimensions only")
    if nbin[0] != nbin[1] or mass[0] != mass[1] or sum(nbin) != npart:
        raise ValueError("unpolarised equal masses only")
    # setpsi
    nodes = r.word()
    a2 = r.real()
    if nodes != "bcs":
        raise ValueError("only bcs nodes are supported")
    if a2 != 0.0:
        raise ValueError("a guiding function (a2 /= 0) is not supported")
    rho = r.real()
    if r.int() != 0 or r.int() != 0:
        raise ValueError("unpaired orbitals are not supported")
    norb = r.int()
    if not 1 <= norb <= 16:
        raise ValueError("1 to 16 pairing shells")
    vorb = tuple(r.real() for _ in range(norb))
    nqex = [int(v) for v in r.tokens(3)]
    v0ex, v0exjas, numq = r.real(), r.real(), r.int()
    r.real(), r.real()  # long-range Jastrow beta and gamma
    if any(nqex) or v0ex != 0.0 or v0exjas != 0.0 or numq != 0:
        raise ValueError("external potential and long-range Jastrow are not supported")
    bbeta, dbeta, bfact = r.real(), r.real(), r.real()
    if bfact < 0.0:
        raise ValueError("beta tables from file (bfact < 0) are not supported")
    # setphi
    v0, amu, cjas1, cjas2, djas, v0op, v0eq = (r.real() for _ in range(7))
    if cjas2 != 0.0:
        raise ValueError("equal-spin Jastrow (cjas(2) /= 0) is not supported")
    # setoperators
    run["ntabg"], run["ntabq"], run["ntabk"], ntabqeuc = (r.int() for _ in range(4))
    if ntabqeuc > 0:
        raise ValueError("the euclidean density response is not supported")
    if run["idmc"] in (-1, -2):
        run["optiter"] = r.int()
        run["optjas"] = r.logical()
        run["optnodes"] = r.logical()
        run["delta"] = r.real()
        run["eps"] = r.real()
    if run["idmc"] in (0, -1):
        # variational runs (step1 with idmc = 0, step.f90:78-122), as the Fortran host maps them
        device = dict(device, movetype=L.DMC_VMC, precision=L.DMC_FP64)
    elif run["idmc"] == 2:
        # Metropolis DMC (step1 with idmc = 2, step.f90:161-226)
        device = dict(device, movetype=L.DMC_METROPOLIS, precision=L.DMC_FP64)
    elif run["idmc"] != 1:
        raise ValueError("the engine runs DMC (idmc = 1, 2) and VMC (idmc = 0, -1)")
    p = Params(npart=npart, norb=norb, ntab=ntab, fn=fn, units=units, nwalk=nwalk, rho=rho,
               hbar=hbar, mass=mass[0], dt=dt, etrial=etrial, dcut=dcut, wtcut=wtcut,
               popup=popup, popdn=popdn, bbeta=bbeta, dbeta=dbeta, bfact=bfact, v0=v0,
               amu=amu, cjas1=cjas1, cjas2=cjas2, djas=djas, v0op=v0op, v0eq=v0eq,
               vorb=vorb, **device)
    return p, run


def write_walkers(path, x, weight, psil, irn, tau, psigl=None):
    """fort.9 of dmcbcs.f90:548-567: x [nw][npart][3]."""
    x = np.asarray(x, np.float64)
    psigl = psil if psigl is None else psigl
    with open(path, "w") as fh:
        fh.write(f"{len(x):10d}{tau:15.8f}\n")
        for k in range(len(x)):
            flat = x[k].reshape(-1)
            for i in range(0, len(flat), 6):
                fh.write("".join(fortran_e(v) for v in flat[i:i + 6]) + "\n")
            fh.write(fortran_e(weight[k]) + fortran_e(psil[k], 25, 15) +
                     fortran_e(psigl[k], 25, 15) + "\n")
            fh.write(f"{int(irn[k]):20d}\n")


def read_walkers(path, npart):
    """fort.9 of dmcbcs.f90:253-293 -> (x [nw][npart][3], weight, psil, psigl, irn, tau)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    n, tau = int(lines[0][:10]), float(lines[0][10:25])
    per = (3 * npart + 5) // 6
    x = np.zeros((n, npart, 3))
    w, pl, pg = np.zeros(n), np.zeros(n), np.zeros(n)
    irn = np.zeros(n, np.int64)
    k = 1
    for j in range(n):
        vals = []
        for line in lines[k:k + per]:
            vals += [_real(line[i:i + 15]) for i in range(0, len(line.rstrip()), 15)]
        x[j] = np.array(vals[:3 * npart]).reshape(npart, 3)
        k += per
        line = lines[k]
        w[j], pl[j], pg[j] = _real(line[:15]), _real(line[15:40]), _real(line[40:65])
        irn[j] = int(lines[k + 1][:20])
        k += 2
    return x, w, pl, pg, irn, tau
