"""Input decks in the reference's positional stdin format.

The order of the lines follows the reads of the reference: dmcbcs.f90:43-66 (run header),
wavefunction.f90:36-37 (nodes, a2), psislaterbcs.f90:47-68 (pairing), betafun.f90:19-21
(beta), coshpot.f90:35-41 (interaction/Jastrow), operators.f90:33-41 (four table sizes) and,
for optimisation runs, optimizer.f90:23-27.
"""

# s-wave pairing amplitudes alpha_1..alpha_10 of the production deck (dmcbcs.dat:29-38)
VORB10 = [
    0.00197519728006648144, 0.00249764752028349595, 0.00194383232328550247,
    0.00081339556934447224, 0.00033078733782252270, 0.00019008071412562492,
    0.00019983385861822818, 0.00016699298789053479, 0.00016360289711864207,
    0.00012098992317746450,
]

DEFAULTS = dict(
    seed=1, idmc=1, isite=True, nwalk=1000, neq=0, nav=5, nstep=100, dt=0.002,
    etrial=0.40, units=True, dcut=-1.0, fn=True, popup=1.1, popdn=0.9, dobal=0.5,
    wtcut=10.0, ntab=1024, hbar=0.5, ndim=3, npart=14, mass=1.0, a2=0.0,
    rho=0.0337737278, norb=3, vorb=None, bbeta=0.5, dbeta=5.0, bfact=1.0, v0=1.0,
    amu=40.0, cjas1=1.0, cjas2=0.0, djas=1.15, v0op=0.0, v0eq=0.0,
    optiter=1, optjas=False, optnodes=True, delta=0.7, eps=1.0e-4, ntabg=0, ntabq=0,
    ntabk=0,
)


def _l(v):
    return ".true." if v else ".false."


def make_deck(**kw):
    p = dict(DEFAULTS)
    p.update(kw)
    vorb = p["vorb"] if p["vorb"] is not None else VORB10[: p["norb"]]
    n = p["npart"]
    lines = [
        str(p["seed"]), str(p["idmc"]), _l(p["isite"]), str(p["nwalk"]), str(p["neq"]),
        str(p["nav"]), str(p["nstep"]), repr(float(p["dt"])), repr(float(p["etrial"])),
        _l(p["units"]), repr(float(p["dcut"])), _l(p["fn"]), repr(float(p["popup"])),
        repr(float(p["popdn"])), repr(float(p["dobal"])), repr(float(p["wtcut"])),
        str(p["ntab"]), repr(float(p["hbar"])), str(p["ndim"]), str(n),
        f"{n // 2} {float(p['mass'])!r}", f"{n - n // 2} {float(p['mass'])!r}",
        "bcs", repr(float(p["a2"])), repr(float(p["rho"])), "0", "0", str(len(vorb)),
    ]
    lines += [repr(float(v)) for v in vorb]
    lines += [
        "0 0 0", "0.0", "0.0", "0", "0.0", "0.0",
        repr(float(p["bbeta"])), repr(float(p["dbeta"])), repr(float(p["bfact"])),
        repr(float(p["v0"])), repr(float(p["amu"])), repr(float(p["cjas1"])),
        repr(float(p["cjas2"])), repr(float(p["djas"])), repr(float(p["v0op"])),
        repr(float(p["v0eq"])), str(p["ntabg"]), str(p["ntabq"]), str(p["ntabk"]), "0",
    ]
    if p["idmc"] in (-1, -2):
        lines += [str(p["optiter"]), _l(p["optjas"]), _l(p["optnodes"]),
                  repr(float(p["delta"])), repr(float(p["eps"]))]
    return "\n".join(lines).replace("e-", "d-").replace("e+", "d+") + "\n", p
