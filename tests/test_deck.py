"""The reference's deck and walker file on the Python host (dmcbs_project_amd/deck.py):
decks written in the reference's read order (oracle/refdeck.py, whose order the reference
driver itself reads) come back with every value; unsupported options are refused as the
Fortran host refuses them; fort.9 round-trips in the reference's formats."""
import numpy as np
import pytest

from dmcbs_project_amd import deck


@pytest.mark.parametrize("kw", [dict(npart=14, norb=3), dict(npart=66, norb=5, nwalk=32768,
                                 dt=0.001, dcut=2.0, fn=False, ntabg=100, ntabq=10, ntabk=20),
                                dict(npart=38, norb=4, idmc=-1, optiter=3, delta=0.5)])
def test_deck_values_come_back(kw):
    from oracle import refdeck
    text, p = refdeck.make_deck(**kw)
    params, run = deck.read_deck(text)
    for k in ("npart", "norb", "ntab", "nwalk", "dt", "etrial", "dcut", "wtcut", "popup",
              "popdn", "bbeta", "dbeta", "bfact", "v0", "amu", "cjas1", "cjas2", "djas",
              "v0op", "v0eq", "rho", "hbar", "mass"):
        assert getattr(params, k) == p[k], k
    assert params.fn == p["fn"] and params.units == p["units"]
    assert params.vorb == tuple(p["vorb"] if p["vorb"] is not None else
                                refdeck.VORB10[:p["norb"]])
    for k in ("seed", "idmc", "isite", "neq", "nav", "nstep", "ntabg", "ntabq", "ntabk"):
        assert run[k] == p[k], k
    if p["idmc"] == -1:
        assert run["optiter"] == p["optiter"] and run["delta"] == p["delta"]
        from dmcbs_project_amd import DMC_VMC, DMC_FP64
        assert params.movetype == DMC_VMC and params.precision == DMC_FP64


def test_deck_with_comments_and_fortran_exponents():
    lines = ["17             # seed", "1              # dmc", ".true.         # sites",
             "1000", "0", "10", "200", "5.0d-5         # dt", "0.40", "T", "-1.0", ".true.",
             "1.1", "0.9", "0.5", "10.0", "10000", "0.5", "3", "66", "33    1.0",
             "33    1.0", "bcs            # wavefunction", "0.0", "0.0337737278", "0", "0",
             "2", "0.00197519728006648144", "0.00249764752028349595",
             "0 0 0          # q external potential", "0.0", "0.0", "0", "0.0", "0.0", "0.5",
             "5.0", "1.0", "1.0", "24.0", "1.0", "0.0", "1.15", "0.0", "-10.0", "0", "0",
             "0", "0"]
    params, run = deck.read_deck("\n".join(lines) + "\n")
    assert params.dt == 5.0e-5 and params.units and params.ntab == 10000
    assert params.vorb == (0.00197519728006648144, 0.00249764752028349595)
    assert params.v0eq == -10.0 and run["seed"] == 17 and run["nstep"] == 200


@pytest.mark.parametrize("line,value", [(22, "slater"), (23, "0.5"), (26, "3"),
                                        (43, "0.5")])
def test_unsupported_options_are_refused(line, value):
    from oracle import refdeck
    text, _ = refdeck.make_deck(npart=14, norb=3)
    lines = text.split("\n")
    lines[line] = value
    with pytest.raises(ValueError):
        deck.read_deck("\n".join(lines))


def test_walker_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    nw, n = 5, 14
    x = rng.normal(size=(nw, n, 3)) * 3.0
    w = rng.uniform(0.5, 2.0, nw)
    pl = rng.normal(size=nw) * 40.0
    irn = rng.integers(1, 1 << 48, nw)
    path = tmp_path / "fort.9"
    deck.write_walkers(path, x, w, pl, irn, tau=0.25)
    lines = path.read_text().splitlines()
    assert len(lines[0]) == 25 and len(lines[1]) == 90 and len(lines[8]) == 65
    assert len(lines) == 1 + nw * (7 + 2)
    x2, w2, pl2, pg2, irn2, tau = deck.read_walkers(path, n)
    assert tau == 0.25
    # e15.7 keeps 7 significant digits: half a unit of the 7th digit is <= 5e-7 relative
    np.testing.assert_allclose(x2, x, rtol=5.01e-7)
    np.testing.assert_allclose(w2, w, rtol=5.01e-7)
    np.testing.assert_allclose(pl2, pl, rtol=1e-14)
    np.testing.assert_array_equal(pg2, pl2)
    np.testing.assert_array_equal(irn2, irn)
