"""The Fortran host (dmcbs_project_amd/fortran): modules step/stack/wavefunction/estimator
over the C-ABI and the DMC driver dmcbcs_gpu, built with amdflang by build().  On a machine
without a HIP device the driver reads the whole deck and then fails loudly at dmc_create."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

EXE = os.path.join(ROOT, "dmcbs_project_amd", "dmcbcs_gpu")


def have_device():
    return os.path.exists("/dev/kfd")


def test_fortran_driver_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() (make -C dmcbs_project_amd/fortran)"


@pytest.mark.skipif(have_device(), reason="checks the no-device error path")
def test_fortran_driver_fails_loudly_without_device(tmp_path):
    deck = str(golden("dmc_N14_short.npz")["deck"])
    r = subprocess.run([EXE], input=deck, capture_output=True, text=True, cwd=tmp_path,
                       timeout=60)
    assert r.returncode == 2
    assert "no HIP device" in r.stdout


def test_fortran_driver_rejects_unsupported_deck(tmp_path):
    """Options outside the engine (here a guiding function a2 /= 0) stop at set-up, before
    any device call, as the reference stops on bad input."""
    deck = str(golden("dmc_N14_short.npz")["deck"]).split("\n")
    i = deck.index("bcs")
    deck[i + 1] = "0.5"
    r = subprocess.run([EXE], input="\n".join(deck), capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode != 0
    assert "a2 /= 0" in (r.stdout + r.stderr)


FRAGMENT = os.path.join(ROOT, "tests", "fortran", "walkers_fragment.f90")


def build_fragment(tmp_path):
    """tests/fortran/walkers_fragment.f90 compiled with the host modules from their sources
    (amdflang) and linked to the engine library"""
    fdir = os.path.join(ROOT, "dmcbs_project_amd", "fortran")
    objs = []
    for src in (os.path.join(fdir, "dmc_gpu.f90"), os.path.join(fdir, "host_modules.f90"),
                FRAGMENT):
        o = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["/opt/rocm/bin/amdflang", "-O2", "-J", str(tmp_path), "-c", src, "-o", o],
                       check=True, capture_output=True)
        objs.append(o)
    exe = str(tmp_path / "walkers_fragment")
    lib = os.path.join(ROOT, "dmcbs_project_amd")
    subprocess.run(["/opt/rocm/bin/amdflang", "-o", exe, *objs, f"-L{lib}", "-ldmc_gpu",
                    f"-Wl,-rpath,{lib}"], check=True, capture_output=True)
    return exe


def fragment_deck(npart=14, norb=3, nwalk=24, seed=1):
    from oracle import refdeck
    deck, _ = refdeck.make_deck(npart=npart, norb=norb, nwalk=nwalk, seed=seed, idmc=-1,
                                dt=0.002)
    return deck


def test_fragment_generates_the_reference_walkers(tmp_path, oracle_mod):
    """A driver fragment written as dmcbcs.f90:43-66,237-301,548-567 against the host
    modules (deck reads, setpsi/setphi/setupopt/setstep, create, the walker record's x,
    weight, irn, x0, psil0, psigl0 fields, setrn/randn/ran2/gaussian/savern, the walker file)
    compiles and, without touching the device, generates the reference's initial walkers:
    positions and RNG states equal the oracle's (pinned bit for bit to the reference's), the
    walker file reads back in the reference's formats, and gaussian(7) from seed 7 equals
    the oracle's (pinned to the reference's gaussian)."""
    from dmcbs_project_amd import deck as deckmod
    exe = build_fragment(tmp_path)
    r = subprocess.run([exe, "gen"], input=fragment_deck(), capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    s = oracle_mod.System(npart=14, norb=3)
    x, irn = s.init_walkers(1, 24)
    xf, wf, _, _, irnf, tau = deckmod.read_walkers(str(tmp_path / "walkers.out"), 14)
    assert tau == 0.0
    np.testing.assert_array_equal(irnf, irn)
    # the file holds e15.7 positions: 8 significant digits
    np.testing.assert_allclose(xf, x, rtol=1e-7, atol=1e-6)
    np.testing.assert_array_equal(wf, np.ones(24))
    lines = open(tmp_path / "gauss.out").read().split()
    st, g = oracle_mod.gaussian(7, 7)
    np.testing.assert_array_equal(np.array([float(v) for v in lines[:7]]), g)
    assert int(lines[7]) == st
