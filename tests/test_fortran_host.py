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
