"""The C-ABI library: builds, loads and exports every entry point include/dmc_gpu.h
declares (no compute calls without a GPU)."""
import os
import re
import subprocess

from conftest import ROOT


def declared():
    txt = open(os.path.join(ROOT, "include", "dmc_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(dmc_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    so = os.path.join(ROOT, "dmcbs_project_amd", "libdmc_gpu.so")
    assert os.path.exists(so), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (dmc_\w+)$", out, re.M))
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from dmcbs_project_amd import lib
    assert set(declared()) == set(lib.SIGNATURES)
    lib.load()


def test_no_device_is_reported_not_faked():
    """Without a GPU the product path must fail loudly (no CPU fallback)."""
    import pytest
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from dmcbs_project_amd import Engine, DMCError
    with pytest.raises(DMCError):
        Engine(npart=14, norb=3)
