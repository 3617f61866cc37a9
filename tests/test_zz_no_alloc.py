"""dmc_opt_accumulate_noalloc (the optimiser sums with the parameter derivatives in the
library's fixed device buffer instead of an allocated one) against dmc_opt_accumulate, called
through the C-ABI: two calls in a row, the second on a smaller population, and the size
limit refused."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(e, fn):
    P = e.params.norb + 2
    n = 3 + 2 * P + P * P
    out = np.zeros(n)
    e.lib.dmc_opt_sums(e.h, out, 1)  # read and reset
    from dmcbs_project_amd import lib
    lib.check(getattr(e.lib, fn)(e.h))
    lib.check(e.lib.dmc_opt_sums(e.h, out, 1))
    return out


def test_noalloc_matches_allocating_entry_point():
    from dmcbs_project_amd import DMC_MIXED, DMC_SINGLEPARTICLE, Engine
    e = Engine(npart=66, norb=5, nwalk=4096, dt=0.001, dcut=2.0, capacity=2 * 4096,
               precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=66)
    e.init_walkers(7)
    e.step()
    for call in range(2):
        if call == 1:
            assert e.branch(0.25) < 0.5 * 4096  # the second call on a smaller population
        print(f"no-alloc call {call}: {e.numstack()} walkers", flush=True)
        ref = _sums(e, "dmc_opt_accumulate")
        got = _sums(e, "dmc_opt_accumulate_noalloc")
        assert ref[0] == e.numstack()
        np.testing.assert_array_equal(got, ref)
    e.close()


def test_noalloc_refuses_populations_beyond_its_buffer():
    from dmcbs_project_amd import DMCError, Engine, lib
    e = Engine(npart=2, norb=3, nwalk=220000, dt=0.002, capacity=220000)
    e.init_walkers(3)
    assert e.numstack() * (3 + 2) > (1 << 20)
    with pytest.raises(DMCError) as ei:
        lib.check(e.lib.dmc_opt_accumulate_noalloc(e.h))
    assert ei.value.code == 1 and "device buffer" in str(ei.value)
    e.close()
