"""g(r), density profile and S(q) (k_obs through dmc_gofr / dmc_sofq) and n(k) (k_nofk
through dmc_nofk) on the device against the oracle's addgofr / addsofq / addenk on the same
walkers.  The oracle is pinned to the reference's own gofr.f90 / sofq.f90 / momentum3d.f90
tables (tests/test_observables.py).

The device sums one weighted count per bin and walker in another order than the reference's
sequential loop, and S(q) forms exp(-i q.x) from per-axis powers instead of one sincos per
(q, particle): sums agree to 1e-10 relative to the largest entry of each table.  n(k) takes
its plane-wave ratio terms through projections on the inverse and its cosines by the
Chebyshev recurrence: 1e-9.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rtol * scale, (np.abs(a - b).max(), scale)


def _check_population(oracle_mod, e, ntab, ntq):
    w = e.walkers()
    wt = w["weight"] * w["is_"] * w["wt0"]
    print(f"population {len(wt)}, weights {wt.min():.3f}..{wt.max():.3f}", flush=True)
    dev = e.addgofr(ntab)
    ref = oracle_mod.addgofr(w["x"], wt, e.el, ntab)
    for a, b in zip(dev[:4], ref[:4]):
        _close(a, b, 1e-11)
    assert dev[4][0] == pytest.approx(ref[4], rel=1e-12)
    numsh, ak = e.sofq_kvec(ntq)
    numsh_o, ak_o = oracle_mod.kshells(e.el, ntq)
    np.testing.assert_array_equal(numsh, numsh_o)
    np.testing.assert_array_equal(ak, ak_o)
    dev = e.addsofq(ntq)
    ref = oracle_mod.addsofq(w["x"], wt, ak_o)
    _close(dev[0], ref[0], 1e-10)
    _close(dev[1], ref[1], 1e-10)
    assert dev[2][0] == pytest.approx(ref[2], rel=1e-12)
    return wt


def test_obs_dmc_population_fp64_allparticle(oracle_mod):
    """N = 14, all-particle moves without fixed node, so that is * wt0 can turn negative."""
    from dmcbs_project_amd import Engine
    e = Engine(npart=14, norb=3, nwalk=700, fn=False, dt=0.01)
    e.init_walkers(4242)
    for _ in range(40):
        e.step()
    wt = _check_population(oracle_mod, e, ntab=40, ntq=6)
    assert (wt < 0).any() and (wt > 0).any()


def test_obs_dmc_population_mixed_singleparticle(oracle_mod):
    """N = 66, mixed-precision single-particle moves, 2048 walkers (two or more walkers per
    workgroup, uneven ranges)."""
    from dmcbs_project_amd import Engine, DMC_MIXED, DMC_SINGLEPARTICLE
    e = Engine(npart=66, norb=5, nwalk=2048, dt=0.001, dcut=2.0, capacity=4 * 2048,
               precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=33)
    e.init_walkers(977)
    for _ in range(3):
        e.step()
    _check_population(oracle_mod, e, ntab=60, ntq=8)


def test_obs_reference_configurations(oracle_mod):
    """The configurations of the reference's g(r)/S(q) tables, loaded as walkers (weight 1):
    device sums equal the oracle's, and the device box and q vectors are the reference's."""
    from dmcbs_project_amd import Engine
    g = golden("obs_N66.npz")
    e = Engine(npart=66, norb=5, nwalk=len(g["x"]))
    assert e.el == pytest.approx(float(g["el"]), rel=1e-15)
    e.push_walkers(g["x"], np.arange(1, len(g["x"]) + 1, dtype=np.int64))
    wt = _check_population(oracle_mod, e, ntab=int(g["ntg"]), ntq=int(g["ntq"]))
    np.testing.assert_array_equal(wt, np.ones(len(g["x"])))


def test_obs_accumulates_and_rejects_oversize():
    """Calls add into the caller's accumulators (as addgofr adds into the module's); bin,
    q or shell counts the kernels do not take are refused before any launch."""
    from dmcbs_project_amd import DMCError, Engine
    e = Engine(npart=14, norb=3, nwalk=37)
    e.init_walkers(5)
    one = e.addgofr(12)
    two = e.addgofr(12, e.addgofr(12))
    for a, b in zip(one, two):
        np.testing.assert_allclose(b, 2 * a, rtol=1e-15)
    s1 = e.addsofq(4)
    s2 = e.addsofq(4, e.addsofq(4))
    np.testing.assert_allclose(s2[0], 2 * s1[0], rtol=1e-15)
    with pytest.raises(DMCError):
        e.addgofr(4000)
    with pytest.raises(DMCError):
        e.addgofr(0)
    with pytest.raises(DMCError):
        e.addsofq(0)
    with pytest.raises(DMCError):
        e.addenk(64)  # components up to 8: beyond the kernel's cosine table
    k1 = e.addenk(9)
    k2 = e.addenk(9, e.addenk(9))
    np.testing.assert_allclose(k2[0], 2 * k1[0], rtol=1e-15)


def _check_nofk(oracle_mod, e, s, ntk):
    w = e.walkers()
    wt = w["weight"] * w["is_"] * w["wt0"]
    print(f"n(k): population {len(wt)}, ntab {ntk}", flush=True)
    dev = e.addenk(ntk)
    ref = s.addenk(w["x"], w["irn"], wt, ntk)
    _close(dev[0], ref[0], 1e-9)
    _close(dev[1], ref[1], 1e-9)
    assert dev[2][0] == pytest.approx(ref[2], rel=1e-12)
    kmag, nk = e.nofk_shells(ntk)
    ik, indk, nk_o = oracle_mod.enk_shells(ntk)
    np.testing.assert_array_equal(nk, nk_o)
    np.testing.assert_allclose(kmag, 2 * np.pi / e.el * np.sqrt((ik[indk] ** 2).sum(axis=1)),
                               rtol=1e-15)


def test_nofk_dmc_population_n14(oracle_mod):
    """addenk (momentum3d.f90:89-120) of a DMC population on the device (k_nofk: FP64
    inverse, projections of the partner phases, one lane per displacement) against the
    oracle on the read-back walkers, RNG states and weights."""
    from dmcbs_project_amd import Engine
    e = Engine(npart=14, norb=3, nwalk=400, fn=False, dt=0.01)
    e.init_walkers(31)
    for _ in range(10):
        e.step()
    _check_nofk(oracle_mod, e, oracle_mod.System(npart=14, norb=3), 20)


def test_nofk_reference_configurations_n66(oracle_mod):
    """The configurations and RNG states of the reference's n(k) table at N = 66, loaded as
    walkers (weight 1), and a short DMC population: device sums equal the oracle's."""
    from dmcbs_project_amd import Engine, DMC_MIXED, DMC_SINGLEPARTICLE
    g = golden("enk_N66.npz")
    s = oracle_mod.System(npart=66, norb=5)
    e = Engine(npart=66, norb=5, nwalk=len(g["x"]))
    e.push_walkers(g["x"], g["irn"])
    _check_nofk(oracle_mod, e, s, int(g["ntk"]))
    e = Engine(npart=66, norb=5, nwalk=96, dt=0.001, dcut=2.0, capacity=4 * 96,
               precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=33)
    e.init_walkers(5)
    for _ in range(2):
        e.step()
    _check_nofk(oracle_mod, e, s, 20)


@pytest.mark.parametrize("n,norb,nw", [(2, 1, 50), (38, 4, 160)])
def test_observables_other_particle_numbers(oracle_mod, n, norb, nw):
    """N = 2 and N = 38 (the other kernel instantiations): g(r), S(q) and n(k) of a short
    DMC population against the oracle."""
    from dmcbs_project_amd import Engine
    e = Engine(npart=n, norb=norb, nwalk=nw, dt=0.002, capacity=8 * nw)
    e.init_walkers(17 + n)
    for _ in range(3):
        e.step()
    _check_population(oracle_mod, e, ntab=30, ntq=6)
    _check_nofk(oracle_mod, e, oracle_mod.System(npart=n, norb=norb), 12)


def test_gofr_large_table_keeps_sums_outside_lds(oracle_mod):
    """ntab = 600 (the partial sums no longer fit in LDS and stay in each block's row of
    the output) against the oracle, on a population large enough for several walkers per
    workgroup."""
    from dmcbs_project_amd import Engine
    e = Engine(npart=14, norb=3, nwalk=2500, dt=0.01)
    e.init_walkers(99)
    for _ in range(4):
        e.step()
    w = e.walkers()
    wt = w["weight"] * w["is_"] * w["wt0"]
    dev = e.addgofr(600)
    ref = oracle_mod.addgofr(w["x"], wt, e.el, 600)
    for a, b in zip(dev[:4], ref[:4]):
        _close(a, b, 1e-11)
    assert dev[4][0] == pytest.approx(ref[4], rel=1e-12)


def test_gofr_sofq_full_population(oracle_mod):
    """g(r) (100 bins) and S(q) (10 shells) of the whole measured population (N = 66,
    32768 walkers after 3 mixed-precision steps) against the oracle on every walker."""
    from dmcbs_project_amd import DMC_MIXED, DMC_SINGLEPARTICLE, Engine
    nw = 32768
    e = Engine(npart=66, norb=5, nwalk=nw, dt=0.001, dcut=2.0, capacity=2 * nw,
               precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=66)
    e.init_walkers(7)
    for _ in range(3):
        e.step()
    _check_population(oracle_mod, e, ntab=100, ntq=10)


def test_sofq_large_q_set_keeps_sums_outside_lds(oracle_mod):
    """S(q) with 70 shells at N = 14 (1536 q vectors: 6 partial sums per q no longer fit in
    a workgroup's 64 KB of LDS with the axis powers, so each block keeps them in its own row
    of the output) against the oracle's addsofq."""
    from dmcbs_project_amd import Engine
    e = Engine(npart=14, norb=3, nwalk=600, dt=0.01)
    e.init_walkers(31)
    for _ in range(3):
        e.step()
    w = e.walkers()
    wt = w["weight"] * w["is_"] * w["wt0"]
    numsh, ak = e.sofq_kvec(70)
    assert len(ak) == 1536
    numsh_o, ak_o = oracle_mod.kshells(e.el, 70, maxk=4096)
    np.testing.assert_array_equal(numsh, numsh_o)
    np.testing.assert_array_equal(ak, ak_o)
    dev = e.addsofq(70)
    ref = oracle_mod.addsofq(w["x"], wt, ak_o)
    _close(dev[0], ref[0], 1e-10)
    _close(dev[1], ref[1], 1e-10)
    assert dev[2][0] == pytest.approx(ref[2], rel=1e-12)


def test_gofr_N66_near_largest_table(oracle_mod):
    """g(r) and the density profile at N = 66 with ntab = 1000 (close to the largest table
    the kernel's LDS holds, 1065) against the oracle's addgofr."""
    from dmcbs_project_amd import DMC_MIXED, DMC_SINGLEPARTICLE, Engine
    e = Engine(npart=66, norb=5, nwalk=512, dt=0.001, dcut=2.0, capacity=4 * 512,
               precision=DMC_MIXED, movetype=DMC_SINGLEPARTICLE, krefresh=66)
    e.init_walkers(13)
    for _ in range(2):
        e.step()
    w = e.walkers()
    wt = w["weight"] * w["is_"] * w["wt0"]
    dev = e.addgofr(1000)
    ref = oracle_mod.addgofr(w["x"], wt, e.el, 1000)
    for a, b in zip(dev[:4], ref[:4]):
        _close(a, b, 1e-11)
    assert dev[4][0] == pytest.approx(ref[4], rel=1e-12)
