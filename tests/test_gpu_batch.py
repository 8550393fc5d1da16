"""Batched replicas through the C ABI: periodic S(k) rebuilds inside a batch
sweep, and a captured batch that follows its members' box changes.

Bar: a replica advanced inside a batch equals the same replica run alone
through yamc_mc_run with the same rebuild interval, bit for bit (positions,
energies, accepted-move counts, RNG counters).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

SPECS = [(40, 0.5, 300.0), (64, 0.0, 290.0), (27, 1.0, 340.0), (50, 0.3, 320.0), (33, 0.8, 285.0)]


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


def pairs(precision, block=256, seed0=40):
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC
    systems = [replica_system(n, c, t, seed=seed0 + i) for i, (n, c, t) in enumerate(SPECS)]
    a = [EwaldMC(s, precision=precision) for s in systems]
    b = [EwaldMC(s, precision=precision) for s in systems]
    for i, (x, y) in enumerate(zip(a, b)):
        for r in (x, y):
            r.set_block_size(block)
            r.mc_init()
            r.set_rng(23, i)
    return a, b


def same_state(x, y, tag):
    for p, q in zip(x.positions(), y.positions()):
        assert np.array_equal(p, q), tag
    ex, ey = x.energies(), y.energies()
    assert all(ex[k] == ey[k] for k in ex), (tag, ex, ey)
    assert np.array_equal(x.sums(), y.sums()), tag
    assert x.rng_counter() == y.rng_counter(), tag


@pytest.mark.parametrize("precision", [1, 0])
def test_batch_with_rebuilds_equals_single_runs(lib, precision):
    """resync_every = 37 moves (not a divisor of any replica's cycle, so
    segments resume inside cycles and replicas end their sweeps in different
    segments): the eager batch equals each replica's yamc_mc_run with the
    same interval."""
    from yamc_amd.ewald import batch_run
    a, b = pairs(precision)
    nacc, _ = batch_run(a, 3, resync_every=37)
    for i, y in enumerate(b):
        n, _ = y.run(3, resync_every=37)
        assert n == nacc[i], i
        same_state(a[i], y, i)


def test_captured_batch_with_rebuilds_replays_equal_single_runs(lib):
    """The captured sweep with its rebuilds, replayed three times, equals
    three yamc_mc_run calls of each replica (FP32C, rebuild every 50 moves)."""
    from yamc_amd.ewald import Batch
    a, b = pairs(1, seed0=60)
    g = Batch(a, 2, resync_every=50)
    for rep in range(3):
        na, _ = g.launch()
        for i, y in enumerate(b):
            n, _ = y.run(2, resync_every=50)
            assert n == na[i], (rep, i)
            same_state(a[i], y, (rep, i))
    g.close()


def test_rebuild_keeps_fp32c_sums_next_to_fp64_rebuild(lib):
    """Right after a batch sweep whose last rebuild is at most 36 moves old,
    every replica's FP32C S(k) equals a fresh FP64 rebuild to 1e-9."""
    from yamc_amd.ewald import batch_run
    a, _ = pairs(1, seed0=80)
    batch_run(a, 4, resync_every=37)
    for r in a:
        s0 = r.sums()
        r.resync()
        s1 = r.sums()
        assert np.max(np.abs(s0 - s1)) <= 1e-9 * np.max(np.abs(s1))


def test_captured_batch_follows_member_box_change(lib):
    """A member's box changes between two replays (yamc_set_box: new box
    parameters, positions, beta and state): the next replay runs that member
    in its new box, equal to the member's own run after the same change."""
    from yamc_amd.ewald import Batch
    a, b = pairs(0, seed0=100)
    g = Batch(a, 1)
    g.launch()
    for y in b:
        y.run(1)
    for x in (a[2], b[2]):
        L = x.box_length()
        assert not x.set_box(L * 1.015)
    na, _ = g.launch()
    for i, y in enumerate(b):
        n, _ = y.run(1)
        assert n == na[i], i
        same_state(a[i], y, i)
    assert abs(a[2].box_length() - b[2].box_length()) == 0.0
    g.close()


def test_captured_batch_refuses_member_without_state(lib):
    """A member whose positions were replaced (state not rebuilt) makes the
    replay fail with ERROR_INITIALIZATION instead of running on a stale
    state; after yamc_mc_init the replay runs again."""
    from yamc_amd import YamcError
    from yamc_amd.ewald import Batch
    a, _ = pairs(0, seed0=120)
    g = Batch(a, 1)
    g.launch()
    a[1].set_positions(*a[1].positions())
    with pytest.raises(YamcError) as ei:
        g.launch()
    assert ei.value.code == 4
    a[1].mc_init()
    na, _ = g.launch()
    assert (na >= 0).all()
    g.close()
