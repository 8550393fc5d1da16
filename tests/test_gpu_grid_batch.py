"""Batches of grid chains: members set to several workgroups per chain
(yamc_set_chain_groups) run as one cooperative launch of all their
workgroups.

Bar: a member with G >= 2 equals the same replica run alone as a grid chain
with the same G, bit for bit (positions, energies, S(k), accepted-move
counts, RNG counters); a member with G = 1 (the grid kernel over one
workgroup) takes the one-workgroup kernel's decisions, with states equal to
the summation-order bar.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_batch import pairs, same_state  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


@pytest.mark.parametrize("precision", [0, 1])
def test_grid_batch_equals_grid_chains_alone(lib, precision):
    """Five uneven replicas at 2, 3, 2, 4 and 2 workgroups per chain: two
    cycles eagerly (yamc_batch_run), then one through yamc_batch_create and
    launch."""
    from yamc_amd.ewald import Batch, batch_run
    a, b = pairs(precision, seed0=140)
    for G, x, y in zip((2, 3, 2, 4, 2), a, b):
        x.set_chain_groups(G)
        y.set_chain_groups(G)
    print(f"grid batch precision={precision}", flush=True)
    nacc, _ = batch_run(a, 2)
    for i, y in enumerate(b):
        n, _ = y.run(2)
        assert n == nacc[i], i
        same_state(a[i], y, i)
    g = Batch(a, 1)
    na, _ = g.launch()
    for i, y in enumerate(b):
        n, _ = y.run(1)
        assert n == na[i], i
        same_state(a[i], y, i)
    g.close()


def test_grid_batch_member_on_one_workgroup(lib):
    """A G = 1 member next to grid members: the same decisions as the
    replica run alone on the one-workgroup kernel, positions within 1e-12
    box lengths and energies within 1e-10 relative."""
    from yamc_amd.ewald import batch_run
    a, b = pairs(0, seed0=150)
    for G, x in zip((1, 2, 1, 3, 2), a):
        x.set_chain_groups(G)
    print("grid batch with one-workgroup members", flush=True)
    nacc, _ = batch_run(a, 2)
    for i in (0, 2):
        n, _ = b[i].run(2)
        assert n == nacc[i], i
        for p, q in zip(a[i].positions(), b[i].positions()):
            assert np.max(np.abs(p - q)) < 1e-12, i
        ea, eb = a[i].energies(), b[i].energies()
        for k in eb:
            assert abs(ea[k] - eb[k]) <= 1e-10 * max(1.0, abs(eb[k])), (i, k)
        assert a[i].rng_counter() == b[i].rng_counter(), i


@pytest.mark.parametrize("precision", [1, 0])
def test_grid_batch_with_rebuilds_equals_grid_chains_alone(lib, precision):
    """resync_every = 37 moves with members at 2, 3, 2, 4 and 2 workgroups
    per chain: each member's workgroups rebuild its S(k) inside the launch,
    each over its own k range; the batch equals every replica run alone as
    a grid chain by yamc_mc_run with the same interval (segments of 37 moves
    with the launch-level rebuild between them), bit for bit, over two
    sweeps."""
    from yamc_amd.ewald import batch_run
    a, b = pairs(precision, seed0=160)
    for G, x, y in zip((2, 3, 2, 4, 2), a, b):
        x.set_chain_groups(G)
        y.set_chain_groups(G)
    print(f"grid batch with rebuilds, precision={precision}", flush=True)
    for sweep in range(2):
        nacc, _ = batch_run(a, 2, resync_every=37)
        for i, y in enumerate(b):
            n, _ = y.run(2, resync_every=37)
            assert n == nacc[i], (sweep, i)
            same_state(a[i], y, (sweep, i))
