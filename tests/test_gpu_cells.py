"""Cell list of the real-space scan (yamc_set_cell_div): cells of width
rc / div with a sphere-pruned stencil.  The scan must find exactly the
candidates of the full scan, so a chain's run is the same to the bit at
every cell size (decisions, per-move energies, final positions, per-atom
arrays, S(k)), and against the oracle it keeps the usual bar.  Checked on
the one-workgroup chain (both launch shapes, FP64 and FP32C), in a batch,
and on the grid chain (there to the summation-order bar, since its
workgroups split the candidates differently with and without cells), at
cell counts from 3 to 16 per axis."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_parity import make  # noqa: E402
from test_gpu_grid import check_records, check_state  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


def run_div(name, div, precision=0, block=256, groups=1, ncyc=1, seed=5):
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config(name, seed=seed)
    g = EwaldMC(s, precision=precision)
    g.set_block_size(block)
    g.set_chain_groups(groups)
    g.set_cell_div(div)
    g.mc_init()
    g.set_rng(1234, 3)
    print(f"cells {name} div={div} prec={precision} bs={block} G={groups}", flush=True)
    n, r = g.run(ncyc, records=True)
    out = dict(n=n, rec=r, pos=g.positions(), real=g.real_arrays(), sums=g.sums(), e=g.energies())
    g.close()
    return out


def same(a, b, what):
    assert a["n"] == b["n"], what
    for k in ("imol", "accepted", "overlap", "du_total", "du_real", "du_kc", "du_klj", "f_new", "t_new"):
        assert np.array_equal(a["rec"][k], b["rec"][k]), (what, k)
    for x, y in zip(a["pos"], b["pos"]):
        assert np.array_equal(x, y), what
    for x, y in zip(a["real"], b["real"]):
        assert np.array_equal(x, y), what
    assert np.array_equal(a["sums"], b["sums"]), what
    assert a["e"] == b["e"], what


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("block", [256, 512])
def test_cell_sizes_give_the_full_scan_run(lib, precision, block):
    """c2 (6224 atoms, rc = 0.305 L): no cells, 3, 6, 9 and 13 cells per axis."""
    ref = run_div("c2", 0, precision, block)
    for div in (1, 2, 3, 4):
        same(run_div("c2", div, precision, block), ref, ("c2", div, precision, block))


def test_cell_sizes_small_cells_system(lib):
    """The 'cells' system (L = 25 A, rc = 6 A): 4, 8, 12, 16 cells per axis,
    several cycles so molecules change cells."""
    ref = run_div("cells", 0, ncyc=4)
    for div in (1, 2, 3, 4):
        same(run_div("cells", div, ncyc=4), ref, ("cells", div))


def test_cell_sizes_match_oracle(lib):
    """Against the oracle at the finest cell size (16 per axis)."""
    s, g, o = make("cells", seed=21)
    g.set_cell_div(4)
    g.mc_init()
    o.mc_init()
    g.set_rng(99, 1)
    o.set_rng(99, 1)
    ng, rg = g.run(2, records=True)
    no, ro = o.run(2)
    assert ng == no
    check_records(rg, ro, "cells div 4")
    check_state(g, o, "cells div 4")


def test_grid_chain_cell_sizes(lib):
    """The grid chain (8 workgroups) splits the candidates among its
    workgroups by atom range without cells and by the visited cells' atoms
    with them, so the partial sums are added in another order: the same
    decisions, per-move energies within 1e-10 of the sum of their parts,
    positions within 1e-12 box lengths."""
    ref = run_div("cells", 0, groups=8, ncyc=2)
    for div in (1, 3):
        got = run_div("cells", div, groups=8, ncyc=2)
        assert got["n"] == ref["n"], div
        ra, rb = got["rec"], ref["rec"]
        for k in ("imol", "accepted", "overlap"):
            assert np.array_equal(ra[k], rb[k]), (div, k)
        ok = rb["overlap"] == 0
        scale = (np.abs(rb["du_real"]) + np.abs(rb["du_kc"]) + np.abs(rb["du_klj"]) + np.abs(rb["du_ext"]))[ok]
        for k in ("du_total", "du_real", "du_kc", "du_klj"):
            assert np.max(np.abs(ra[k][ok] - rb[k][ok]) / scale) < 1e-10, (div, k)
        for x, y in zip(got["pos"], ref["pos"]):
            assert np.max(np.abs(x - y)) < 1e-12, div


def test_batch_with_cells_equals_single_runs(lib):
    """A batch of three FP32C replicas with 9 cells per axis equals each
    replica run alone."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC, batch_run
    reps, alone = [], []
    for i in range(3):
        for lst in (reps, alone):
            g = EwaldMC(config("c2", seed=40 + i), precision=1)
            g.set_cell_div(3)
            g.mc_init()
            g.set_rng(77, i)
            lst.append(g)
    print("cells batch", flush=True)
    nacc, _ = batch_run(reps, 1)
    for i, (a, b) in enumerate(zip(reps, alone)):
        n, _ = b.run(1)
        assert n == nacc[i], i
        for x, y in zip(a.positions(), b.positions()):
            assert np.array_equal(x, y), i
        assert np.array_equal(a.sums(), b.sums()), i
    for g in reps + alone:
        g.close()
