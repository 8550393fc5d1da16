"""One chain over several workgroups (yamc_set_chain_groups): the grid
chain kernel splits every k sweep and S(k) commit by k range and the
real-space pass by atom range over G co-resident workgroups.

Bar: the same as the one-workgroup chain.  Against the oracle on one RNG
stream: identical molecules, overlaps and decisions; per-move dE and its
parts within 1e-10 of sum |parts|; forces per component within 1e-8; the
state after the run (energies 1e-10, positions 1e-12 box lengths, per-atom
arrays and S(k) 1e-10).  Against the reference's golden movemc trace
through the constant-uniform mode.  At c2 and c5 against the one-workgroup
chain on the same stream (decisions identical, dE within 1e-10)."""
import os
import sys
import time

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_parity import RTOL, make, rel  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


def check_records(rg, ro, what):
    assert np.array_equal(rg["imol"], ro["imol"]), what
    assert np.array_equal(rg["overlap"], ro["overlap"]), what
    assert np.array_equal(rg["accepted"], ro["accepted"]), what
    ok = ro["overlap"] == 0
    scale = (np.abs(ro["du_real"]) + np.abs(ro["du_kc"]) + np.abs(ro["du_klj"]) + np.abs(ro["du_ext"]))[ok]
    for k in ("du_total", "du_real", "du_kc", "du_klj", "du_ext"):
        assert rel(rg[k][ok], ro[k][ok], scale) < RTOL, (what, k)
    for k in ("f_old", "t_old", "f_new", "t_new"):
        a, b = rg[k][ok], ro[k][ok]
        den = np.maximum(np.abs(b), 1e-4 * np.linalg.norm(b, axis=1, keepdims=True))
        assert float(np.max(np.abs(a - b) / np.maximum(den, 1e-300))) < 1e-8, (what, k)


def check_state(g, o, what):
    eg, eo = g.energies(), o.energies()
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, (what, k)
    for a, b in zip(g.positions(), o.positions()):
        assert rel(a, b, 1.0) < 1e-12, what
    for a, b in zip(g.real_arrays(), o.real_arrays()):
        assert rel(a, b, np.abs(b).max()) < RTOL, what
    assert rel(g.sums(), o.sums(), np.abs(o.sums()).max()) < RTOL, what


@pytest.mark.parametrize("groups", [2, 3, 8])
@pytest.mark.parametrize("name,ncyc", [("small", 20), ("small_nobias", 20), ("c1", 1), ("cells", 1)])
def test_grid_trace_matches_oracle(lib, name, ncyc, groups):
    s, g, o = make(name, seed=21)
    g.set_chain_groups(groups)
    g.mc_init()
    o.mc_init()
    g.set_rng(99, 1)
    o.set_rng(99, 1)
    print(f"grid chain {name} G={groups}", flush=True)
    ng, rg = g.run(ncyc, records=True)
    no, ro = o.run(ncyc)
    assert ng == no
    check_records(rg, ro, (name, groups))
    assert g.rng_counter() == o.rng_counter()
    check_state(g, o, (name, groups))


@pytest.mark.parametrize("name,ncyc,psm,kind", [("small", 12, 1.0, None), ("small", 10, 0.0, "all")])
def test_grid_npt_and_exchange_cycles_match_oracle(lib, name, ncyc, psm, kind):
    """Launches that stop at a volume or exchange step and resume
    (max_moves, mv_begin) on the grid chain."""
    import golden_cases as gc
    s, g, o = make(name, seed=31)
    g.set_chain_groups(4)
    if kind:
        xc = np.ones(s.nmol)
        g.set_frequencies(None, xc)
        o.set_frequencies(None, xc)
    g.mc_init()
    o.mc_init()
    g.set_rng(123, 2)
    o.set_rng(123, 2)
    print(f"grid chain {name} npt/exchange", flush=True)
    if kind:
        ng, rg = g.run(ncyc, records=True)
        no, ro = o.run(ncyc)
        assert ng == no
        for k in ("imol", "imol2", "overlap", "accepted"):
            assert np.array_equal(rg[k], ro[k]), k
    else:
        p = gc.pressure_kT_A3(1e5, s.params.temp)
        rg = g.run_npt(ncyc, psm, p, float(np.log(1.03)))
        ro = o.run_npt(ncyc, psm, p, float(np.log(1.03)))
        assert rg == ro and ro[1] > 0
        assert abs(g.box_length() - o.box_length()) <= 1e-13 * o.box_length()
    assert g.rng_counter() == o.rng_counter()
    check_state(g, o, name)


@pytest.mark.parametrize("path", __import__("golden_cases").cases(), ids=lambda p: p.split("/")[-1])
def test_grid_matches_reference_golden_trace(lib, path):
    """The reference's own movemc trace (RND_CONSTANT mode, golden fixture)
    on a chain of 4 workgroups."""
    import golden_cases as gc
    from yamc_amd.ewald import EwaldMC
    sysm, gd = gc.load(path)
    g = EwaldMC(sysm)
    g.set_chain_groups(4)
    g.mc_init()

    def step(imol, u):
        n, _ = g.run_const([imol], [u], records=False)
        return n

    gc.check_trace(g, sysm, gd, step)


@pytest.mark.parametrize("name,groups,ncyc", [("c2", 8, 1), ("c5", 32, 1)])
def test_grid_large_systems_match_one_workgroup(lib, name, groups, ncyc):
    """c2 (6,224 atoms, nk 1,068) and c5 (99,504 atoms, nk 7,912, cell-list
    scan) in FP64: one run_mc cycle on one workgroup and on a grid of
    workgroups, same stream: identical decisions, dE and forces within the
    bar, the same state after the cycle."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config(name, seed=20261210)
    recs, chains, rates = {}, {}, {}
    for G in (1, groups):
        g = EwaldMC(s, precision=0)
        g.set_chain_groups(G)
        g.mc_init()
        g.set_rng(20261210, 0)
        print(f"{name} G={G}", flush=True)
        t0 = time.perf_counter()
        _, recs[G] = g.run(ncyc, records=True)
        rates[G] = len(recs[G]) / (time.perf_counter() - t0)
        chains[G] = g
    print(f"{name}: trial moves/s {rates}", flush=True)
    check_records(recs[groups], recs[1], name)
    a, b = chains[groups], chains[1]
    ea, eb = a.energies(), b.energies()
    for k in eb:
        assert rel(ea[k], eb[k], max(1.0, abs(eb[k]))) < RTOL, k
    for x, y in zip(a.positions(), b.positions()):
        assert rel(x, y, 1.0) < 1e-12
    assert rel(a.sums(), b.sums(), np.abs(b.sums()).max()) < RTOL
    for G in chains:
        chains[G].close()


def test_grid_fp32c_equals_one_workgroup(lib):
    """FP32C: a k vector's S(k) update is the same arithmetic whichever
    workgroup owns it, so a grid chain (with host resyncs between segments)
    takes the decisions of the one-workgroup chain with dE within 1e-10."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c1", seed=23)
    recs, state = {}, {}
    for G in (1, 4):
        g = EwaldMC(s, precision=1)
        g.set_chain_groups(G)
        g.mc_init()
        g.set_rng(5, 5)
        print(f"fp32c G={G}", flush=True)
        _, recs[G] = g.run(2, resync_every=100, records=True)
        state[G] = (g.energies(), g.positions(), g.sums())
        g.close()
    check_records(recs[4], recs[1], "fp32c")
    (ea, pa, sa), (eb, pb, sb) = state[4], state[1]
    for k in eb:
        assert rel(ea[k], eb[k], max(1.0, abs(eb[k]))) < RTOL, k
    for x, y in zip(pa, pb):
        assert rel(x, y, 1.0) < 1e-12
    assert rel(sa, sb, np.abs(sb).max()) < 1e-9
