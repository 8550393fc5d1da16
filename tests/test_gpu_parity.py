"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Bar (FP64 form): every energy component, beta table, structure-factor sum,
per-move dE and force within 1e-10 relative of the oracle; identical
accept/reject decisions on the same RNG stream.  Relative errors of dE use
the scale sum |terms| (dE_real, dE_kC, dE_kLJ, dE_ext) because dE itself can
vanish.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def rel(a, b, scale=None):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    s = np.abs(b) if scale is None else np.asarray(scale, dtype=np.float64)
    s = np.maximum(s, 1e-300)
    return float(np.max(np.abs(a - b) / s)) if a.size else 0.0


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


def make(name, seed=11, precision=0):
    from oracle import pyoracle
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config(name, seed=seed)
    return s, EwaldMC(s, precision=precision), pyoracle.Oracle(s)


@pytest.mark.parametrize("name", ["small", "c1"])
def test_init_energies_and_tables(lib, name):
    s, g, o = make(name)
    eg, ovg = g.mc_init()
    eo, ovo = o.mc_init()
    assert ovg == ovo == False
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, (k, eg[k], eo[k])
    bg, bo = g.beta(), o.beta()
    for a, b in zip(bg, bo):
        assert rel(a, b, np.abs(b) + 1e-300) < RTOL
    sg, so = g.sums(), o.sums()
    assert rel(sg, so, np.abs(so).max()) < RTOL
    gg = g.grid()
    from oracle import pyoracle
    go = pyoracle.grid(s.kmax)
    for k in ("kx", "ky", "kz", "ip", "iq"):
        assert np.array_equal(gg[k], go[k]), k
    assert rel(g.mu(), o.mu(), np.abs(o.mu()).max()) < RTOL
    for a, b in zip(g.tails(), o.tails()):
        assert rel(a, b) < RTOL


@pytest.mark.parametrize("name", ["small", "c1"])
def test_real_space_full_and_partial(lib, name):
    s, g, o = make(name)
    cg, ug, fxg, fyg, fzg, ovg = g.rspace_full()
    co, uo, fxo, fyo, fzo, ovo = o.rspace_full()
    assert ovg == ovo
    assert rel(cg, co, np.abs(co).max()) < RTOL
    assert rel(ug, uo, np.abs(uo).max()) < RTOL
    for a, b in ((fxg, fxo), (fyg, fyo), (fzg, fzo)):
        assert rel(a, b, np.abs(b).max()) < RTOL
    rng = np.random.default_rng(3)
    for imol in rng.choice(s.nmol, 4, replace=False):
        f, l = s.mol_first[imol], s.mol_last[imol]
        d = rng.uniform(-0.01, 0.01, 3)
        x, y, z = s.x[f:l + 1] + d[0], s.y[f:l + 1] + d[1], s.z[f:l + 1] + d[2]
        pg = g.rspace_partial(imol, x, y, z)
        po = o.rspace_partial(imol, x, y, z)
        assert pg["overlap"] == po["overlap"]
        if po["overlap"]:
            continue
        sc = max(np.abs(po["uu"]).max(), 1.0)
        assert rel(pg["uu"], po["uu"], sc) < RTOL
        assert rel(pg["comps"][:3], po["comps"][:3], np.abs(po["comps"][:3]).max()) < RTOL
        for k in ("fx", "fy", "fz", "uu_back", "fx_back", "fy_back", "fz_back"):
            assert rel(pg[k], po[k], max(np.abs(po[k]).max(), 1e-12)) < RTOL, k


@pytest.mark.parametrize("name", ["small", "c1"])
def test_kspace_delta_and_forces(lib, name):
    s, g, o = make(name)
    g.mc_init()
    o.mc_init()
    rng = np.random.default_rng(5)
    for imol in rng.choice(s.nmol, 5, replace=False):
        f, l = s.mol_first[imol], s.mol_last[imol]
        d = rng.uniform(-0.02, 0.02, 3)
        x, y, z = s.x[f:l + 1] + d[0], s.y[f:l + 1] + d[1], s.z[f:l + 1] + d[2]
        dg = g.kspace_delta(imol, x, y, z)
        do = o.kspace_delta(imol, x, y, z)
        sc = abs(do[0]) + abs(do[1])
        assert rel(dg, do, sc) < RTOL
        fg = np.stack(g.kspace_forces(f, x, y, z), 1)
        fo = np.stack(o.kspace_forces(f, x, y, z), 1)
        assert rel(fg, fo, np.abs(fo).max()) < RTOL


@pytest.mark.parametrize("block", [256, 512, 64, 128])
@pytest.mark.parametrize("name,ncyc", [("small", 20), ("small_nobias", 20), ("c1", 1), ("cells", 1)])
def test_mc_trace_matches_oracle(lib, name, ncyc, block):
    """Every launch shape of the chain kernel (workgroup of 256 or 512 per
    chain, or one wave per chain) against the oracle on one RNG stream."""
    s, g, o = make(name, seed=21)
    g.set_block_size(block)
    g.mc_init()
    o.mc_init()
    g.set_rng(99, 1)
    o.set_rng(99, 1)
    ng, rg = g.run(ncyc, records=True)
    no, ro = o.run(ncyc)
    assert np.array_equal(rg["imol"], ro["imol"])
    assert np.array_equal(rg["overlap"], ro["overlap"])
    assert np.array_equal(rg["accepted"], ro["accepted"]), "accept/reject decisions differ"
    assert ng == no
    ok = ro["overlap"] == 0
    scale = (np.abs(ro["du_real"]) + np.abs(ro["du_kc"]) + np.abs(ro["du_klj"]) + np.abs(ro["du_ext"]))[ok]
    assert rel(rg["du_total"][ok], ro["du_total"][ok], scale) < RTOL
    for k in ("du_real", "du_kc", "du_klj", "du_ext"):
        assert rel(rg[k][ok], ro[k][ok], scale) < RTOL, k
    for k in ("f_old", "t_old", "f_new", "t_new"):
        # per component, relative to the component or, for components much
        # smaller than their vector, to 1e-4 of the vector's length
        a, b = rg[k][ok], ro[k][ok]
        den = np.maximum(np.abs(b), 1e-4 * np.linalg.norm(b, axis=1, keepdims=True))
        assert float(np.max(np.abs(a - b) / np.maximum(den, 1e-300))) < 1e-8, k
    assert g.rng_counter() == o.rng_counter()
    # state after the run
    eg, eo = g.energies(), o.energies()
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, k
    for a, b in zip(g.positions(), o.positions()):
        assert rel(a, b, 1.0) < 1e-12
    for a, b in zip(g.real_arrays(), o.real_arrays()):
        assert rel(a, b, np.abs(b).max()) < RTOL
    assert rel(g.sums(), o.sums(), np.abs(o.sums()).max()) < RTOL


def test_mc_incremental_state_matches_fresh_sums(lib):
    """Size-independent property at the c2 workload: after device moves the
    incrementally updated energies equal a from-scratch recompute."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c2")
    g = EwaldMC(s)
    g.mc_init()
    g.set_rng(5)
    nacc, _ = g.run(1)
    assert 0 < nacc <= s.nmol
    e_inc = g.energies()
    x, y, z = g.positions()
    h = EwaldMC(s)
    h.set_positions(x, y, z)
    e_new, ov = h.mc_init()
    assert not ov
    for k in e_new:
        assert abs(e_inc[k] - e_new[k]) <= 1e-9 * max(1.0, abs(e_new[k])), (k, e_inc[k], e_new[k])


@pytest.mark.parametrize("block", [256, 64, 128])
def test_batch_equals_single(lib, block):
    """Uneven replicas (sizes, compositions, temperatures) advanced in one
    launch equal the same replicas run one by one, bit for bit."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC, batch_run
    systems = [replica_system(n, c, t, seed=i) for i, (n, c, t) in enumerate([(40, 0.5, 300.0), (64, 0.0, 290.0),
                                                                             (27, 1.0, 340.0), (50, 0.3, 320.0),
                                                                             (33, 0.8, 285.0)])]
    a = [EwaldMC(s) for s in systems]
    b = [EwaldMC(s) for s in systems]
    for x in a + b:
        x.set_block_size(block)
    for i, (x, y) in enumerate(zip(a, b)):
        x.mc_init(); y.mc_init()
        x.set_rng(17, i); y.set_rng(17, i)
    nacc, _ = batch_run(a, 2)
    for i, y in enumerate(b):
        n, _ = y.run(2)
        assert n == nacc[i]
        for p, q in zip(a[i].positions(), y.positions()):
            assert np.array_equal(p, q)


FP32C_WAVE_SKIP = ("the FP32C wave-per-replica kernel made an illegal memory access on its first GPU launch; "
                   "its cause is not found, so it is switched off (FP32C chains run the workgroup kernel)")


@pytest.mark.parametrize("block,precision", [
    (256, 1), (64, 0), (128, 0),
    pytest.param(64, 1, marks=pytest.mark.skip(reason=FP32C_WAVE_SKIP)),
])
def test_graph_batch_replays_equal_eager_batch(lib, block, precision):
    """A captured sweep replayed several times (eager host work and device
    reads between replays) equals the same sweeps launched eagerly, replay
    by replay, for uneven replicas."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import Batch, EwaldMC, batch_run
    specs = [(40, 0.5, 300.0), (64, 0.0, 290.0), (27, 1.0, 340.0), (50, 0.3, 320.0), (33, 0.8, 285.0)]
    systems = [replica_system(n, c, t, seed=10 + i) for i, (n, c, t) in enumerate(specs)]
    a = [EwaldMC(s, precision=precision) for s in systems]
    b = [EwaldMC(s, precision=precision) for s in systems]
    for i, (x, y) in enumerate(zip(a, b)):
        for r in (x, y):
            r.set_block_size(block)
            r.mc_init()
            r.set_rng(31, i)
    g = Batch(a, 1)
    for rep in range(3):
        na, _ = g.launch()
        nb, _ = batch_run(b, 1)
        assert np.array_equal(na, nb), rep
        for x, y in zip(a, b):
            for p, q in zip(x.positions(), y.positions()):
                assert np.array_equal(p, q), rep
            ex, ey = x.energies(), y.energies()
            assert all(ex[k] == ey[k] for k in ex), rep
    g.close()


def test_fp32c_resync_every_n_moves_keeps_the_stream(lib):
    """FP32C with a from-scratch S(k) every 37 moves (not a divisor of the
    cycle, so segments resume inside a cycle): the same molecules, draws and
    decisions as the FP64 run of the same stream, the same RNG counter at the
    end, and S(k) right after the run equal to a fresh rebuild (the last
    resync is at most 36 moves old)."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC
    s = replica_system(40, 0.5, 300.0, seed=77)
    a, b = EwaldMC(s, precision=0), EwaldMC(s, precision=1)
    for x in (a, b):
        x.mc_init()
        x.set_rng(5, 3)
    na, ra = a.run(3, records=True)
    nb, rb = b.run(3, resync_every=37, records=True)
    assert s.nmol % 37 != 0
    assert na == nb
    assert a.rng_counter() == b.rng_counter()
    assert np.array_equal(ra["imol"], rb["imol"])
    assert np.array_equal(ra["accepted"], rb["accepted"])
    ok = ra["overlap"] == 0
    scale = (np.abs(ra["du_real"]) + np.abs(ra["du_kc"]) + np.abs(ra["du_klj"]) + np.abs(ra["du_ext"]))[ok]
    assert rel(rb["du_total"][ok], ra["du_total"][ok], scale) < 1e-6
    for p, q in zip(a.positions(), b.positions()):
        assert rel(p, q, 1.0) < 1e-9
    s0 = b.sums()
    b.resync()
    assert rel(s0, b.sums(), np.abs(b.sums()).max()) < 1e-9
    a.close()
    b.close()


def test_uneven_replicas_fp32c_state_stays_consistent(lib):
    """Batched uneven replicas (sizes, NaCl concentration, temperature), FP32C:
    after two sweeps every replica's incrementally updated energies agree with
    a from-scratch FP64 recompute of its own state."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC, batch_run
    rng = np.random.default_rng(20261206)
    reps = []
    for i in range(6):
        s = replica_system(int(rng.integers(64, 256)), float(rng.uniform(0, 1)), float(rng.uniform(280, 360)),
                           seed=100 + i)
        r = EwaldMC(s, precision=1)
        r.set_block_size(64)
        r.mc_init()
        r.set_rng(7, i)
        reps.append(r)
    nacc, _ = batch_run(reps, 2)
    assert (nacc > 0).all()
    for r in reps:
        inc = r.energies()
        r.resync()
        fresh_k = r.energies()
        # k-space energies: incremental FP32C sums vs the FP64 rebuild
        for k in ("E_kC", "E_kLJ"):
            assert abs(inc[k] - fresh_k[k]) <= 1e-6 * max(1.0, abs(fresh_k[k])), (k, inc[k], fresh_k[k])


def test_run_const_matches_oracle(lib):
    s, g, o = make("small_nobias", seed=4)
    g.mc_init()
    o.mc_init()
    rng = np.random.default_rng(8)
    imol = rng.integers(0, s.nmol, 30).astype(np.int32)
    u = rng.uniform(0.3, 0.7, 30)
    u[np.abs(u - 0.5) < 1e-3] = 0.4
    ng, rg = g.run_const(imol, u)
    no, ro = o.run_const(imol, u)
    assert ng == no
    assert np.array_equal(rg["accepted"], ro["accepted"])
    ok = ro["overlap"] == 0
    scale = (np.abs(ro["du_real"]) + np.abs(ro["du_kc"]) + np.abs(ro["du_klj"]) + np.abs(ro["du_ext"]))[ok]
    assert rel(rg["du_total"][ok], ro["du_total"][ok], scale) < RTOL


def test_forces_reference_convention_and_gradient(lib):
    s, g, o = make("small", seed=2)
    g.mc_init()
    o.mc_init()
    fg, fr = g.forces()
    fo = o.forces_refconv()
    assert rel(fr, fo, np.abs(fo).max()) < 1e-8
    # central differences of the total energy on a few atoms (h = 1e-4 A)
    from yamc_amd.ewald import EwaldMC
    h = 1e-4 / s.params.box_length
    x0, y0, z0 = s.positions()
    probe = EwaldMC(s)
    worst = 0.0
    for i in (0, 1, 5, s.natom - 1):
        for d, arr in enumerate((x0, y0, z0)):
            vals = []
            for sgn in (1, -1):
                xs = [x0.copy(), y0.copy(), z0.copy()]
                xs[d][i] += sgn * h
                probe.set_positions(*xs)
                e, _ = probe.mc_init()
                vals.append(e["U_total"])
            fd = -(vals[0] - vals[1]) / (2 * h)
            worst = max(worst, abs(fd - fg[i, d]) / max(np.abs(fg).max(), 1.0))
    assert worst < 1e-5, worst


def test_fp32c_error_is_bounded(lib):
    """FP32-compensated form against FP64 on the same stream (256 waters, no
    force bias; the force-biased 2128-molecule case is in
    test_gpu_configs.py): identical decisions over the whole sweep, per-move
    dE within 1e-6 kJ/mol, forces within 1e-5."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c1", seed=3)
    a, b = EwaldMC(s, precision=0), EwaldMC(s, precision=1)
    a.mc_init(); b.mc_init()
    ea, eb = a.energies(), b.energies()
    for k in ea:
        assert abs(ea[k] - eb[k]) <= 1e-9 * max(1.0, abs(ea[k])), k
    a.set_rng(3); b.set_rng(3)
    _, ra = a.run(1, records=True)
    _, rb = b.run(1, records=True)
    assert np.array_equal(ra["imol"], rb["imol"])
    assert np.array_equal(ra["accepted"], rb["accepted"]), "decisions differ"
    n = len(ra)
    ok = ra["overlap"][:n] == 0
    d = np.abs(ra["du_total"][:n] - rb["du_total"][:n])[ok]
    # per-move dE within 1e-6 kJ/mol = 4.04e-7 kT at 298 K
    assert np.max(d) < 4.04e-7, np.max(d)
    fa = np.concatenate([ra["f_old"][:n][ok], ra["f_new"][:n][ok]])
    fb = np.concatenate([rb["f_old"][:n][ok], rb["f_new"][:n][ok]])
    fe = np.linalg.norm(fa - fb, axis=1) / np.linalg.norm(fa, axis=1)
    assert fe.max() < 1e-5, fe.max()


def test_autograd_harness_gradient(lib):
    """torch.autograd over the adjoint: dU_total/dr from backward equals
    central differences of U_total (h = 1e-4 A) and minus the library's
    gradient forces."""
    import torch
    from yamc_amd import config
    from yamc_amd.autograd import ewald_energy
    from yamc_amd.ewald import EwaldMC
    s = config("c1", seed=9)
    mc = EwaldMC(s)
    x, y, z = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in s.positions())
    e = ewald_energy(mc, x, y, z)
    e.backward()
    grad = torch.stack([x.grad, y.grad, z.grad], 1).numpy()
    g, _ = mc.forces()
    assert np.array_equal(grad, -g)
    probe = EwaldMC(s)
    h = 1e-4 / s.params.box_length
    x0, y0, z0 = s.positions()
    rng = np.random.default_rng(9)
    worst = 0.0
    for i in rng.choice(s.natom, 8, replace=False):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                xs = [x0.copy(), y0.copy(), z0.copy()]
                xs[d][i] += sgn * h
                probe.set_positions(*xs)
                ev, _ = probe.mc_init()
                vals.append(ev["U_total"])
            worst = max(worst, abs((vals[0] - vals[1]) / (2 * h) - grad[i, d]) / np.abs(grad).max())
    assert worst < 1e-5, worst


@pytest.mark.parametrize("case", __import__("golden_cases").cases(), ids=lambda p: p.split("/")[-1])
def test_gpu_matches_reference_golden(lib, case):
    """The HIP path against the compiled reference's own outputs."""
    import golden_cases as gc
    from yamc_amd.ewald import EwaldMC
    sysm, g = gc.load(case)
    gpu = EwaldMC(sysm)
    gpu.mc_init()
    gc.check_init(gpu, sysm, g)

    def step(imol, u):
        n, recs = gpu.run_const([imol], [u])
        return n

    gc.check_trace(gpu, sysm, g, step)

    def vstep(u, p, lmvs):
        return gpu.volume_move(p, lmvs, u=u)

    assert gc.check_volume_steps(gpu, sysm, g, vstep) == 4

    def xstep(i1, i2, u):
        acc, rec = gpu.xchange(i1, i2, u=u)
        return acc

    assert gc.check_xchange_steps(gpu, sysm, g, xstep) == 8


@pytest.mark.parametrize("case", __import__("golden_cases").runmc_cases(), ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("block", [256, 64])
def test_gpu_replays_reference_run_mc(lib, case, block):
    """run_mc cycles on the device (chain kernel: pressure draw, exchange
    draw, pick, proposal rejection loops, acceptance; host: volume and
    exchange steps) consuming the reference's own recorded uniform stream:
    identical step counts, every energy component after every cycle within
    1e-10, the final state, and the same number of uniforms drawn as the
    reference's run_mc.  Shape 64 runs the move cycles in the wave kernel
    (FP64, no exchange weights) or the workgroup kernel."""
    import golden_cases as gc
    from yamc_amd.ewald import EwaldMC
    sysm, g = gc.load(case)
    gpu = EwaldMC(sysm)
    gpu.set_block_size(block)
    e, _ = gpu.mc_init()
    ev = gc.energies_vec(e, gpu.mu(), gpu.tails())
    assert gc.rel(ev, g["e_init"], g["e_init"]) < gc.RTOL
    psm, p, lmvs = gc.runmc_setup(gpu, g)
    ncyc, _ = gc.check_runmc(gpu, sysm, g, lambda: gpu.run_npt(1, psm, p, lmvs))
    assert ncyc == len(g["cycle_e"])


def test_replay_list_overrun_is_an_error(lib):
    """A run that needs more uniforms than the replayed list holds fails
    with ERROR_LIMITS instead of inventing draws."""
    from yamc_amd import YamcError, config
    from yamc_amd.ewald import EwaldMC
    g = EwaldMC(config("small", seed=5))
    g.mc_init()
    g.set_rng_list(np.full(20, 0.37))
    with pytest.raises(YamcError) as ei:
        g.run(1)
    assert ei.value.code == 2


def test_rdf_counts_match_oracle(lib):
    """yamc_rdf (COM pair counts, mcrdf2.f90 count_distances) equals the CPU
    restatement exactly on the 2128-molecule ionic system, before and after
    device moves."""
    from oracle import pyoracle
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c2", seed=12)
    g = EwaldMC(s)
    g.mc_init()
    L = s.params.box_length
    for step in range(2):
        c = g.rdf(s.mol_type, 3, dr=0.1)
        x, y, z = g.positions()
        s2 = config("c2", seed=12)
        s2.x, s2.y, s2.z = x, y, z
        want = pyoracle.rdf_counts(s2, s.mol_type, 3, 0.1, 0.5 * L, c.shape[2])
        assert np.array_equal(c, want), step
        assert int(c.sum()) > 0
        g.set_rng(3)
        g.run(1)


def test_fp32c_rdf_within_statistical_error(lib):
    """The FP32C form's COM RDFs equal the FP64 form's within statistical
    error: 32 independent replicas per form (own Philox streams) of 256
    SPC/E + 5 NaCl, 20 sweeps then 40 frames one sweep apart; z-scores of
    the replica means of g(r) per bin (water-water and water-ion pairs)."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC, batch_run, rdf_normalize
    s = replica_system(256, 1.08, 298.0, seed=77)
    nsp = int(s.mol_type.max()) + 1
    n_per = [int((s.mol_type == k).sum()) for k in range(nsp)]
    V = s.params.box_length ** 3
    gs = {}
    for form, prec, base in (("fp64", 0, 0), ("fp32c", 1, 1000)):
        reps = [EwaldMC(s, precision=prec) for _ in range(32)]
        for i, r in enumerate(reps):
            r.set_block_size(64)
            r.mc_init()
            r.set_rng(2026, base + i)
        batch_run(reps, 20)
        counts = [None] * len(reps)
        for _ in range(40):
            batch_run(reps, 1)
            for i, r in enumerate(reps):
                counts[i] = r.rdf(s.mol_type, nsp, dr=0.2, counts=counts[i])
        gs[form] = np.stack([rdf_normalize(c, 40, n_per, V, 0.2) for c in counts])
    for (i, j) in [(0, 0), (0, 1), (0, 2)][: nsp]:
        a, b = gs["fp64"][:, i, j], gs["fp32c"][:, i, j]
        se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
        m = se > 0
        z = (a.mean(0) - b.mean(0))[m] / se[m]
        assert np.abs(z).max() < 6.0, (i, j, np.abs(z).max())
        assert (z ** 2).mean() < 2.5, (i, j, (z ** 2).mean())


@pytest.mark.parametrize("name,ncyc,psm", [("small", 24, 1.0), ("c1", 4, 1.0), ("cells", 4, 1.0)])
def test_npt_cycles_match_oracle(lib, name, ncyc, psm):
    """run_mc cycles with volume steps (runmc.f90:204-268, mcvol
    MC.f90:701-825) on one RNG stream: identical move and volume decisions,
    box length, energies, positions and sums after the run."""
    import golden_cases as gc
    s, g, o = make(name, seed=31)
    g.mc_init()
    o.mc_init()
    g.set_rng(123, 2)
    o.set_rng(123, 2)
    p = gc.pressure_kT_A3(1e5, s.params.temp)
    lmvs = float(np.log(1.03))
    rg = g.run_npt(ncyc, psm, p, lmvs)
    ro = o.run_npt(ncyc, psm, p, lmvs)
    assert rg == ro, (rg, ro)
    assert ro[1] > 0, "no volume step drawn"
    assert g.rng_counter() == o.rng_counter()
    assert abs(g.box_length() - o.box_length()) <= 1e-13 * o.box_length()
    eg, eo = g.energies(), o.energies()
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, k
    for a, b in zip(g.positions(), o.positions()):
        assert rel(a, b, 1.0) < 1e-12
    assert rel(g.sums(), o.sums(), np.abs(o.sums()).max()) < RTOL


def test_volume_step_rejected_restores_state(lib):
    """A volume step that cannot be accepted (huge pressure, u near 1 so the
    box grows) leaves positions, box, energies and sums as they were."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("small", seed=5)
    g = EwaldMC(s)
    g.mc_init()
    e0, x0, S0, L0 = g.energies(), g.positions(), g.sums(), g.box_length()
    acc, L = g.volume_move(1e3, float(np.log(1.05)), u=0.99)
    assert not acc and L == L0
    e1 = g.energies()
    for k in e0:
        assert e1[k] == e0[k], k
    for a, b in zip(g.positions(), x0):
        assert np.array_equal(a, b)
    assert np.array_equal(g.sums(), S0)
    # the chain continues from the restored state like a fresh context
    g.set_rng(3)
    h = EwaldMC(s)
    h.mc_init()
    h.set_rng(3)
    assert g.run(2)[0] == h.run(2)[0]
    assert g.energies() == h.energies()


def test_fp32c_npt_state_matches_fresh_sums(lib):
    """FP32C form: after cycles with volume steps the carried state (box,
    scaled beta, energies) equals a from-scratch rebuild of the same
    configuration in the same context (mc_init: beta from the box's own
    xclb, sums and energies anew)."""
    import golden_cases as gc
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c1", seed=8)
    g = EwaldMC(s, precision=1)
    g.mc_init()
    g.set_rng(77)
    # 8 cycles on this stream draw 3 volume steps, one accepted (as the oracle runs them in FP64)
    _, nv, nva = g.run_npt(8, 1.0, gc.pressure_kT_A3(1e5, s.params.temp), float(np.log(1.005)))
    assert nv > 0 and nva > 0, (nv, nva)
    assert g.box_length() != s.params.box_length
    e_carried, b_carried = g.energies(), g.beta()
    e_fresh, ov = g.mc_init()
    assert not ov
    b_fresh = g.beta()
    for a, b in zip(b_carried, b_fresh):
        assert rel(a, b, np.abs(b).max()) < 1e-12
    # carried FP32C energies drift by at most the per-move bound times the moves since the last rebuild
    assert abs(e_carried["U_total"] - e_fresh["U_total"]) < 8 * 256 * 4.04e-7


def _exchange_weights(s, kind):
    mt = np.asarray(s.mol_type)
    if kind == "ions":
        return (mt > 0).astype(np.float64)
    if kind == "all":
        return np.ones(s.nmol)
    w = np.zeros(s.nmol)
    w[::16] = 1.0
    return w


@pytest.mark.parametrize("name,ncyc,kind", [("small", 20, "ions"), ("small", 10, "all"), ("c1", 2, "every16")])
def test_exchange_cycles_match_oracle(lib, name, ncyc, kind):
    """run_mc cycles with exchange steps (runmc.f90:223-266, mc_xchange
    MC.f90:1325-1954) on one RNG stream: the device runs the moves and hands
    each exchange step to the host; identical step kinds, molecules and
    decisions, and the same state after the run."""
    s, g, o = make(name, seed=41)
    xc = _exchange_weights(s, kind)
    g.set_frequencies(None, xc)
    o.set_frequencies(None, xc)
    g.mc_init()
    o.mc_init()
    g.set_rng(55, 3)
    o.set_rng(55, 3)
    ng, rg = g.run(ncyc, records=True)
    no, ro = o.run(ncyc)
    for k in ("imol", "imol2", "overlap", "accepted"):
        assert np.array_equal(rg[k], ro[k]), k
    xs = ro["imol2"] >= 0
    assert xs.sum() > 0, "no exchange step drawn"
    if kind != "ions":  # Na/Cl swaps in the 8-water box are all rejected on this stream
        assert ro["accepted"][xs].sum() > 0, "no accepted exchange step"
    assert ng == no
    ok = ro["overlap"] == 0
    scale = (np.abs(ro["du_real"]) + np.abs(ro["du_kc"]) + np.abs(ro["du_klj"]) + np.abs(ro["du_ext"]))[ok]
    for k in ("du_total", "du_real", "du_kc", "du_klj", "du_ext"):
        assert rel(rg[k][ok], ro[k][ok], scale) < RTOL, k
    assert g.rng_counter() == o.rng_counter()
    eg, eo = g.energies(), o.energies()
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, k
    for a, b in zip(g.positions(), o.positions()):
        assert rel(a, b, 1.0) < 1e-12
    for a, b in zip(g.real_arrays(), o.real_arrays()):
        assert rel(a, b, np.abs(b).max()) < RTOL
    assert rel(g.sums(), o.sums(), np.abs(o.sums()).max()) < RTOL


@pytest.mark.parametrize("block", [256, 64])
def test_weighted_move_picker_matches_oracle(lib, block):
    """Non-uniform move weights (BiasedRandom intervals with a zero weight)
    pick the same molecules on the device as in the oracle."""
    s, g, o = make("small", seed=43)
    mv = np.where(np.asarray(s.mol_type) > 0, 5.0, 1.0)
    mv[2] = 0.0
    g.set_block_size(block)
    g.set_frequencies(mv, None)
    o.set_frequencies(mv, None)
    g.mc_init()
    o.mc_init()
    g.set_rng(9, 1)
    o.set_rng(9, 1)
    ng, rg = g.run(10, records=True)
    no, ro = o.run(10)
    assert np.array_equal(rg["imol"], ro["imol"])
    assert not np.any(ro["imol"] == 2)
    assert np.array_equal(rg["accepted"], ro["accepted"]) and ng == no
    assert g.rng_counter() == o.rng_counter()


def test_fp32c_exchange_state_matches_fresh_sums(lib):
    """FP32C form with exchange steps: the carried energies equal a rebuild
    of the final configuration within the FP32C bound."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c1", seed=8)
    g = EwaldMC(s, precision=1)
    g.set_frequencies(None, _exchange_weights(s, "every16"))
    g.mc_init()
    g.set_rng(78)
    _, recs = g.run(2, records=True)
    assert int((recs["imol2"] >= 0).sum()) > 0
    e_carried = g.energies()
    e_fresh, ov = g.mc_init()
    assert not ov
    assert abs(e_carried["U_total"] - e_fresh["U_total"]) < 2 * 256 * 4.04e-7


def test_batch_refuses_exchange_contexts(lib):
    from yamc_amd import _lib as L
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC, batch_run
    s = config("small", seed=5)
    g = EwaldMC(s)
    g.set_frequencies(None, _exchange_weights(s, "ions"))
    g.mc_init()
    with pytest.raises(L.YamcError):
        batch_run([g], 1)


def test_runmc_driver_npt_files(lib, tmp_path):
    """The run_mc host driver on the reference's ions100 parameter and
    composition files (300 NPT cycles): output files consistent with the
    run's counters and final state."""
    from yamc_amd import formats as fm
    from yamc_amd import system as ys
    from yamc_amd.runmc import run
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ions100")
    prm = fm.read_parameters(os.path.join(d, "parameters.prm"))
    prm.update(n_store_energy_interval=50.0, n_store_traj_interval=100.0)
    species, counts = fm.read_composition(os.path.join(d, "ions100.composition"))
    sysm = ys.lattice_system(fm.params_from(prm, sum(counts)), species, counts, 3)
    res = run(sysm, prm, 300, 0, seed=3, out_dir=str(tmp_path))
    assert res["n_vol"] > 0 and res["n_vol_acc"] > 0
    assert res["n_mv"] + res["n_vol"] * sysm.nmol == 300 * sysm.nmol
    e = fm.read_table(str(tmp_path / "energy.dat"))
    assert e.shape == (6, 9) and np.array_equal(e[:, 0], np.arange(50, 301, 50))
    b = fm.read_table(str(tmp_path / "boxlength.dat"))
    assert len(b) == res["n_vol_acc"] and abs(b[-1, 1] - res["box"][-1]) < 1e-12 * b[-1, 1]
    raw = open(tmp_path / "traj.moltraj", "rb").read()
    assert len(raw) == 3 * sysm.nmol * 12
    fr = fm.read_table(str(tmp_path / "frames.dat"))
    assert fr.shape[0] == 3 and fr[-1, 4] == res["n_vol"]
    pos, _ = fm.moltab_decode(raw, sysm.nmol, 2, 2, offset=2 * sysm.nmol * 12)
    assert np.all(np.abs(pos) <= 0.5 + 1e-4)
    # the last energy row is the final state
    assert abs(e[-1, 1] - res["final_energies"]["U_total"]) < 1e-9 * abs(e[-1, 1])
    assert abs(e[-1, 1] - e[-1, 2] - e[-1, 3]) < 1e-9 * abs(e[-1, 1])


@pytest.mark.parametrize("precision,cells,big", [(0, False, False), (1, False, False), (1, True, False),
                                                 (1, False, True)])
def test_batch_two_chains_per_cu_equals_single(lib, precision, cells, big):
    """More 256-thread chains than CUs, so two chains share a CU: every
    replica still equals its own run alone, where it has a CU to itself (also
    with the cell-list scan: 3 A cut-off, 4 cells per axis).  The scan's
    packed position copy (4 bytes per atom) sits in LDS for the small
    replicas; the big ones (3000 waters, 9000 atoms: 36 KB next to the
    kernel's static LDS is more than half of a CU's 160 KB) read it from HBM
    in the batch and from LDS alone."""
    import torch
    from yamc_amd import replica_system
    from yamc_amd.ewald import EwaldMC, batch_run
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    n = ncu + 8
    kw = dict(rc_angstrom=3.0, alpha_per_angstrom=1.0667, sk=3.0) if cells else {}
    if big:
        systems = [replica_system(3000, 0.0, 298.0, seed=100 + i) for i in range(n)]
    else:
        systems = [replica_system(64 if cells else 27 + (i % 5) * 4, 0.5 * (i % 3), 290.0 + i % 40, seed=100 + i,
                                  **kw) for i in range(n)]
    ncyc = 1 if big else 2
    a = [EwaldMC(s, precision=precision) for s in systems]
    for i, x in enumerate(a):
        x.set_block_size(256)
        x.mc_init()
        x.set_rng(23, i)
    nacc, _ = batch_run(a, ncyc)
    for i in ((0, n - 1) if big else (0, 1, n // 2, n - 1)):
        y = EwaldMC(systems[i], precision=precision)
        y.set_block_size(256)
        y.mc_init()
        y.set_rng(23, i)
        m, _ = y.run(ncyc)
        assert m == nacc[i]
        for p, q in zip(a[i].positions(), y.positions()):
            assert np.array_equal(p, q)
        y.close()
    for x in a:
        x.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_host_driven_moves_commit_consistently(lib, precision):
    """A host-side move loop over the per-function surface: yamc_kspace_delta
    for dU_k, yamc_mc_commit_move for accepted displacements.  After 30 moves
    the carried S(k), energy components and per-atom real-space arrays equal
    a rebuild at the final positions, and dU_k equals the change of E_k."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("small", seed=12)
    g = EwaldMC(s, precision=precision)
    g.mc_init()
    rng = np.random.default_rng(4)
    ncommit = 0
    for _ in range(30):
        im = int(rng.integers(0, s.nmol))
        f, l = int(s.mol_first[im]), int(s.mol_last[im])
        x, y, z = (a[f:l + 1] for a in g.positions())
        d = rng.uniform(-0.01, 0.01, 3)
        nx, ny, nz = x + d[0], y + d[1], z + d[2]
        dc, dl = g.kspace_delta(im, nx, ny, nz)
        e0 = g.energies()
        if g.commit_move(im, nx, ny, nz, dc, dl):
            ncommit += 1
            e1 = g.energies()
            assert abs((e1["E_kC"] - e0["E_kC"]) - dc) <= 1e-12 * max(1.0, abs(e0["E_kC"]))
    assert ncommit > 20
    h = EwaldMC(s, precision=precision)
    h.set_positions(*g.positions())
    e_new, ov = h.mc_init()
    assert not ov
    tol = 1e-10 if precision == 0 else 1e-6
    eg = g.energies()
    for k in e_new:
        assert abs(eg[k] - e_new[k]) <= tol * max(1.0, abs(e_new[k])), (k, eg[k], e_new[k])
    for a, b in zip(g.real_arrays(), h.real_arrays()):
        assert rel(a, b, np.abs(b).max()) < 1e-10
    assert rel(g.sums(), h.sums(), np.abs(h.sums()).max()) < (1e-10 if precision == 0 else 1e-7)


def test_set_box_matches_accepted_volume_step(lib):
    """yamc_set_box with scaled positions reproduces the state of an accepted
    mcvol step of the oracle to the same new box length (and a second call
    back to the old box restores the energies)."""
    import golden_cases as gc
    s, g, o = make("small", seed=21)
    g.mc_init()
    o.mc_init()
    e0 = g.energies()
    acc, L, ov = o.volume_move(gc.pressure_kT_A3(1e5, s.params.temp), float(np.log(1.02)), u=0.52)
    assert acc and not ov
    assert not g.set_box(L, True)
    assert abs(g.box_length() - L) < 1e-13 * L
    eg, eo = g.energies(), o.energies()
    for k in eo:
        assert rel(eg[k], eo[k], max(1.0, abs(eo[k]))) < RTOL, k
    for a, b in zip(g.positions(), o.positions()):
        assert rel(a, b, 1.0) < 1e-13
    assert not g.set_box(s.params.box_length, True)
    e2 = g.energies()
    for k in e0:
        assert rel(e2[k], e0[k], max(1.0, abs(e0[k]))) < 1e-9, k
