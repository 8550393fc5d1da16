"""The benchmark workloads at their full sizes on the device.

c2  2048 SPC/E + 40 Na + 40 Cl, force bias: the FP32-compensated form
    against FP64 on one stream over a whole sweep (2128 moves).
c3  8192 SPC/E + 160 Na + 160 Cl (24,896 atoms): reference-convention
    forces against the oracle per component, and the adjoint gradient
    against central differences on 64 seeded atoms.
c4  64 uneven replicas (128-2048 waters, 0-1 M NaCl, 280-360 K) in one
    captured FP32C sweep with rebuilds: replicas equal their single runs bit
    for bit, and every replica's carried state equals a from-scratch FP64
    recompute.
c5  32768 SPC/E + 600 Na + 600 Cl (99,504 atoms): FP32C with an FP64
    rebuild every 1000 moves against FP64 on one stream.

Bounds: per-move dE within 1e-6 kJ/mol (4.04e-7 kT at 298 K) and forces
within 1e-5 relative for FP32C against FP64; reference-convention forces
within 1e-8 relative per component (below a floor of 1e-4 of the atom's
force magnitude, relative to that floor) for FP64 against the oracle.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

DE_KT = 4.04e-7  # 1e-6 kJ/mol at 298 K in kT


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


def component_rel(a, b, floor_frac):
    """max over atoms and components of |a - b| / max(|b|, floor_frac |b_atom|)."""
    a, b = np.asarray(a), np.asarray(b)
    floor = floor_frac * np.linalg.norm(b, axis=-1, keepdims=True)
    return float(np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(b), floor), 1e-300)))


def compare_streams(ra, rb, what, n_bound=None):
    """FP64 records ra against FP32C records rb of the same stream:
    identical picks and decisions throughout, dE and force bounds over the
    first n_bound moves (all by default)."""
    assert np.array_equal(ra["imol"], rb["imol"]), what
    assert np.array_equal(ra["accepted"], rb["accepted"]), what
    assert np.array_equal(ra["overlap"], rb["overlap"]), what
    ra, rb = ra[:n_bound], rb[:n_bound]
    ok = ra["overlap"] == 0
    d = np.abs(ra["du_total"] - rb["du_total"])[ok]
    assert d.max() < DE_KT, (what, d.max())
    for k in ("f_old", "f_new", "t_old", "t_new"):
        fa, fb = ra[k][ok], rb[k][ok]
        e = np.linalg.norm(fa - fb, axis=1) / np.maximum(np.linalg.norm(fa, axis=1), 1e-300)
        assert e.max() < 1e-5, (what, k, e.max())
    return d.max()


def test_c2_fp32c_force_biased_sweep(lib):
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c2", seed=41)
    a, b = EwaldMC(s, precision=0), EwaldMC(s, precision=1)
    ea, _ = a.mc_init()
    eb, _ = b.mc_init()
    for k in ea:
        assert abs(ea[k] - eb[k]) <= 1e-9 * max(1.0, abs(ea[k])), k
    a.set_rng(41, 2)
    b.set_rng(41, 2)
    _, ra = a.run(1, records=True)
    _, rb = b.run(1, resync_every=1000, records=True)
    assert len(ra) == s.nmol == 2128
    compare_streams(ra, rb, "c2")
    assert a.rng_counter() == b.rng_counter()


def test_c3_reference_forces_match_oracle(lib):
    """yamc_forces' reference convention (real space j > i, k space,
    external; mc_calc_forces.f90:230-316) at 24,896 atoms, FP64, against the
    oracle's forces_refconv, per component."""
    from oracle import pyoracle
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c3", seed=20261205)
    g = EwaldMC(s, precision=0)
    g.mc_init()
    _, r = g.forces()
    o = pyoracle.Oracle(s)
    o.mc_init()
    fo = o.forces_refconv()
    assert component_rel(r, fo, 1e-4) < 1e-8


def test_c3_adjoint_gradient_matches_central_differences(lib):
    """The adjoint -dU_total/dr (FP32C form) on 64 seeded atoms against
    central differences of U_total with h = 1e-4 A, relative to the largest
    force.  A probe whose atom has a partner within 2h of the cut-off is left
    out: the reference energy is truncated there (EwaldSumRealSpace.f90:437)
    and no gradient reproduces the jump."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c3", seed=20261205)
    g = EwaldMC(s, precision=1)
    g.mc_init()
    grad, _ = g.forces()  # -dU/dr in box units
    probe = EwaldMC(s, precision=1)
    h = 1e-4 / s.params.box_length
    rc = float(np.sqrt(s.derived["rmax2"]))
    x0, y0, z0 = s.positions()
    atoms = np.random.default_rng(20261205).choice(s.natom, 64, replace=False)
    scale = np.abs(grad).max()
    worst, used = 0.0, 0
    for i in atoms:
        dx = [a - a[i] for a in (x0, y0, z0)]
        dx = [v - np.round(v) for v in dx]
        r = np.sqrt(dx[0] ** 2 + dx[1] ** 2 + dx[2] ** 2)
        r[i] = 0.0
        if np.any(np.abs(r - rc) < 2 * h):
            continue
        used += 1
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                xs = [x0.copy(), y0.copy(), z0.copy()]
                xs[d][i] += sgn * h
                probe.set_positions(*xs)
                e, _ = probe.mc_init()
                vals.append(e["U_total"])
            fd = -(vals[0] - vals[1]) / (2 * h)
            worst = max(worst, abs(fd - grad[i, d]) / scale)
    assert used >= 40
    assert worst < 1e-4, worst


def test_c4_uneven_replicas_batch(lib):
    """64 seeded uneven replicas, FP32C, one captured sweep with an FP64
    rebuild every 1000 moves: the smallest, the largest and a middle replica
    equal their own yamc_mc_run bit for bit, and every replica's carried
    energies equal a from-scratch FP64 recompute of its configuration."""
    from yamc_amd import replica_system
    from yamc_amd.ewald import Batch, EwaldMC
    rng = np.random.default_rng(20261206)
    specs = [(int(rng.integers(128, 2049)), float(rng.uniform(0.0, 1.0)), float(rng.uniform(280.0, 360.0)))
             for _ in range(64)]
    systems = [replica_system(nw, c, t, seed=1000 + i) for i, (nw, c, t) in enumerate(specs)]
    reps = []
    for i, s in enumerate(systems):
        r = EwaldMC(s, precision=1)
        r.mc_init()
        r.set_rng(20261206, i)
        reps.append(r)
    sizes = [s.natom for s in systems]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argsort(sizes)[32])})
    bat = Batch(reps, 1, resync_every=1000)
    nacc, _ = bat.launch()
    bat.close()
    assert (nacc > 0).all()
    for i in picks:
        t = EwaldMC(systems[i], precision=1)
        t.mc_init()
        t.set_rng(20261206, i)
        n, _ = t.run(1, resync_every=1000)
        assert n == nacc[i], i
        for p, q in zip(reps[i].positions(), t.positions()):
            assert np.array_equal(p, q), i
        assert np.array_equal(reps[i].sums(), t.sums()), i
        t.close()
    for i, r in enumerate(reps):
        inc = r.energies()
        f = EwaldMC(systems[i], precision=0)
        f.set_positions(*r.positions())
        fresh, ov = f.mc_init()
        assert not ov
        for k in fresh:
            assert abs(inc[k] - fresh[k]) <= 1e-8 * max(1.0, abs(fresh[k])), (i, k, inc[k], fresh[k])
        f.close()


def test_c5_fp32c_with_rebuilds_against_fp64(lib):
    """99,504 atoms, one stream: one run_mc cycle (33,968 moves) in FP64 and
    in FP32C with an FP64 S(k) rebuild every 1000 moves.  Identical picks
    and decisions over the whole cycle; per-move dE within 1e-6 kJ/mol and
    forces within 1e-5 over the first 3000 moves (further along, the two
    force-biased chains drift apart at the 1e-9 level in their positions, so
    later moves no longer compare the same configuration); then on one
    configuration (the FP64 chain's final state) the k-space dE and forces of
    both forms for 200 trial displacements."""
    from yamc_amd import config
    from yamc_amd.ewald import EwaldMC
    s = config("c5", seed=20261207)
    recs, pos = {}, None
    for prec, rs in ((0, 0), (1, 1000)):
        g = EwaldMC(s, precision=prec)
        g.mc_init()
        g.set_rng(20261207, 0)
        _, recs[prec] = g.run(1, resync_every=rs, records=True)
        if prec == 0:
            pos = g.positions()
        g.close()
    assert len(recs[0]) == s.nmol
    compare_streams(recs[0], recs[1], "c5", n_bound=3000)
    a64, a32 = EwaldMC(s, precision=0), EwaldMC(s, precision=1)
    for h in (a64, a32):
        h.set_positions(*pos)
        h.mc_init()
    rng = np.random.default_rng(20261208)
    x, y, z = pos
    worst, fworst = 0.0, 0.0
    for _ in range(200):
        im = int(rng.integers(0, s.nmol))
        f, l = int(s.mol_first[im]), int(s.mol_last[im])
        dv = rng.uniform(-1.0, 1.0, 3) * s.derived["dr"]
        nx, ny, nz = x[f:l + 1] + dv[0], y[f:l + 1] + dv[1], z[f:l + 1] + dv[2]
        c64, l64 = a64.kspace_delta(im, nx, ny, nz)
        c32, l32 = a32.kspace_delta(im, nx, ny, nz)
        worst = max(worst, abs((c32 + l32) - (c64 + l64)))
        g64 = np.stack(a64.kspace_forces(f, nx, ny, nz), 1)
        g32 = np.stack(a32.kspace_forces(f, nx, ny, nz), 1)
        fworst = max(fworst, float(np.max(np.linalg.norm(g32 - g64, axis=1) / np.linalg.norm(g64, axis=1))))
    assert worst < DE_KT, worst
    assert fworst < 1e-5, fworst
