"""Loads the golden fixtures written from the compiled reference
(tests/golden/make_golden.py) and compares an implementation against them.

Used by the CPU oracle test and by the GPU parity test, so both are pinned to
the reference's own outputs on the same inputs.
"""
import glob
import os

import numpy as np

from yamc_amd import system as ys

HERE = os.path.dirname(os.path.abspath(__file__))
SPECIES = {"SPCE": ys.SPCE, "Na": ys.NA_DANG, "Cl": ys.CL_DANG}
RTOL = 1e-10


def cases():
    return sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))


def runmc_cases():
    """Fixtures of the reference's own run_mc with its own random stream."""
    return sorted(glob.glob(os.path.join(HERE, "golden", "runmc_*.npz")))


def load(path):
    g = dict(np.load(path, allow_pickle=False))
    p = g["params"]
    params = ys.Params(box_length=p[0], temp=p[1], sr=p[2], sk=p[3], alpha=p[4], external_permutivity=p[5],
                       dr_a=p[6], d_angle_degree=p[7], lambda_f=p[8], lambda_c=p[9])
    species = [SPECIES[str(s)] for s in g["species"]]
    sysm = ys.build_system(params, species, [int(c) for c in g["counts"]], g["mol_type"], g["mol_pos"],
                           g["mol_angles"])
    return sysm, g


def rel(a, b, scale):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(scale))), 1e-300)) if a.size else 0.0


def energies_vec(e, mu, tails):
    """Reference energy record layout (ref_driver dump_energies)."""
    return np.array([e["U_total"], e["U_total_ew"], e["U_total_lj"], e["uu_ew"], e["uu_lj6"], e["uu_lj12"],
                     e["E_kC"], e["E_kLJ"], e["U_tail"], e["U_ext"], tails[0], tails[1], tails[2], mu[0], mu[1],
                     mu[2]])


def check_staging(sysm, g):
    for k, mine in (("xx", sysm.x), ("yy", sysm.y), ("zz", sysm.z), ("charge", sysm.charge),
                    ("sigma", sysm.sigma), ("epsilon", sysm.epsilon), ("hard_core", sysm.hard_core)):
        assert rel(mine, g[k], g[k]) < 1e-14, k
    assert np.array_equal(sysm.molnum, g["molnum"])
    assert np.array_equal(sysm.lj_type, g["lj_type"])
    assert np.array_equal(sysm.mol_first, g["first"]) and np.array_equal(sysm.mol_last, g["last"])
    T = sysm.ntype
    for k, mine in (("lj6", sysm.lj6), ("four_eps", sysm.four_eps), ("sigma6", sysm.sigma6)):
        assert rel(mine, g[k], g[k]) < 1e-14, k
    up = np.array([sysm.lj12[t1 * T + t2] for t1 in range(T) for t2 in range(t1, T)])
    assert rel(up, g["lj12_upper"], g["lj12_upper"]) < 1e-14
    d = sysm.derived
    mine = [d["alpha"], d["rmax2"], d["xclb"], d["kext"], d["dr"], d["d_angle"]]
    assert rel(mine, g["derived"], g["derived"]) < 1e-14
    assert sysm.kmax == int(g["kmax"])


def check_init(impl, sysm, g):
    """impl: object with the EwaldMC / Oracle surface after mc_init."""
    grid = impl.grid() if hasattr(impl, "grid") else None
    if grid is not None:
        for k in ("kx", "ky", "kz", "ip", "iq"):
            assert np.array_equal(grid[k], g[k]), k
    b, b6, b12, bl = impl.beta()
    assert rel(b, g["beta"], g["beta"]) < RTOL
    assert rel(b6, g["beta6"], g["beta6"]) < RTOL
    assert rel(b12, g["beta12"], g["beta12"]) < RTOL
    assert rel(bl, g["beta_lj"], g["beta_lj"]) < RTOL
    assert rel(impl.sums(), g["sums"], g["sums"]) < RTOL
    e = impl.energies()
    ev = energies_vec(e, impl.mu(), impl.tails())
    assert rel(ev, g["e_init"], g["e_init"]) < RTOL, np.abs(ev - g["e_init"]) / np.abs(g["e_init"]).max()
    uu, fx, fy, fz = impl.real_arrays()
    assert rel(uu, g["uu"], g["uu"]) < RTOL
    assert rel(np.stack([fx, fy, fz]), np.stack([g["fx"], g["fy"], g["fz"]]), np.stack([g["fx"], g["fy"], g["fz"]])) < RTOL
    fk = g["fk"]
    kx, ky, kz = impl.kspace_forces(0, sysm.x, sysm.y, sysm.z)
    want = np.stack([fk[:, 0] + fk[:, 3], fk[:, 1] + fk[:, 4], fk[:, 2] + fk[:, 5]], 1)
    assert rel(np.stack([kx, ky, kz], 1), want, want) < RTOL


def check_trace(impl, sysm, g, step):
    """step(imol, u) runs one constant-uniform move and returns accepted."""
    for k, (im, u) in enumerate(zip(g["move_imol"], g["move_u"])):
        acc = step(int(im), float(u))
        assert int(acc) == int(g["move_accepted"][k]), f"decision differs at move {k}"
        e = impl.energies()
        ev = energies_vec(e, impl.mu(), impl.tails())
        ref = g["move_e"][k]
        # mu is recomputed from scratch each move by the reference; all other
        # components are the incremental state after the move
        assert rel(ev[:13], ref[:13], ref[:13]) < RTOL, (k, ev[:13] - ref[:13])
        x, y, z = impl.positions()
        f, l = sysm.mol_first[im], sysm.mol_last[im]
        pos = np.stack([x[f:l + 1], y[f:l + 1], z[f:l + 1]], 1)
        assert rel(pos, g["move_pos"][k][: l - f + 1], 1.0) < 1e-13
    uu, fx, fy, fz = impl.real_arrays()
    assert rel(uu, g["uu_end"], g["uu_end"]) < RTOL
    fe = np.stack([g["fx_end"], g["fy_end"], g["fz_end"]])
    assert rel(np.stack([fx, fy, fz]), fe, fe) < RTOL
    assert rel(impl.sums()[0], g["sums_coulomb_end"], g["sums_coulomb_end"]) < RTOL


def pressure_kT_A3(pressure_pa, temp):
    """parameters.f90:411: pressure [J/m^3] / kT [J] * 1e-30."""
    return pressure_pa / (ys.BOLTZ * temp) * 1e-30


def check_volume_steps(impl, sysm, g, vstep):
    """vstep(u, pressure_kT_A3, log_max_volume_scaling) runs one mcvol step
    with a constant uniform and returns (accepted, box length); compared
    with the reference after each step (run after the move trace)."""
    if "vol_u" not in g:
        return 0
    T = sysm.params.temp
    for k in range(len(g["vol_u"])):
        acc, L = vstep(float(g["vol_u"][k]), pressure_kT_A3(float(g["vol_pressure"][k]), T),
                       float(np.log(g["vol_mvs"][k])))
        assert int(acc) == int(g["vol_accepted"][k]), f"volume decision differs at step {k}"
        assert abs(L - g["vol_box"][k]) <= 1e-12 * g["vol_box"][k], (k, L, g["vol_box"][k])
        e = impl.energies()
        ev = energies_vec(e, impl.mu(), impl.tails())
        ref = g["vol_e"][k]
        assert rel(ev[:13], ref[:13], ref[:13]) < RTOL, (k, ev[:13] - ref[:13])
        x, y, z = impl.positions()
        pos = np.stack([x, y, z], 1)
        assert rel(pos, g["vol_pos"][k], 1.0) < 1e-13, k
    return len(g["vol_u"])


def check_xchange_steps(impl, sysm, g, xstep):
    """xstep(imol1, imol2, u) runs one mc_xchange step with a constant uniform
    and returns accepted; compared with the reference after each step (run
    after the move trace and the volume steps), then the per-atom real-space
    arrays and the Coulomb sums after the last step."""
    if "xchg_u" not in g:
        return 0
    for k in range(len(g["xchg_u"])):
        acc = xstep(int(g["xchg_imol1"][k]), int(g["xchg_imol2"][k]), float(g["xchg_u"][k]))
        assert int(acc) == int(g["xchg_accepted"][k]), f"exchange decision differs at step {k}"
        e = impl.energies()
        ev = energies_vec(e, impl.mu(), impl.tails())
        ref = g["xchg_e"][k]
        assert rel(ev[:13], ref[:13], ref[:13]) < RTOL, (k, ev[:13] - ref[:13])
        x, y, z = impl.positions()
        assert rel(np.stack([x, y, z], 1), g["xchg_pos"][k], 1.0) < 1e-13, k
    uu, fx, fy, fz = impl.real_arrays()
    assert rel(uu, g["uu_xchg"], g["uu_xchg"]) < RTOL
    fe = np.stack([g["fx_xchg"], g["fy_xchg"], g["fz_xchg"]])
    assert rel(np.stack([fx, fy, fz]), fe, fe) < RTOL
    assert rel(impl.sums()[0], g["sums_coulomb_xchg"], g["sums_coulomb_xchg"]) < RTOL
    return len(g["xchg_u"])


def check_runmc(impl, sysm, g, cycle):
    """Replays the reference's recorded uniform stream (g["stream"]) through
    run_mc cycles: cycle() runs one run_mc cycle (runmc.f90:204-268: the
    pressure-step draw, then a volume step or nmol steps of exchange draw,
    pick and movemc or mc_xchange) and returns (accepted steps, volume steps,
    accepted volume steps).  After every cycle the counters, the box length
    and every energy component equal the reference's; at the end the
    positions, per-atom real-space arrays, Coulomb S(k) and the number of
    uniforms consumed equal the reference's."""
    acc = nv = nva = 0
    nmol = sysm.nmol
    for k in range(len(g["cycle_e"])):
        a, v, va = cycle()
        acc, nv, nva = acc + a, nv + v, nva + va
        cnt = g["cycle_counts"][k]
        steps = nmol * (k + 1 - nv)
        assert (steps, acc, nv, nva) == (int(cnt[0]), int(cnt[1]), int(cnt[2]), int(cnt[3])), (k, steps, acc, nv, nva,
                                                                                              cnt)
        L = impl.box_length()
        assert abs(L - g["cycle_box"][k]) <= 1e-12 * g["cycle_box"][k], (k, L, g["cycle_box"][k])
        ev = energies_vec(impl.energies(), impl.mu(), impl.tails())
        ref = g["cycle_e"][k]
        assert rel(ev[:13], ref[:13], ref[:13]) < RTOL, (k, ev[:13] - ref[:13])
    x, y, z = impl.positions()
    dpos = rel(np.stack([x, y, z]), np.stack([g["xx_end"], g["yy_end"], g["zz_end"]]), 1.0)
    assert dpos < RTOL, dpos  # box units
    uu, fx, fy, fz = impl.real_arrays()
    assert rel(uu, g["uu_end"], g["uu_end"]) < RTOL
    fe = np.stack([g["fx_end"], g["fy_end"], g["fz_end"]])
    assert rel(np.stack([fx, fy, fz]), fe, fe) < RTOL
    assert rel(impl.sums()[0], g["sums_coulomb_end"], g["sums_coulomb_end"]) < RTOL
    assert impl.rng_counter() == int(g["stream_used"]), (impl.rng_counter(), int(g["stream_used"]))
    return len(g["cycle_e"]), int(g["cycle_counts"][-1][4])


def runmc_setup(impl, g):
    """Frequencies and the recorded stream of a run_mc fixture."""
    impl.set_frequencies(g["move_weights"], g["xchange_weights"])
    impl.set_rng_list(g["stream"])
    T = float(g["params"][1])
    return float(g["run_psm"]), pressure_kT_A3(float(g["run_pressure"]), T), float(np.log(g["run_mvs"]))
