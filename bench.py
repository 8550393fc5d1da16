#!/usr/bin/env python3
"""Benchmark: Ewald Monte Carlo trial moves/s on one MI355X.

Workload (the metric's configuration): 2048 SPC/E + 40 Na+ + 40 Cl- (6224
atoms) in a 39.4 A box, T = 298 K, rc = 12 A, kmax = 12 (nk = 1068), Coulomb
+ LJ-Ewald, force-biased single-molecule moves (lambda_f = lambda_c = 0.5).
A step is one MC sweep: nmol = 2128 trial moves of every chain, run on the
device (molecule pick, proposal, dE, Metropolis test and commit all inside
the chain kernel).

Forms timed:
  fp64      plain FP64, one chain (the reference algorithm on one replica)
  fp32c     FP32-compensated S(k), one chain
  batch     FP32-compensated, R independent replicas of the same system (own
            Philox streams) advanced by one captured launch per sweep, launch
            shape chosen by yamc_amd.tune -> `value`
  batch_fp64  the same batch in the plain FP64 form
The per-move dE and force errors of fp32c against fp64 are measured on one
stream (moves up to the first differing decision).  The CPU baseline is the
reference's own Fortran compiled from its sources (oracle/_ref, when built;
else the plain-C oracle), timed on this host on one core, and, as
cpu_baseline_all_cores, as one copy per core (up to 16) at once.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RESYNC = 1000  # moves between from-scratch FP64 rebuilds of S(k) in the FP32C forms


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with its family/model/stepping
    numbers, the logical CPUs this process may use and the clock it reports."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (t.strip() for t in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count() or 0
    extra = [f"family {info.get('cpu family', '?')} model {info.get('model', '?')} stepping {info.get('stepping', '?')}",
             f"{ncpu} logical CPUs available"]
    if "cpu MHz" in info:
        extra.append(f"{float(info['cpu MHz']):.0f} MHz")
    return f"{name} ({', '.join(extra)})"


def batch_error(reps, base, kT_kjmol):
    """Error of the FP32C batch against FP64 on the batch's own states:
    for each replica after the timed sweeps, (i) its incrementally carried
    energy components against a from-scratch FP64 recompute of the same
    configuration, and (ii) one more sweep from that configuration run in
    FP32C and in FP64 on the same RNG stream: per-move dE and force errors
    over the moves before the first differing decision."""
    import numpy as np

    from yamc_amd import YAMC_FP32C, YAMC_FP64
    from yamc_amd.ewald import EwaldMC
    state, de, fr, nmv, agree = 0.0, 0.0, 0.0, 0, []
    for i, r in enumerate(reps):
        inc = r.energies()
        x, y, z = r.positions()
        L = r.box_length()
        twins = []
        for prec in (YAMC_FP64, YAMC_FP32C):
            t = EwaldMC(base, precision=prec)
            t.set_positions(x, y, z)
            e, _ = t.mc_init()
            t.set_rng(99, i)
            twins.append((t, e))
        e64 = twins[0][1]
        state = max(state, max(abs(inc[k] - e64[k]) for k in e64))
        _, ra = twins[0][0].run(1, records=True)
        _, rb = twins[1][0].run(1, records=True)
        same = ra["accepted"] == rb["accepted"]
        n = int(np.argmin(same)) if not same.all() else len(ra)
        ok = ra["overlap"][:n] == 0
        if ok.any():
            de = max(de, float(np.abs(ra["du_total"][:n] - rb["du_total"][:n])[ok].max()))
            fa = np.concatenate([ra["f_old"][:n][ok], ra["f_new"][:n][ok]])
            fb = np.concatenate([rb["f_old"][:n][ok], rb["f_new"][:n][ok]])
            fr = max(fr, float((np.linalg.norm(fa - fb, axis=1) / np.maximum(np.linalg.norm(fa, axis=1), 1e-300)).max()))
        nmv += n
        agree.append(float(same.mean()))
        for t, _ in twins:
            t.close()
        assert abs(L - base.params.box_length) < 1e-12
    return dict(replicas_checked=len(reps), max_state_energy_err_kT=state,
                max_dE_kJmol=de * kT_kjmol, max_dE_kT=de, max_rel_force_err_norm=fr, moves_compared=nmv,
                decision_agreement=float(np.mean(agree)))


def algorithmic_bytes_per_move(sysm, nk_info, w, p_acc):
    """k-space bytes per trial move:
        w [2 (1+T) n_s + (1 + T |t_M|) nk] + 2 w' 2 (1 + |t_M|) n_s P_acc
    (the S(k) cos/sin of the Coulomb set and all T LJ sets over the nonzero
    sign slots, plus beta_C and the beta_LJ rows of the moved molecule's type,
    plus the read-modify-write of the touched sets on acceptance; w = 8 (FP64)
    or 4 (FP32 hi read), w' = 8 (FP64 or the FP32 hi+lo pair); |t_M| = 1 for
    every species here)."""
    ky, kz = nk_info["ky"], nk_info["kz"]
    n_s = int(sum(1 + (a > 0) + (b > 0) + ((a > 0) and (b > 0)) for a, b in zip(ky, kz)))
    T = sysm.ntype
    nk = len(ky)
    sweep = w * (2 * (1 + T) * n_s + (1 + T * 1) * nk)
    commit = 2 * 8 * 2 * (1 + 1) * n_s * p_acc
    return sweep + commit, n_s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--block", type=int, default=0, help="batch launch shape (64, 128, 256, 512); 0 = autotune")
    ap.add_argument("--cell-div", type=int, default=-1,
                    help="batch cell division of the real-space scan (0 = no cell list, 1..4); -1 = autotune")
    ap.add_argument("--cpu-moves", type=int, default=0, help="moves for the CPU baseline (0 = auto ~15 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,fp32c,batch,batch_fp64")
    ap.add_argument("--forces-check", action="store_true",
                    help="instead of the headline line: forward energy + adjoint forces of the 24,896-atom "
                         "workload (FP32C), timed, with a central-difference check on 64 seeded atoms and the "
                         "FP64 forces as reference")
    ap.add_argument("--large-check", type=int, default=0, metavar="CYCLES",
                    help="instead of the headline line: CYCLES run_mc cycles (33,968 moves each) of the "
                         "99,504-atom workload in FP64 and in FP32C (FP64 resync every 1000 moves) on one "
                         "RNG stream; rates, per-move dE error until the first differing decision, drift")
    ap.add_argument("--large-groups", type=int, default=32,
                    help="workgroups per chain of the --large-check runs (1 = one workgroup)")
    ap.add_argument("--replica-check", action="store_true",
                    help="instead of the headline line: 64 uneven FP32C replicas (128-2048 waters, 0-1 M NaCl, "
                         "280-360 K, seeded) advanced by one captured launch per sweep, launch shape autotuned")
    ap.add_argument("--npt-check", type=int, default=0, metavar="CYCLES",
                    help="instead of the headline line: the reference's own NPT run (100 SPC/E + 5 Na + 5 Cl, "
                         "its parameters.prm) for CYCLES run_mc cycles on the device, compared with the "
                         "reference's energy.dat, boxlength.dat and RDF_*.dat (means from cycle 2000 on)")
    ap.add_argument("--npt-precision", choices=("fp64", "fp32c"), default="fp64")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain fp32 streaming copy of MIB MiB (3 launches) and print its byte count; "
                         "used to scale the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(replicas would shard across GPUs with no communication). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.rand(n, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)  # one elementwise kernel: reads 4n bytes, writes 4n bytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"kernel": "elementwise mul by 1", "launches": 3,
                                               "read_bytes": 4 * n, "write_bytes": 4 * n}}))
        return

    from yamc_amd import YAMC_FP32C, YAMC_FP64, config, system
    from yamc_amd.ewald import Batch, EwaldMC
    from yamc_amd.tune import k_per_wave, tune_batch, tune_cells, tune_cells_batch, tune_groups, tune_single

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.forces_check:
        forces_check(args)
        return
    if args.large_check:
        large_check(args)
        return
    if args.replica_check:
        replica_check(args)
        return
    if args.npt_check:
        npt_check(args)
        return
    kT_kjmol = system.BOLTZ * 298.0 / system.KCAL_MOL * 4.184

    forms = args.forms.split(",")
    base = config(args.workload, seed=20261203)
    nmol = base.nmol
    res = {}

    def sync():
        torch.cuda.synchronize()

    # ---------------------------------------------------------- single chains
    for form, prec in (("fp64", YAMC_FP64), ("fp32c", YAMC_FP32C)):
        if form not in forms:
            continue
        g = EwaldMC(base, precision=prec)
        g.mc_init()
        g.set_rng(20261203, 0)
        shape1, _ = tune_single(g, shapes=(256, 512))  # two ordinary sweeps, one per shape
        groups1, grates = tune_groups(g)  # one ordinary sweep per workgroup count
        div1, crates = tune_cells(g)  # one ordinary sweep per cell size
        for _ in range(args.warmup):
            g.run(1)
        sync()
        t0 = time.perf_counter()
        kms = 0.0
        nacc = 0
        for _ in range(args.steps):
            n, _ = g.run(1, resync_every=RESYNC if prec == YAMC_FP32C else 0)
            kms += g.last_kernel_ms()
            nacc += n
        sync()
        dt = time.perf_counter() - t0
        moves = args.steps * nmol
        res[form] = dict(moves_per_s=moves / dt, ms_per_step=1e3 * dt / args.steps, kernel_ms=kms,
                         acceptance=nacc / moves, launch_shape=shape1, workgroups_per_chain=groups1,
                         k_per_wave=round(k_per_wave(g.nk, 256 if groups1 > 1 else shape1, groups1), 1),
                         cell_div=div1,
                         tuner_moves_per_s_by_workgroups={str(k): round(v, 1) for k, v in grates.items()},
                         tuner_moves_per_s_by_cell_div={str(k): round(v, 1) for k, v in crates.items()})
        g.close()

    # ------------------------------------------------- fp32c error vs fp64
    err = None
    if "fp64" in forms and "fp32c" in forms:
        a, b = EwaldMC(base, precision=YAMC_FP64), EwaldMC(base, precision=YAMC_FP32C)
        a.mc_init(); b.mc_init()
        a.set_rng(7, 0); b.set_rng(7, 0)
        _, ra = a.run(1, records=True)
        _, rb = b.run(1, records=True)
        same = ra["accepted"] == rb["accepted"]
        n = int(np.argmin(same)) if not same.all() else len(ra)
        ok = ra["overlap"][:n] == 0
        d = np.abs(ra["du_total"][:n] - rb["du_total"][:n])[ok]
        fa = np.concatenate([ra["f_old"][:n][ok], ra["f_new"][:n][ok]])
        fb = np.concatenate([rb["f_old"][:n][ok], rb["f_new"][:n][ok]])
        f_norm = np.linalg.norm(fa - fb, axis=1) / np.maximum(np.linalg.norm(fa, axis=1), 1e-300)
        f_comp = np.abs(fa - fb) / np.maximum(np.abs(fa), 1e-300)
        err = dict(moves_compared=int(n), max_dE_kT=float(d.max()) if d.size else 0.0,
                   max_dE_kJmol=float(d.max() * kT_kjmol) if d.size else 0.0,
                   rms_dE_kT=float(np.sqrt((d ** 2).mean())) if d.size else 0.0,
                   max_rel_force_err_norm=float(f_norm.max()) if f_norm.size else 0.0,
                   max_rel_force_err_component=float(f_comp.max()) if f_comp.size else 0.0,
                   decision_agreement_first_sweep=float(same.mean()))
        a.close(); b.close()

    # ------------------------------------------------------ batched replicas
    grid = None
    tuned = None
    batch_err = None
    if "batch" in forms:
        R = args.replicas
        reps = []
        for i in range(R):
            r = EwaldMC(base, precision=YAMC_FP32C)
            r.mc_init()
            r.set_rng(20261203, i)
            reps.append(r)
        grid = reps[0].grid()
        if args.block:
            for r in reps:
                r.set_block_size(args.block)
            shape, tuned = args.block, None
        else:
            shape, tuned = tune_batch(reps)  # three ordinary sweeps, one per launch shape
        if args.cell_div >= 0:
            for r in reps:
                r.set_cell_div(args.cell_div)
            div, ctuned = args.cell_div, {}
        else:
            div, ctuned = tune_cells_batch(reps)  # one ordinary sweep per cell size
        # one captured sweep per step; every replica's S(k) is rebuilt from
        # scratch in FP64 every RESYNC moves inside the sweep (batched)
        bat = Batch(reps, 1, resync_every=RESYNC)
        for _ in range(args.warmup):
            bat.launch()
        sync()
        t0 = time.perf_counter()
        kms = 0.0
        nacc = 0
        for _ in range(args.steps):
            acc, ms = bat.launch()
            kms += ms
            nacc += int(acc.sum())
        sync()
        dt = time.perf_counter() - t0
        moves = args.steps * nmol * R
        res["batch"] = dict(replicas=R, moves_per_s=moves / dt, ms_per_step=1e3 * dt / args.steps, kernel_ms=kms,
                            acceptance=nacc / moves, launch_shape=shape, resync_every=RESYNC,
                            k_per_wave=round(k_per_wave(grid["kx"].size, shape), 1), cell_div=div,
                            tuner_moves_per_s={str(k): v for k, v in tuned.items()} if tuned else None,
                            tuner_moves_per_s_by_cell_div={str(k): round(v, 1) for k, v in ctuned.items()})
        bat.close()
        batch_err = batch_error(reps[:4], base, kT_kjmol)
        for r in reps:
            r.close()

    # ------------------------------------- the same batch in the plain FP64 form
    if "batch_fp64" in forms:
        R = args.replicas
        reps = []
        for i in range(R):
            r = EwaldMC(base, precision=YAMC_FP64)
            r.mc_init()
            r.set_rng(20261203, i)
            reps.append(r)
        shape64, tuned64 = tune_batch(reps)
        bat = Batch(reps, 1)
        for _ in range(args.warmup):
            bat.launch()
        sync()
        t0 = time.perf_counter()
        nacc = 0
        for _ in range(args.steps):
            acc, _ = bat.launch()
            nacc += int(acc.sum())
        sync()
        dt = time.perf_counter() - t0
        moves = args.steps * nmol * R
        res["batch_fp64"] = dict(replicas=R, moves_per_s=moves / dt, ms_per_step=1e3 * dt / args.steps,
                                 acceptance=nacc / moves, launch_shape=shape64,
                                 tuner_moves_per_s={str(k): v for k, v in tuned64.items()})
        bat.close()
        for r in reps:
            r.close()

    # ------------------------------------------------------------ roofline
    if grid is None:
        g0 = EwaldMC(base)
        grid = g0.grid()
        g0.close()
    main_form = "batch" if "batch" in res else ("fp32c" if "fp32c" in res else "fp64")
    w = 4 if main_form in ("batch", "fp32c") else 8
    r0 = res[main_form]
    bpm, n_s = algorithmic_bytes_per_move(base, grid, w, r0["acceptance"])
    moves_per_launch = nmol * r0.get("replicas", 1)
    launch_s = r0["kernel_ms"] / 1e3 / args.steps
    achieved = bpm * moves_per_launch / launch_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if os.path.exists(tpath):
        try:
            t = json.load(open(tpath))
            key = f"{main_form}_{r0.get('launch_shape')}_{r0.get('replicas', 1)}"
            if key in t:  # HBM bytes per launch from the PMC runs of this exact launch
                traffic = t[key]["bytes_per_launch"]
        except Exception:
            traffic = None

    # --------------------------------------------------------- CPU baseline
    cpu = None
    cpu_all = None
    if not args.no_cpu:
        from oracle import pyoracle, refcase
        have_ref = refcase.available()
        o = pyoracle.Oracle(base)
        o.mc_init()
        o.set_rng(20261203, 0)
        im = np.random.default_rng(1).integers(0, nmol, 64).astype(np.int32)
        t0 = time.perf_counter()
        o.run_const(im, np.full(64, 0.4))
        per = (time.perf_counter() - t0) / 64
        nmv = args.cpu_moves or max(64, min(20000, int((5.0 if have_ref else 15.0) / per)))
        o2 = pyoracle.Oracle(base)
        o2.mc_init()
        o2.set_rng(20261203, 0)
        ncyc = max(1, nmv // nmol)
        t0 = time.perf_counter()
        if nmv >= nmol:
            o2.run(ncyc, records=False)
            done = ncyc * nmol
        else:
            im = np.random.default_rng(2).integers(0, nmol, nmv).astype(np.int32)
            o2.run_const(im, np.full(nmv, 0.4))
            done = nmv
        dtc = time.perf_counter() - t0
        port = dict(value=done / dtc, unit="trial moves/s", cores=1, kind="port", cpu_model=cpu_model(),
                    sample=f"{done} FP64 trial moves of the same 6224-atom workload, plain-C oracle, 1 thread")
        cpu = port
        if have_ref:
            try:
                ref = time_reference(base, pyoracle, refcase)
                if "all_cores" in ref:
                    ac = ref["all_cores"]
                    cpu_all = dict(value=ac["moves_per_s"], unit="trial moves/s", cores=ac["cores"], kind="reference",
                                   cpu_model=cpu_model(),
                                   sample=f"{ac['cores']} copies of the reference's Fortran driver at once, one per "
                                                  f"core, each {ref['cycles']} run_mc cycles ({ref['moves']} "
                                          f"force-biased trial moves) of the same 6224-atom workload with its own "
                                          f"seed (independent replicas), aggregate rate")
                cpu = dict(value=ref["moves_per_s"], unit="trial moves/s", cores=1, kind="reference",
                           cpu_model=cpu_model(),
                           sample=f"{ref['cycles']} run_mc cycles ({ref['moves']} force-biased trial moves) "
                                  f"of the same 6224-atom workload by the reference's own Fortran run_mc "
                                  f"(compiled from its sources with amdflang -O2, its own random_number "
                                  f"stream, set-up time subtracted), 1 thread; the plain-C port of the "
                                  f"same path: {port['value']:.0f} trial moves/s")
            except Exception as e:  # the port stays the baseline
                print(f"reference CPU timing failed ({e}); reporting the port", file=sys.stderr)

    value = res[main_form]["moves_per_s"]
    out = {
        "metric": "MC trial moves/s at 2048 SPC/E+NaCl, FP32-comp vs FP64 GPU; dE err vs FP64",
        "value": value,
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[main_form]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 arithmetic, f32 hi/lo S(k) storage" if main_form != "fp64" else "f64",
        "data": "synthetic: seeded lattice start, seeded orientations",
        "config": {
            "workload": f"{args.workload}: 2048 SPC/E + 40 Na + 40 Cl, L=39.4 A, rc=12 A, kmax=12, "
                        f"force-biased moves; step = 1 sweep (2128 moves) per chain",
            "form": main_form,
            "replicas": res[main_form].get("replicas", 1),
            "launch_shape": res[main_form].get("launch_shape"),
            "nk": len(grid["kx"]),
            "n_s": n_s,
        },
        "forms": res,
        "dE_error_fp32c_vs_fp64": err,
        "batch_error_fp32c_vs_fp64": batch_err,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_move": bpm,
            "kernel": "yk_wchain" if (res[main_form].get("launch_shape") in (64, 128) and main_form == "fp64") else "yk_chain",
        },
        "cpu_baseline": cpu,
    }
    if cpu_all is not None:
        out["cpu_baseline_all_cores"] = cpu_all
    # atom x k-vector evaluations: each trial move evaluates the moved molecule's
    # atoms at every nonzero sign slot, at the old and at the new positions
    # (2 n_m n_s); n_m averaged over the uniformly picked molecules
    out["atom_k_evals_per_s"] = 2.0 * (base.natom / nmol) * n_s * value
    if "fp64" in res and "batch" in res:
        out["batch_vs_fp64_single"] = res["batch"]["moves_per_s"] / res["fp64"]["moves_per_s"]
    if "batch_fp64" in res and "batch" in res:
        out["batch_fp32c_vs_batch_fp64"] = res["batch"]["moves_per_s"] / res["batch_fp64"]["moves_per_s"]
    print(json.dumps(out))


def forces_check(args):
    """8192 SPC/E + 160 Na + 160 Cl (24,896 atoms), L = 62.6 A, rc = 12 A,
    kmax = 16, FP32C: time forward U_total (full recompute) + backward
    (adjoint forces, sin/cos recomputed) through the torch.autograd harness,
    check the gradient on 64 seeded atoms against central differences of
    U_total (h = 1e-4 A), and compare all forces with the FP64 form."""
    import numpy as np
    import torch

    from yamc_amd import YAMC_FP32C, YAMC_FP64, config
    from yamc_amd.autograd import ewald_energy
    from yamc_amd.ewald import EwaldMC

    s = config("c3", seed=20261205)
    f32 = EwaldMC(s, precision=YAMC_FP32C)
    x, y, z = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in s.positions())
    ewald_energy(f32, x, y, z).backward()  # warm-up
    fw, bw = [], []
    for _ in range(max(1, args.steps)):
        for t in (x, y, z):
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = ewald_energy(f32, x, y, z)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        e.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        bw.append(t2 - t1)
    grad = torch.stack([x.grad, y.grad, z.grad], 1).numpy()  # dU/dr, box units
    f64 = EwaldMC(s, precision=YAMC_FP64)
    f64.mc_init()
    g64, r64 = f64.forces()
    g32 = -grad
    norm_err = np.linalg.norm(g32 - g64, axis=1) / np.maximum(np.linalg.norm(g64, axis=1), 1e-300)
    h = 1e-4 / s.params.box_length
    rng = np.random.default_rng(20261205)
    atoms = rng.choice(s.natom, 64, replace=False)
    x0, y0, z0 = s.positions()
    probe = EwaldMC(s, precision=YAMC_FP32C)
    # The reference energy is truncated at rc (EwaldSumRealSpace.f90:437):
    # a pair whose distance is within h of rc jumps across the cut-off inside
    # the difference stencil, which no gradient reproduces.  Such probes are
    # counted and reported separately.
    rc = float(np.sqrt(s.derived["rmax2"]))
    errs, jumps = [], []
    scale = np.abs(g64).max()
    for k, i in enumerate(atoms):
        dx = [a - a[i] for a in (x0, y0, z0)]
        dx = [v - np.round(v) for v in dx]
        r = np.sqrt(dx[0] ** 2 + dx[1] ** 2 + dx[2] ** 2)
        r[i] = 0.0
        crosses = bool(np.any(np.abs(r - rc) < 2 * h))
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                xs = [x0.copy(), y0.copy(), z0.copy()]
                xs[d][i] += sgn * h
                probe.set_positions(*xs)
                ev, _ = probe.mc_init()
                vals.append(ev["U_total"])
            fd = (vals[0] - vals[1]) / (2 * h)  # dU/dr
            (jumps if crosses else errs).append(abs(fd - grad[i, d]) / scale)
        if k % 8 == 7:
            print(f"forces-check: {k + 1}/64 atoms, worst |fd - grad| / max|grad| = {max(errs or [0]):.3e}",
                  file=sys.stderr, flush=True)
    worst = max(errs) if errs else 0.0
    out = {
        "metric": "forward U_total + adjoint forces, 24,896 atoms, FP32C",
        "workload": "c3: 8192 SPC/E + 160 Na + 160 Cl, L=62.6 A, rc=12 A, kmax=16",
        "forward_ms": 1e3 * float(np.median(fw)),
        "backward_ms": 1e3 * float(np.median(bw)),
        "forward_backward_ms": 1e3 * float(np.median(np.array(fw) + np.array(bw))),
        "fd_atoms": 64, "fd_h_angstrom": 1e-4,
        "fd_max_err_rel_to_max_force": worst,
        "fd_median_err_rel_to_max_force": float(np.median(errs)) if errs else 0.0,
        "fd_probes": len(errs),
        "fd_probes_with_a_pair_at_the_cutoff": len(jumps),
        "fd_max_err_cutoff_probes": max(jumps) if jumps else None,
        "fp32c_vs_fp64_force_max_rel_err_norm": float(norm_err.max()),
        "fp32c_vs_fp64_force_median_rel_err_norm": float(np.median(norm_err)),
        "dtype": "f64 arithmetic, f32 hi/lo S(k) storage",
        "data": "synthetic: seeded lattice start, seeded orientations",
    }
    print(json.dumps(out))


def replica_check(args):
    """64 independent replicas of uneven size, composition and temperature
    (seeded: N_w in [128, 2048], NaCl uniform 0-1 M, T uniform 280-360 K,
    density 0.0334 /A^3, alpha = 0.2667 /A, kmax from sk = 3.2), FP32C,
    one captured launch per sweep."""
    import numpy as np

    from yamc_amd import YAMC_FP32C, YAMC_FP64, replica_system
    from yamc_amd.ewald import Batch, EwaldMC
    from yamc_amd.tune import tune_batch, tune_cells_batch

    rng = np.random.default_rng(20261206)
    reps, sizes, kmaxes, temps = [], [], [], []
    for i in range(64):
        nw, conc, T = int(rng.integers(128, 2049)), float(rng.uniform(0.0, 1.0)), float(rng.uniform(280.0, 360.0))
        s = replica_system(nw, conc, T, seed=1000 + i)
        r = EwaldMC(s, precision=YAMC_FP32C)
        r.mc_init()
        r.set_rng(20261206, i)
        reps.append(r)
        sizes.append(s.natom)
        kmaxes.append(int(s.derived["kmax"]))
        temps.append(T)
    shape, tuned = tune_batch(reps)
    div, ctuned = tune_cells_batch(reps, divs=(0, 1, 2, 3))
    bat = Batch(reps, 1, resync_every=RESYNC)
    bat.launch()
    torch_sync()
    moves_per_sweep = sum(r.sys.nmol for r in reps)
    t0 = time.perf_counter()
    kms = 0.0
    nacc = 0
    for _ in range(args.steps):
        acc, ms = bat.launch()
        kms += ms
        nacc += int(acc.sum())
    torch_sync()
    dt = time.perf_counter() - t0
    bat.close()
    drift = 0.0
    for r in reps[:8]:
        inc = r.energies()
        r.resync()
        fresh = r.energies()
        drift = max(drift, max(abs(inc[k] - fresh[k]) / max(1.0, abs(fresh[k])) for k in ("E_kC", "E_kLJ")))
    # The same replicas with the workgroups per chain tuned over the batch:
    # one 512-thread workgroup per replica, or grid chains of natom // A
    # workgroups (1 to 8), all in one launch; in FP64 (no rebuild needed)
    # and in FP32C with the rebuild every RESYNC moves (inside the launch,
    # each member's workgroups over their own k ranges).
    from yamc_amd.tune import tune_groups_batch

    def grid_leg(precision, resync):
        rng = np.random.default_rng(20261206)
        rs = []
        for i in range(64):
            nw, conc, T = int(rng.integers(128, 2049)), float(rng.uniform(0.0, 1.0)), float(rng.uniform(280.0, 360.0))
            s = replica_system(nw, conc, T, seed=1000 + i)
            r = EwaldMC(s, precision=precision)
            r.set_block_size(512)
            r.mc_init()
            r.set_rng(20261206, i)
            rs.append(r)
        apg, apg_rates = tune_groups_batch(rs)
        bg = Batch(rs, 1, resync_every=resync)
        bg.launch()
        torch_sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bg.launch()
        torch_sync()
        dtg = time.perf_counter() - t0
        bg.close()
        res = {"atoms_per_workgroup": apg, "workgroups": sum(getattr(r, "groups", 1) for r in rs),
               "tuner_moves_per_s_by_atoms_per_workgroup": {str(k): round(v, 1) for k, v in apg_rates.items()},
               "moves_per_s": moves_per_sweep * args.steps / dtg, "ms_per_sweep": 1e3 * dtg / args.steps,
               "resync_every": resync}
        for r in rs:
            r.close()
        return res

    fp64 = grid_leg(YAMC_FP64, 0)
    fp32c_grid = grid_leg(YAMC_FP32C, RESYNC)
    out = {
        "metric": "batched uneven replicas: trial moves/s, FP32C, one captured launch per sweep",
        "workload": "64 replicas, seeded N_w 128-2048, NaCl 0-1 M, T 280-360 K, rho 0.0334/A^3, rc 12 A",
        "replicas": 64, "atoms_min": min(sizes), "atoms_max": max(sizes), "atoms_total": sum(sizes),
        "kmax_min": min(kmaxes), "kmax_max": max(kmaxes), "T_min": min(temps), "T_max": max(temps),
        "launch_shape": shape, "tuner_moves_per_s": {str(k): v for k, v in tuned.items()},
        "cell_div": div, "tuner_moves_per_s_by_cell_div": {str(k): round(v, 1) for k, v in ctuned.items()},
        "moves_per_sweep": moves_per_sweep, "steps": args.steps,
        "moves_per_s": moves_per_sweep * args.steps / dt, "ms_per_sweep": 1e3 * dt / args.steps,
        "kernel_ms_per_sweep": kms / args.steps, "acceptance": nacc / (moves_per_sweep * args.steps),
        "kspace_drift_rel_after_run": drift,
        "fp64_batch": fp64,
        "fp32c_grid_batch": fp32c_grid,
    }
    print(json.dumps(out))


def npt_check(args):
    """The reference's ions100 NPT run (tests/golden/ions100: parameters.prm,
    ions100.composition and its output files) repeated on the device from a
    seeded lattice start: mean energy and box length from cycle 2000 on with
    block standard errors, and the centre-of-mass RDFs (frames every 100
    cycles, r < 7 A), against the reference's files."""
    import numpy as np

    from yamc_amd import formats as fm
    from yamc_amd import system as ys
    from yamc_amd.ewald import rdf_normalize
    from yamc_amd.runmc import run

    d = os.path.join(ROOT, "tests", "golden", "ions100")
    prm = fm.read_parameters(os.path.join(d, "parameters.prm"))
    species, counts = fm.read_composition(os.path.join(d, "ions100.composition"))
    params = fm.params_from(prm, sum(counts))
    sysm = ys.lattice_system(params, species, counts, 20261207)
    ncyc, start, dr, rmax = args.npt_check, 2000, 0.1, 7.0
    res = run(sysm, prm, ncyc, 1 if args.npt_precision == "fp32c" else 0, seed=20261207,
              rdf=(sysm.mol_type, 3, dr, rmax, 100, start), progress=max(1, ncyc // 20))

    def block_mean(v, nb=8):
        v = np.asarray(v, dtype=np.float64)
        m = v[: len(v) // nb * nb].reshape(nb, -1).mean(1)
        return float(v.mean()), float(m.std(ddof=1) / np.sqrt(nb))

    e_ref = fm.read_table(os.path.join(d, "energy.dat"))
    b_ref = fm.read_table(os.path.join(d, "boxlength.dat"))
    u_ref = e_ref[e_ref[:, 0] >= start, 1]
    grid = np.arange(1, int(e_ref[-1, 0]) + 1)
    idx = np.searchsorted(b_ref[:, 0], grid, side="right") - 1
    L_ref = np.where(idx >= 0, b_ref[np.maximum(idx, 0), 1], params.box_length)[start - 1:]
    e = res["energy"]
    u_mine = e[e[:, 0] >= start, 1]
    L_mine = res["box"][start - 1:]
    (ur, ur_e), (um, um_e) = block_mean(u_ref, 4), block_mean(u_mine, 4)
    (lr, lr_e), (lm, lm_e) = block_mean(L_ref), block_mean(L_mine)
    g = rdf_normalize(res["rdf_counts"].astype(np.float64), res["rdf_frames"], counts, 1.0 / res["rdf_inv_vol"], dr)
    names = ["water", "Na", "Cl"]
    rdfs = {}
    for i in range(3):
        for j in range(i, 3):
            ref = fm.read_table(os.path.join(d, f"RDF_{names[i]}_{names[j]}.dat"))
            nb = min(g.shape[2], len(ref))
            r = ref[:nb, 0]
            sel = (r > 2.0) & (r < rmax)
            cnt = res["rdf_counts"][i, j, :nb].astype(np.float64)
            # Poisson error of this run's bins, doubled for the reference's own noise
            err = np.where(cnt > 0, g[i, j, :nb] / np.sqrt(np.maximum(cnt, 1.0)), np.nan) * np.sqrt(2.0)
            dg = g[i, j, :nb] - ref[:nb, 1]
            z = np.abs(dg[sel]) / err[sel]
            pk = int(np.argmax(ref[:nb, 1] * sel))
            rdfs[f"{names[i]}_{names[j]}"] = {
                "rms_diff_2_7A": float(np.sqrt(np.mean(dg[sel] ** 2))),
                "median_z": float(np.nanmedian(z)),
                "peak_r_ref": float(r[pk]), "peak_g_ref": float(ref[pk, 1]),
                "peak_r": float(r[int(np.argmax(g[i, j, :nb] * sel))]), "peak_g": float(np.max(g[i, j, :nb] * sel))}
    out = {
        "check": "reference ions100 NPT run repeated on the device (its parameters.prm and composition)",
        "precision": args.npt_precision, "cycles": ncyc, "seconds": res["seconds"],
        "moves": res["n_mv"], "moves_accepted": res["n_mv_acc"], "volume_steps": res["n_vol"],
        "volume_steps_accepted": res["n_vol_acc"],
        "U_total_mean_from_2000": {"this": um, "this_err": um_e, "reference": ur, "reference_err": ur_e,
                                   "z": (um - ur) / float(np.hypot(um_e, ur_e))},
        "box_length_mean_from_2000": {"this": lm, "this_err": lm_e, "reference": lr, "reference_err": lr_e,
                                      "z": (lm - lr) / float(np.hypot(lm_e, lr_e))},
        "rdf_frames": res["rdf_frames"], "rdf": rdfs,
    }
    print(json.dumps(out))


def time_reference(sysm, pyoracle, refcase, seconds=12.0):
    """The reference's own run_mc (oracle/_ref/ref_driver, RUNMC section:
    runmc.f90:178-334 with movemc's force-biased proposals and the
    reference's own random_number stream) on sysm's configuration: driver
    runs with 0 and with K cycles, per-move time = difference / (K nmol).
    The same workload the device runs: force-biased single-molecule moves of
    the same system, one cycle = nmol trial moves."""
    import tempfile

    def run_cfg(ncycles, seed):
        return dict(ncycles=ncycles, seed=seed, psm=0.0, pressure=1.0e5, mvs=1.05, nstream=4, freq=[])

    def timed(ncycles, seed=20261203):
        with tempfile.TemporaryDirectory() as td:
            refcase.write_case(td, sysm, [], [], runmc=run_cfg(ncycles, seed))
            _, dt = refcase.run(td, timeout=600)
        return dt

    t_setup = timed(0)
    t_probe = timed(1)
    per_cycle = max(t_probe - t_setup, 1e-6)
    K = int(max(1, min(200, round(seconds / per_cycle))))
    t_k = timed(K)
    m = K * sysm.nmol
    out = {"moves": m, "cycles": K, "moves_per_s": m / max(t_k - t_setup, 1e-9), "setup_s": t_setup}

    # All cores: P copies of the same driver run at once, one per core (at
    # most 16, a GPU host's share), each an independent replica (its own
    # seed); aggregate rate = P m / (wall time - set-up time).
    import subprocess
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    P = max(1, min(16, ncpu))
    with tempfile.TemporaryDirectory() as td:
        dirs = []
        for i in range(P):
            d = os.path.join(td, f"r{i}")
            os.makedirs(d)
            refcase.write_case(d, sysm, [], [], runmc=run_cfg(K, 20261203 + i))
            dirs.append(d)
        t0 = time.perf_counter()
        procs = [subprocess.Popen([refcase.DRIVER, "case.in"], cwd=d, stdout=subprocess.DEVNULL,
                                  stderr=subprocess.PIPE) for d in dirs]
        rcs = [pr.wait(timeout=900) for pr in procs]
        wall = time.perf_counter() - t0
    if all(rc == 0 for rc in rcs):
        out["all_cores"] = {"cores": P, "moves": P * m, "moves_per_s": P * m / max(wall - t_setup, 1e-9)}
    return out


def torch_sync():
    import torch
    torch.cuda.synchronize()


def large_check(args):
    """32768 SPC/E + 600 Na + 600 Cl (99,504 atoms), L = 99.4 A, rc = 12 A,
    kmax = 24, force-biased moves: the FP32C form (S(k) resync in FP64 every
    1000 moves) against the plain FP64 form on one Philox stream."""
    import numpy as np

    from yamc_amd import YAMC_FP32C, YAMC_FP64, config, system
    from yamc_amd.ewald import EwaldMC

    kT_kjmol = system.BOLTZ * 298.0 / system.KCAL_MOL * 4.184
    s = config("c5", seed=20261207)
    out = {"metric": "large system: FP32C (resync every 1000 moves) vs FP64, one RNG stream",
           "workload": "c5: 32768 SPC/E + 600 Na + 600 Cl, L=99.4 A, rc=12 A, kmax=24",
           "cycles": args.large_check, "moves": args.large_check * s.nmol}
    recs = {}
    final_pos = None
    # the cell size of the real-space scan, tuned on a scratch chain (one
    # cycle per size) so that both forms start from the same state
    from yamc_amd.tune import tune_cells
    t = EwaldMC(s, precision=YAMC_FP64)
    t.set_chain_groups(args.large_groups)
    t.mc_init()
    t.set_rng(20261209, 0)
    cell_div, by_div = tune_cells(t)
    t.close()
    out["cell_div"] = cell_div
    out["tuner_moves_per_s_by_cell_div"] = {str(k): round(v, 1) for k, v in by_div.items()}
    print(f"large-check: cell div {cell_div} ({out['tuner_moves_per_s_by_cell_div']})", file=sys.stderr, flush=True)
    for form, prec, rs in (("fp64", YAMC_FP64, 0), ("fp32c", YAMC_FP32C, 1000)):
        g = EwaldMC(s, precision=prec)
        g.set_chain_groups(args.large_groups)
        g.set_cell_div(cell_div)
        g.mc_init()
        g.set_rng(20261207, 0)
        t0 = time.perf_counter()
        n, r = g.run(args.large_check, resync_every=rs, records=True)
        dt = time.perf_counter() - t0
        inc = g.energies()
        g.resync()
        fresh = g.energies()
        recs[form] = r
        out[form] = dict(moves_per_s=len(r) / dt, seconds=dt, acceptance=n / len(r), workgroups_per_chain=args.large_groups,
                         kspace_drift_rel=max(abs(inc[k] - fresh[k]) / max(1.0, abs(fresh[k]))
                                              for k in ("E_kC", "E_kLJ")))
        print(f"large-check: {form} done, {len(r)} moves in {dt:.1f} s", file=sys.stderr, flush=True)
        final_pos = g.positions()
        g.close()
    # k-space dE and forces of the two forms on one configuration (the
    # FP32C run's final state) for 2000 seeded trial displacements: the
    # per-move error of the FP32C sums themselves, free of the slow
    # divergence of the two trajectories (force-biased proposals follow the
    # forms' forces, so the chains drift apart at the 1e-9 level over 10^5
    # moves while keeping identical decisions)
    a64, a32 = EwaldMC(s, precision=YAMC_FP64), EwaldMC(s, precision=YAMC_FP32C)
    for h in (a64, a32):
        h.set_positions(*final_pos)
        h.mc_init()
    rng = np.random.default_rng(20261208)
    x, y, z = final_pos
    errs, scales, ferr = [], [], []
    for _ in range(2000):
        im = int(rng.integers(0, s.nmol))
        f, l = int(s.mol_first[im]), int(s.mol_last[im])
        dv = rng.uniform(-1.0, 1.0, 3) * s.derived["dr"]
        nx, ny, nz = x[f:l + 1] + dv[0], y[f:l + 1] + dv[1], z[f:l + 1] + dv[2]
        c64, l64 = a64.kspace_delta(im, nx, ny, nz)
        c32, l32 = a32.kspace_delta(im, nx, ny, nz)
        errs.append(abs((c32 + l32) - (c64 + l64)))
        scales.append(abs(c64) + abs(l64))
        g64 = np.stack(a64.kspace_forces(f, nx, ny, nz), 1)
        g32 = np.stack(a32.kspace_forces(f, nx, ny, nz), 1)
        ferr.append(float(np.max(np.linalg.norm(g32 - g64, axis=1) / np.maximum(np.linalg.norm(g64, axis=1), 1e-300))))
    a64.close(); a32.close()
    errs = np.array(errs)
    out["same_configuration_kspace"] = {
        "trial_moves": len(errs), "max_dE_kT": float(errs.max()), "max_dE_kJmol": float(errs.max() * kT_kjmol),
        "max_rel_dE_vs_scale": float((errs / np.maximum(scales, 1e-300)).max()),
        "max_rel_force_err_norm": float(max(ferr))}
    a, b = recs["fp64"], recs["fp32c"]
    same = (a["accepted"] == b["accepted"]) & (a["imol"] == b["imol"])
    n = int(np.argmin(same)) if not same.all() else len(a)
    ok = a["overlap"][:n] == 0
    d = np.abs(a["du_total"][:n] - b["du_total"][:n])[ok]
    scale = (np.abs(a["du_real"]) + np.abs(a["du_kc"]) + np.abs(a["du_klj"]) + np.abs(a["du_ext"]))[:n][ok]
    # along the two chains (same stream, diverging at the 1e-9 level)
    out["moves_before_first_differing_decision"] = n
    out["max_dE_kT"] = float(d.max()) if d.size else 0.0
    out["max_dE_kJmol"] = float(d.max() * kT_kjmol) if d.size else 0.0
    if d.size:
        k = int(np.argmax(d))
        rel = d / np.maximum(scale, 1e-300)
        out["at_max_dE"] = {"dE_fp64_kT": float(a["du_total"][:n][ok][k]), "scale_kT": float(scale[k]),
                            "dE_real_kT": float(a["du_real"][:n][ok][k]), "dE_kC_kT": float(a["du_kc"][:n][ok][k]),
                            "dE_kLJ_kT": float(a["du_klj"][:n][ok][k]),
                            "kC_err": float(abs(a["du_kc"][:n][ok][k] - b["du_kc"][:n][ok][k])),
                            "kLJ_err": float(abs(a["du_klj"][:n][ok][k] - b["du_klj"][:n][ok][k])),
                            "real_err": float(abs(a["du_real"][:n][ok][k] - b["du_real"][:n][ok][k]))}
        out["max_rel_dE_vs_scale"] = float(rel.max())
        out["dE_err_quantiles_kT"] = {q: float(np.quantile(d, float(q))) for q in ("0.5", "0.99", "0.9999")}
        out["moves_over_1e-6_kJmol"] = int((d * kT_kjmol > 1e-6).sum())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
