"""run_mc as a host program over the device library (mc_main.f90 + runmc.f90).

    python -m yamc_amd.runmc parameters.prm system.composition [--out DIR]
        [--cycles N] [--precision fp64|fp32c] [--seed S] [--freq FILE]

reads the reference's parameter and composition files (and .mol files next
to the composition), stages the molecules on a seeded simple-cubic lattice
at the parameter file's box length or density, and runs max_cycle cycles of
run_mc (runmc.f90:178-334): per cycle one draw chooses a volume step (mcvol,
when pressure_step_multiplier > 0) or nmol steps of moves and exchange
steps (exchange weights from freq_file).  Every cycle runs on the device
through yamc_mc_run_npt; the host writes the reference's output files
(energy.dat every n_store_energy_interval cycles, boxlength.dat after each
accepted volume step, a moltab trajectory frame and a frames.dat row every
n_store_traj_interval cycles).
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
from typing import Dict, Optional

import numpy as np

from . import formats as fm
from . import system as ys
from .ewald import EwaldMC


def pressure_kT_A3(pressure_pa: float, temp: float) -> float:
    """pressure_angstr_kT (parameters.f90:411): P / (k_B T) * 1e-30."""
    return pressure_pa / (ys.BOLTZ * temp) * 1e-30


def run(sysm: ys.System, prm: Dict[str, object], cycles: Optional[int] = None, precision: int = 0, seed: int = 1,
        out_dir: Optional[str] = None, freq=None, rdf=None, progress=None):
    """Runs the cycles; returns a dict of per-cycle box lengths, the energy
    rows (every n_store_energy_interval cycles) and the step counters.
    rdf = (mol_species, nspecies, dr, rmax, every, first_cycle) accumulates
    centre-of-mass pair counts on the device every `every` cycles from
    first_cycle on (mcrdf2's cnt4 over the frames) with the frames' sum of
    1/V, returned as rdf_counts, rdf_frames, rdf_inv_vol."""
    g = EwaldMC(sysm, precision=precision)
    if freq is not None:
        g.set_frequencies(*freq)
    e0, ov = g.mc_init()
    if ov:
        raise RuntimeError("overlap in the starting configuration")
    g.set_rng(seed)
    T = sysm.params.temp
    psm = float(prm["pressure_step_multiplier"])
    P = pressure_kT_A3(float(prm["pressure"]), T)
    lmvs = math.log(float(prm["max_volume_scaling"]))
    e_int = float(prm["n_store_energy_interval"])
    t_int = float(prm["n_store_traj_interval"])
    ncyc = int(prm["max_cycle"]) if cycles is None else int(cycles)
    rf = fm.RunFiles(prm, out_dir) if out_dir else None
    n_mv = n_mv_acc = n_vol = n_vol_acc = 0
    box = np.empty(ncyc)
    rows = []
    counts, nfr, inv_vol = None, 0, 0.0
    t0 = time.time()
    for i in range(1, ncyc + 1):
        acc, nv, nva = g.run_npt(1, psm, P, lmvs)
        if nv:
            n_vol += nv
            n_vol_acc += nva
            if nva and rf:
                rf.boxlength(i, g.box_length())
        else:
            n_mv += sysm.nmol
            n_mv_acc += acc
        box[i - 1] = g.box_length()
        done = n_mv / sysm.nmol + n_vol
        if math.fmod(done, e_int) < 1e-10:
            e = g.energies()
            rows.append([i] + [e[k] for k in ("U_total", "U_total_ew", "U_total_lj", "uu_ew", "uu_lj6", "uu_lj12",
                                              "E_kC", "E_kLJ")])
            if rf:
                rf.energy(i, e)
        if rf and math.fmod(done, t_int) < 1e-10:
            x, y, z = g.positions()
            pos, ang = fm.positions_orientations(sysm, x, y, z, g.box_length())
            rf.frame(i, g.box_length(), [n_mv, n_mv_acc, n_vol, n_vol_acc, 0, 0], pos, ang, int(time.time()))
        if rdf is not None and i >= rdf[5] and i % rdf[4] == 0:
            counts = g.rdf(rdf[0], rdf[1], rdf[2], rdf[3], counts)
            nfr += 1
            inv_vol += 1.0 / g.box_length() ** 3
        if progress and i % progress == 0:
            print(f"cycle {i}/{ncyc}  L={g.box_length():.4f}  {time.time() - t0:.1f}s", flush=True)
    out = dict(box=box, energy=np.array(rows), n_mv=n_mv, n_mv_acc=n_mv_acc, n_vol=n_vol, n_vol_acc=n_vol_acc,
               seconds=time.time() - t0, final_energies=g.energies(), rdf_counts=counts, rdf_frames=nfr,
               rdf_inv_vol=inv_vol / nfr if nfr else 0.0)
    g.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("parameters")
    ap.add_argument("composition")
    ap.add_argument("--out", default=".")
    ap.add_argument("--cycles", type=int, default=None)
    ap.add_argument("--precision", choices=("fp64", "fp32c"), default="fp64")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--freq", default=None, help="frequency file (overrides freq_file)")
    a = ap.parse_args(argv)
    prm = fm.read_parameters(a.parameters)
    species, counts = fm.read_composition(a.composition)
    nmol = sum(counts)
    params = fm.params_from(prm, nmol)
    seed = a.seed if a.seed is not None else int(prm["random_seed"]) or 1
    sysm = ys.lattice_system(params, species, counts, seed)
    ff = a.freq or (None if prm["freq_file"] == "none" else str(prm["freq_file"]))
    freq = fm.read_freq_file(ff, nmol) if ff else None
    os.makedirs(a.out, exist_ok=True)
    res = run(sysm, prm, a.cycles, 1 if a.precision == "fp32c" else 0, seed, a.out, freq,
              progress=max(1, (a.cycles or int(prm["max_cycle"])) // 20))
    print(f"{res['n_mv']} moves ({res['n_mv_acc']} accepted), {res['n_vol']} volume steps "
          f"({res['n_vol_acc']} accepted) in {res['seconds']:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
