"""Case files for the compiled reference driver -- TEST INFRASTRUCTURE ONLY.

oracle/_ref/ref_driver (built by oracle/build_ref.sh from the reference's
own Fortran sources) reads a composition file, .mol files and a case file
(parameters, molecule table, a constant-uniform move list, optional volume
and exchange steps; see oracle/ref_driver.f90).  This module writes those
inputs for a staged system and runs the driver: tests/golden/make_golden.py
uses it to write the golden fixtures, and bench.py's CPU-baseline leg uses it
to time the reference's own movemc on the benchmark system.
"""
from __future__ import annotations

import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "ref_driver")
# the same driver and the reference's MC code over the drop-in modules of
# fortran/shim/ and the device library (build_shim.sh)
SHIM_DRIVER = os.path.join(HERE, "_ref", "shim", "shim_driver")


def available() -> bool:
    return os.path.isfile(DRIVER) and os.access(DRIVER, os.X_OK)


def write_mol(path, mol):
    with open(path, "w") as f:
        for r in mol.atoms:
            f.write("%s %.15f %.15f %.15f %.6f %.6f %.6f %.4f %.4f\n" % r)


def write_case(td, sysm, imol, u, vols=(), xchgs=(), runmc=None):
    """Writes the inputs of one driver run into directory td; imol is
    0-based.  runmc (optional): dict(ncycles, seed, psm, pressure [Pa],
    mvs, nstream, freq=[(first, last, move_weight, xchange_weight), ...]
    with 1-based molecule intervals, [] = the 'none' file) for the run_mc
    section.  Returns the case file name."""
    p = sysm.params
    names = []
    for k, mol in enumerate(sysm.species):
        fn = f"species{k}.mol"
        write_mol(os.path.join(td, fn), mol)
        names.append(fn)
    with open(os.path.join(td, "case.composition"), "w") as f:
        f.write(f"{len(sysm.species)}\n")
        for fn, cnt in zip(names, sysm.counts):
            f.write(f"{fn} {cnt}\n")
    with open(os.path.join(td, "case.in"), "w") as f:
        f.write("case.composition\n")
        f.write(" ".join("%.17g" % v for v in (p.box_length, p.temp, p.sr, p.sk, p.alpha, p.external_permutivity,
                                              p.dr_a, p.d_angle_degree, p.lambda_f, p.lambda_c)) + "\n")
        f.write(f"{sysm.nmol}\n")
        for i in range(sysm.nmol):
            f.write("%d %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                sysm.mol_type[i] + 1, *sysm.mol_pos[i], *sysm.mol_angles[i]))
        f.write(f"{len(imol)}\n")
        for a, b in zip(imol, u):
            f.write("%d %.17g\n" % (a + 1, b))
        f.write(f"{len(vols)}\n")
        for vu, vp, vm in vols:
            f.write("%.17g %.17g %.17g\n" % (vu, vp, vm))
        f.write(f"{len(xchgs)}\n")
        for a, b, xu in xchgs:
            f.write("%d %d %.17g\n" % (a + 1, b + 1, xu))
        if runmc:
            fr = runmc.get("freq", [])
            f.write("%d %d %.17g %.17g %.17g %d %d\n" % (runmc["ncycles"], runmc["seed"], runmc["psm"],
                                                       runmc["pressure"], runmc["mvs"], runmc["nstream"], len(fr)))
            for a, b, wm, wx in fr:
                f.write("%d %d %.17g %.17g\n" % (a, b, wm, wx))
    return "case.in"


def run(td, timeout=600, driver=DRIVER):
    """Runs the driver (default: the reference build) in td; returns
    (stdout, seconds)."""
    t0 = time.perf_counter()
    r = subprocess.run([driver, "case.in"], cwd=td, capture_output=True, text=True, timeout=timeout)
    dt = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
    return r.stdout, dt


def safe_moves(sysm, nmoves, seed, oracle_mod, u_range=(0.25, 0.75), all_at_once=False):
    """Molecules and constant uniforms whose proposals finish (checked with
    the oracle: with one constant uniform per move a rejection loop can spin
    forever in the reference).  A stalling move gets a new uniform, the first
    one per pass or, with all_at_once, every one."""
    rng = np.random.default_rng(seed)
    imol = rng.integers(0, sysm.nmol, nmoves).astype(np.int32)
    u = rng.uniform(u_range[0], u_range[1], nmoves)
    for _ in range(200):
        o = oracle_mod.Oracle(sysm)
        o.mc_init()
        _, recs = o.run_const(imol, u)
        bad = np.nonzero(recs["overlap"] == 2)[0]
        if not len(bad):
            return imol, u
        if all_at_once:
            u[bad] = rng.uniform(u_range[0], u_range[1], len(bad))
        else:
            u[bad[0]] = rng.uniform(u_range[0], u_range[1])
    raise RuntimeError("could not find non-stalling uniforms")
