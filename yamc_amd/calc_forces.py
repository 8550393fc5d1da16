"""mc_calc_forces (mc_calc_forces.f90) as a host program over the device library.

    python -m yamc_amd.calc_forces parameters.prm system.composition forces.ftraj
        [--dir DIR] [--precision fp64|fp32c]

reads a run's frames file (frames_file: frame number and box length per row,
read_traj_frame, runmc.f90:418-425) and its binary trajectory (traj_file,
output_nbytes_xyz / output_nbytes_ang bytes per sample,
MoleculeTable_load_binary, MoleculeTable.f90:240-319) from DIR (default:
next to the parameters file), and for every frame stages the molecules
(MoleculeTable_fillAtomicData), computes every atom's force on the device and
appends the frame to forces.ftraj as store_forces does
(mc_calc_forces.f90:348-368): per atom F/L of x, y, z as 24+8-bit floats
(write_float, FloatingPoint.f90:166-205), frames one after the other.

The box of each frame follows the reference's load_frame: Parameters_recalc
at the frame's box length and LJTypes_scale_sigma from the previous frame's
box (yamc_set_box), then the new molecule table, a from-scratch recompute
(calc_forces calls initBeta and rebuilds the sums each frame) and
yamc_forces.

The force written is the reference's convention for F (real-space pair forces
of pairs j > i as movemc keeps them in RealSpace_TotalSum%F, plus
ForceKSpace_coulomb/_LJ and the external-field force).  calc_forces itself
calls the full real-space sum with first = 1 (mc_calc_forces.f90:246), which
also visits j = i (a 0/0 pair) and every pair twice; that sum is not
reproduced.
"""
from __future__ import annotations

import argparse
import dataclasses
import os
import sys
from typing import Iterator, List, Sequence, Tuple

import numpy as np

from . import formats as fm
from . import system as ys
from .ewald import EwaldMC


def read_frames(frames_path: str) -> List[Tuple[int, float]]:
    """(frame number, box length) of every row of a frames file."""
    out = []
    with open(frames_path) as f:
        for line in f:
            tok = line.split()
            if len(tok) >= 2:
                out.append((int(float(tok[0])), float(tok[1])))
    return out


def trajectory(traj_path: str, nmol: int, nframes: int, nbytes_xyz: int,
               nbytes_ang: int) -> Iterator[Tuple[np.ndarray, np.ndarray]]:
    """Molecule positions (box units) and ZYZ angles of each frame."""
    rec = nmol * 3 * (nbytes_xyz + nbytes_ang)
    with open(traj_path, "rb") as f:
        buf = f.read()
    if len(buf) < nframes * rec:
        raise ValueError(f"{traj_path}: {len(buf)} bytes hold fewer than {nframes} frames of {nmol} molecules")
    for k in range(nframes):
        yield fm.moltab_decode(buf, nmol, nbytes_xyz, nbytes_ang, offset=k * rec)


def frame_forces(species: Sequence[ys.Molecule], counts: Sequence[int], params: ys.Params,
                 frames: Sequence[Tuple[float, np.ndarray, np.ndarray]], precision: int = 0) -> Iterator[np.ndarray]:
    """Forces F [natom, 3] (kT per box length) of each (box length,
    positions, angles) frame, reference convention."""
    mol_type = np.concatenate([np.full(c, t, dtype=np.int32) for t, c in enumerate(counts)])
    g = None
    for L, pos, ang in frames:
        p = dataclasses.replace(params, box_length=float(L))
        sysm = ys.build_system(p, species, counts, mol_type, pos, ang)
        if g is None:
            g = EwaldMC(sysm, precision=precision)
        else:
            g.set_box(float(L), scale_positions=False)
            g.set_positions(sysm.x, sysm.y, sysm.z)
        g.mc_init()
        _, ref = g.forces()
        yield ref
    if g is not None:
        g.close()


def ftraj_frame(F: np.ndarray, box_length: float) -> bytes:
    """One frame of a force file: F/L per atom, x, y, z (store_forces)."""
    return fm.float_encode((np.asarray(F, np.float64) / box_length).reshape(-1))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("parameters")
    ap.add_argument("composition")
    ap.add_argument("output")
    ap.add_argument("--dir", default=None, help="directory of the frames and trajectory files")
    ap.add_argument("--precision", choices=("fp64", "fp32c"), default="fp64")
    a = ap.parse_args(argv)
    prm = fm.read_parameters(a.parameters)
    species, counts = fm.read_composition(a.composition)
    nmol = sum(counts)
    params = fm.params_from(prm, nmol)
    d = a.dir or os.path.dirname(os.path.abspath(a.parameters))
    rows = read_frames(os.path.join(d, str(prm["frames_file"])))
    traj = trajectory(os.path.join(d, str(prm["traj_file"])), nmol, len(rows), int(prm["output_nbytes_xyz"]),
                      int(prm["output_nbytes_ang"]))
    frames = [(L, pos, ang) for (_, L), (pos, ang) in zip(rows, traj)]
    with open(a.output, "wb") as out:
        for k, F in enumerate(frame_forces(species, counts, params, frames, 1 if a.precision == "fp32c" else 0)):
            out.write(ftraj_frame(F, frames[k][0]))
            print(f"Fram: {k + 1}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
