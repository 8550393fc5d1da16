"""mc_calc_forces on the device (yamc_amd.calc_forces): force files in the
reference's store_forces layout (per atom F/L x, y, z as 24+8-bit floats,
mc_calc_forces.f90:348-368), pinned against the frame the compiled reference
wrote from its own forces of the golden configurations, and end to end over
a run's trajectory (frames.dat + binary moltab frames at changing box
lengths) against the oracle's forces of every decoded frame."""
import os
import shutil
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import golden_cases as gc  # noqa: E402
from test_formats import ftraj_compare  # noqa: E402

pytestmark = pytest.mark.gpu

IONS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ions100")


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib
    return _lib.load()


@pytest.mark.parametrize("path", gc.cases(), ids=lambda p: p.split("/")[-1])
def test_device_force_frame_matches_reference(lib, path):
    from yamc_amd.calc_forces import frame_forces, ftraj_frame
    sysm, g = gc.load(path)
    frames = [(sysm.params.box_length, sysm.mol_pos, sysm.mol_angles)]
    F = next(frame_forces(sysm.species, sysm.counts, sysm.params, frames))
    ftraj_compare(ftraj_frame(F, sysm.params.box_length), g["ftraj_init"].tobytes(), 3 * sysm.natom)


def test_calc_forces_over_a_run_trajectory(lib, tmp_path, oracle_mod):
    """The reference's ions100 system (its parameters.prm, composition and
    .mol files, a volume step in half of the cycles): 12 NPT run_mc cycles
    on the device writing a trajectory frame every 2 cycles, then the
    mc_calc_forces program over those files.  Every frame of the force file equals the oracle's forces of the decoded
    frame at that frame's box length."""
    from yamc_amd import calc_forces as cf
    from yamc_amd import formats as fm
    from yamc_amd import system as ys
    from yamc_amd.runmc import run
    for fn in ("parameters.prm", "ions100.composition", "SPCE.mol", "Na_Dang.mol", "Cl_Dang.mol"):
        shutil.copy(os.path.join(IONS, fn), tmp_path / fn)
    prm = fm.read_parameters(str(tmp_path / "parameters.prm"))
    prm["n_store_traj_interval"] = 2.0
    prm["pressure_step_multiplier"] = 1.0  # half of the cycles are volume steps
    species, counts = fm.read_composition(str(tmp_path / "ions100.composition"))
    params = fm.params_from(prm, sum(counts))
    sysm = ys.lattice_system(params, species, counts, 20261209)
    res = run(sysm, prm, 12, 0, seed=20261209, out_dir=str(tmp_path))
    rows = cf.read_frames(str(tmp_path / str(prm["frames_file"])))
    assert len(rows) == 6
    out = str(tmp_path / "forces.ftraj")
    assert cf.main([str(tmp_path / "parameters.prm"), str(tmp_path / "ions100.composition"), out]) == 0
    data = open(out, "rb").read()
    n = 3 * sysm.natom
    assert len(data) == len(rows) * 4 * n
    traj = cf.trajectory(str(tmp_path / str(prm["traj_file"])), sysm.nmol, len(rows), 2, 2)
    mol_type = sysm.mol_type
    boxes = set()
    for k, ((_, L), (pos, ang)) in enumerate(zip(rows, traj)):
        boxes.add(L)
        p = ys.Params(**{**params.__dict__, "box_length": L})
        fr = ys.build_system(p, species, counts, mol_type, pos, ang)
        o = oracle_mod.Oracle(fr)
        o.mc_init()
        ftraj_compare(data[k * 4 * n:(k + 1) * 4 * n], cf.ftraj_frame(o.forces_refconv(), L), n)
    assert res["n_vol"] > 0 and len(boxes) > 1  # frames at more than one box length
