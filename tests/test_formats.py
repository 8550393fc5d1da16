"""The on-disk formats (yamc_amd.formats) against the compiled reference's
own bytes and values (tests/golden/*.npz: MoleculeTable_save_binary at 1-3
bytes per sample, write_float of forces and special values,
MoleculeTable_calcPositionsOrientations of the final configuration), and the
reference's own parameter/composition/output files kept as fixtures
(tests/golden/ions100/)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import golden_cases as gc  # noqa: E402

from yamc_amd import formats as fm  # noqa: E402
from yamc_amd import system as ys  # noqa: E402

CASES = gc.cases()
IONS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ions100")


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1])
def test_moltab_bytes_match_reference(path):
    sysm, g = gc.load(path)
    mt = g["moltab_final"]
    for nb in (1, 2, 3):
        mine = np.frombuffer(fm.moltab_encode(mt[:, :3], mt[:, 3:], nb, nb), dtype=np.uint8)
        assert np.array_equal(mine, g[f"moltab{nb}"]), nb
        pos, ang = fm.moltab_decode(g[f"moltab{nb}"].tobytes(), sysm.nmol, nb, nb)
        assert np.max(np.abs(pos - mt[:, :3])) <= 0.5 / 256 ** nb + 1e-15
        # angles come back modulo 2 pi, to within half a quantum
        d = np.abs((ang - mt[:, 3:] + np.pi) % (2 * np.pi) - np.pi)
        assert np.max(d) <= np.pi / 256 ** nb + 1e-12


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1])
def test_float24_bytes_match_reference(path):
    sysm, g = gc.load(path)
    L = float(g["moltab_box"])
    vals = np.concatenate([g["fx_xchg"] / L, g["float_specials"]])
    mine = np.frombuffer(fm.float_encode(vals), dtype=np.uint8)
    assert np.array_equal(mine, g["floats"])
    back = fm.float_decode(g["floats"].tobytes(), len(vals))
    big = np.abs(vals) > 1e-14
    ok = big & (np.abs(vals) < 1e14)
    assert np.all(np.abs(back[ok] - vals[ok]) <= np.abs(vals[ok]) * 2.0 ** -22)


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1])
def test_positions_orientations_rebuild_the_atoms(path):
    """Centres of mass equal the reference's; the extracted angles may be
    another representation of the same rotation, so they are pinned by
    restaging the molecules and comparing the atoms."""
    sysm, g = gc.load(path)
    x, y, z = (g["xchg_pos"][-1][:, d] for d in range(3))
    L = float(g["moltab_box"])
    pos, ang = fm.positions_orientations(sysm, x, y, z, L)
    assert np.max(np.abs(pos - g["moltab_final"][:, :3])) < 1e-13
    p2 = ys.Params(**{**sysm.params.__dict__, "box_length": L})
    re = ys.build_system(p2, sysm.species, sysm.counts, sysm.mol_type, pos, ang)
    assert max(np.max(np.abs(re.x - x)), np.max(np.abs(re.y - y)), np.max(np.abs(re.z - z))) < 1e-12
    # the reference's own angles restage to the same atoms too
    re2 = ys.build_system(p2, sysm.species, sysm.counts, sysm.mol_type, g["moltab_final"][:, :3],
                          g["moltab_final"][:, 3:])
    assert np.max(np.abs(re2.x - x)) < 1e-12


def test_reference_parameter_and_composition_files():
    prm = fm.read_parameters(os.path.join(IONS, "parameters.prm"))
    assert prm["sr"] == 4.0 and prm["sk"] == 4.0 and prm["alphaL"] == 0.0 and prm["temp"] == 300.0
    assert prm["pressure"] == 100000.0 and prm["pressure_step_multiplier"] == 5.0
    assert prm["traj_file"] == "traj.moltraj" and prm["output_nbytes_xyz"] == 2
    species, counts = fm.read_composition(os.path.join(IONS, "ions100.composition"))
    assert [m.name for m in species] == ["SPCE", "Na_Dang", "Cl_Dang"] and counts == [100, 5, 5]
    p = fm.params_from(prm, sum(counts))
    der = p.derived()
    # parameters.out of the reference run
    assert abs(p.box_length - 15.420216697275805) < 1e-12
    assert der["kmax"] == 10 and abs(der["xclb"] - 36.121565224020877) < 1e-12
    assert abs(der["kext"] - 1.6096346658716105) < 1e-14
    assert abs(der["dr"] - 2.5939972689921119e-2) < 1e-16


def test_reference_output_tables_parse():
    e = fm.read_table(os.path.join(IONS, "energy.dat"))
    assert e.shape == (100, 9) and np.all(np.diff(e[:, 0]) == 100)
    # U_total = U_total_ew + U_total_lj
    assert np.max(np.abs(e[:, 1] - e[:, 2] - e[:, 3])) < 1e-9 * np.max(np.abs(e[:, 1]))
    b = fm.read_table(os.path.join(IONS, "boxlength.dat"))
    assert b.shape[1] == 2 and np.all(np.diff(b[:, 0]) > 0)
    r = fm.read_table(os.path.join(IONS, "RDF_Na_Cl.dat"))
    assert r.shape == (121, 2)


def test_freq_file_roundtrip(tmp_path):
    f = tmp_path / "freq.txt"
    f.write_text("# comment\n2\n1-100 1 0\n101-110 1 0.5\n")
    mv, xc = fm.read_freq_file(str(f), 110)
    assert np.all(mv == 1) and np.all(xc[:100] == 0) and np.all(xc[100:] == 0.5)
    f.write_text("1\n1-100 1 0\n")
    with pytest.raises(ValueError):
        fm.read_freq_file(str(f), 110)


def test_run_files(tmp_path):
    prm = dict(fm.DEFAULTS)
    rf = fm.RunFiles(prm, str(tmp_path))
    e = {k: float(i) for i, k in enumerate(["U_total", "U_total_ew", "U_total_lj", "uu_ew", "uu_lj6", "uu_lj12",
                                            "E_kC", "E_kLJ"])}
    rf.energy(100, e)
    rf.boxlength(7, 15.5)
    pos = np.array([[0.1, -0.2, 0.3]])
    ang = np.array([[1.0, 2.0, 3.0]])
    rf.frame(1, 15.5, [1, 2, 3, 4, 5, 6], pos, ang, 0)
    t = fm.read_table(rf.path("energy_file"))
    assert t.shape == (1, 9) and t[0, 0] == 100 and t[0, 8] == 7.0
    raw = open(rf.path("traj_file"), "rb").read()
    p2, a2 = fm.moltab_decode(raw, 1, 2, 2)
    assert np.max(np.abs(p2 - pos)) < 1e-5 and np.max(np.abs(a2 - ang)) < 1e-4


def ftraj_compare(mine: bytes, ref: bytes, n: int):
    """Two force files of n values: the same layout and size, every value
    within the 24-bit mantissa's quantum (values that differ by rounding
    in the last place of FP64 can land on either side of a quantum step),
    and nearly all 4-byte records identical."""
    assert len(mine) == len(ref) == 4 * n
    a, b = fm.float_decode(mine, n), fm.float_decode(ref, n)
    assert np.all(np.abs(a - b) <= np.abs(b) * 2.0 ** -22 + 1e-300)
    same = np.all(np.frombuffer(mine, np.uint8).reshape(n, 4) == np.frombuffer(ref, np.uint8).reshape(n, 4), axis=1)
    assert same.mean() > 0.99, same.mean()


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1])
def test_force_file_frame_matches_reference(oracle_mod, path):
    """A force-file frame (store_forces layout, mc_calc_forces.f90:348-368:
    per atom F/L x, y, z as 24+8-bit floats) of the oracle's forces in the
    reference convention (real space j > i + k space + external) against the
    frame the compiled reference wrote with write_float from its own forces
    of the same configuration."""
    from yamc_amd.calc_forces import ftraj_frame
    sysm, g = gc.load(path)
    o = oracle_mod.Oracle(sysm)
    o.mc_init()
    F = o.forces_refconv()
    ftraj_compare(ftraj_frame(F, sysm.params.box_length), g["ftraj_init"].tobytes(), 3 * sysm.natom)
