"""The reference's own Monte Carlo code on the device library.

oracle/_ref/shim/shim_driver (oracle/build_shim.sh) is the golden-vector
driver and the reference's MC.f90 / runmc.f90, compiled unchanged, with its
FourierGrid, SumSinCosKR, RhoSquared, EwaldSumKSpace, ForceKSpace,
RealSumLocal, EwaldSumRealSpace and EwaldSumTails modules replaced by the
drop-in modules of fortran/shim/ over libyamc_ewald.so.  On the inputs of
every golden case it must write what the pure reference build wrote: the
set-up (k grid, beta tables, S(k) sums, energies, per-atom real-space uu and
F, per-atom k-space forces), every move of the constant-uniform movemc trace,
the volume and exchange steps, the final molecule table and force-file
frame, and the reference's own run_mc cycles with its own random stream.

Bounds: integer data and every accept/reject decision identical; real data
within 1e-10 of the largest magnitude of each array (the device sums in
another order than the reference's loops); move positions within 1e-12 box
lengths; force-file values within the 24-bit mantissa quantum.
"""
import os
import sys
import tempfile
import time

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import golden_cases as gc  # noqa: E402

pytestmark = pytest.mark.gpu

RTOL = gc.RTOL


@pytest.fixture(scope="module")
def driver():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from oracle import refcase
    if not (os.path.isfile(refcase.SHIM_DRIVER) and os.access(refcase.SHIM_DRIVER, os.X_OK)):
        pytest.skip("oracle/_ref/shim/shim_driver is not built (oracle/build_shim.sh needs the reference sources)")
    return refcase


def rel(a, b):
    return gc.rel(a, b, b)


def run_case(refcase, sysm, imol, u, vols=(), xchgs=(), runmc=None):
    with tempfile.TemporaryDirectory() as td:
        refcase.write_case(td, sysm, imol, u, vols, xchgs, runmc=runmc)
        print(f"shim_driver: {sysm.natom} atoms, {len(imol)} moves", flush=True)
        t0 = time.perf_counter()
        out, _ = refcase.run(td, timeout=300, driver=refcase.SHIM_DRIVER)
        dt = time.perf_counter() - t0
        blobs = {}
        for fn in ("forces.ftraj",):
            p = os.path.join(td, fn)
            if os.path.exists(p):
                blobs[fn] = open(p, "rb").read()
    return out, dt, blobs


@pytest.mark.parametrize("path", gc.cases(), ids=lambda p: p.split("/")[-1])
def test_reference_mc_through_shims_matches_reference(driver, path):
    import make_golden as mg
    from test_formats import ftraj_compare
    sysm, g = gc.load(path)
    vols = list(zip(g["vol_u"], g["vol_pressure"], g["vol_mvs"])) if "vol_u" in g else []
    xchgs = list(zip(g["xchg_imol1"], g["xchg_imol2"], g["xchg_u"])) if "xchg_u" in g else []
    out, dt, blobs = run_case(driver, sysm, g["move_imol"], g["move_u"], vols, xchgs)
    print(f"  {dt:.2f} s", flush=True)
    s = mg.parse(out)
    natom, nk, nt = s["SIZES"]
    assert (natom, nk) == (sysm.natom, len(g["kx"]))
    for k in ("kx", "ky", "kz", "ip", "iq"):
        assert np.array_equal(np.array(s[k.upper()]), g[k]), k
    for k in ("beta", "beta6", "beta12"):
        assert rel(s[k.upper()], g[k]) < RTOL, k
    assert rel(np.array(s["BETALJ"]).reshape(nt, nt, nk), g["beta_lj"]) < RTOL
    assert rel(np.array(s["SUMS"][:(1 + nt) * 8]).reshape(1 + nt, 8, nk), g["sums"]) < RTOL
    assert rel(s["E_INIT"], g["e_init"]) < RTOL
    assert rel(s["UU"], g["uu"]) < RTOL
    f, fr = np.stack([s["FX"], s["FY"], s["FZ"]]), np.stack([g["fx"], g["fy"], g["fz"]])
    assert rel(f, fr) < RTOL
    assert rel(s["FK"], g["fk"]) < RTOL
    moves = s["MOVES"]
    assert [m["accepted"] for m in moves] == [bool(a) for a in g["move_accepted"]]
    for k, m in enumerate(moves):
        assert rel(m["E"], g["move_e"][k]) < RTOL, k
        p = np.array(m["pos"])
        assert np.abs(p - g["move_pos"][k][:len(p)]).max() < 1e-12, k
    assert rel(s["UUEND"], g["uu_end"]) < RTOL
    fe = np.stack([g["fx_end"], g["fy_end"], g["fz_end"]])
    assert rel(np.stack([s["FXEND"], s["FYEND"], s["FZEND"]]), fe) < RTOL
    ns = (1 + nt) * 8
    assert rel(np.array(s["SUMS"][ns:ns + 8]).reshape(8, nk), g["sums_coulomb_end"]) < RTOL
    if vols:
        vs = s["VOLS"]
        assert [v["accepted"] for v in vs] == [bool(a) for a in g["vol_accepted"]]
        for k, v in enumerate(vs):
            assert abs(v["box"] - g["vol_box"][k]) <= 1e-12 * g["vol_box"][k], k
            assert rel(v["E"], g["vol_e"][k]) < RTOL, k
            assert np.abs(np.array(v["pos"]) - g["vol_pos"][k]).max() < 1e-12, k
    if xchgs:
        xs = s["XCHGS"]
        assert [x["accepted"] for x in xs] == [bool(a) for a in g["xchg_accepted"]]
        for k, x in enumerate(xs):
            assert rel(x["E"], g["xchg_e"][k]) < RTOL, k
            assert np.abs(np.array(x["pos"]) - g["xchg_pos"][k]).max() < 1e-12, k
        assert rel(s["UUX"], g["uu_xchg"]) < RTOL
        fx = np.stack([g["fx_xchg"], g["fy_xchg"], g["fz_xchg"]])
        assert rel(np.stack([s["FXX"], s["FYX"], s["FZX"]]), fx) < RTOL
        assert rel(np.array(s["SUMS"][ns + 8:ns + 16]).reshape(8, nk), g["sums_coulomb_xchg"]) < RTOL
    assert np.abs(np.array(s["MOLTAB"]) - g["moltab_final"]).max() < 1e-10
    assert s["MOLTAB_L"] == pytest.approx(float(g["moltab_box"]), rel=1e-12)
    ftraj_compare(blobs["forces.ftraj"], g["ftraj_init"].tobytes(), 3 * natom)


@pytest.mark.parametrize("path", gc.runmc_cases(), ids=lambda p: p.split("/")[-1])
def test_reference_run_mc_through_shims_matches_reference(driver, path):
    """run_mc (runmc.f90:178-334) with the reference's own random stream:
    the same stream, the same step counters after every cycle, the same
    number of uniforms consumed and the same final state."""
    import make_golden as mg
    sysm, g = gc.load(path)
    freq = mg.freq_intervals(g["move_weights"], g["xchange_weights"]) if int(g["run_weighted"]) else []
    nc = len(g["cycle_counts"])
    run = dict(ncycles=nc, seed=int(g["run_seed"]), psm=float(g["run_psm"]), pressure=float(g["run_pressure"]),
               mvs=float(g["run_mvs"]), nstream=len(g["stream"]), freq=freq)
    out, dt, _ = run_case(driver, sysm, [], [], runmc=run)
    head, _, _ = out.partition("\nSTREAM")
    s0 = mg.parse(head)
    r = mg.parse_runmc(out)
    steps = g["cycle_counts"][-1][0] + g["cycle_counts"][-1][2]
    print(f"  {dt:.2f} s for {steps:.0f} steps ({1e3 * dt / steps:.3f} ms per step)", flush=True)
    assert rel(s0["E_INIT"], g["e_init"]) < RTOL
    assert np.array_equal(np.array(r["STREAM"]), g["stream"])
    cyc = r["CYCLES"]
    assert len(cyc) == nc
    assert np.array_equal(np.array([c["counts"] for c in cyc]), g["cycle_counts"])
    assert rel([c["box"] for c in cyc], g["cycle_box"]) < 1e-12
    for k, c in enumerate(cyc):
        assert rel(c["E"], g["cycle_e"][k]) < RTOL, k
    st, tail = g["stream"], np.array(r["TAIL"])
    used = [k for k in range(len(st) - 3) if np.array_equal(st[k:k + 4], tail)]
    assert used and used[0] == int(g["stream_used"])
    for k in ("xx", "yy", "zz"):
        e = np.abs(np.array(r["RUN" + k.upper()]) - g[k + "_end"]).max()
        assert e < 1e-10, (k, e)
    assert rel(r["RUNUU"], g["uu_end"]) < RTOL
    fe = np.stack([g["fx_end"], g["fy_end"], g["fz_end"]])
    assert rel(np.stack([r["RUNFX"], r["RUNFY"], r["RUNFZ"]]), fe) < RTOL
    nk = len(g["sums_coulomb_end"][0])
    assert rel(np.array(r["SUMS"][-8:]).reshape(8, nk), g["sums_coulomb_end"]) < RTOL
