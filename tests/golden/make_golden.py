#!/usr/bin/env python3
"""Writes the golden fixtures tests/golden/<case>.npz from the compiled
reference (oracle/_ref/ref_driver, built by oracle/build_ref.sh from the YAMC
Fortran sources).  Run in the development container; the fixtures are data
(inputs and reference outputs) and are committed.

Each case stages a small system with yamc_amd.system, writes the .mol and
composition files and a case file, runs the reference driver, and stores
the inputs (molecule table, parameters, move list) with the reference's
outputs: staged atoms, LJ tables, k grid, beta tables, S(k) sums, energy
components, per-atom real-space uu/F, per-atom k-space forces, and the state
after every move of a constant-uniform move trace (movemc with RND_CONSTANT).

The constant uniform of each move is chosen so that no rejection-sampling
loop of the proposal spins forever (checked with the oracle first).
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle, refcase  # noqa: E402
from yamc_amd import system as ys  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def parse(text):
    lines = text.split("\n")
    i = 0
    out = {"BETALJ": [], "SUMS": [], "MOVES": [], "VOLS": [], "XCHGS": []}
    cur = None  # record (move or volume step) that E/positions attach to
    while i < len(lines):
        ln = lines[i].strip()
        if not ln:
            i += 1
            continue
        tok = ln.split()
        if tok[0] == "SIZES":
            out["SIZES"] = [int(v) for v in tok[1:]]
        elif tok[0] == "KMAX":
            out["KMAX"] = int(tok[1])
        elif tok[0] in ("PARAMS", "PARAMS2"):
            out[tok[0]] = [float(v) for v in tok[1:]]
        elif tok[0] in ("R", "I"):
            tag, n = tok[1], int(tok[2])
            vals = [float(v) if tok[0] == "R" else int(v) for v in lines[i + 1:i + 1 + n]]
            i += n
            if tag in ("BETALJ", "SUMS"):
                out[tag].append(vals)
            else:
                out[tag] = vals
        elif tok[0] == "LJ12UPPER":
            n = int(tok[1])
            out["LJ12UPPER"] = [float(v) for v in lines[i + 1:i + 1 + n]]
            i += n
        elif tok[0] == "E":
            vals = " ".join(lines[i + 1:i + 4]).split()
            e = [float(v) for v in vals]
            if tok[1] == "INIT":
                out["E_INIT"] = e
            else:
                cur["E"] = e
            i += 3
        elif tok[0] == "FK":
            n = int(tok[1])
            out["FK"] = [[float(v) for v in lines[i + 1 + k].split()] for k in range(n)]
            i += n
        elif tok[0] == "MOVE":
            cur = {"imol": int(tok[2]), "accepted": tok[3] == "T"}
            out["MOVES"].append(cur)
        elif tok[0] == "VOL":
            cur = {"accepted": tok[2] == "T", "box": float(tok[3])}
            out["VOLS"].append(cur)
        elif "XCHG" in tok[0] and tok[0] != "XCHGS":  # after an overlap the reference prints OX first
            cur = {"imol1": int(tok[2]) - 1, "imol2": int(tok[3]) - 1, "accepted": tok[4] == "T"}
            out["XCHGS"].append(cur)
        elif tok[0] in ("MOVES", "VOLS", "XCHGS"):
            pass
        elif tok[0] == "MOLTAB":
            n = int(tok[1])
            out["MOLTAB_L"] = float(tok[2])
            out["MOLTAB"] = [[float(v) for v in lines[i + 1 + k].split()] for k in range(n)]
            i += n
        elif tok[0] == "END":
            break
        elif len(tok) == 1 and cur is not None and "E" in cur and "pos" not in cur:
            na = int(tok[0])
            cur["pos"] = [[float(v) for v in lines[i + 1 + k].split()] for k in range(na)]
            i += na
        i += 1
    return out


def pick_moves(sysm, nmoves, seed):
    """Molecules and constant uniforms such that every proposal finishes."""
    return refcase.safe_moves(sysm, nmoves, seed, pyoracle)


def make_case(name, sysm, nmoves, seed, vols=(), xchgs=()):
    """vols: (u, pressure [Pa], max_volume_scaling) of mcvol steps run after
    the move trace; xchgs: (imol1, imol2, u) of mc_xchange steps (0-based
    molecules) run after the volume steps."""
    imol, u = pick_moves(sysm, nmoves, seed)
    p = sysm.params
    with tempfile.TemporaryDirectory() as td:
        refcase.write_case(td, sysm, imol, u, vols, xchgs)
        stdout, _ = refcase.run(td, timeout=120)
        blobs = {}
        for fn in ("moltab1.bin", "moltab2.bin", "moltab3.bin", "floats.bin", "forces.ftraj"):
            with open(os.path.join(td, fn), "rb") as fb:
                blobs[fn] = np.frombuffer(fb.read(), dtype=np.uint8).copy()
    g = parse(stdout)
    natom, nk, nt = g["SIZES"]
    moves = g["MOVES"]
    assert len(moves) == nmoves
    arrays = dict(
        # inputs
        params=np.array([p.box_length, p.temp, p.sr, p.sk, p.alpha, p.external_permutivity, p.dr_a,
                         p.d_angle_degree, p.lambda_f, p.lambda_c]),
        species=np.array([m.name for m in sysm.species]), counts=np.array(sysm.counts, np.int32),
        mol_type=sysm.mol_type, mol_pos=sysm.mol_pos, mol_angles=sysm.mol_angles,
        move_imol=imol, move_u=u,
        # reference outputs
        kmax=np.int32(g["KMAX"]), derived=np.array(g["PARAMS"] + g["PARAMS2"]),
        xx=np.array(g["XX"]), yy=np.array(g["YY"]), zz=np.array(g["ZZ"]), charge=np.array(g["CHARGE"]),
        sigma=np.array(g["SIGMA"]), epsilon=np.array(g["EPSILON"]), hard_core=np.array(g["HARDCORE"]),
        molnum=np.array(g["MOLNUM"], np.int32), lj_type=np.array(g["TYPE"], np.int32),
        first=np.array(g["FIRST"], np.int32) - 1, last=np.array(g["LAST"], np.int32) - 1,
        lj6=np.array(g["LJ6"]), lj12_upper=np.array(g["LJ12UPPER"]), four_eps=np.array(g["FOUREPS"]),
        sigma6=np.array(g["SIGMA6"]),
        kx=np.array(g["KX"], np.int32), ky=np.array(g["KY"], np.int32), kz=np.array(g["KZ"], np.int32),
        ip=np.array(g["IP"], np.int32), iq=np.array(g["IQ"], np.int32),
        beta=np.array(g["BETA"]), beta6=np.array(g["BETA6"]), beta12=np.array(g["BETA12"]),
        beta_lj=np.array(g["BETALJ"]).reshape(nt, nt, nk),
        sums=np.array(g["SUMS"][:(1 + nt) * 8]).reshape(1 + nt, 8, nk),
        e_init=np.array(g["E_INIT"]),
        uu=np.array(g["UU"]), fx=np.array(g["FX"]), fy=np.array(g["FY"]), fz=np.array(g["FZ"]),
        fk=np.array(g["FK"]),
        move_accepted=np.array([m["accepted"] for m in moves], np.int32),
        move_e=np.array([m["E"] for m in moves]),
        move_pos=np.array([np.pad(np.array(m["pos"]), ((0, 4 - len(m["pos"])), (0, 0))) for m in moves]),
        uu_end=np.array(g["UUEND"]), fx_end=np.array(g["FXEND"]), fy_end=np.array(g["FYEND"]),
        fz_end=np.array(g["FZEND"]),
        sums_coulomb_end=np.array(g["SUMS"][(1 + nt) * 8:(1 + nt) * 8 + 8]).reshape(8, nk),
    )
    if vols:
        vs = g["VOLS"]
        assert len(vs) == len(vols)
        arrays.update(
            vol_u=np.array([v[0] for v in vols]), vol_pressure=np.array([v[1] for v in vols]),
            vol_mvs=np.array([v[2] for v in vols]),
            vol_accepted=np.array([v["accepted"] for v in vs], np.int32),
            vol_box=np.array([v["box"] for v in vs]), vol_e=np.array([v["E"] for v in vs]),
            vol_pos=np.array([v["pos"] for v in vs]))
    if xchgs:
        xs = g["XCHGS"]
        assert len(xs) == len(xchgs)
        ns = (1 + nt) * 8 + 8
        arrays.update(
            xchg_imol1=np.array([x[0] for x in xchgs], np.int32), xchg_imol2=np.array([x[1] for x in xchgs], np.int32),
            xchg_u=np.array([x[2] for x in xchgs]),
            xchg_accepted=np.array([x["accepted"] for x in xs], np.int32),
            xchg_e=np.array([x["E"] for x in xs]), xchg_pos=np.array([x["pos"] for x in xs]),
            uu_xchg=np.array(g["UUX"]), fx_xchg=np.array(g["FXX"]), fy_xchg=np.array(g["FYX"]),
            fz_xchg=np.array(g["FZX"]), sums_coulomb_xchg=np.array(g["SUMS"][ns:ns + 8]).reshape(8, nk))
    # formats of the final configuration (after the exchange steps)
    arrays.update(
        moltab_final=np.array(g["MOLTAB"]), moltab_box=np.float64(g["MOLTAB_L"]),
        moltab1=blobs["moltab1.bin"], moltab2=blobs["moltab2.bin"], moltab3=blobs["moltab3.bin"],
        floats=blobs["floats.bin"], ftraj_init=blobs["forces.ftraj"],
        float_specials=np.array([0.0, 1.0, -1.0, 0.75, -3.5, 2.0, 1e-15, -1e15, 123456.789, -2.5e-7]))
    path = os.path.join(OUT, f"{name}.npz")
    np.savez_compressed(path, **arrays)
    print(f"{path}: natom={natom} nk={nk} ntype={nt} moves={nmoves} accepted={arrays['move_accepted'].sum()}")


def parse_runmc(text):
    """The run_mc section of the driver's output: the uniform stream, per
    cycle the step counters, box length and energies, the four uniforms
    after the run, the final positions, per-atom arrays and Coulomb S(k)."""
    lines = text.split("\n")
    out = {"CYCLES": [], "SUMS": []}
    i = 0
    cur = None
    while i < len(lines):
        tok = lines[i].split()
        if not tok:
            i += 1
            continue
        if tok[0] == "STREAM":
            n = int(tok[1])
            out["STREAM"] = [float(v) for v in lines[i + 1:i + 1 + n]]
            i += n
        elif tok[0] == "CYCLE":
            cur = {"cycle": int(tok[1]), "counts": [float(v) for v in tok[2:8]], "box": float(tok[8])}
            out["CYCLES"].append(cur)
        elif tok[0] == "E" and cur is not None and "E" not in cur:
            cur["E"] = [float(v) for v in " ".join(lines[i + 1:i + 4]).split()]
            i += 3
        elif tok[0] == "TAIL":
            out["TAIL"] = [float(v) for v in lines[i + 1:i + 5]]
            i += 4
            cur = None
        elif tok[0] == "R" and "TAIL" in out:
            tag, n = tok[1], int(tok[2])
            vals = [float(v) for v in lines[i + 1:i + 1 + n]]
            i += n
            if tag == "SUMS":
                out["SUMS"].append(vals)
            else:
                out[tag] = vals
        elif tok[0] == "MOLTAB":
            break
        i += 1
    return out


def freq_intervals(mv, xc):
    """(first, last, move weight, exchange weight) of the maximal runs of
    equal weights, 1-based molecules: the lines of a frequency file."""
    freq, a, n = [], 0, len(mv)
    for b in range(1, n + 1):
        if b == n or mv[b] != mv[a] or xc[b] != xc[a]:
            freq.append((a + 1, b, float(mv[a]), float(xc[a])))
            a = b
    return freq


def make_runmc_case(name, sysm, ncycles, seed, psm, pressure, mvs, weights=None, nstream=60000):
    """The reference's own run_mc (runmc.f90:178-334) for ncycles cycles with
    its standard random stream: records the stream it draws, the state after
    every cycle and the number of uniforms the run consumed.  weights:
    per-molecule (move, exchange) weights written as a frequency file (None:
    the 'none' file)."""
    freq = freq_intervals(*weights) if weights is not None else []
    run = dict(ncycles=ncycles, seed=seed, psm=psm, pressure=pressure, mvs=mvs, nstream=nstream, freq=freq)
    with tempfile.TemporaryDirectory() as td:
        refcase.write_case(td, sysm, [], [], (), (), runmc=run)
        stdout, _ = refcase.run(td, timeout=600)
    head, _, _ = stdout.partition("\nSTREAM")
    g = parse(head)
    r = parse_runmc(stdout)
    natom, nk, nt = g["SIZES"]
    stream = np.array(r["STREAM"])
    tail = np.array(r["TAIL"])
    # uniforms consumed: the first position where the four uniforms drawn
    # after the run continue the recorded stream
    used = [k for k in range(len(stream) - 3) if np.array_equal(stream[k:k + 4], tail)]
    assert len(used) >= 1, "stream too short or not found"
    used = used[0]
    cyc = r["CYCLES"]
    assert len(cyc) == ncycles
    arrays = dict(
        params=np.array([sysm.params.box_length, sysm.params.temp, sysm.params.sr, sysm.params.sk,
                         sysm.params.alpha, sysm.params.external_permutivity, sysm.params.dr_a,
                         sysm.params.d_angle_degree, sysm.params.lambda_f, sysm.params.lambda_c]),
        species=np.array([m.name for m in sysm.species]), counts=np.array(sysm.counts, np.int32),
        mol_type=sysm.mol_type, mol_pos=sysm.mol_pos, mol_angles=sysm.mol_angles,
        run_psm=np.float64(psm), run_pressure=np.float64(pressure), run_mvs=np.float64(mvs),
        run_seed=np.int64(seed), run_weighted=np.int32(weights is not None),
        move_weights=np.asarray(weights[0] if weights is not None else np.ones(sysm.nmol), np.float64),
        xchange_weights=np.asarray(weights[1] if weights is not None else np.zeros(sysm.nmol), np.float64),
        stream=stream[:used + 64], stream_used=np.int64(used),
        e_init=np.array(g["E_INIT"]),
        cycle_counts=np.array([c["counts"] for c in cyc]), cycle_box=np.array([c["box"] for c in cyc]),
        cycle_e=np.array([c["E"] for c in cyc]),
        xx_end=np.array(r["RUNXX"]), yy_end=np.array(r["RUNYY"]), zz_end=np.array(r["RUNZZ"]),
        uu_end=np.array(r["RUNUU"]), fx_end=np.array(r["RUNFX"]), fy_end=np.array(r["RUNFY"]),
        fz_end=np.array(r["RUNFZ"]), sums_coulomb_end=np.array(r["SUMS"][-8:]).reshape(8, nk),
    )
    path = os.path.join(OUT, f"{name}.npz")
    np.savez_compressed(path, **arrays)
    cc = arrays["cycle_counts"][-1]
    print(f"{path}: natom={natom} nk={nk} cycles={ncycles} uniforms={used} moves+exchanges={cc[0]:.0f} "
          f"accepted={cc[1]:.0f} volume steps={cc[2]:.0f} ({cc[3]:.0f} accepted) exchanges={cc[4]:.0f} "
          f"({cc[5]:.0f} accepted)")


def runmc_cases():
    """Ions with force bias, exchange weights and volume steps; and plain
    NVT water without force bias (the 'none' frequency file)."""
    small = ys.config("small", seed=31)
    t = np.asarray(small.mol_type)
    mv = np.where(t == 0, 1.0, 2.0)
    xc = np.where(t == 0, 0.0, 0.5)
    xc[np.nonzero(t == 0)[0][:3]] = 0.25  # three waters take part in exchanges too
    make_runmc_case("runmc_small_bias", small, 60, 1234567, 2.0, 1.0e5, 1.05, (mv, xc))
    w = ys.lattice_system(
        ys.Params(box_length=12.42, temp=298.0, sr=3.2, sk=3.0, alpha=6.4, external_permutivity=70.0, dr_a=0.3,
                  d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5), [ys.SPCE, ys.NA_DANG, ys.CL_DANG],
        [60, 2, 2], 34)
    make_runmc_case("runmc_ions64_nvt", w, 12, 7654321, 0.0, 1.0e5, 1.05, None)


def main():
    if not refcase.available():
        raise SystemExit("oracle/_ref/ref_driver is not built (oracle/build_ref.sh)")
    if len(sys.argv) > 1 and sys.argv[1] == "runmc":
        runmc_cases()
        return
    # volume steps: shrink and grow at 1 bar and at 5 kbar, 2 % maximum scaling
    vols = [(0.30, 1.0e5, 1.02), (0.85, 1.0e5, 1.02), (0.10, 5.0e8, 1.02), (0.65, 1.0e5, 1.05)]
    small = ys.config("small", seed=31)
    make_case("ref_small_bias", small, 40, 1, vols, xchange_pairs(small, 11))
    small_nb = ys.config("small_nobias", seed=32)
    make_case("ref_small_nobias", small_nb, 40, 2, vols, xchange_pairs(small_nb, 12))
    w64 = ys.lattice_system(
        ys.Params(box_length=12.42, temp=298.0, sr=3.2, sk=3.0, alpha=6.4, external_permutivity=70.0, dr_a=0.3,
                  d_angle_degree=15.0, lambda_f=0.0, lambda_c=0.0), [ys.SPCE], [64], 33)
    make_case("ref_water64", w64, 30, 3, vols, xchange_pairs(w64, 13))
    runmc_cases()


def xchange_pairs(sysm, seed, n=8):
    """Exchange steps over every species pairing present (ion-ion,
    water-ion, water-water), with small and large constant uniforms."""
    rng = np.random.default_rng(seed)
    t = np.asarray(sysm.mol_type)
    by = {k: np.nonzero(t == k)[0] for k in sorted(set(t.tolist()))}
    kinds = sorted(by)
    combos = [(a, b) for a in kinds for b in kinds if a <= b and (a != b or len(by[a]) > 1)]
    out = []
    for k in range(n):
        a, b = combos[k % len(combos)]
        i1 = int(rng.choice(by[a]))
        i2 = int(rng.choice([m for m in by[b] if m != i1]))
        out.append((i1, i2, 0.05 if k % 2 == 0 else 0.9))
    return out


if __name__ == "__main__":
    main()
