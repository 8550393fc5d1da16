"""The FP64 oracle against golden vectors of the compiled reference Fortran
(tests/golden/*.npz, written by tests/golden/make_golden.py from
oracle/_ref/ref_driver).  Pins staging, beta tables, S(k), every energy
component, per-atom real-space and k-space forces, a movemc trace, mcvol
volume steps and mc_xchange exchange steps with identical decisions, all
within 1e-10 relative (staging within 1e-14)."""
import numpy as np
import pytest

import os, sys
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import golden_cases as gc  # noqa: E402

CASES = gc.cases()


RUNMC = gc.runmc_cases()


def test_fixtures_present():
    assert len(CASES) >= 3
    assert len(RUNMC) >= 2


@pytest.mark.parametrize("path", RUNMC, ids=lambda p: p.split("/")[-1])
def test_oracle_replays_reference_run_mc(oracle_mod, path):
    """run_mc's random-number consumption (runmc.f90:204-268: a pressure
    draw per cycle, an exchange draw per step, BiasedRandom_choose picks,
    the rejection loops of the proposals, the acceptance draws of movemc,
    mcvol and mc_xchange) pinned against the reference's own run: its
    recorded uniform stream, replayed by the oracle, gives the same step
    counts, energies after every cycle and final state, and uses up exactly
    as many uniforms."""
    sysm, g = gc.load(path)
    o = oracle_mod.Oracle(sysm)
    e, _ = o.mc_init()
    ev = gc.energies_vec(e, o.mu(), o.tails())
    assert gc.rel(ev, g["e_init"], g["e_init"]) < gc.RTOL
    psm, p, lmvs = gc.runmc_setup(o, g)
    ncyc, nx = gc.check_runmc(o, sysm, g, lambda: o.run_npt(1, psm, p, lmvs))
    assert ncyc == len(g["cycle_e"])


@pytest.mark.parametrize("path", CASES, ids=lambda p: p.split("/")[-1])
def test_oracle_matches_reference(oracle_mod, path):
    sysm, g = gc.load(path)
    gc.check_staging(sysm, g)
    o = oracle_mod.Oracle(sysm)
    o.mc_init()

    class View:  # oracle surface in the layout check_init expects
        def grid(self_):
            return oracle_mod.grid(sysm.kmax)

        def __getattr__(self_, k):
            return getattr(o, k)

    v = View()
    gc.check_init(v, sysm, g)

    def step(imol, u):
        acc, recs = o.run_const(np.array([imol], np.int32), np.array([u]))
        return acc

    gc.check_trace(v, sysm, g, step)

    def vstep(u, p, lmvs):
        acc, L, ov = o.volume_move(p, lmvs, u=u)
        return acc, L

    assert gc.check_volume_steps(v, sysm, g, vstep) == 4

    def xstep(i1, i2, u):
        acc, rec = o.xchange(i1, i2, u=u)
        return acc

    assert gc.check_xchange_steps(v, sysm, g, xstep) == 8
