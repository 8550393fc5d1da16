"""Launch-shape autotuner of the chain kernels.

The library has four launch shapes for a batch of replicas (and for a single
chain): a workgroup of 256 or 512 threads per chain, or one wavefront per
replica with four replicas per workgroup and registers sized for four or
eight replicas per CU (yamc_set_block_size 256, 512, 64, 128).
Which one is fastest depends on the replica count and size: one chain per
workgroup spreads every sum over the workgroup (short moves, few chains per
CU), one wave per replica trades a longer move for more chains in flight.

A single chain can also spread over G workgroups of a cooperative launch
(yamc_set_chain_groups): each takes a range of nk / G k vectors (nk / 4G per
wavefront) and of the atoms, with two grid barriers per move.  More
workgroups shorten the sweeps and the real-space pass but add barrier and
reduction time, so the best G grows with nk and the atom count;
tune_groups() times one cycle per candidate G and keeps the fastest.  In a
batch the members can be grid chains too (one cooperative launch of all
their workgroups); tune_groups_batch() times the batch with every member at
natom // A workgroups for a few A and keeps the fastest.

tune_batch() times one sweep of the batch itself with each shape and keeps
the fastest.  The timed sweeps are ordinary Monte Carlo sweeps of the chains
(every shape runs the same algorithm on the same RNG streams), so tuning
doubles as warm-up; the choice is applied to every context of the batch.

The shape and the workgroup count fix how many k vectors each wavefront
sweeps per move: nk / (G * shape / 64) (k_per_wave()).  The cell size of the
real-space scan's cell list is the other tile: tune_cells_batch() and
tune_cells() time the cell divisions (cells of width rc / div; div = 0: no
cell list) that give distinct cell counts and keep the fastest.  The scan
finds the same candidates at every cell size, so results do not depend on
it.
"""
from __future__ import annotations

from typing import Dict, Sequence, Tuple

from .ewald import EwaldMC, batch_run

SHAPES = (256, 64, 128, 512)


def shapes_for(precision: int) -> Tuple[int, ...]:
    """Launch shapes that run distinct kernels for this precision (FP32C
    chains run the wave shapes as workgroups of 256)."""
    return (256, 512) if precision == 1 else SHAPES


def tune_batch(reps: Sequence[EwaldMC], shapes: Sequence[int] = None, cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """Returns (best shape, {shape: trial moves/s}) and sets the best shape on
    every replica.  Rates come from the HIP-event time of each launch."""
    if shapes is None:
        shapes = shapes_for(reps[0].precision)
    moves = sum(r.sys.nmol for r in reps) * cycles
    rates = {}
    for bs in shapes:
        for r in reps:
            r.set_block_size(bs)
        _, ms = batch_run(reps, cycles)
        rates[bs] = moves / (ms * 1e-3)
    best = max(rates, key=rates.get)
    for r in reps:
        r.set_block_size(best)
    return best, rates


def tune_single(rep: EwaldMC, shapes: Sequence[int] = None, cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """Same for one chain (yamc_mc_run)."""
    if shapes is None:
        shapes = shapes_for(rep.precision)
    rates = {}
    for bs in shapes:
        rep.set_block_size(bs)
        rep.run(cycles)
        rates[bs] = rep.sys.nmol * cycles / (rep.last_kernel_ms() * 1e-3)
    best = max(rates, key=rates.get)
    rep.set_block_size(best)
    return best, rates


def groups_for(nk: int, natom: int) -> Tuple[int, ...]:
    """Workgroup counts worth timing for one chain: 1, and powers of two
    that leave at least 64 k vectors and 64 atoms per workgroup."""
    out = [1]
    g = 2
    while g <= 64 and nk // g >= 64 and natom // g >= 64:
        out.append(g)
        g *= 2
    return tuple(out)


def tune_groups(rep: EwaldMC, groups: Sequence[int] = None, cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """Workgroups per chain for yamc_mc_run: times `cycles` ordinary cycles
    of the chain at each G (HIP-event kernel time), sets and returns the
    fastest with {G: trial moves/s}."""
    if groups is None:
        groups = groups_for(rep.nk, rep.sys.natom)
    rates = {}
    for g in groups:
        rep.set_chain_groups(g)
        rep.run(cycles)
        rates[g] = rep.sys.nmol * cycles / (rep.last_kernel_ms() * 1e-3)
    best = max(rates, key=rates.get)
    rep.set_chain_groups(best)
    return best, rates


def tune_groups_batch(reps: Sequence[EwaldMC], atoms_per_group: Sequence[int] = (0, 600, 800, 1000, 1200, 1500),
                      gmax: int = 8, cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """Workgroups per chain inside a batch (a batch of grid chains, one
    cooperative launch): for each A in atoms_per_group, every replica takes
    G = natom // A workgroups (1 to gmax; A = 0: one workgroup each, the
    batch's launch shape), and `cycles` cycles of the whole batch are timed
    (HIP-event kernel time, no rebuilds).  Sets and returns the fastest A
    with {A: trial moves/s}."""
    def apply(a: int) -> None:
        for r in reps:
            r.set_chain_groups(1 if a == 0 else max(1, min(gmax, r.sys.natom // a)))
    moves = sum(r.sys.nmol for r in reps) * cycles
    rates = {}
    for a in atoms_per_group:
        apply(a)
        batch_run(reps, cycles)
        rates[a] = moves / (reps[0].last_kernel_ms() * 1e-3)
    best = max(rates, key=rates.get)
    apply(best)
    return best, rates


def k_per_wave(nk: int, shape: int, groups: int = 1) -> float:
    """k vectors each wavefront sweeps per move for a launch shape and a
    workgroup count (one wave per replica for the shapes 64 and 128)."""
    waves = 1 if shape in (64, 128) else shape // 64
    return nk / float(groups * waves)


def cell_divs_for(rep: EwaldMC) -> Tuple[int, ...]:
    """Cell divisions worth timing: 0 (no cell list) and the divisions 1..4
    that give a distinct cell count of at least 3 per axis."""
    rc = float(rep.sys.derived["rmax2"]) ** 0.5
    out, seen = [0], set()
    for div in (1, 2, 3, 4):
        nc = min(40, int(div / (rc * (1.0 + 1e-9))))
        if nc >= 3 and nc not in seen:
            seen.add(nc)
            out.append(div)
    return tuple(out)


def tune_cells_batch(reps: Sequence[EwaldMC], divs: Sequence[int] = None,
                     cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """Cell division of a batch (the same for every replica): times one
    sweep per candidate with the replicas' current launch shape and keeps
    the fastest.  Returns (best div, {div: trial moves/s})."""
    if divs is None:
        divs = cell_divs_for(reps[0])
    moves = sum(r.sys.nmol for r in reps) * cycles
    rates = {}
    for div in divs:
        for r in reps:
            r.set_cell_div(div)
        _, ms = batch_run(reps, cycles)
        rates[div] = moves / (ms * 1e-3)
    best = max(rates, key=rates.get)
    for r in reps:
        r.set_cell_div(best)
    return best, rates


def tune_cells(rep: EwaldMC, divs: Sequence[int] = None, cycles: int = 1) -> Tuple[int, Dict[int, float]]:
    """The same for one chain (yamc_mc_run with its current shape and
    workgroup count)."""
    if divs is None:
        divs = cell_divs_for(rep)
    rates = {}
    for div in divs:
        rep.set_cell_div(div)
        rep.run(cycles)
        rates[div] = rep.sys.nmol * cycles / (rep.last_kernel_ms() * 1e-3)
    best = max(rates, key=rates.get)
    rep.set_cell_div(best)
    return best, rates
