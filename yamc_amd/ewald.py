"""Host-side wrapper of one replica context of libyamc_ewald.so.

Method names follow the YAMC module procedures each one replaces (cited per
method); all arithmetic runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import YAMC_FP32C, YAMC_FP64, MOVE_DTYPE, check, dptr, iptr
from .system import System


def _desc(sys: System, precision: int):
    d = sys.derived
    keep = dict(
        charge=np.ascontiguousarray(sys.charge, dtype=np.float64),
        hard_core=np.ascontiguousarray(sys.hard_core, dtype=np.float64),
        mass=np.ascontiguousarray(sys.mass, dtype=np.float64),
        molnum=np.ascontiguousarray(sys.molnum, dtype=np.int32),
        lj_type=np.ascontiguousarray(sys.lj_type, dtype=np.int32),
        first=np.ascontiguousarray(sys.mol_first, dtype=np.int32),
        last=np.ascontiguousarray(sys.mol_last, dtype=np.int32),
        lj6=np.ascontiguousarray(sys.lj6, dtype=np.float64) if sys.ntype else np.zeros(1),
        lj12=np.ascontiguousarray(sys.lj12, dtype=np.float64) if sys.ntype else np.zeros(1),
        four_eps=np.ascontiguousarray(sys.four_eps, dtype=np.float64) if sys.ntype else np.zeros(1),
        sigma6=np.ascontiguousarray(sys.sigma6, dtype=np.float64) if sys.ntype else np.zeros(1),
        tail=np.ascontiguousarray(sys.tail_counts, dtype=np.int32) if sys.ntype else np.zeros(1, np.int32),
    )
    desc = _lib.SystemDesc(
        natom=sys.natom, nmol=sys.nmol, ntype=sys.ntype, kmax=int(d["kmax"]), precision=precision, pad=0,
        box_length=sys.params.box_length, alpha=d["alpha"], rmax2=d["rmax2"], xclb=d["xclb"], kext=d["kext"],
        dr=d["dr"], d_angle=d["d_angle"], lambda_f=sys.params.lambda_f, lambda_c=sys.params.lambda_c,
        charge=dptr(keep["charge"]), hard_core_angstr=dptr(keep["hard_core"]), mass=dptr(keep["mass"]),
        molnum=iptr(keep["molnum"]), lj_type=iptr(keep["lj_type"]), mol_first=iptr(keep["first"]),
        mol_last=iptr(keep["last"]), lj6=dptr(keep["lj6"]), lj12=dptr(keep["lj12"]),
        four_eps=dptr(keep["four_eps"]), sigma6=dptr(keep["sigma6"]), tail_counts=iptr(keep["tail"]),
    )
    return desc, keep


ENERGY_NAMES = ("U_total", "U_total_ew", "U_total_lj", "uu_ew", "uu_lj6", "uu_lj12", "E_kC", "E_kLJ", "U_tail",
                "U_ext")


class EwaldMC:
    """One Monte Carlo replica on the GPU (the MC module's state, MC.f90:18-71)."""

    def __init__(self, sys: System, precision: int = YAMC_FP64):
        self.lib = _lib.load()
        self.sys = sys
        self.precision = precision
        desc, self._keep = _desc(sys, precision)
        ctx = C.c_void_p()
        check(self.lib.yamc_ctx_create(C.byref(desc), C.byref(ctx)))
        self.ctx = ctx
        nk = C.c_int32()
        check(self.lib.yamc_ctx_info(self.ctx, None, C.byref(nk)))
        self.nk = nk.value
        self.set_positions(sys.x, sys.y, sys.z)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.yamc_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def set_positions(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        check(self.lib.yamc_set_positions(self.ctx, dptr(x), dptr(y), dptr(z)))

    def positions(self):
        n = self.sys.natom
        x, y, z = np.empty(n), np.empty(n), np.empty(n)
        check(self.lib.yamc_get_positions(self.ctx, dptr(x), dptr(y), dptr(z)))
        return x, y, z

    def grid(self):
        """FourierGrid_init (FourierGrid.f90:118-189)."""
        out = [np.empty(self.nk, np.int32) for _ in range(5)]
        check(self.lib.yamc_grid(self.ctx, *(iptr(a) for a in out)))
        return dict(zip(("kx", "ky", "kz", "ip", "iq"), out))

    # -------------------------------------------------------------- k space
    def kspace_init(self, scale_beta: float = 0.0) -> Tuple[float, float]:
        """EwaldSumKSpace_init (EwaldSumKSpace.f90:197-223)."""
        ec, el = C.c_double(), C.c_double()
        check(self.lib.yamc_kspace_init(self.ctx, C.c_double(scale_beta), C.byref(ec), C.byref(el)))
        return ec.value, el.value

    def beta(self):
        """initBeta / initBetaLJ (EwaldSumKSpace.f90:263-386)."""
        T = self.sys.ntype
        b, b6, b12 = np.empty(self.nk), np.empty(self.nk), np.empty(self.nk)
        bl = np.empty(max(T * T, 1) * self.nk)
        check(self.lib.yamc_kspace_get_beta(self.ctx, dptr(b), dptr(b6), dptr(b12), dptr(bl)))
        return b, b6, b12, bl[: T * T * self.nk].reshape(T, T, self.nk)

    def sums(self) -> np.ndarray:
        """TRhoSquared sums as [(1+T), 8, nk] (RhoSquared.f90:9-45)."""
        s = np.empty((1 + self.sys.ntype) * 8 * self.nk)
        check(self.lib.yamc_kspace_get_sums(self.ctx, dptr(s)))
        return s.reshape(1 + self.sys.ntype, 8, self.nk)

    def kspace_delta(self, imol: int, x, y, z) -> Tuple[float, float]:
        """calc_dU_coulomb + calc_dU_LJ (EwaldSumKSpace.f90:528-753)."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        dc, dl = C.c_double(), C.c_double()
        check(self.lib.yamc_kspace_delta(self.ctx, int(imol), dptr(x), dptr(y), dptr(z), C.byref(dc), C.byref(dl)))
        return dc.value, dl.value

    def kspace_forces(self, first: int, x, y, z):
        """ForceKSpace_coulomb + ForceKSpace_LJ (ForceKSpace.f90:59-302)."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        n = x.shape[0]
        fx, fy, fz = np.empty(n), np.empty(n), np.empty(n)
        check(self.lib.yamc_kspace_forces(self.ctx, int(first), n, dptr(x), dptr(y), dptr(z), dptr(fx), dptr(fy),
                                          dptr(fz)))
        return fx, fy, fz

    # ----------------------------------------------------------- real space
    def rspace_full(self):
        """EwaldSumRealSpace_calc, full sum (EwaldSumRealSpace.f90:233-500)."""
        n = self.sys.natom
        comps = np.empty(6)
        uu, fx, fy, fz = (np.empty(n) for _ in range(4))
        ov = C.c_int32()
        check(self.lib.yamc_rspace_full(self.ctx, dptr(comps), dptr(uu), dptr(fx), dptr(fy), dptr(fz), C.byref(ov)))
        return comps, uu, fx, fy, fz, bool(ov.value)

    def rspace_partial(self, imol: int, x, y, z, back: bool = True):
        """EwaldSumRealSpace_set_molecule + _calc, partial sum."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        na, n = x.shape[0], self.sys.natom
        comps = np.empty(6)
        uu, fx, fy, fz = (np.empty(na) for _ in range(4))
        ub, fxb, fyb, fzb = (np.empty(n) for _ in range(4))
        ov = C.c_int32()
        bp = (dptr(ub), dptr(fxb), dptr(fyb), dptr(fzb)) if back else (None, None, None, None)
        check(self.lib.yamc_rspace_partial(self.ctx, int(imol), dptr(x), dptr(y), dptr(z), dptr(comps), dptr(uu),
                                           dptr(fx), dptr(fy), dptr(fz), *bp, C.byref(ov)))
        return dict(comps=comps, uu=uu, fx=fx, fy=fy, fz=fz, uu_back=ub, fx_back=fxb, fy_back=fyb, fz_back=fzb,
                    overlap=bool(ov.value))

    # ------------------------------------------------------------ constants
    def tails(self):
        """ewald_sum_coulomb_tail, ewald_sum_lj_tails (EwaldSumTails.f90:18-127)."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.lib.yamc_tails(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def mu(self):
        """EwaldSumExternal_calc_mu (EwaldSumExternal.f90:165-183)."""
        m = np.empty(3)
        check(self.lib.yamc_external_mu(self.ctx, dptr(m)))
        return m

    # ---------------------------------------------------------- Monte Carlo
    def mc_init(self):
        """mc_calc_total_sums + mc_total_energy (MC.f90:477-585, 681-698)."""
        e = np.empty(10)
        ov = C.c_int32()
        check(self.lib.yamc_mc_init(self.ctx, dptr(e), C.byref(ov)))
        return dict(zip(ENERGY_NAMES, e)), bool(ov.value)

    def energies(self):
        e = np.empty(10)
        check(self.lib.yamc_mc_energies(self.ctx, dptr(e)))
        return dict(zip(ENERGY_NAMES, e))

    def real_arrays(self):
        n = self.sys.natom
        uu, fx, fy, fz = (np.empty(n) for _ in range(4))
        check(self.lib.yamc_mc_real_arrays(self.ctx, dptr(uu), dptr(fx), dptr(fy), dptr(fz)))
        return uu, fx, fy, fz

    def set_rng(self, seed: int, stream: int = 0, counter: int = 0):
        check(self.lib.yamc_mc_set_rng(self.ctx, C.c_uint64(seed), C.c_uint64(stream), C.c_uint64(counter)))

    def set_rng_list(self, u: Sequence[float]) -> None:
        """Replay mode: the recorded uniforms u are drawn in order instead
        of the Philox stream (yamc_mc_set_rng_list)."""
        uu = np.ascontiguousarray(u, dtype=np.float64)
        check(self.lib.yamc_mc_set_rng_list(self.ctx, dptr(uu), C.c_int64(len(uu))))

    def rng_counter(self) -> int:
        v = C.c_uint64()
        check(self.lib.yamc_mc_get_rng(self.ctx, C.byref(v)))
        return v.value

    def run(self, ncycles: int, resync_every: int = 0, records: bool = False):
        """run_mc cycles (runmc.f90:204-268) on the device."""
        nacc = C.c_int64()
        recs = np.zeros(ncycles * self.sys.nmol, dtype=MOVE_DTYPE) if records else None
        rp = recs.ctypes.data_as(C.c_void_p) if records else None
        check(self.lib.yamc_mc_run(self.ctx, int(ncycles), int(resync_every), rp, C.byref(nacc)))
        return nacc.value, recs

    def run_const(self, imol: Sequence[int], u: Sequence[float], records: bool = True):
        """movemc per move with a given molecule and constant uniform."""
        im = np.ascontiguousarray(imol, dtype=np.int32)
        uu = np.ascontiguousarray(u, dtype=np.float64)
        nacc = C.c_int64()
        recs = np.zeros(len(im), dtype=MOVE_DTYPE) if records else None
        rp = recs.ctypes.data_as(C.c_void_p) if records else None
        check(self.lib.yamc_mc_run_const(self.ctx, len(im), iptr(im), dptr(uu), rp, C.byref(nacc)))
        return nacc.value, recs

    def resync(self):
        check(self.lib.yamc_resync(self.ctx))

    def set_frequencies(self, move: Optional[Sequence[float]] = None, xchange: Optional[Sequence[float]] = None):
        """init_freq (runmc.f90:26-66) with per-molecule move and exchange
        weights (None: the reference's 'none' file, moves 1, exchanges 0)."""
        mv = None if move is None else np.ascontiguousarray(move, dtype=np.float64)
        xc = None if xchange is None else np.ascontiguousarray(xchange, dtype=np.float64)
        for a in (mv, xc):
            if a is not None and a.shape != (self.sys.nmol,):
                raise ValueError("one weight per molecule")
        check(self.lib.yamc_mc_set_frequencies(self.ctx, None if mv is None else dptr(mv),
                                               None if xc is None else dptr(xc)))

    def xchange(self, imol1: int, imol2: int, u: Optional[float] = None):
        """mc_xchange (MC.f90:1325-1954): returns (accepted, record)."""
        acc = C.c_int32()
        rec = np.zeros(1, dtype=MOVE_DTYPE)
        up = C.byref(C.c_double(u)) if u is not None else None
        check(self.lib.yamc_mc_xchange(self.ctx, int(imol1), int(imol2), up, C.byref(acc),
                                       rec.ctypes.data_as(C.c_void_p)))
        return bool(acc.value), rec[0]

    def volume_move(self, pressure_kT_A3: float, log_max_volume_scaling: float, u: Optional[float] = None):
        """mcvol (MC.f90:701-825): returns (accepted, box length in Angstrom)."""
        acc = C.c_int32()
        L = C.c_double()
        up = C.byref(C.c_double(u)) if u is not None else None
        check(self.lib.yamc_mc_volume_move(self.ctx, C.c_double(pressure_kT_A3), C.c_double(log_max_volume_scaling),
                                           up, C.byref(acc), C.byref(L)))
        return bool(acc.value), L.value

    def run_npt(self, ncycles: int, pressure_step_multiplier: float, pressure_kT_A3: float,
                log_max_volume_scaling: float):
        """run_mc cycles with volume steps (runmc.f90:204-268): returns
        (accepted moves, volume steps, accepted volume steps)."""
        a, v, va = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.yamc_mc_run_npt(self.ctx, int(ncycles), C.c_double(pressure_step_multiplier),
                                       C.c_double(pressure_kT_A3), C.c_double(log_max_volume_scaling),
                                       C.byref(a), C.byref(v), C.byref(va)))
        return a.value, v.value, va.value

    def set_box(self, box_length: float, scale_positions: bool = True) -> bool:
        """The box change of mcvol without its acceptance test; returns the
        recompute's overlap flag."""
        ov = C.c_int32()
        check(self.lib.yamc_set_box(self.ctx, C.c_double(box_length), int(bool(scale_positions)), C.byref(ov)))
        return bool(ov.value)

    def commit_move(self, imol: int, x, y, z, du_c: float, du_lj: float) -> bool:
        """movemc's commit (MC.f90:1242-1320) of molecule imol moved to
        (x, y, z); returns False (nothing committed) if the place overlaps."""
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        ov = C.c_int32()
        check(self.lib.yamc_mc_commit_move(self.ctx, int(imol), dptr(x), dptr(y), dptr(z), C.c_double(du_c),
                                           C.c_double(du_lj), C.byref(ov)))
        return not ov.value

    def box_length(self) -> float:
        L = C.c_double()
        check(self.lib.yamc_box_length(self.ctx, C.byref(L)))
        return L.value

    def forces(self):
        """-dE/dr of the full energy and the reference force convention."""
        n = self.sys.natom
        g = [np.empty(n) for _ in range(3)]
        r = [np.empty(n) for _ in range(3)]
        check(self.lib.yamc_forces(self.ctx, *(dptr(a) for a in g), *(dptr(a) for a in r)))
        return np.stack(g, 1), np.stack(r, 1)

    PHASES = ("pick", "mu_tables", "old_sweep", "proposal", "rs_new", "new_sweep", "accept", "commit",
              "finalize")

    def phase_times(self):
        """Cumulative device microseconds per chain-kernel phase."""
        us = np.empty(9)
        check(self.lib.yamc_phase_times(self.ctx, dptr(us)))
        return dict(zip(self.PHASES, us))

    def rdf(self, mol_species, nspecies: int, dr: float = 0.1, rmax: Optional[float] = None,
            counts: Optional[np.ndarray] = None) -> np.ndarray:
        """yamc_rdf: adds this configuration's COM pair counts (mcrdf2.f90
        count_distances, cnt4) into counts[nspecies, nspecies, nbins]."""
        rmax = 0.5 * self.box_length() if rmax is None else rmax
        nbins = int((rmax - 1e-10) / dr) + 1
        if counts is None:
            counts = np.zeros((nspecies, nspecies, nbins), dtype=np.uint64)
        sp = np.ascontiguousarray(mol_species, dtype=np.int32)
        check(self.lib.yamc_rdf(self.ctx, iptr(sp), int(nspecies), C.c_double(dr), C.c_double(rmax),
                                int(counts.shape[2]), counts.ctypes.data_as(C.POINTER(C.c_uint64))))
        return counts

    def set_block_size(self, block: int) -> None:
        """Workgroup size of the chain kernel (256 or 512)."""
        check(self.lib.yamc_set_block_size(self.ctx, int(block)))

    def set_chain_groups(self, groups: int) -> None:
        """Workgroups per chain of the single-chain runs and of this chain in
        a batch (1: one workgroup; G > 1: one chain spread over G co-resident
        workgroups)."""
        check(self.lib.yamc_set_chain_groups(self.ctx, int(groups)))
        self.groups = int(groups)

    def set_cell_div(self, div: int) -> None:
        """Cells per cut-off of the real-space scan's cell list (0: none,
        the scan visits every atom; 1..4: cells of width >= rc / div)."""
        check(self.lib.yamc_set_cell_div(self.ctx, int(div)))

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        check(self.lib.yamc_last_kernel_ms(self.ctx, C.byref(v)))
        return v.value


def batch_run(reps: Sequence[EwaldMC], ncycles: int, resync_every: int = 0):
    """yamc_batch_run: all replicas advance in one persistent launch (or in
    segments of resync_every moves with batched S(k) rebuilds between)."""
    lib = _lib.load()
    arr = (C.c_void_p * len(reps))(*[r.ctx.value for r in reps])
    nacc = np.zeros(len(reps), dtype=np.int64)
    check(lib.yamc_batch_run(arr, len(reps), int(ncycles), int(resync_every),
                             nacc.ctypes.data_as(C.POINTER(C.c_int64))))
    ms = reps[0].last_kernel_ms()
    return nacc, ms


class Batch:
    """yamc_batch_create/launch/destroy: one sweep of ncycles cycles of all
    replicas captured in a HIP graph and replayed per launch()."""

    def __init__(self, reps: Sequence[EwaldMC], ncycles: int = 1, resync_every: int = 0):
        self.lib = _lib.load()
        self.reps = list(reps)
        arr = (C.c_void_p * len(self.reps))(*[r.ctx.value for r in self.reps])
        h = C.c_void_p()
        check(self.lib.yamc_batch_create(arr, len(self.reps), int(ncycles), int(resync_every), C.byref(h)))
        self.h = h

    def launch(self):
        """Replays the graph; returns (accepted per replica, launch ms)."""
        nacc = np.zeros(len(self.reps), dtype=np.int64)
        check(self.lib.yamc_batch_launch(self.h, nacc.ctypes.data_as(C.POINTER(C.c_int64))))
        return nacc, self.reps[0].last_kernel_ms()

    def close(self):
        if getattr(self, "h", None):
            self.lib.yamc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rdf_normalize(counts: np.ndarray, nframes: int, n_per_species: Sequence[int], volume: float,
                  dr: float) -> np.ndarray:
    """mcrdf2.f90:440-470 (rdf1 column) at constant volume: g_ij(r) =
    cnt / N_frame / <1/V> / N_i / N_j / dV(r) / (1 if i == j else 2), with
    dV = 4 pi dr (r^2 + r dr) + 4/3 pi dr^3 and r the bin's lower edge."""
    ns, _, nb = counts.shape
    r = np.arange(nb) * dr
    dV = 4.0 * np.pi * dr * (r ** 2 + r * dr) + 4.0 / 3.0 * np.pi * dr ** 3
    g = np.zeros(counts.shape)
    for i in range(ns):
        for j in range(i, ns):
            f = 1.0 if i == j else 2.0
            g[i, j] = counts[i, j] / nframes * volume / n_per_species[i] / n_per_species[j] / dV / f
    return g
