"""System staging for the Ewald Monte Carlo path.

Host-side restatement of the reference's set-up steps, so that a system built
here carries exactly the values the reference modules read:

* molecule records of the .mol format (Molecule.f90:142-188; hard core
  defaults to sigma*sqrt(0.4), Molecule.f90:179; centred on the mass-weighted
  centre, Molecule.f90:190-209),
* derived parameters (Parameters_recalc, parameters.f90:355-413, with the
  constants of constants.f90:4-13),
* atom staging in box units (MoleculeTable_fillAtomicData,
  MoleculeTable.f90:584-705: rotation psi about z, theta about y, phi about z),
* LJ type detection and Lorentz-Berthelot tables (LJTypes_fill,
  LJTypes.f90:134-308),
* composition counts used as LJ-type counts by the tails
  (EwaldSumTails.f90:88,113).

Coordinates are box units, box = [-0.5, 0.5)^3.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# constants.f90:5-13 (the reference's literals, not CODATA)
PI = 3.14159265358979323846
BOLTZ = 1.38065e-23
ELEC = 1.602176e-19
EPSI0 = 8.854187e-12
KCAL_MOL = 6.9477e-21


@dataclasses.dataclass
class Molecule:
    """One .mol record set: name, x, y, z (A), sigma (A), epsilon (kcal/mol),
    charge (e), mass, hard core (A)."""

    name: str
    atoms: List[Tuple[str, float, float, float, float, float, float, float, float]]

    @property
    def natoms(self) -> int:
        return len(self.atoms)

    def arrays(self):
        a = np.array([r[1:] for r in self.atoms], dtype=np.float64).reshape(-1, 8)
        x, y, z, sig, eps, q, m, hc = (a[:, i].copy() for i in range(8))
        hc = np.where(hc < 1e-9, sig * math.sqrt(0.4), hc)  # Molecule.f90:179
        tm = m.sum()
        if tm > 0:  # Molecule_centrate (Molecule.f90:190-209)
            cx, cy, cz = (m * x).sum() / tm, (m * y).sum() / tm, (m * z).sum() / tm
            sx = sy = sz = 0.0
            for i in range(len(m)):  # sequential sums, as center_of_mass does
                sx += m[i] * x[i]
                sy += m[i] * y[i]
                sz += m[i] * z[i]
            cx, cy, cz = sx / tm, sy / tm, sz / tm
            x, y, z = x - cx, y - cy, z - cz
        return x, y, z, sig, eps, q, m, hc


# SPC/E water and Dang ion parameters (the reference's SPCE.mol, Na_Dang.mol,
# Cl_Dang.mol records).
SPCE = Molecule("SPCE", [
    ("O", 0.0, 0.0, 0.0, 3.1655, 0.1554, -0.84760, 16.0, 2.0),
    ("H", 1.0, 0.0, 0.0, 0.0, 0.0, 0.42380, 0.0, 0.0),
    ("H", -0.325568154457156, 0.945518575599317, 0.0, 0.0, 0.0, 0.42380, 0.0, 0.0),
])
NA_DANG = Molecule("Na", [("Na", 0.0, 0.0, 0.0, 2.592, 0.10361, 1.0, 1.0, 2.0)])
CL_DANG = Molecule("Cl", [("Cl", 0.0, 0.0, 0.0, 4.41, 0.10361, -1.0, 1.0, 2.0)])


@dataclasses.dataclass
class Params:
    """Inputs of Parameters_recalc (parameters.f90:355-413)."""

    box_length: float
    temp: float = 298.15
    sr: float = 4.0
    sk: float = 4.0
    alpha: float = 0.0          # alpha * L; 0 -> 2*sr (rmax = L/2)
    external_permutivity: float = 50.0
    dr_a: float = 0.4
    d_angle_degree: float = 40.0
    lambda_f: float = 0.5
    lambda_c: float = 0.5

    def derived(self) -> Dict[str, float]:
        alpha = self.alpha if self.alpha >= 1e-9 else 2.0 * self.sr
        rmax = self.sr / alpha
        kmax = int(_nint(self.sk * alpha / PI))
        kT_kcal_mol = BOLTZ * self.temp / KCAL_MOL
        dbjr_a = ELEC ** 2 / (4.0 * PI * EPSI0 * BOLTZ * self.temp) / 1.0e-10
        xclb = dbjr_a / self.box_length
        return dict(
            alpha=alpha, rmax=rmax, rmax2=rmax ** 2, kmax=kmax, kT_kcal_mol=kT_kcal_mol,
            xclb=xclb, dr=self.dr_a / self.box_length, d_angle=self.d_angle_degree * PI / 180.0,
            kext=2.0 * xclb * PI / (2.0 * self.external_permutivity + 1.0),
        )


def _nint(v: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


@dataclasses.dataclass
class System:
    """A staged system: everything yamc_system_desc carries plus positions."""

    params: Params
    derived: Dict[str, float]
    species: List[Molecule]
    counts: List[int]
    mol_type: np.ndarray      # [nmol] species index
    mol_pos: np.ndarray       # [nmol, 3] box units
    mol_angles: np.ndarray    # [nmol, 3] theta, phi, psi
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    charge: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray
    hard_core: np.ndarray
    mass: np.ndarray
    molnum: np.ndarray        # 1-based
    mol_first: np.ndarray     # 0-based
    mol_last: np.ndarray
    lj_type: np.ndarray       # 0 = none
    ntype: int
    lj6: np.ndarray
    lj12: np.ndarray
    four_eps: np.ndarray
    sigma6: np.ndarray
    sigma2: np.ndarray
    tail_counts: np.ndarray

    @property
    def natom(self) -> int:
        return int(self.x.shape[0])

    @property
    def nmol(self) -> int:
        return int(self.mol_first.shape[0])

    @property
    def kmax(self) -> int:
        return int(self.derived["kmax"])

    def positions(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        return self.x.copy(), self.y.copy(), self.z.copy()


def _rot(x, y, z, cc, ss, axis):
    """geometry.f90:28-63 rot_vect."""
    if axis == 2:
        return cc * x + ss * z, y, -ss * x + cc * z
    return cc * x - ss * y, ss * x + cc * y, z  # axis 3


def lj_types_fill(sigma: np.ndarray, epsilon: np.ndarray):
    """LJTypes.f90:134-308: type ids by (sigma, epsilon) within 1e-8 (sigma < 1e-8
    has no LJ), then Lorentz-Berthelot tables indexed (t1-1)*NType + (t2-1)."""
    st: List[float] = []
    et: List[float] = []
    for s, e in zip(sigma, epsilon):
        if s < 1e-8:
            continue
        if not any(abs(s - a) < 1e-8 and abs(e - b) < 1e-8 for a, b in zip(st, et)):
            st.append(float(s))
            et.append(float(e))
    typ = np.zeros(len(sigma), dtype=np.int32)
    for i, (s, e) in enumerate(zip(sigma, epsilon)):
        for t, (a, b) in enumerate(zip(st, et)):
            if abs(s - a) < 1e-8 and abs(e - b) < 1e-8:
                typ[i] = t + 1
                break
    T = len(st)
    lj6 = np.zeros(T * T); lj12 = np.zeros(T * T); fe = np.zeros(T * T)
    s6 = np.zeros(T * T); s2 = np.zeros(T * T)
    for i in range(T):
        for j in range(i, T):
            if i == j:
                sig12, eps12 = st[i], et[i]
            else:
                sig12, eps12 = (st[i] + st[j]) / 2, math.sqrt(et[i] * et[j])
            p2 = sig12 ** 2
            p6 = p2 ** 3
            v6 = 4 * eps12 * p6
            for o in {i * T + j, j * T + i}:
                lj6[o] = v6
                lj12[o] = v6 * p6
                s2[o] = p2
                s6[o] = p6
                fe[o] = 4 * eps12
    return T, typ, lj6, lj12, fe, s6, s2


def build_system(params: Params, species: Sequence[Molecule], counts: Sequence[int], mol_type: np.ndarray,
                 mol_pos: np.ndarray, mol_angles: np.ndarray) -> System:
    """Stage atoms from a molecule table (MoleculeTable_fillAtomicData)."""
    der = params.derived()
    L = params.box_length
    kT = der["kT_kcal_mol"]
    xs, ys, zs, qs, sg, ep, hc, ms, mn, first, last = [], [], [], [], [], [], [], [], [], [], []
    sp_arrays = [m.arrays() for m in species]
    n = 0
    for imol in range(len(mol_type)):
        mx, my, mz, msig, meps, mq, mm, mhc = sp_arrays[int(mol_type[imol])]
        theta, phi, psi = (float(v) for v in mol_angles[imol])
        ct, st_, cp, sp, cs, ss = math.cos(theta), math.sin(theta), math.cos(phi), math.sin(phi), math.cos(psi), math.sin(psi)
        first.append(n)
        for a in range(len(mx)):
            x1, y1, z1 = _rot(mx[a], my[a], mz[a], cs, ss, 3)
            x1, y1, z1 = _rot(x1, y1, z1, ct, st_, 2)
            x1, y1, z1 = _rot(x1, y1, z1, cp, sp, 3)
            xs.append(x1 / L + mol_pos[imol, 0])
            ys.append(y1 / L + mol_pos[imol, 1])
            zs.append(z1 / L + mol_pos[imol, 2])
            qs.append(mq[a]); sg.append(msig[a] / L); hc.append(mhc[a]); ep.append(meps[a] / kT)
            ms.append(mm[a]); mn.append(imol + 1)
            n += 1
        last.append(n - 1)
    sigma = np.array(sg); epsilon = np.array(ep)
    T, typ, lj6, lj12, fe, s6, s2 = lj_types_fill(sigma, epsilon)
    tail_counts = np.array([counts[t] if t < len(counts) else 0 for t in range(T)], dtype=np.int32)
    return System(
        params=params, derived=der, species=list(species), counts=list(counts),
        mol_type=np.asarray(mol_type, dtype=np.int32), mol_pos=np.asarray(mol_pos, dtype=np.float64),
        mol_angles=np.asarray(mol_angles, dtype=np.float64),
        x=np.array(xs), y=np.array(ys), z=np.array(zs), charge=np.array(qs), sigma=sigma, epsilon=epsilon,
        hard_core=np.array(hc), mass=np.array(ms), molnum=np.array(mn, dtype=np.int32),
        mol_first=np.array(first, dtype=np.int32), mol_last=np.array(last, dtype=np.int32),
        lj_type=typ, ntype=T, lj6=lj6, lj12=lj12, four_eps=fe, sigma6=s6, sigma2=s2, tail_counts=tail_counts,
    )


def lattice_system(params: Params, species: Sequence[Molecule], counts: Sequence[int], seed: int,
                   shuffle_species: bool = True) -> System:
    """Molecules on a simple-cubic lattice with seeded orientations.

    Sites (i+0.5)/n - 0.5 per axis, n = ceil(nmol^(1/3)); the first nmol sites
    are used.  Species are placed on seeded random sites (shuffle_species) or
    in composition order.  Orientations follow the reference's random
    orientation rule (theta = pi u, phi = psi = 2 pi u).
    """
    rng = np.random.default_rng(seed)
    nmol = int(sum(counts))
    ns = 1
    while ns ** 3 < nmol:
        ns += 1
    sites = np.array([[(i + 0.5) / ns - 0.5, (j + 0.5) / ns - 0.5, (k + 0.5) / ns - 0.5]
                      for i in range(ns) for j in range(ns) for k in range(ns)])[:nmol]
    mol_type = np.concatenate([np.full(c, t, dtype=np.int32) for t, c in enumerate(counts)])
    if shuffle_species:
        perm = rng.permutation(nmol)
        pos = sites[perm]
        # keep the molecule table in composition order (species blocks), sites shuffled
    else:
        pos = sites
    u = rng.random((nmol, 3))
    angles = np.stack([PI * u[:, 0], 2 * PI * u[:, 1], 2 * PI * u[:, 2]], axis=1)
    return build_system(params, species, counts, mol_type, pos, angles)


# --------------------------------------------------------------------------
# Workloads (box sizes, cut-offs and k limits of the benchmark configurations)

def config(name: str, seed: int = 20261203, precision_scale: float = 1.0) -> System:
    """Named workloads.

    c1: 256 SPC/E, L=19.72 A, rc=9 A (sr 3.2, alphaL 7.0116), kmax=7 (sk 3.1364),
        T=298 K, dr 0.3 A, 15 deg, no force bias.
    c2: 2048 SPC/E + 40 Na + 40 Cl, L=39.4 A, rc=12 A (alphaL 10.5067), kmax=12
        (sk 3.5881), force bias lambda 0.5.
    c3: 8192 SPC/E + 160 Na + 160 Cl, L=62.6 A, rc=12 A (alphaL 16.6933), kmax=16.
    c5: 32768 SPC/E + 600 Na + 600 Cl, L=99.4 A, rc=12 A (alphaL 26.5067), kmax=24.
    small: 8 SPC/E + 1 Na + 1 Cl in L=12 A (unit tests).
    cells: 512 SPC/E + 5 Na + 5 Cl, L=25.0 A, rc=6 A (4 cut-off cells per axis,
        the cell-list scan), kmax=13, force bias 0.5 (unit tests).
    """
    if name == "c1":
        p = Params(box_length=19.72, temp=298.0, sr=3.2, sk=3.1364, alpha=7.0116, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.0, lambda_c=0.0)
        return lattice_system(p, [SPCE], [256], seed, shuffle_species=False)
    if name == "c2":
        p = Params(box_length=39.4, temp=298.0, sr=3.2, sk=3.5881, alpha=10.5067, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5)
        return lattice_system(p, [SPCE, NA_DANG, CL_DANG], [2048, 40, 40], seed)
    if name == "c3":
        p = Params(box_length=62.6, temp=298.0, sr=3.2, sk=3.0111, alpha=16.6933, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5)
        return lattice_system(p, [SPCE, NA_DANG, CL_DANG], [8192, 160, 160], seed)
    if name == "c5":
        p = Params(box_length=99.4, temp=298.0, sr=3.2, sk=2.8445, alpha=26.5067, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5)
        return lattice_system(p, [SPCE, NA_DANG, CL_DANG], [32768, 600, 600], seed)
    if name == "small":
        p = Params(box_length=12.0, temp=298.0, sr=3.2, sk=3.0, alpha=6.4, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5)
        return lattice_system(p, [SPCE, NA_DANG, CL_DANG], [8, 1, 1], seed)
    if name == "small_nobias":
        p = Params(box_length=12.0, temp=298.0, sr=3.2, sk=3.0, alpha=6.4, external_permutivity=50.0,
                   dr_a=0.3, d_angle_degree=15.0, lambda_f=0.0, lambda_c=0.0)
        return lattice_system(p, [SPCE, NA_DANG, CL_DANG], [8, 1, 1], seed)
    if name == "cells":
        return replica_system(512, 0.5, 300.0, seed, rc_angstrom=6.0, alpha_per_angstrom=0.5333, sk=3.0)
    raise ValueError(f"unknown workload {name!r}")


def replica_system(n_water: int, conc_molar: float, temp: float, seed: int, rc_angstrom: float = 12.0,
                   alpha_per_angstrom: float = 0.2667, sk: float = 3.2) -> System:
    """One uneven replica: density 0.0334 molecules/A^3, NaCl at conc_molar
    (pairs = round(conc * N_w / 55.5)), fixed alpha in 1/A so alphaL = alpha*L,
    kmax = nint(sk*alphaL/pi) (parameters.f90:388), sr = alpha*rc capped at L/2."""
    n_pair = int(round(conc_molar * n_water / 55.5))
    nmol = n_water + 2 * n_pair
    L = (nmol / 0.0334) ** (1.0 / 3.0)
    alphaL = alpha_per_angstrom * L
    sr = min(alpha_per_angstrom * rc_angstrom, alphaL / 2.0)
    p = Params(box_length=L, temp=temp, sr=sr, sk=sk, alpha=alphaL, external_permutivity=50.0, dr_a=0.3,
               d_angle_degree=15.0, lambda_f=0.5, lambda_c=0.5)
    counts = [n_water, n_pair, n_pair] if n_pair > 0 else [n_water]
    species = [SPCE, NA_DANG, CL_DANG] if n_pair > 0 else [SPCE]
    return lattice_system(p, species, counts, seed)


def write_host_case(sysm: System, path: str, precision: int = 0, seed: int = 1, stream: int = 0,
                    ncycles: int = 1) -> None:
    """Writes the text case file read by the Fortran host fortran/mc_host
    (layout documented at the top of fortran/mc_host.f90)."""
    d = sysm.derived
    T = sysm.ntype
    with open(path, "w") as f:
        f.write(f"{sysm.natom} {sysm.nmol} {T} {int(d['kmax'])} {precision}\n")
        f.write(" ".join("%.17g" % v for v in (sysm.params.box_length, d["alpha"], d["rmax2"], d["xclb"],
                                              d["kext"], d["dr"], d["d_angle"], sysm.params.lambda_f,
                                              sysm.params.lambda_c)) + "\n")
        for i in range(sysm.natom):
            f.write("%.17g %.17g %.17g %d %d %.17g %.17g %.17g\n" % (
                sysm.charge[i], sysm.hard_core[i], sysm.mass[i], sysm.molnum[i], sysm.lj_type[i], sysm.x[i],
                sysm.y[i], sysm.z[i]))
        for m in range(sysm.nmol):
            f.write(f"{sysm.mol_first[m]} {sysm.mol_last[m]}\n")
        for k in range(T * T):
            f.write("%.17g %.17g %.17g %.17g\n" % (sysm.lj6[k], sysm.lj12[k], sysm.four_eps[k], sysm.sigma6[k]))
        for t in range(T):
            f.write(f"{sysm.tail_counts[t]}\n")
        f.write(f"{seed} {stream} {ncycles}\n")
