"""The reference's on-disk formats around the Monte Carlo path.

* parameters file (parameters.f90:85-305): ``name = value  # comment`` lines,
  the reference's defaults, BoxLength from the density when it is 0;
* composition file (composition.f90:84-142) and ``.mol`` molecule files
  (Molecule.f90:142-209);
* frequency file (runmc.f90:69-176): per-molecule move and exchange weights;
* binary molecule tables / trajectory frames (MoleculeTable.f90:170-319):
  per molecule x, y, z as ``nint((x + 0.5) * 256**nb)`` and theta, phi, psi as
  ``nint(angle / 2pi * 256**nb)``, little-endian, nb = 1..3 bytes per sample;
* 24+8-bit floats (FloatingPoint.f90:18-247) of the force files written by
  mc_calc_forces (mc_calc_forces.f90:348-368);
* molecule positions and ZYZ orientations from atom positions
  (MoleculeTable_calcPositionsOrientations, MoleculeTable.f90:707-769);
* run output: energy.dat, boxlength.dat, frames.dat (runmc.f90:336-417) and
  the RDF tables of mcrdf2 (two columns r, g(r)).

All of it is host I/O; nothing here touches the device.
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import system as ys

TWO_PI = 2.0 * math.pi
LN2 = 0.6931471805599453  # constants.f90:15

# parameters.f90:17-80 defaults (pressure has none in the reference)
DEFAULTS = {
    "sr": 4.0, "sk": 4.0, "alphaL": 0.0, "temp": 298.15, "external_permutivity": 50.0, "BoxLength": 0.0,
    "density": 33.33, "dr_a": 0.4, "d_angle_degree": 40.0, "xlambda_c": 0.5, "xlambda_f": 0.5,
    "random_seed": 0, "pressure": 0.0, "pressure_step_multiplier": 0.0, "max_volume_scaling": 1.05,
    "max_cycle": 1000000, "n_store_traj_interval": 100.0, "n_store_energy_interval": 100.0,
    "input_nbytes_xyz": 2, "input_nbytes_ang": 2, "output_nbytes_xyz": 2, "output_nbytes_ang": 2,
    "traj_file": "traj.moltab", "energy_file": "energy.dat", "boxlength_file": "boxlength.dat",
    "frames_file": "frames.dat", "param_out_file": "parameters.out", "freq_file": "none",
}
_INTS = {"random_seed", "max_cycle", "input_nbytes_xyz", "input_nbytes_ang", "output_nbytes_xyz",
         "output_nbytes_ang"}
_STRS = {"traj_file", "energy_file", "boxlength_file", "frames_file", "param_out_file", "freq_file"}


def read_parameters(path: str) -> Dict[str, object]:
    """Parameters_read (parameters.f90:266-305): '=' is a blank, the first two
    words of a line are name and value, '#' starts a comment line; unknown
    names are ignored (read_parameter)."""
    out = dict(DEFAULTS)
    with open(path) as f:
        for line in f:
            words = line.replace("=", " ").split()
            if len(words) < 2 or words[0].startswith("#"):
                continue
            nam, val = words[0], words[1]
            if nam not in DEFAULTS:
                continue
            if nam in _STRS:
                out[nam] = val
            elif nam in _INTS:
                out[nam] = int(val)
            else:
                out[nam] = float(val.replace("d", "e").replace("D", "e"))
    return out


def box_length_from_density(density: float, nmol: int) -> float:
    """parameters.f90:307-320: (nmol / density [nm^-3])^(1/3) * 10 A."""
    return (nmol / density) ** (1.0 / 3.0) * 10.0


def params_from(prm: Dict[str, object], nmol: int) -> ys.Params:
    """Parameters_init (parameters.f90:330-350) + the Params of the staging."""
    L = float(prm["BoxLength"])
    if L == 0.0:
        if abs(float(prm["density"])) < 1e-9:
            raise ValueError("both BoxLength and density are zero")
        L = box_length_from_density(float(prm["density"]), nmol)
    return ys.Params(box_length=L, temp=float(prm["temp"]), sr=float(prm["sr"]), sk=float(prm["sk"]),
                     alpha=float(prm["alphaL"]), external_permutivity=float(prm["external_permutivity"]),
                     dr_a=float(prm["dr_a"]), d_angle_degree=float(prm["d_angle_degree"]),
                     lambda_f=float(prm["xlambda_f"]), lambda_c=float(prm["xlambda_c"]))


def read_mol(path: str) -> ys.Molecule:
    """A .mol file (Molecule.f90:142-209): one atom per line, name x y z [A]
    sigma [A] epsilon [kcal/mol] charge [e] mass [hard core [A]]."""
    atoms = []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w or w[0].startswith("#"):
                continue
            v = [float(t) for t in w[1:]]
            if len(v) < 7:
                raise ValueError(f"{path}: atom line needs x y z sigma epsilon charge mass")
            hc = v[7] if len(v) > 7 else 0.0
            atoms.append((w[0], v[0], v[1], v[2], v[3], v[4], v[5], v[6], hc))
    name = os.path.splitext(os.path.basename(path))[0]
    return ys.Molecule(name, atoms)


def read_composition(path: str, mol_dir: Optional[str] = None) -> Tuple[List[ys.Molecule], List[int]]:
    """composition.f90:84-142: the number of species, then 'file.mol count'
    per species; .mol files are looked up next to the composition file."""
    base = mol_dir or os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    n = int(lines[0][0])
    species, counts = [], []
    for w in lines[1:1 + n]:
        species.append(read_mol(os.path.join(base, w[0])))
        counts.append(int(w[1]))
    return species, counts


def read_freq_file(path: str, nmol: int) -> Tuple[np.ndarray, np.ndarray]:
    """read_freq_file (runmc.f90:69-176): the number of records, then lines
    'first-last move_weight xchange_weight' (1-based, inclusive); every
    molecule must be covered exactly once."""
    with open(path) as f:
        lines = [ln for ln in f if ln.strip() and "#" not in ln.split()[0]]
    nrec = int(lines[0].split()[0])
    mv = np.full(nmol, -1.0)
    xc = np.full(nmol, -1.0)
    for k in range(nrec):
        w = lines[1 + k].replace("-", " ", 1).split()
        first, last, pm, px = int(w[0]), int(w[1]), float(w[2]), float(w[3])
        if last > nmol:
            raise ValueError(f"{path} is inconsistent with the composition")
        for j in range(first - 1, last):
            if mv[j] > 0:
                raise ValueError(f"{path}: crossing intervals at record {k + 1}")
            mv[j] = pm
            xc[j] = px
    if np.any(mv < 0):
        raise ValueError(f"{path}: not all frequencies defined")
    return mv, xc


# --- binary molecule tables ---------------------------------------------------

def _le_bytes(vals: np.ndarray, n: int) -> np.ndarray:
    """write_little_endian (io.f90:187-219): n bytes of each value, byte k =
    mod(v, 256) (sign of v, stored as its low 8 bits) then v = v / 256
    truncated toward zero, as the reference's integer arithmetic does."""
    v = np.asarray(vals, dtype=np.int64).copy()
    out = []
    for _ in range(n):
        r = np.fmod(v, 256)
        out.append((r & 255).astype(np.uint8))
        v = np.sign(v) * (np.abs(v) // 256)
    return np.stack(out, axis=-1)


def _le_values(raw: np.ndarray, n: int) -> np.ndarray:
    """read_little_endian (io.f90:221-262): unsigned little-endian n-byte values."""
    r = raw.astype(np.int64)
    return sum(r[..., b] << (8 * b) for b in range(n))


def _nint(v: np.ndarray) -> np.ndarray:
    """Fortran nint: nearest integer, halves away from zero."""
    return (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)


def moltab_encode(pos: np.ndarray, angles: np.ndarray, nbytes_xyz: int = 2, nbytes_ang: int = 2) -> bytes:
    """MoleculeTable_save_binary (MoleculeTable.f90:170-214): pos [nmol,3] in
    box units, angles [nmol,3] = theta, phi, psi."""
    if not (1 <= nbytes_xyz <= 3 and 1 <= nbytes_ang <= 3):
        raise ValueError("1, 2 or 3 bytes per sample")
    px, pa = 256 ** nbytes_xyz, 256 ** nbytes_ang
    xi = _nint((np.asarray(pos, np.float64) + 0.5) * px)
    ai = _nint(np.asarray(angles, np.float64) / TWO_PI * pa)
    rec = np.concatenate([_le_bytes(xi, nbytes_xyz).reshape(len(xi), -1),
                          _le_bytes(ai, nbytes_ang).reshape(len(ai), -1)], axis=1)
    return rec.tobytes()


def moltab_decode(buf: bytes, nmol: int, nbytes_xyz: int = 2, nbytes_ang: int = 2, offset: int = 0):
    """MoleculeTable_load_binary (MoleculeTable.f90:243-319): returns pos
    [nmol,3] and angles [nmol,3]."""
    rec = 3 * (nbytes_xyz + nbytes_ang)
    raw = np.frombuffer(buf, dtype=np.uint8, count=nmol * rec, offset=offset).reshape(nmol, rec)
    px, pa = float(256 ** nbytes_xyz), float(256 ** nbytes_ang)
    xi = _le_values(raw[:, :3 * nbytes_xyz].reshape(nmol, 3, nbytes_xyz), nbytes_xyz)
    ai = _le_values(raw[:, 3 * nbytes_xyz:].reshape(nmol, 3, nbytes_ang), nbytes_ang)
    return xi / px - 0.5, ai / pa * TWO_PI


# --- 24+8-bit floats ----------------------------------------------------------

_TINY = float(np.float32(1e-14))  # default-real literals of mantisse_exponent_real
_HUGE = float(np.float32(1e14))


def _float_bits(a: float, mw: int, ew: int) -> Tuple[int, int]:
    """mantisse_exponent_real + mantisse_exponent_bits (FloatingPoint.f90:18-96)."""
    if abs(a) < _TINY:
        m, n = math.copysign(1.0, a), -1000
    elif abs(a) > _HUGE:
        m, n = math.copysign(1.0, a), 1000
    else:
        t = math.log(abs(a)) / LN2
        n = int(t) if t > 0 else int(t - 1.0 + 1e-14)  # myint: aint rounds toward zero
        m = a / 2.0 ** n
    m_scale, n_scale = 2 ** (mw - 1), 2 ** (ew - 1)
    sgn_m = 0 if m > 0 else 1
    sgn_n = 0 if n > 0 else 1
    if abs(n) >= n_scale:
        return sgn_m * m_scale, sgn_n * n_scale + n_scale - 1
    mm = sgn_m * m_scale + int((abs(m) - 1.0) * (m_scale - 1))
    return mm, sgn_n * n_scale + abs(n)


def float_encode(values: Sequence[float], mantisse_width: int = 24, exp_width: int = 8) -> bytes:
    """write_float (FloatingPoint.f90:166-205) for each value."""
    mb = int((mantisse_width + 0.999) / 8)
    eb = int((exp_width + 0.999) / 8)
    out = bytearray()
    for a in values:
        mm, nn = _float_bits(float(a), mantisse_width, exp_width)
        out += _le_bytes(np.array([mm]), mb).tobytes() + _le_bytes(np.array([nn]), eb).tobytes()
    return bytes(out)


def float_decode(buf: bytes, count: int, mantisse_width: int = 24, exp_width: int = 8) -> np.ndarray:
    """read_float + construct_float (FloatingPoint.f90:98-163, 207-247)."""
    mb = int((mantisse_width + 0.999) / 8)
    eb = int((exp_width + 0.999) / 8)
    raw = np.frombuffer(buf, dtype=np.uint8, count=count * (mb + eb)).reshape(count, mb + eb)
    mms = _le_values(raw[:, :mb], mb)
    nns = _le_values(raw[:, mb:], eb)
    m_scale, n_scale = 2 ** (mantisse_width - 1), 2 ** (exp_width - 1)
    out = np.empty(count)
    for i, (mm, nn) in enumerate(zip(mms.tolist(), nns.tolist())):
        mm1, sign_m = (mm - m_scale, -1.0) if mm >= m_scale else (mm, 1.0)
        nn1, sign_n = (nn - n_scale, -1) if nn >= n_scale else (nn, 1)
        if nn1 == n_scale - 1:
            out[i] = 0.0 if sign_n == -1 else 1e14
            continue
        m = mm1 / (m_scale - 1.0) + 1.0
        out[i] = sign_m * m * 2.0 ** (sign_n * nn1)
    return out


# --- molecule positions and orientations ----------------------------------------

def _zyz_angles(R: np.ndarray) -> Tuple[float, float, float]:
    """theta, phi, psi of R = Rz(phi) Ry(theta) Rz(psi), the staging rotation
    (psi about z, theta about y, phi about z; geometry.f90:28-63 sign
    convention), with theta in [0, pi] and phi, psi in [0, 2pi)."""
    ct = min(1.0, max(-1.0, R[2, 2]))
    theta = math.acos(ct)
    if math.sin(theta) > 1e-12:
        phi = math.atan2(R[1, 2], R[0, 2])
        psi = math.atan2(R[2, 1], -R[2, 0])
    else:  # rotation about z only: put it all into phi
        phi = math.atan2(R[1, 0], R[0, 0]) if ct > 0 else math.atan2(-R[1, 0], -R[0, 0])
        psi = 0.0
    return theta, phi % TWO_PI, psi % TWO_PI


def _frame(v: np.ndarray) -> Optional[np.ndarray]:
    """Two non-collinear atom vectors and their cross product as columns (the
    three-point frame of xyz_to_angles, geometry.f90:352-459)."""
    nz = [i for i in range(len(v)) if np.abs(v[i]).sum() > 1e-8]
    if not nz:
        return None
    i1 = nz[0]
    for i in nz[1:]:
        c = np.cross(v[i1], v[i])
        if np.abs(c).sum() > 1e-8:
            return np.stack([v[i1], v[i], c], axis=1)
    return np.stack([v[i1]], axis=1)


def positions_orientations(sysm: ys.System, x: np.ndarray, y: np.ndarray, z: np.ndarray,
                           box_length: float) -> Tuple[np.ndarray, np.ndarray]:
    """MoleculeTable_calcPositionsOrientations (MoleculeTable.f90:707-769):
    centre of mass of each molecule (box units) and the ZYZ angles that carry
    the species template onto its atoms."""
    nmol = sysm.nmol
    pos = np.zeros((nmol, 3))
    ang = np.zeros((nmol, 3))
    tmpl = []
    for mol in sysm.species:
        a = mol.arrays()
        tmpl.append(np.stack([a[0], a[1], a[2]], axis=1))  # centred template (A)
    for i in range(nmol):
        f, l = int(sysm.mol_first[i]), int(sysm.mol_last[i])
        m = sysm.mass[f:l + 1]
        xs, ysl, zs = x[f:l + 1], y[f:l + 1], z[f:l + 1]
        tm = 0.0
        c = [0.0, 0.0, 0.0]
        for k in range(l - f + 1):  # geometry.f90:170-201
            c[0] += m[k] * xs[k]; c[1] += m[k] * ysl[k]; c[2] += m[k] * zs[k]
            tm += m[k]
        c = [v / tm for v in c]
        pos[i] = c
        cur = np.stack([(xs - c[0]) * box_length, (ysl - c[1]) * box_length, (zs - c[2]) * box_length], axis=1)
        T = tmpl[int(sysm.mol_type[i])]
        X = _frame(T)
        if X is None:
            continue
        Y = _frame(cur)
        if X.shape[1] == 3:
            R = Y @ np.linalg.inv(X)
            ang[i] = _zyz_angles(R)
        else:  # linear molecule: psi = 0, theta/phi carry the axis direction
            u, w = X[:, 0] / np.linalg.norm(X[:, 0]), Y[:, 0] / np.linalg.norm(Y[:, 0])
            # rotate u to the z axis first is not needed for templates along z; general case via axis-angle
            ax = np.cross(u, w)
            s, co = np.linalg.norm(ax), float(np.dot(u, w))
            if s < 1e-12:
                R = np.eye(3) if co > 0 else np.diag([1.0, -1.0, -1.0])
            else:
                k = ax / s
                K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
                R = np.eye(3) + s * K + (1 - co) * K @ K
            ang[i] = _zyz_angles(R)
    return pos, ang


# --- run output -----------------------------------------------------------------

def _ffmt(v: float) -> str:
    return "%26.17G" % v


class RunFiles:
    """Appends run_mc's output records (runmc.f90:336-417): energy.dat rows
    'cycle U_total U_total_ew U_total_lj uu_ew uu_lj6 uu_lj12 E_kC E_kLJ' (kT),
    boxlength.dat rows 'cycle L', frames.dat rows 'frame L n_mv n_mv_acc
    n_vol n_vol_acc n_xchg n_xchg_acc time', and moltab trajectory frames."""

    def __init__(self, prm: Dict[str, object], out_dir: str = "."):
        self.dir = out_dir
        self.prm = prm
        for k in ("energy_file", "boxlength_file", "frames_file", "traj_file"):
            p = self.path(k)
            if os.path.exists(p):
                os.remove(p)

    def path(self, key: str) -> str:
        return os.path.join(self.dir, str(self.prm[key]))

    def energy(self, cycle: int, e: Dict[str, float]) -> None:
        vals = [e["U_total"], e["U_total_ew"], e["U_total_lj"], e["uu_ew"], e["uu_lj6"], e["uu_lj12"], e["E_kC"],
                e["E_kLJ"]]
        with open(self.path("energy_file"), "a") as f:
            f.write("%12d" % cycle + "".join(_ffmt(v) for v in vals) + "\n")

    def boxlength(self, cycle: int, L: float) -> None:
        with open(self.path("boxlength_file"), "a") as f:
            f.write("%12d%26.17G\n" % (cycle, L))

    def frame(self, frame: int, L: float, counts: Sequence[float], pos: np.ndarray, angles: np.ndarray,
              wall: int) -> None:
        with open(self.path("traj_file"), "ab") as f:
            f.write(moltab_encode(pos, angles, int(self.prm["output_nbytes_xyz"]), int(self.prm["output_nbytes_ang"])))
        with open(self.path("frames_file"), "a") as f:
            f.write("%12d%15.10f" % (frame, L) + "".join("%12.0f" % float(c) for c in counts) + "%12d\n" % wall)


def read_table(path: str) -> np.ndarray:
    """Whitespace-separated numeric rows (energy.dat, boxlength.dat, RDF files)."""
    return np.loadtxt(path, ndmin=2)
