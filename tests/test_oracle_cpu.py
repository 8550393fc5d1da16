"""CPU checks that pin the FP64 oracle (no GPU).

Each check is independent of the oracle's code: published known-answer
vectors, closed forms, direct sums over the full k sphere computed with numpy,
invariants (dE = E_after - E_before, incremental state = from-scratch state,
Ewald energy independent of the splitting parameter, Madelung constant), and
brute-force enumeration.  The bit-level anchor is the compiled reference
Fortran (test_golden_reference.py).
"""
import math

import numpy as np
import pytest
from scipy import special

from yamc_amd import system as ys


def test_philox_known_answers(oracle_mod):
    # Random123 kat_vectors for philox4x32-10
    assert oracle_mod.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle_mod.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle_mod.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("kmax,nk", [(7, 238), (10, 647), (12, 1068), (16, 2445), (24, 7912)])
def test_grid_enumeration(oracle_mod, kmax, nk):
    g = oracle_mod.grid(kmax)
    assert len(g["kx"]) == nk
    # brute force: every k with kx,ky,kz >= 0, 0 < |k| and integer kz <= floor(sqrt(kmax^2-kx^2-ky^2))
    want = [(a, b, c) for a in range(kmax + 1) for b in range(int(math.isqrt(kmax * kmax - a * a)) + 1)
            for c in range(int(math.isqrt(kmax * kmax - a * a - b * b)) + 1) if a * a + b * b + c * c > 0]
    got = list(zip(g["kx"], g["ky"], g["kz"]))
    assert got == want
    iq = np.where(g["ky"] == 0, np.where(g["kz"] == 0, 1, 3), np.where(g["kz"] == 0, 2, 4))
    assert np.array_equal(iq, g["iq"])
    assert g["ip"][-1] == 3


def test_erfc_luc_accuracy(oracle_mod):
    lib = oracle_mod.load()
    for x in np.linspace(0.01, 6.0, 400):
        e = math.exp(-x * x)
        v = lib.yo_erfc_luc(x, x * x, e)
        w = lib.yo_erf_luc(x, x * x, e)
        # Abramowitz-Stegun 7.1.26: |eps| <= 1.5e-7; Taylor / asymptotic branches are tighter
        assert abs(v - math.erfc(x)) < 1.6e-7
        assert abs(w - math.erf(x)) < 1.6e-7
        assert abs(v + w - 1.0) < 1e-15


def _B6(b):
    return math.pi ** 1.5 / 3.0 * (b ** 3 * math.sqrt(math.pi) * special.erfc(b) + (0.5 - b * b) * math.exp(-b * b))


def _B12(b):
    return math.pi ** 1.5 / (945 * 120) * (-16 * b ** 9 * math.sqrt(math.pi) * special.erfc(b) +
                                           (16 * b ** 8 - 8 * b ** 6 + 12 * b ** 4 - 30 * b * b + 105) * math.exp(-b * b))


@pytest.mark.parametrize("b", [0.2, 0.45, 0.6, 1.0, 2.5, 4.0, 7.9, 8.5, 10.0])
def test_beta6_beta12_closed_form(oracle_mod, b):
    erfk, b6, b12 = oracle_mod.erfc_bet6_bet12(b)
    assert abs(erfk - special.erfc(b)) <= 2e-7 * max(special.erfc(b), 1e-300) + 2e-7 * math.exp(-b * b)
    # series / asymptotic branches are exact to ~1e-7, the middle branch is a trapezoid rule with dx ~ 1e-3
    assert abs(b6 - _B6(b)) <= 1e-6 * abs(_B6(b)) + 1e-12
    assert abs(b12 - _B12(b)) <= 1e-5 * abs(_B12(b)) + 1e-14


def test_shi_sans_exp(oracle_mod):
    lib = oracle_mod.load()
    for x in (0.3, 1.0, 2.0, 4.9, 7.5, 12.0, 17.0, 25.0):
        shi, _ = special.shichi(x)
        assert abs(lib.yo_shi_sans_exp(x) - shi * math.exp(-x)) <= 2e-4 * shi * math.exp(-x), x


def _direct_rho(sysm, kmax):
    """sum_j q_j exp(2 pi i k . r_j) for every k != 0 with |k| <= kmax (full sphere)."""
    r = np.stack([sysm.x, sysm.y, sysm.z], 1)
    ks = np.array([(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1)
                   for c in range(-kmax, kmax + 1) if 0 < a * a + b * b + c * c <= kmax * kmax], dtype=np.float64)
    ph = 2 * np.pi * ks @ r.T
    return ks, ph


def test_kspace_coulomb_energy_direct_full_sphere(oracle_mod):
    s = ys.config("small", seed=5)
    o = oracle_mod.Oracle(s)
    ec, elj = o.kspace_init()
    d = s.derived
    ks, ph = _direct_rho(s, s.kmax)
    rho = (s.charge[None, :] * np.exp(1j * ph)).sum(1)
    k2 = (ks ** 2).sum(1)
    w = d["xclb"] * np.exp(-np.pi ** 2 * k2 / d["alpha"] ** 2) / (2 * np.pi * k2)
    assert abs(ec - (w * np.abs(rho) ** 2).sum()) < 1e-11 * abs(ec)
    # LJ-Ewald k-space: sum_k beta_LJ(t1,t2) rho_t1 conj(rho_t2) over all type pairs, half-weighted
    b, b6, b12, bl = o.beta()
    kx = oracle_mod.grid(s.kmax)["kx"]
    fac = np.where(kx > 0, 2.0, 1.0)
    erf_b6 = b6 / fac
    erf_b12 = b12 / fac
    # map full-sphere k back to |k| index via k^2
    g = oracle_mod.grid(s.kmax)
    gk2 = g["kx"] ** 2 + g["ky"] ** 2 + g["kz"] ** 2
    lut6 = {int(k): v for k, v in zip(gk2, erf_b6)}
    lut12 = {int(k): v for k, v in zip(gk2, erf_b12)}
    B6 = np.array([lut6[int(v)] for v in k2])
    B12 = np.array([lut12[int(v)] for v in k2])
    T = s.ntype
    rt = [(np.where(s.lj_type == t + 1, 1.0, 0.0)[None, :] * np.exp(1j * ph)).sum(1) for t in range(T)]
    elj_direct = 0.0
    for t1 in range(T):
        for t2 in range(T):
            A6, A12 = s.lj6[t1 * T + t2], s.lj12[t1 * T + t2]
            elj_direct += ((A12 * B12 - A6 * B6) * (rt[t1] * np.conj(rt[t2])).real).sum()
    assert abs(elj - elj_direct) < 1e-11 * max(abs(elj), 1e-12)


def test_sincos_kr_direct(oracle_mod):
    s = ys.config("small", seed=9)
    o = oracle_mod.Oracle(s)
    g = oracle_mod.grid(s.kmax)
    x, y, z, q = 0.123, -0.377, 0.481, -0.8476
    out = o.sincos_kr(x, y, z, q)
    kx, ky, kz = g["kx"], g["ky"], g["kz"]
    for slot, (sy, sz) in enumerate([(1, 1), (1, -1), (-1, 1), (-1, -1)]):
        th = 2 * np.pi * (kx * x + sy * ky * y + sz * kz * z)
        mask = np.ones_like(kx, dtype=bool)
        if sy < 0:
            mask &= ky > 0
        if sz < 0:
            mask &= kz > 0
        # slot order in memory: ppp, ppm, pmp, pmm -> (sy,sz) = (+,+), (+,-), (-,+), (-,-)
        assert np.allclose(out[2 * slot][mask], q * np.cos(th[mask]), atol=1e-13)
        assert np.allclose(out[2 * slot + 1][mask], q * np.sin(th[mask]), atol=1e-13)
        assert np.all(out[2 * slot][~mask] == 0) and np.all(out[2 * slot + 1][~mask] == 0)


def test_kspace_delta_invariant_and_forces_fd(oracle_mod):
    s = ys.config("small", seed=3)
    o = oracle_mod.Oracle(s)
    ec0, el0 = o.kspace_init()
    x, y, z = s.positions()
    f, l = s.mol_first[4], s.mol_last[4]
    xn, yn, zn = x[f:l + 1] + 0.013, y[f:l + 1] - 0.021, z[f:l + 1] + 0.004
    duc, dul = o.kspace_delta(4, xn, yn, zn)
    x2, y2, z2 = x.copy(), y.copy(), z.copy()
    x2[f:l + 1], y2[f:l + 1], z2[f:l + 1] = xn, yn, zn
    o.set_positions(x2, y2, z2)
    ec1, el1 = o.kspace_init()
    assert abs(duc - (ec1 - ec0)) < 1e-10 * (abs(ec0) + abs(ec1))
    assert abs(dul - (el1 - el0)) < 1e-10 * (abs(el0) + abs(el1))
    # k-space force = -dE_k/dr (exact gradient of the k-space energy)
    o.set_positions(x, y, z)
    o.kspace_init()
    fx, fy, fz = o.kspace_forces(0, x[:4], y[:4], z[:4])
    h = 1e-6
    for i in range(4):
        for d, arr in enumerate((x, y, z)):
            vals = []
            for sg in (1, -1):
                pos = [x.copy(), y.copy(), z.copy()]
                pos[d][i] += sg * h
                o.set_positions(*pos)
                a, b = o.kspace_init()
                vals.append(a + b)
            fd = -(vals[0] - vals[1]) / (2 * h)
            fa = (fx, fy, fz)[d][i]
            assert abs(fd - fa) < 1e-6 * max(1.0, abs(fa)), (i, d, fd, fa)


def test_madelung_rocksalt(oracle_mod):
    """Ewald Coulomb energy of a NaCl rock-salt crystal (4x4x4 conventional
    cells) reproduces the Madelung constant 1.747565 for two splittings."""
    n = 8
    Lr = float(n)  # box length in units of the Na-Cl distance a
    coords, types = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                coords.append(((i + 0.5) / n - 0.5, (j + 0.5) / n - 0.5, (k + 0.5) / n - 0.5))
                types.append((i + j + k) % 2)
    nion = len(coords)
    counts = [nion // 2, nion // 2]
    mol_type = np.array(types, dtype=np.int32)
    order = np.argsort(mol_type, kind="stable")
    pos = np.array(coords)[order]
    mt = mol_type[order]
    a_angstrom = 2.82
    for alphaL, sk in ((9.0, 3.8), (12.0, 4.0)):
        p = ys.Params(box_length=a_angstrom * Lr, temp=298.0, sr=3.6, sk=sk, alpha=alphaL, external_permutivity=1e30,
                      dr_a=0.3, d_angle_degree=15.0, lambda_f=0.0, lambda_c=0.0)
        s = ys.build_system(p, [ys.NA_DANG, ys.CL_DANG], counts, mt, pos, np.zeros((nion, 3)))
        o = oracle_mod.Oracle(s)
        e, _ = o.mc_init()
        coul = e["uu_ew"] + e["E_kC"] + (e["U_tail"] - (o.tails()[2] - o.tails()[1]))
        per_ion = coul / nion
        madelung = -2 * per_ion / (s.derived["xclb"] * Lr)
        assert abs(madelung - 1.747564594633) < 1e-6, (alphaL, madelung)


def test_lj_tails_direct(oracle_mod):
    s = ys.config("small", seed=1)
    o = oracle_mod.Oracle(s)
    self_c, t6, t12 = o.tails()
    d = s.derived
    a = d["alpha"]
    assert abs(self_c - (-d["xclb"] * a / math.sqrt(math.pi) * (s.charge ** 2).sum())) < 1e-12 * abs(self_c)
    T = s.ntype
    typ = s.lj_type[s.lj_type > 0] - 1
    A6 = s.lj6.reshape(T, T)[typ][:, typ]
    A12 = s.lj12.reshape(T, T)[typ][:, typ]
    w6 = math.pi ** 1.5 * a ** 3 / 6 * A6.sum() - a ** 6 / 12 * np.trace(A6)
    w12 = math.pi ** 1.5 * a ** 9 / 1080 * A12.sum() - a ** 12 / 1440 * np.trace(A12)
    assert abs(t6 - w6) < 1e-12 * abs(w6)
    assert abs(t12 - w12) < 1e-12 * abs(w12)


def test_incremental_state_equals_fresh(oracle_mod):
    s = ys.config("small", seed=2)
    o = oracle_mod.Oracle(s)
    o.mc_init()
    o.set_rng(42)
    nacc, recs = o.run(10)
    assert 0 < nacc < len(recs)
    e1 = o.energies()
    o2 = oracle_mod.Oracle(s)
    o2.set_positions(*o.positions())
    e2, _ = o2.mc_init()
    for k in e1:
        assert abs(e1[k] - e2[k]) < 1e-10 * max(1.0, abs(e2[k])), k
    uu1, fx1, fy1, fz1 = o.real_arrays()
    uu2, fx2, fy2, fz2 = o2.real_arrays()
    assert np.allclose(uu1, uu2, atol=1e-10) and np.allclose(fx1, fx2, atol=1e-9)


def test_metropolis_without_bias(oracle_mod):
    """lambda = 0: moves stay within dr of the old centre, the acceptance
    factor is exp(-dU), and the run is reproducible from the RNG counter."""
    s = ys.config("small_nobias", seed=6)
    o = oracle_mod.Oracle(s)
    o.mc_init()
    o.set_rng(5, 3, 0)
    _, r1 = o.run(5)
    c1 = o.rng_counter()
    dr = s.derived["dr"]
    assert np.all(np.sqrt((r1["delta"] ** 2).sum(1)) <= dr * (1 + 1e-12))
    assert np.all(np.abs(r1["angle"]) <= s.derived["d_angle"] * (1 + 1e-12))
    o2 = oracle_mod.Oracle(s)
    o2.mc_init()
    o2.set_rng(5, 3, 0)
    _, r2 = o2.run(5)
    assert np.array_equal(r1, r2) and o2.rng_counter() == c1
    acc = r1["accepted"] == 1
    assert np.all(r1["du_total"][r1["du_total"] < 0][:, None] < 0)
    assert np.all(acc[(r1["du_total"] < 0) & (r1["overlap"] == 0)])


def test_overlap_rejects_without_acceptance_draw(oracle_mod):
    """Two LJ sites closer than the mean hard core: the move is rejected and
    no acceptance uniform is consumed (MC.f90:1080-1084)."""
    s = ys.config("small_nobias", seed=6)
    # push molecule 0's oxygen onto molecule 1's oxygen
    f0, f1 = s.mol_first[0], s.mol_first[1]
    dx = s.x[f1] - s.x[f0] - 0.05 / s.params.box_length
    dy, dz = s.y[f1] - s.y[f0], s.z[f1] - s.z[f0]
    l0 = s.mol_last[0]
    o = oracle_mod.Oracle(s)
    o.mc_init()
    x, y, z = s.positions()
    pr = o.rspace_partial(0, x[f0:l0 + 1] + dx, y[f0:l0 + 1] + dy, z[f0:l0 + 1] + dz)
    assert pr["overlap"]
