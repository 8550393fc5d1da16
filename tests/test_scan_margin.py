"""The chain kernel's real-space scan tests packed FP32 positions (10 bits of
the fractional box coordinate per axis, the bin centre standing for the
position) against the molecule's sites with the nearest image d - rint(d)
and an r^2 bound grown by the packing error.  It must keep every atom that
the exact FP64 pass counts (single-wrap image, r^2 < rmax2,
EwaldSumRealSpace.f90:394-401): a lost atom would drop pair terms.

This restates the device arithmetic (yk_pack_pos, yk_unpack_pos,
yk_scan_r2, yk_near in yamc_kernels.hip) in float32 numpy and checks the
bound on random atoms, including positions outside [0, 1) and pairs right
at the cut-off; the GPU parity tests check the kernel itself.
"""
import numpy as np

SLACK = np.float32(8.5e-4)  # yk_scan_r2: > sqrt(3)/2048, the packing error


def pack_bins(v):
    i = np.floor((v - np.floor(v)) * 1024.0).astype(np.int64)
    return np.clip(i, 0, 1023)


def unpack(i):
    return ((i.astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / 1024.0)).astype(np.float32)


def single_wrap(d):
    d = np.where(d > 0.5, d - 1.0, d)
    return np.where(d < -0.5, d + 1.0, d)


def scan_passes(site, pos, rmax2):
    r = np.float32(np.sqrt(np.float32(rmax2))) + SLACK
    r2f = r * r + np.float32(1e-5)
    e = site.astype(np.float32) - unpack(pack_bins(pos))
    e = (e - np.rint(e)).astype(np.float32)
    return (e * e).sum(axis=1, dtype=np.float32) < r2f


def test_packing_error_is_within_the_slack():
    v = np.random.default_rng(0).uniform(-0.1, 1.1, 200000)
    err = np.abs(unpack(pack_bins(v)) - (v - np.floor(v)))
    err = np.minimum(err, 1.0 - err)  # modulo the box
    assert err.max() <= 0.5 / 1024 + 1e-7
    assert np.sqrt(3.0) * (0.5 / 1024 + 1e-7) < float(SLACK)


def test_scan_keeps_every_atom_the_exact_pass_counts():
    rng = np.random.default_rng(5)
    kept = extra = 0
    for _ in range(100):
        rc = rng.uniform(0.05, 0.45)
        rmax2 = rc * rc
        pos = rng.uniform(-0.05, 1.05, (20000, 3))
        site = rng.uniform(-0.05, 1.05, 3)
        d = single_wrap(site - pos)
        inside = (d * d).sum(axis=1) < rmax2
        passed = scan_passes(site, pos, rmax2)
        assert not np.any(inside & ~passed)
        kept += int(inside.sum())
        extra += int((passed & ~inside).sum())
    assert kept > 100000
    assert extra < 0.02 * kept  # the slack adds about 1 % candidates


def test_scan_keeps_pairs_at_the_cutoff():
    rng = np.random.default_rng(7)
    rc = 0.305
    u = rng.normal(size=(50000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    site = rng.uniform(0.0, 1.0, 3)
    pos = site + u * (rc * (1.0 - 1e-12))  # just inside the cut-off, across the box edges
    d = single_wrap(site - pos)
    inside = (d * d).sum(axis=1) < rc * rc
    assert inside.all()
    assert scan_passes(site, pos, rc * rc).all()
