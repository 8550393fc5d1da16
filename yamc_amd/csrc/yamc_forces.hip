// yamc_forces.hip -- hand-written adjoint of the total Ewald energy.
//
// One thread per atom i computes -dE/dr_i for
//   E = E_real (all pairs, intramolecular -erf terms included) + E_k(Coulomb)
//     + E_k(LJ-Ewald) + U_ext,
// and, from the same terms, the reference force convention of movemc /
// mc_calc_forces (intramolecular real-space pairs carry energy but no force,
// EwaldSumRealSpace.f90:442-494; k-space self terms included,
// ForceKSpace.f90:59-302; external -2 kext q mu, EwaldSumExternal.f90:215-232).
// sin/cos(k.r_i) are rebuilt per k from the atom's axis tables (recomputed
// in the backward pass, never stored per k); S(k) is read with every lane of
// a wavefront on the same k, so each load is a broadcast from L1/L2.
#include "yamc_device.h"

// One thread per atom i and group y (blockIdx.y of ngroups): the group's
// share of the partners j (a contiguous range) and of the k vectors (a
// contiguous range) gives partial sums of the four force parts, written to
// part[y][12][N] (inter- and intra-molecular real space, Coulomb and LJ
// k-space, 3 components each); yk_forces_reduce adds the groups in order and
// forms the forces (rounding-level regrouping of the per-atom sums).
__global__ void __launch_bounds__(256) yk_forces(YdChain c, const double *sig6, const double *feps,
                                                 const double *tab, int ngroups, double *part) {
    const int N = c.p.natom, T = c.p.ntype, nk = c.p.nk;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int gi = blockIdx.y;
    const int jc = (N + ngroups - 1) / ngroups, j0 = gi * jc, j1 = j0 + jc < N ? j0 + jc : N;
    const int kc = (nk + ngroups - 1) / ngroups, k0 = gi * kc, k1 = k0 + kc < nk ? k0 + kc : nk;
    int im = c.molnum[i], it = c.type[i];
    double qi = c.q[i], xi = c.x[i], yi = c.y[i], zi = c.z[i];
    // real space
    double fe[3] = {0, 0, 0}, fa[3] = {0, 0, 0};  // inter-molecular, intra-molecular
    for (int j = j0; j < j1; ++j) {
        if (j == i) continue;
        int jm = c.molnum[j], jt = c.type[j];
        double xij = yd_wrap(xi - c.x[j]), yij = yd_wrap(yi - c.y[j]), zij = yd_wrap(zi - c.z[j]);
        double r2 = xij * xij + yij * yij + zij * zij;
        if (r2 >= c.p.rmax2) continue;
        bool lj = it != 0 && jt != 0;
        int off = lj ? (it - 1) * T + (jt - 1) : 0;
        YdPair pr;
        yd_pair(c.p, r2, im == jm, qi, c.q[j], lj, sig6[off], feps[off], pr);
        double *f = im == jm ? fa : fe;
        f[0] += xij * pr.FR;
        f[1] += yij * pr.FR;
        f[2] += zij * pr.FR;
    }
    // reciprocal space
    const double *ti = tab + (size_t)i * 6 * YD_KP;
    double fc[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
    for (int kk = k0; kk < k1; ++kk) {
        int kxyz = c.kxyz[kk], kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        double krc[8], krl[8], C[8];
        yd_kr_atom(ti, qi, it, kx, ky, kz, krc, krl);
        if (qi != 0.0) {
            for (int m = 0; m < 8; ++m)
                C[m] = c.p.precision == YAMC_FP32C
                           ? (double)c.Shi[(size_t)m * nk + kk] + (double)c.Slo[(size_t)m * nk + kk]
                           : c.S[(size_t)m * nk + kk];
            yd_force_k(C, krc, kx, ky, kz, c.beta[kk], fc);
        }
        if (it > 0) {
            for (int t = 0; t < T; ++t) {
                size_t base = (size_t)(1 + t) * 8 * nk + kk;
                for (int m = 0; m < 8; ++m)
                    C[m] = c.p.precision == YAMC_FP32C
                               ? (double)c.Shi[base + (size_t)m * nk] + (double)c.Slo[base + (size_t)m * nk]
                               : c.S[base + (size_t)m * nk];
                yd_force_k(C, krl, kx, ky, kz, c.beta_lj[((size_t)(it - 1) * T + t) * nk + kk], fl);
            }
        }
    }
    double *pp = part + (size_t)gi * 12 * N + i;
    for (int d = 0; d < 3; ++d) {
        pp[(size_t)d * N] = fe[d];
        pp[(size_t)(3 + d) * N] = fa[d];
        pp[(size_t)(6 + d) * N] = fc[d];
        pp[(size_t)(9 + d) * N] = fl[d];
    }
}

__global__ void yk_forces_reduce(YdChain c, const double *part, int ngroups, double *g, double *r) {
    const int N = c.p.natom;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double v[12];
    for (int q = 0; q < 12; ++q) v[q] = 0.0;
    for (int gi = 0; gi < ngroups; ++gi)
        for (int q = 0; q < 12; ++q) v[q] = v[q] + part[((size_t)gi * 12 + q) * N + i];
    const double qi = c.q[i];
    double f[3];
    for (int d = 0; d < 3; ++d) {
        f[d] = v[d] + YD_FOUR_PI * v[6 + d] + YD_FOUR_PI * v[9 + d] + c.p.m2kext * qi * c.st->mu[d];
        r[(size_t)d * N + i] = f[d];
        g[(size_t)d * N + i] = f[d] + v[3 + d];
    }
}

int yk_launch_forces(const YdChain &c, const double *sig6, const double *feps, const double *tab, double *g,
                     double *r, hipStream_t s) {
    const int N = c.p.natom;
    // groups of at least 64 partners, about 2^19 (atom, group) threads, at most 64
    int ng = (1 << 19) / (N > 0 ? N : 1);
    ng = ng < 1 ? 1 : ng > 64 ? 64 : ng;
    if (ng > N / 64) ng = N / 64 > 1 ? N / 64 : 1;
    double *part = nullptr;
    if (hipMallocAsync((void **)&part, sizeof(double) * 12 * (size_t)N * ng, s) != hipSuccess) return -1;
    yk_forces<<<dim3((N + 255) / 256, ng), 256, 0, s>>>(c, sig6, feps, tab, ng, part);
    yk_forces_reduce<<<(N + 255) / 256, 256, 0, s>>>(c, part, ng, g, r);
    if (hipFreeAsync(part, s) != hipSuccess) return -1;
    return 0;
}
