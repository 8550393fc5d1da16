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

__global__ void __launch_bounds__(256) yk_forces(YdChain c, const double *sig6, const double *feps,
                                                 const double *tab, double *g, double *r) {
    const int N = c.p.natom, T = c.p.ntype, nk = c.p.nk;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int im = c.molnum[i], it = c.type[i];
    double qi = c.q[i], xi = c.x[i], yi = c.y[i], zi = c.z[i];
    // real space
    double fe[3] = {0, 0, 0}, fa[3] = {0, 0, 0};  // inter-molecular, intra-molecular
    for (int j = 0; j < N; ++j) {
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
    for (int kk = 0; kk < nk; ++kk) {
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
    double fx = fe[0] + YD_FOUR_PI * fc[0] + YD_FOUR_PI * fl[0] + c.p.m2kext * qi * c.st->mu[0];
    double fy = fe[1] + YD_FOUR_PI * fc[1] + YD_FOUR_PI * fl[1] + c.p.m2kext * qi * c.st->mu[1];
    double fz = fe[2] + YD_FOUR_PI * fc[2] + YD_FOUR_PI * fl[2] + c.p.m2kext * qi * c.st->mu[2];
    r[i] = fx;
    r[N + i] = fy;
    r[2 * N + i] = fz;
    g[i] = fx + fa[0];
    g[N + i] = fy + fa[1];
    g[2 * N + i] = fz + fa[2];
}

void yk_launch_forces(const YdChain &c, const double *sig6, const double *feps, const double *tab, double *g,
                      double *r, hipStream_t s) {
    int N = c.p.natom;
    yk_forces<<<(N + 255) / 256, 256, 0, s>>>(c, sig6, feps, tab, g, r);
}
