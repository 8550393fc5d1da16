// yamc_kernels.hip -- HIP kernels of the Ewald Monte Carlo path for gfx950.
//
// Hot path: yk_chain, one workgroup per Monte Carlo chain (replica), looping
// over trial moves on the device.  Per move it runs the phases of movemc
// (MC.f90:831-1322): old k-space forces of the moved molecule (one sweep over
// S(k)), the force-biased proposal and the new real-space partial sum, the new
// sweep (delta S, dU_k, new k-space forces), the Metropolis test, and on
// acceptance the S(k) commit and the O(N) update of the per-atom arrays.
// Scalar work (RNG, proposal, acceptance) runs on thread 0 in FP64; sweeps are
// spread over the workgroup with k (or atom j) fastest for coalesced loads.
// Nothing here is a dense contraction, so no MFMA is used.
#include <stddef.h>

#include "yamc_device.h"

// Largest workgroup of the chain kernel (8 waves) and the size of the
// single-call kernels; the chain kernel is instantiated for 256 and 512 and
// the launch picks one (yamc_tune measures both).
#define YK_BS 512
#define YK_NW (YK_BS / 64)
#define YK_MAXP 2            // distinct LJ types inside one molecule
// Bound on every rejection-sampling loop of the proposal.  The reference
// loops until success; with a constant uniform per move (RND_CONSTANT) some
// loops never succeed, so a move that exceeds the bound is rejected without
// an acceptance draw and recorded with overlap = 2.
#define YK_TRIES 65536
// Candidate neighbours of one move kept in LDS: atoms within the cut-off of
// a new or an old site of the moved molecule.
#define YK_NBCAP YD_NBCAP
// Candidate bitmap over atom indices (order-independent marking, then a
// deterministic ascending compaction into the list).
#define YK_BMWORDS 4096  // 131072 atoms
#define YK_MAXCL 512     // cells visited by one cell-list scan
// Positions each thread loads ahead in the real-space scan (independent
// loads in flight per thread when the packed positions are read from HBM).
#define YK_SCANU 8

// ------------------------------------------------------------------------
// Functions.f90:26-111 erfc_Luc_bet6_bet12 on one wavefront: the series and
// asymptotic branches on every lane (the same values), the trapezoid
// branch's up to 7,500 steps spread over the lanes (lane l takes steps
// 2 + l, 66 + l, ...), the lanes' partial sums added by a fixed butterfly;
// the sums differ from the serial loop's by rounding only.  Every lane
// returns the same values.
__device__ void yk_erfc_bet6_bet12_wave(double x, double x2, double e2, double *erfk, double *bet6, double *bet12) {
    const int lane = threadIdx.x & 63;
    double pi = 4.0 * atan(1.0);
    double pi12 = sqrt(pi);
    double pi15 = pow(pi, 1.5);
    const double xmin = 0.5, xmax = 8.0;
    if (x <= xmin) {
        *erfk = yd_erfc_luc(x, x2, e2);
        *bet6 = pi15 / 3.0 * yd_p3(x) * (pi12 * *erfk + (0.5 / x2 - 1.0) / x * e2);
        *bet12 = pi15 / 1080.0 * yd_p9(x) *
                 (-16.0 * pi12 / 105.0 * *erfk +
                  (((((1.0 / x2 - 2.0 / 7.0) / x2 + 4.0 / 35.0) / x2 - 8.0 / 105.0) / x2 + 16.0 / 105.0)) / x * e2);
    } else if (x >= xmax) {
        *erfk = e2 / pi12 / x * (1.0 - 0.5 / x2 * (1.0 - 1.5 / x2 * (1.0 - 2.5 / x2)));
        *bet6 = pi15 * e2 / 4.0 / (x * x) * (1.0 - 2.5 / x2 * (1.0 - 3.5 / x2 * (1.0 - 4.5 / x2)));
        *bet12 = pi15 * e2 / 240.0 / (x * x) * (1.0 - 5.5 / x2 * (1.0 - 6.5 / x2 * (1.0 - 7.5 / x2)));
    } else {
        double dx = 1.0e-3;
        int n = (int)((xmax - x) / dx + 2);
        dx = (xmax - x) / (n - 1);
        double x4 = (x * x) * (x * x), x3 = x * x * x, x5 = x4 * x, x8 = x4 * x4, x9 = x8 * x, x10 = x8 * x * x,
               x11 = x10 * x;
        double p0 = 0.0, p6 = 0.0, p12 = 0.0;
        for (int i = 2 + lane; i <= n; i += 64) {
            const double y = x + (double)(i - 1) * dx, e = exp(-(y * y));
            const double y4 = (y * y) * (y * y), y10 = y4 * y4 * (y * y);
            p0 = p0 + dx * e;
            p6 = p6 + dx * e / y4;
            p12 = p12 + dx * e / y10;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            p0 += __shfl_xor(p0, off, 64);
            p6 += __shfl_xor(p6, off, 64);
            p12 += __shfl_xor(p12, off, 64);
        }
        double som = dx * e2 / 2.0 - dx * dx / 12.0 * 2.0 * x * e2 + p0;
        double som6 = dx * e2 / x4 / 2.0 - dx * dx / 12.0 * (2.0 / x3 + 4.0 / x5) * e2 + p6;
        double som12 = dx * e2 / x10 / 2.0 - dx * dx / 12.0 * (2.0 / x9 + 10.0 / x11) * e2 + p12;
        const double y = n >= 2 ? x + (double)(n - 1) * dx : x, e = n >= 2 ? exp(-(y * y)) : e2;
        double y2 = y * y, y3 = y2 * y, y4 = y2 * y2, y5 = y4 * y, y8 = y4 * y4, y9 = y8 * y, y10 = y8 * y2,
               y11 = y10 * y;
        som = som - dx * e / 2.0 + dx * dx / 12.0 * 2.0 * y * e;
        som = som + e / 2.0 / y * (1.0 - 0.5 / y2 * (1.0 - 1.5 / y2 * (1.0 - 2.5 / y2)));
        *erfk = 2.0 / pi12 * som;
        som6 = som6 - dx * e / y4 / 2.0 + dx * dx / 12.0 * (2.0 / y3 + 4.0 / y5) * e;
        som6 = som6 + e / 2.0 / y5 * (1.0 - 2.5 / y2 * (1.0 - 3.5 / y2 * (1.0 - 4.5 / y2)));
        *bet6 = pi15 * x3 / 2.0 * som6;
        som12 = som12 - dx * e / y10 / 2.0 + dx * dx / 12.0 * (2.0 / y9 + 10.0 / y11) * e;
        som12 = som12 + e / 2.0 / y11 * (1.0 - 5.5 / y2 * (1.0 - 6.5 / y2 * (1.0 - 7.5 / y2)));
        *bet12 = pi15 * x9 / 120.0 * som12;
    }
}

// EwaldSumKSpace.f90:263-350 initBeta: one wavefront per k (blocks of 64
// threads, one k each).
__global__ void __launch_bounds__(64) yk_beta(YdChain c, const double *xk2, double *beta6, double *beta12) {
    const int kk = blockIdx.x;
    if (kk >= c.p.nk) return;
    double poa = YD_PI / c.p.alpha, poa2 = poa * poa;
    double k2 = xk2[kk], b2 = poa2 * k2, e = exp(-b2);
    double bt = c.p.xclb * e / k2 / YD_TWO_PI;
    double b = poa * sqrt(k2), erfc_b, B6, B12;
    yk_erfc_bet6_bet12_wave(b, b2, e, &erfc_b, &B6, &B12);
    double b6 = yd_p3(c.p.alpha) * B6, b12 = yd_p9(c.p.alpha) * B12;
    if ((yd_g(c.kxyz)[kk] & 255) > 0) { bt = 2.0 * bt; b6 = 2.0 * b6; b12 = 2.0 * b12; }
    if (threadIdx.x == 0) {
        yd_g(c.beta)[kk] = bt;
        beta6[kk] = b6;
        beta12[kk] = b12;
    }
}

// EwaldSumKSpace.f90:352-386 initBetaLJ, full symmetric rows.
__global__ void yk_beta_lj(YdChain c, const double *lj6, const double *lj12, const double *beta6,
                           const double *beta12) {
    int T = c.p.ntype, nk = c.p.nk;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)T * T * nk) return;
    int kk = (int)(i % nk), pr = (int)(i / nk), t1 = pr / T, t2 = pr % T;
    int a = t1 <= t2 ? t1 : t2, b = t1 <= t2 ? t2 : t1;
    yd_g(c.beta_lj)[i] = lj12[a * T + b] * beta12[kk] - lj6[a * T + b] * beta6[kk];
}

// EwaldSumKSpace.f90:210-211: beta(:) = beta(:) * scale_beta.
__global__ void yk_scale_beta(YdChain c, double s) {
    int kk = blockIdx.x * blockDim.x + threadIdx.x;
    if (kk < c.p.nk) yd_g(c.beta)[kk] = yd_g(c.beta)[kk] * s;
}

// scale_box.f90:6-62: every molecule of more than one atom keeps its centre
// of mass (box units) while its atoms' offsets from it scale by sc = L_old /
// L_new, so that bond lengths in Angstrom are preserved.  Thread per molecule.
__global__ void yk_scale_box(YdChain c, double sc) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.p.nmol) return;
    int f = yd_g(c.first)[m], l = yd_g(c.last)[m];
    if (l == f) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, tm = 0.0;  // geometry.f90 center_of_mass
    for (int a = f; a <= l; ++a) {
        double w = yd_g(c.mass)[a];
        sx = sx + w * yd_g(c.x)[a]; sy = sy + w * yd_g(c.y)[a]; sz = sz + w * yd_g(c.z)[a];
        tm = tm + w;
    }
    double cx = sx / tm, cy = sy / tm, cz = sz / tm;
    for (int a = f; a <= l; ++a) {
        yd_g(c.x)[a] = (yd_g(c.x)[a] - cx) * sc + cx;
        yd_g(c.y)[a] = (yd_g(c.y)[a] - cy) * sc + cy;
        yd_g(c.z)[a] = (yd_g(c.z)[a] - cz) * sc + cz;
    }
}

// SumSinCosKR.f90:181-223 for one atom: tab[i][6][YD_KP].
__device__ __forceinline__ void yk_atom_table1(const YdChain &c, const double *x, const double *y, const double *z,
                                               double *tab, int i) {
    double *t = tab + (size_t)i * 6 * YD_KP;
    yd_sincos_axis(x[i], c.p.kmax, t, t + YD_KP);
    yd_sincos_axis(y[i], c.p.kmax, t + 2 * YD_KP, t + 3 * YD_KP);
    yd_sincos_axis(z[i], c.p.kmax, t + 4 * YD_KP, t + 5 * YD_KP);
}

// SumSinCosKR.f90:181-223 for every atom.
__global__ void yk_atom_tables(YdChain c, const double *x, const double *y, const double *z, double *tab) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.p.natom) return;
    yk_atom_table1(c, x, y, z, tab, i);
}

// RhoSquared_addAtoms (RhoSquared.f90:188-243) from scratch, in two steps:
// one thread per (set, k) and atom group (blockIdx.y: the y-th of ngroups
// equal parts of the set's atom list, atoms in index order) writes a partial
// sum to part[group][set][8][nk]; yk_full_sums_reduce adds the groups in
// order (the reference adds all atoms in index order: rounding-level
// difference).  The set's list holds exactly the atoms the reference adds to
// it (charged atoms, or atoms of the set's LJ type).
__global__ void yk_full_sums(YdChain c, const double *tab, int ngroups, double *part) {
    int nk = c.p.nk, T = c.p.ntype;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1 + T) * nk) return;
    int set = i / nk, kk = i % nk;
    int kxyz = yd_g(c.kxyz)[kk], kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
    const int l0 = yd_g(c.set_off)[set], l1 = yd_g(c.set_off)[set + 1];
    const int chunk = (l1 - l0 + ngroups - 1) / ngroups;
    const int t0 = l0 + (int)blockIdx.y * chunk, t1 = t0 + chunk < l1 ? t0 + chunk : l1;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = t0; t < t1; ++t) {
        const int a = yd_g(c.set_atoms)[t];
        double krc[8], krl[8];
        yd_kr_atom(tab + (size_t)a * 6 * YD_KP, yd_g(c.q)[a], yd_g(c.type)[a], kx, ky, kz, krc, krl);
        const double *kr = set == 0 ? krc : krl;
#pragma unroll
        for (int m = 0; m < 8; ++m) s[m] = s[m] + kr[m];
    }
    size_t base = ((size_t)blockIdx.y * (1 + T) + set) * 8 * nk + kk;
#pragma unroll
    for (int m = 0; m < 8; ++m) part[base + (size_t)m * nk] = s[m];
}

// The atom groups' partial sums of yk_full_sums added in group order, one
// thread per S(k) value (flat index = the S layout's), into S (FP32C: the
// (hi, lo) split).
__global__ void yk_full_sums_reduce(YdChain c, const double *part, int ngroups) {
    const int n = (1 + c.p.ntype) * 8 * c.p.nk;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int g = 0; g < ngroups; ++g) s = s + part[(size_t)g * n + i];
    if (c.p.precision == YAMC_FP32C) {
        float hi = (float)s;
        yd_g(c.Shi)[i] = hi;
        yd_g(c.Slo)[i] = (float)(s - (double)hi);
    } else {
        yd_g(c.S)[i] = s;
    }
}

__device__ __forceinline__ double yk_ldS(const YdChain &c, size_t idx) {
    if (c.p.precision == YAMC_FP32C) return (double)yd_g(c.Shi)[idx] + (double)yd_g(c.Slo)[idx];
    return yd_g(c.S)[idx];
}

// calc_coulomb_energy + calc_LJ_energy (EwaldSumKSpace.f90:389-526): the
// terms of one k added into v[0] (Coulomb) and v[1] (LJ).
__device__ __forceinline__ void yk_kenergy_k(const YdChain &c, int kk, double *v) {
    const int nk = c.p.nk, T = c.p.ntype;
    int kxyz = yd_g(c.kxyz)[kk], ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
    double C[8];
    for (int m = 0; m < 8; ++m) C[m] = yk_ldS(c, (size_t)m * nk + kk);
    double r2 = C[0] * C[0] + C[1] * C[1];
    if (ky > 0) r2 = r2 + C[4] * C[4] + C[5] * C[5];
    if (kz > 0) r2 = r2 + C[2] * C[2] + C[3] * C[3];
    if (ky > 0 && kz > 0) r2 = r2 + C[6] * C[6] + C[7] * C[7];
    v[0] += yd_g(c.beta)[kk] * r2;
    for (int t1 = 0; t1 < T; ++t1) {
        double A[8];
        for (int m = 0; m < 8; ++m) A[m] = yk_ldS(c, ((size_t)(1 + t1) * 8 + m) * nk + kk);
        for (int t2 = t1; t2 < T; ++t2) {
            double B[8];
            for (int m = 0; m < 8; ++m) B[m] = yk_ldS(c, ((size_t)(1 + t2) * 8 + m) * nk + kk);
            double loc = A[0] * B[0] + A[1] * B[1];
            if (ky > 0) loc = loc + A[4] * B[4] + A[5] * B[5];
            if (kz > 0) loc = loc + A[2] * B[2] + A[3] * B[3];
            if (ky > 0 && kz > 0) loc = loc + A[6] * B[6] + A[7] * B[7];
            double term = yd_g(c.beta_lj)[((size_t)t1 * T + t2) * nk + kk] * loc;
            v[1] += (t1 != t2) ? term + term : term;
        }
    }
}

// The k-space energies over all k, one workgroup of YK_KE_THREADS threads
// (thread t adds k = t, t + YK_KE_THREADS, ...; then the block sum); result
// into st->e_kc, st->e_klj.
#define YK_KE_THREADS 1024
__global__ void __launch_bounds__(YK_KE_THREADS) yk_kenergy(YdChain c) {
    __shared__ double red[(YK_KE_THREADS / 64) * 2], out[2];
    double v[2] = {0.0, 0.0};
    for (int kk = threadIdx.x; kk < c.p.nk; kk += YK_KE_THREADS) yk_kenergy_k(c, kk, v);
    yd_block_sum<2>(v, red, out);
    if (threadIdx.x == 0) {
        yd_g(c.st)->e_kc = out[0];
        yd_g(c.st)->e_klj = out[1];
    }
}

// EwaldSumExternal_calc_mu (EwaldSumExternal.f90:165-183), one workgroup.
__global__ void yk_mu(YdChain c, double *mu_out) {
    __shared__ double red[16 * 3], out[3];
    double v[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < c.p.natom; i += blockDim.x) {
        double q = yd_g(c.q)[i];
        v[0] += q * yd_g(c.x)[i];
        v[1] += q * yd_g(c.y)[i];
        v[2] += q * yd_g(c.z)[i];
    }
    yd_block_sum<3>(v, red, out);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) yd_g(c.st)->mu[d] = out[d];
        if (mu_out)
            for (int d = 0; d < 3; ++d) mu_out[d] = out[d];
    }
}

// EwaldSumTails.f90:18-127 (constant terms), single thread.
// ewald_sum_coulomb_tail + ewald_sum_lj_tails (EwaldSumTails.f90:18-127):
// the charge sum over one block of 256 threads (strided partial sums, block
// reduction), the type sums on thread 0.
__global__ void __launch_bounds__(256) yk_tails(YdChain c, const double *lj6, const double *lj12, const int *counts) {
    __shared__ double red[4], out[1];
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < c.p.natom; i += (int)blockDim.x) v[0] += yd_g(c.q)[i] * yd_g(c.q)[i];
    yd_block_sum<1>(v, red, out);
    if (threadIdx.x != 0) return;
    const double S = out[0];
    yd_g(c.st)->tail_c = -c.p.xclb * c.p.alpha_over_sqrt_pi * S;
    int T = c.p.ntype;
    double a3 = yd_p3(c.p.alpha), a6 = a3 * a3, a9 = a3 * a6, a12 = a6 * a6, pspi = pow(YD_PI, 1.5);
    double sij6 = 0.0, sij12 = 0.0, sj6 = 0.0, sj12 = 0.0;
    for (int t1 = 0; t1 < T; ++t1)
        for (int t2 = t1; t2 < T; ++t2) {
            int n12 = counts[t1] * counts[t2];
            double l6 = lj6[t1 * T + t2] * (double)n12, l12 = lj12[t1 * T + t2] * (double)n12;
            if (t1 != t2) { l6 = l6 + l6; l12 = l12 + l12; }
            sij6 = sij6 + l6;
            sij12 = sij12 + l12;
        }
    for (int t = 0; t < T; ++t) {
        sj6 = sj6 + lj6[t * T + t] * (double)counts[t];
        sj12 = sj12 + lj12[t * T + t] * (double)counts[t];
    }
    yd_g(c.st)->tail6 = pspi * a3 / 6.0 * sij6 - a6 / 12.0 * sj6;
    yd_g(c.st)->tail12 = pspi * a9 / 1080.0 * sij12 - a12 / 1440.0 * sj12;
}

// Full real-space sum (EwaldSumRealSpace.f90:338-498, full mode): one thread
// per atom i and partner group (blockIdx.y of ngroups, a contiguous range of
// j) visits the group's j != i and writes its partial uu(i), F(i) to
// pat[group][4][N] (yk_rspace_full_atoms adds the groups in order, no
// atomics); the six energy components count pairs j > i only and are reduced
// per workgroup into part[group][block][6].  ov is set on an LJ hard-core
// overlap.
__global__ void __launch_bounds__(256) yk_rspace_full(YdChain c, const double *lj_sig6, const double *lj_4eps,
                                                      int ngroups, double *part, double *pat, int *ov) {
    __shared__ double red[4 * 6], out[6];
    int N = c.p.natom, T = c.p.ntype;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int jc = (N + ngroups - 1) / ngroups, j0 = (int)blockIdx.y * jc, j1 = j0 + jc < N ? j0 + jc : N;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < N) {
        int im = yd_g(c.molnum)[i], it = yd_g(c.type)[i];
        double qi = yd_g(c.q)[i], dhi = yd_g(c.dh)[i], xi = yd_g(c.x)[i], yi = yd_g(c.y)[i], zi = yd_g(c.z)[i];
        double uu = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        for (int j = j0; j < j1; ++j) {
            if (j == i) continue;
            int jm = yd_g(c.molnum)[j], jt = yd_g(c.type)[j];
            double xij = yd_wrap(xi - yd_g(c.x)[j]), yij = yd_wrap(yi - yd_g(c.y)[j]), zij = yd_wrap(zi - yd_g(c.z)[j]);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (dhi + yd_g(c.dh)[j]) / 2.0;
            if (it > 0 && jt > 0 && im != jm && r2 < d12 * d12) *ov = 1;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, im == jm, qi, yd_g(c.q)[j], lj, lj_sig6[off], lj_4eps[off], pr);
            if (j > i) {
                v[0] += pr.potew;
                v[1] += pr.potlj6;
                v[2] += pr.potlj12;
                if (im == jm) {
                    v[3] += pr.potew;
                    v[4] += pr.potlj6;
                    v[5] += pr.potlj12;
                }
            }
            if (im != jm) {
                uu += pr.pot;
                fx += xij * pr.FR;
                fy += yij * pr.FR;
                fz += zij * pr.FR;
            }
        }
        double *pa = pat + (size_t)blockIdx.y * 4 * N + i;
        pa[0] = uu;
        pa[(size_t)N] = fx;
        pa[2 * (size_t)N] = fy;
        pa[3 * (size_t)N] = fz;
    }
    yd_block_sum<6>(v, red, out);
    if (threadIdx.x < 6) part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = out[threadIdx.x];
}

// uu(i), F(i) of the full real-space sum from the groups' partial sums, added
// in group order.
__global__ void yk_rspace_full_atoms(YdChain c, const double *pat, int ngroups) {
    const int N = c.p.natom;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double v[4] = {0, 0, 0, 0};
    for (int g = 0; g < ngroups; ++g)
        for (int q = 0; q < 4; ++q) v[q] = v[q] + pat[((size_t)g * 4 + q) * N + i];
    yd_g(c.uu)[i] = v[0];
    yd_g(c.Fx)[i] = v[1];
    yd_g(c.Fy)[i] = v[2];
    yd_g(c.Fz)[i] = v[3];
}

__global__ void yk_rspace_full_final(YdChain c, const double *part, int nblk, double *comps_out) {
    __shared__ double red[16 * 6], out[6];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
        for (int m = 0; m < 6; ++m) v[m] += part[b * 6 + m];
    yd_block_sum<6>(v, red, out);
    if (threadIdx.x == 0) {
        yd_g(c.st)->uu_ew = out[0]; yd_g(c.st)->uu_lj6 = out[1]; yd_g(c.st)->uu_lj12 = out[2];
        yd_g(c.st)->uu_ew_intra = out[3]; yd_g(c.st)->uu_lj6_intra = out[4]; yd_g(c.st)->uu_lj12_intra = out[5];
        if (comps_out)
            for (int m = 0; m < 6; ++m) comps_out[m] = out[m];
    }
}

// ========================================================================
// Per-move machinery shared by the chain kernel and the single-call entry
// points.  Molecule data of the move lives in LDS.
struct YkMol {
    int imol, first, na, np;
    int typ[YD_MAXA], lslot[YD_MAXA], ptype[YK_MAXP];
    double q[YD_MAXA], mass[YD_MAXA], dh[YD_MAXA];
    double xo[YD_MAXA], yo[YD_MAXA], zo[YD_MAXA], xn[YD_MAXA], yn[YD_MAXA], zn[YD_MAXA];
    double tab_o[YD_MAXA * 6 * YD_KP], tab_n[YD_MAXA * 6 * YD_KP];
};

// forward declarations of the thread-0 state types (defined below)
struct YkMcm {
    int stall;  // a rejection-sampling loop exceeded YK_TRIES draws
    double fo[3], ftotold, co[3], couplold, umaxold, vmaxold, cold, qold;
    double v[3], d[3], angle;
    double fn[3], ftotnew, cn[3], couplnew, umaxnew, vmaxnew, cnew, qnew;
    double cphi, sphi, ct, st, ca, sa;
};

// Scalar state of the move, written by thread 0 only; kept in LDS so that
// the wide sweeps do not carry it in registers.
struct YkScalar {
    YkMcm mcm;
    YdRng rng;
    YkMcm mcm2;  // the rotation's copy when displacement and rotation run on two waves
    YdRng rng2;
    double U_old, frn[YD_MAXA][3], uun[YD_MAXA], cmp_new[3], cmp_old[3], duc, dulj, du_real, du_ext;
    double mu[3], mu_new[3], muo[3], com[3];
    long long acc_count, nmv, nov;
    // the next move's pick, taken during this move's acceptance (workgroup
    // chain kernel): its molecule and the RNG state after its draws
    YdRng rng_next;
    int prepicked, nimol;
    // the state's carried energy sums (workgroup chain kernel), so that a
    // commit does not wait for their loads
    double e_kc, e_klj, uu_ew, uu_lj6, uu_lj12;
    double qold_pre;  // the rotation's Q_old, computed by wave 2 during the proposal
    double qnew_pre;  // the acceptance's Q_new, computed by wave 2 during the acceptance
    double du_total;  // the move's energy change, between the two parts of the acceptance
};

struct YkShared {
    YkMol m;
    YkScalar t0;
    double red[YK_NW * 32];
    double out[32];
    double fr_old[YD_MAXA][3], uu_old[YD_MAXA];
    int ov, accepted, stall, nbcount;
    int nblist[YK_NBCAP];
    unsigned int bm[YK_BMWORDS];
    int wsum[YK_NW];
    unsigned long long ph[9], t_last;  // phase timers (thread 0)
    // cell-list scan: the cells around the molecule's old and new sites and
    // the running offsets of their buckets; cell_ok = 0 after a bucket
    // overflowed (the scan then visits every atom until the next launch)
    int cl_cell[YK_MAXCL], cl_start[YK_MAXCL + 1], ncl, cell_ok;
};

// Loads molecule imol (old positions from the state) into LDS; thread 0.
__device__ void yk_load_mol(const YdChain &c, YkMol &m, int imol) {
    m.imol = imol;
    m.first = yd_g(c.first)[imol];
    m.na = yd_g(c.last)[imol] - m.first + 1;
    m.np = 0;
    for (int a = 0; a < m.na; ++a) {
        int g = m.first + a;
        m.q[a] = yd_g(c.q)[g];
        m.typ[a] = yd_g(c.type)[g];
        m.mass[a] = yd_g(c.mass)[g];
        m.dh[a] = yd_g(c.dh)[g];
        m.xo[a] = yd_g(c.x)[g];
        m.yo[a] = yd_g(c.y)[g];
        m.zo[a] = yd_g(c.z)[g];
        m.lslot[a] = -1;
        if (m.typ[a] > 0) {
            int s = -1;
            for (int p = 0; p < m.np; ++p)
                if (m.ptype[p] == m.typ[a]) s = p;
            if (s < 0 && m.np < YK_MAXP) { s = m.np; m.ptype[m.np++] = m.typ[a]; }
            m.lslot[a] = s;
        }
    }
}

// Axis tables of the molecule's atoms (old or new positions): 3*na threads.
__device__ void yk_mol_tables(const YdChain &c, YkMol &m, bool newpos, int t) {
    if (t < 3 * m.na) {
        int a = t / 3, d = t % 3;
        double v = newpos ? (d == 0 ? m.xn[a] : d == 1 ? m.yn[a] : m.zn[a])
                          : (d == 0 ? m.xo[a] : d == 1 ? m.yo[a] : m.zo[a]);
        double *tab = (newpos ? m.tab_n : m.tab_o) + a * 6 * YD_KP + 2 * d * YD_KP;
        yd_sincos_axis(v, c.p.kmax, tab, tab + YD_KP);
    }
}

// Delta of the Coulomb sums at k: (sum of new) - (sum of old) over the
// molecule's charged atoms in index order (RhoSquared_addAtom into nulified
// sums, then RhoSquared_sub: MC.f90:946-958, 1096-1125).
// The molecule's dS at one k (new - old, atoms in index order), with the
// slots of ky = 0 or kz = 0 zero as SinCosKR_fill leaves them (set once
// here: the per-atom terms of those slots are summed unzeroed and dropped).
__device__ __forceinline__ void yk_zero_slots(double *d, int ky, int kz) {
    if (kz == 0) { d[2] = 0.0; d[3] = 0.0; d[6] = 0.0; d[7] = 0.0; }
    if (ky == 0) { d[4] = 0.0; d[5] = 0.0; d[6] = 0.0; d[7] = 0.0; }
}
__device__ __forceinline__ void yk_delta_coul(const YkMol &m, int kx, int ky, int kz, double *dc) {
    double ns[8] = {0, 0, 0, 0, 0, 0, 0, 0}, os[8] = {0, 0, 0, 0, 0, 0, 0, 0}, kr[8];
    for (int a = 0; a < m.na; ++a) {
        double q = m.q[a];
        if (q == 0.0) continue;
        yd_kr_nz(m.tab_n + a * 6 * YD_KP, q, kx, ky, kz, kr);
#pragma unroll
        for (int i = 0; i < 8; ++i) ns[i] = ns[i] + kr[i];
        yd_kr_nz(m.tab_o + a * 6 * YD_KP, q, kx, ky, kz, kr);
#pragma unroll
        for (int i = 0; i < 8; ++i) os[i] = os[i] + kr[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) dc[i] = ns[i] - os[i];
    yk_zero_slots(dc, ky, kz);
}
__device__ __forceinline__ void yk_delta_lj(const YkMol &m, int p, int kx, int ky, int kz, double *dl) {
    double ns[8] = {0, 0, 0, 0, 0, 0, 0, 0}, os[8] = {0, 0, 0, 0, 0, 0, 0, 0}, kr[8];
    for (int a = 0; a < m.na; ++a) {
        if (m.lslot[a] != p) continue;
        // yd_kr_atom's LJ values: (q sin/cos) (1/q) of a charged site, else sin/cos
        const double q = m.q[a], inv = q != 0.0 ? 1.0 / q : 1.0, qk = q != 0.0 ? q : 1.0;
        yd_kr_nz(m.tab_n + a * 6 * YD_KP, qk, kx, ky, kz, kr);
#pragma unroll
        for (int i = 0; i < 8; ++i) ns[i] = ns[i] + (q != 0.0 ? kr[i] * inv : kr[i]);
        yd_kr_nz(m.tab_o + a * 6 * YD_KP, qk, kx, ky, kz, kr);
#pragma unroll
        for (int i = 0; i < 8; ++i) os[i] = os[i] + (q != 0.0 ? kr[i] * inv : kr[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) dl[i] = ns[i] - os[i];
    yk_zero_slots(dl, ky, kz);
}

// Loads the 8 components of set at k.  FP32C reads the hi word of the LJ
// sets (their k-space terms are small, the hi word's 2^-25 relative error
// stays far below the dE bound) and hi + lo of the Coulomb set, whose
// 2 beta C dC terms carry the dE error of a hi-only read.
template <bool F32>
__device__ __forceinline__ void yk_ld8(const YdChain &c, int set, int kk, double *C) {
    size_t base = (size_t)set * 8 * c.p.nk + kk;
    if (F32 && set == 0) {
#pragma unroll
        for (int m = 0; m < 8; ++m)
            C[m] = (double)yd_g(c.Shi)[base + (size_t)m * c.p.nk] + (double)yd_g(c.Slo)[base + (size_t)m * c.p.nk];
        return;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        if (F32) C[m] = (double)yd_g(c.Shi)[base + (size_t)m * c.p.nk];
        else C[m] = yd_g(c.S)[base + (size_t)m * c.p.nk];
    }
}

// The same 8 components kept as loaded (FP32C: hi and, for the Coulomb set,
// lo words; FP64: the doubles), so that a sweep can issue the loads of a k
// well before it needs the values: converting at load time would wait for
// them there.
template <bool F32> struct YkRaw8 { double v[8]; };
template <> struct YkRaw8<true> { float h[8], l[8]; };

template <bool F32, bool LO>
__device__ __forceinline__ void yk_raw8_load(const YdChain &c, int set, int kk, YkRaw8<F32> &r) {
    const size_t base = (size_t)set * 8 * c.p.nk + kk;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const size_t idx = base + (size_t)m * c.p.nk;
        if constexpr (F32) {
            r.h[m] = yd_g(c.Shi)[idx];
            if (LO) r.l[m] = yd_g(c.Slo)[idx];
        } else {
            r.v[m] = yd_g(c.S)[idx];
        }
    }
}

// Values of yk_ld8 from a raw load (LO: the Coulomb set's hi + lo).
template <bool F32, bool LO>
__device__ __forceinline__ void yk_raw8_get(const YkRaw8<F32> &r, double *C) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        if constexpr (F32) C[m] = LO ? (double)r.h[m] + (double)r.l[m] : (double)r.h[m];
        else C[m] = r.v[m];
    }
}

// Accumulator layout of the sweeps: [0] dU_C, [1] dU_LJ, [2 + 3a + d] the
// k-space force on atom a (Coulomb and LJ parts summed per k; the total over
// k differs from the reference's separate sums by rounding only).
#define YK_NACC (2 + 3 * YD_MAXA)
#define YK_FK(a, d) (2 + 3 * (a) + (d))

// acc[base + 3a + d] += f[d] for a runtime site index a, written so that acc
// stays in registers (every index is a compile-time constant).
template <int N>
__device__ __forceinline__ void yk_scatter3(double (&acc)[N], int base, int a, const double *f) {
#pragma unroll
    for (int b = 0; b < YD_MAXA; ++b)
        if (b == a) {
            acc[base + 3 * b + 0] += f[0];
            acc[base + 3 * b + 1] += f[1];
            acc[base + 3 * b + 2] += f[2];
        }
}

// One sweep over k.  DELTA = false: k-space forces of the molecule at its
// old positions against S (MC.f90:963-968).  DELTA = true: dU_C, dU_LJ
// against S and forces at new positions against S + delta
// (MC.f90:1136-1173).  Per k, each site's q sin/cos (SinCosKR_fill) is
// evaluated once and reused by the delta sums, the Coulomb force and the
// LJ force of every type (LJ part = Coulomb part / q, RhoSquared.f90:
// 173-176).  Systems with up to YK_TPF LJ types load every S(k) value and
// beta of a k before any arithmetic; larger T streams the LJ sets.
#define YK_TPF 3
template <bool F32, bool DELTA, int NA>
__device__ void yk_sweep(const YdChain &c, const YkMol &m, double (&acc)[YK_NACC], int rank, int size) {
    const int nk = c.p.nk, T = c.p.ntype;
    const double *tabs = DELTA ? m.tab_n : m.tab_o;
    const bool pf = T <= YK_TPF;
    for (int kk = rank; kk < nk; kk += size) {
        int kxyz = yd_g(c.kxyz)[kk], kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        double bt = yd_g(c.beta)[kk];
        double C[8];
        double CP[YK_TPF][8], BL[YK_TPF][YK_MAXP];
        yk_ld8<F32>(c, 0, kk, C);
        if (pf) {
#pragma unroll
            for (int t1 = 0; t1 < YK_TPF; ++t1) {
                if (t1 < T) {
                    yk_ld8<F32>(c, 1 + t1, kk, CP[t1]);
#pragma unroll
                    for (int q = 0; q < YK_MAXP; ++q)
                        BL[t1][q] = q < m.np ? yd_g(c.beta_lj)[((size_t)t1 * T + (m.ptype[q] - 1)) * nk + kk] : 0.0;
                }
            }
        }
        // q sin/cos of every site at the sweep's positions; with DELTA also
        // the new-minus-old sums of the Coulomb set and of each LJ slot, in
        // site order (RhoSquared_addAtom into nulified sums, then
        // RhoSquared_sub: MC.f90:946-958, 1096-1125)
        double kn[NA][8], dc[8], dl[YK_MAXP][8];
        {
            double ns[8], os[8], nl[YK_MAXP][8], ol[YK_MAXP][8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                ns[i] = 0.0; os[i] = 0.0;
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q) { nl[q][i] = 0.0; ol[q][i] = 0.0; }
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const double qa = m.q[a];
                const int ta = m.typ[a];
                const double *tn = tabs + a * 6 * YD_KP;
                if (qa != 0.0) {
                    yd_kr(tn, qa, kx, ky, kz, kn[a]);
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i) kn[a][i] = 0.0;
                }
                if (DELTA) {
                    const double *to = m.tab_o + a * 6 * YD_KP;
                    double ko[8];
                    if (qa != 0.0) {
                        yd_kr(to, qa, kx, ky, kz, ko);
#pragma unroll
                        for (int i = 0; i < 8; ++i) { ns[i] = ns[i] + kn[a][i]; os[i] = os[i] + ko[i]; }
                    }
                    if (ta > 0) {
                        double ln[8], lo[8];
                        if (qa != 0.0) {
                            double inv = 1.0 / qa;
#pragma unroll
                            for (int i = 0; i < 8; ++i) { ln[i] = kn[a][i] * inv; lo[i] = ko[i] * inv; }
                        } else {
                            yd_kr(tn, 1.0, kx, ky, kz, ln);
                            yd_kr(to, 1.0, kx, ky, kz, lo);
                        }
#pragma unroll
                        for (int q = 0; q < YK_MAXP; ++q)
                            if (q == m.lslot[a]) {
#pragma unroll
                                for (int i = 0; i < 8; ++i) { nl[q][i] = nl[q][i] + ln[i]; ol[q][i] = ol[q][i] + lo[i]; }
                            }
                    }
                }
            }
            if (DELTA) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    dc[i] = ns[i] - os[i];
#pragma unroll
                    for (int q = 0; q < YK_MAXP; ++q) dl[q][i] = nl[q][i] - ol[q][i];
                }
            }
        }
        if (DELTA) {
            acc[0] += yd_du_k(C, dc, ky, kz) * bt;
#pragma unroll
            for (int i = 0; i < 8; ++i) C[i] = C[i] + dc[i];
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (m.q[a] == 0.0) continue;
            double f[3] = {0, 0, 0};
            yd_force_k(C, kn[a], kx, ky, kz, bt, f);
            acc[YK_FK(a, 0)] += f[0];
            acc[YK_FK(a, 1)] += f[1];
            acc[YK_FK(a, 2)] += f[2];
        }
        if (T == 0 || m.np == 0) continue;
        // one LJ set t1 against the molecule (the body of calc_dU_LJ's t1
        // loop and of ForceKSpace_LJ's type loop)
        auto lj_set = [&](int t1, double *C1, const double *bl) {
            if (DELTA) {
                int s1 = -1;
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q < m.np && m.ptype[q] == t1 + 1) s1 = q;
                double tw[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    double d1 = 0.0;
#pragma unroll
                    for (int q = 0; q < YK_MAXP; ++q)
                        if (q == s1) d1 = dl[q][i];
                    tw[i] = C1[i] + C1[i] + d1;
                    C1[i] = C1[i] + d1;
                }
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q < m.np) acc[1] += yd_dulj_k(tw, dl[q], ky, kz) * bl[q];
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int ta = m.typ[a];
                if (ta <= 0) continue;
                const double qa = m.q[a];
                double krl[8];
                if (qa != 0.0) {
                    double inv = 1.0 / qa;
#pragma unroll
                    for (int i = 0; i < 8; ++i) krl[i] = kn[a][i] * inv;
                } else {
                    yd_kr(tabs + a * 6 * YD_KP, 1.0, kx, ky, kz, krl);
                }
                double f[3] = {0, 0, 0};
                // beta_LJ is symmetric in its type pair: row (ta-1, t1) = (t1, ta-1)
                double b = 0.0;
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q == m.lslot[a]) b = bl[q];
                if (m.lslot[a] < 0) b = yd_g(c.beta_lj)[((size_t)(ta - 1) * T + t1) * nk + kk];
                yd_force_k(C1, krl, kx, ky, kz, b, f);
                acc[YK_FK(a, 0)] += f[0];
                acc[YK_FK(a, 1)] += f[1];
                acc[YK_FK(a, 2)] += f[2];
            }
        };
        if (pf) {
#pragma unroll
            for (int t1 = 0; t1 < YK_TPF; ++t1)
                if (t1 < T) lj_set(t1, CP[t1], BL[t1]);
        } else {
            for (int t1 = 0; t1 < T; ++t1) {
                double C1[8], bl[YK_MAXP];
                yk_ld8<F32>(c, 1 + t1, kk, C1);
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    bl[q] = q < m.np ? yd_g(c.beta_lj)[((size_t)t1 * T + (m.ptype[q] - 1)) * nk + kk] : 0.0;
                lj_set(t1, C1, bl);
            }
        }
    }
}

// RhoSquared_add on acceptance (MC.f90:1245): S += delta for the Coulomb set
// and the molecule's LJ sets.  FP32C keeps (hi, lo) as a two-sum pair.  All
// touched values of a k are loaded before the first store.
template <bool F32>
__device__ __forceinline__ void yk_commit_load(const YdChain &c, const YkMol &m, int kk, int &kxyz,
                                               YkRaw8<F32> (&v)[1 + YK_MAXP]) {
    kxyz = yd_g(c.kxyz)[kk];
#pragma unroll
    for (int set = 0; set <= YK_MAXP; ++set) {
        if (set > m.np) break;
        yk_raw8_load<F32, true>(c, set == 0 ? 0 : m.ptype[set - 1], kk, v[set]);
    }
}

// The loads of a thread's next k are issued before the current k's stores
// (a different kk, so they never alias), which keeps one k of loads in
// flight while the current one is updated.
template <bool F32, bool RANGE = false>
__device__ void yk_commit(const YdChain &c, const YkMol &m, int rank, int size, int kend = -1) {
    const int nk = c.p.nk, ke = RANGE ? kend : nk;
    int kxyz_n = 0;
    YkRaw8<F32> vn[1 + YK_MAXP];  // the next k's values as loaded, converted where they are used
    if (rank < ke) yk_commit_load<F32>(c, m, rank, kxyz_n, vn);
    for (int kk = rank; kk < ke; kk += size) {
        const int kxyz = kxyz_n, kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        YkRaw8<F32> vr[1 + YK_MAXP];
#pragma unroll
        for (int set = 0; set <= YK_MAXP; ++set) vr[set] = vn[set];
        if (kk + size < ke) yk_commit_load<F32>(c, m, kk + size, kxyz_n, vn);
#pragma unroll
        for (int set = 0; set <= YK_MAXP; ++set) {
            if (set > m.np) break;
            double v[8];
            yk_raw8_get<F32, true>(vr[set], v);
            double d[8];
            if (set == 0) yk_delta_coul(m, kx, ky, kz, d);
            else yk_delta_lj(m, set - 1, kx, ky, kz, d);
            int sidx = set == 0 ? 0 : m.ptype[set - 1];
            size_t base = (size_t)sidx * 8 * nk + kk;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                size_t idx = base + (size_t)i * nk;
                double w = v[i] + d[i];
                if (F32) {
                    float hi = (float)w;
                    yd_g(c.Shi)[idx] = hi;
                    yd_g(c.Slo)[idx] = (float)(w - (double)hi);
                } else {
                    yd_g(c.S)[idx] = w;
                }
            }
        }
    }
}

// Real-space partial sum of the molecule at new positions against all
// other atoms (EwaldSumRealSpace.f90:338-498, partial mode).
// Accumulators: [a] uu, [4 + 3a + d] F, [16..18] uu_ew, uu_lj6, uu_lj12.
#define YK_NRS (4 * YD_MAXA + 3)

// Adds one pair of site a into the partial-sum accumulators (constant
// indices after unrolling, so acc stays in registers).
template <int N>
__device__ __forceinline__ void yk_rs_acc(double (&acc)[N], int a, const YdPair &pr, double xij, double yij,
                                          double zij) {
#pragma unroll
    for (int b = 0; b < YD_MAXA; ++b)
        if (b == a) {
            acc[b] += pr.pot;
            acc[YD_MAXA + 3 * b + 0] += xij * pr.FR;
            acc[YD_MAXA + 3 * b + 1] += yij * pr.FR;
            acc[YD_MAXA + 3 * b + 2] += zij * pr.FR;
        }
    acc[4 * YD_MAXA + 0] += pr.potew;
    acc[4 * YD_MAXA + 1] += pr.potlj6;
    acc[4 * YD_MAXA + 2] += pr.potlj12;
}
template <int NA>
__device__ void yk_rs_new(const YdChain &c, const YkMol &m, const double *sig6, const double *feps,
                          double (&acc)[YK_NRS], int &ov) {
    const int N = c.p.natom, T = c.p.ntype, me = m.imol + 1;
    for (int j = threadIdx.x; j < N; j += (int)blockDim.x) {
        int jm = yd_g(c.molnum)[j];
        if (jm == me) continue;
        int jt = yd_g(c.type)[j];
        double qj = yd_g(c.q)[j], dhj = yd_g(c.dh)[j], xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xij = yd_wrap(m.xn[a] - xj), yij = yd_wrap(m.yn[a] - yj), zij = yd_wrap(m.zn[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (m.dh[a] + dhj) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) ov = 1;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            yk_rs_acc(acc, a, pr, xij, yij, zij);
        }
    }
}

// Fused real-space pass of a trial move over every other atom j: the new
// partial sum (EwaldSumRealSpace.f90:338-498 for the molecule at its new
// sites) and the list of atoms within the cut-off of a new or an old site,
// which are the only atoms whose per-atom arrays change on acceptance
// (MC.f90:1251-1275).  *count may exceed YK_NBCAP; the list is then partial.
template <int NA>
__device__ void yk_rs_move(const YdChain &c, const YkMol &m, const double *sig6, const double *feps,
                           double (&acc)[YK_NRS], int &ov, int *list, int *count, int rank, int size) {
    const int N = c.p.natom, T = c.p.ntype, me = m.imol + 1;
    for (int j = rank; j < N; j += size) {
        int jm = yd_g(c.molnum)[j];
        if (jm == me) continue;
        int jt = yd_g(c.type)[j];
        double qj = yd_g(c.q)[j], dhj = yd_g(c.dh)[j], xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
        bool cand = false;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xo = yd_wrap(m.xo[a] - xj), yo = yd_wrap(m.yo[a] - yj), zo = yd_wrap(m.zo[a] - zj);
            if (xo * xo + yo * yo + zo * zo < c.p.rmax2) cand = true;
            double xij = yd_wrap(m.xn[a] - xj), yij = yd_wrap(m.yn[a] - yj), zij = yd_wrap(m.zn[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (m.dh[a] + dhj) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) ov = 1;
            if (r2 >= c.p.rmax2) continue;
            cand = true;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            yk_rs_acc(acc, a, pr, xij, yij, zij);
        }
        if (cand) {
            int slot = atomicAdd(count, 1);
            if (slot < YK_NBCAP) list[slot] = j;
        }
    }
}

// Accepted move, one atom j: its per-atom arrays take the molecule's new
// minus old contribution, summed over the molecule's sites in index order as
// the back arrays of the two partial sums hold them; acc[0..2] gather the
// old partial sum's uu_ew, uu_lj6, uu_lj12.
template <bool AOS>
__device__ __forceinline__ void yk_rs_apply(const YdChain &c, const YkMol &m, const double *sig6,
                                            const double *feps, int j, double (&acc)[3]) {
    const int T = c.p.ntype;
    int jt = yd_g(c.type)[j];
    double qj = yd_g(c.q)[j], xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
    double nb[4] = {0, 0, 0, 0}, ob[4] = {0, 0, 0, 0};
    for (int pass = 0; pass < 2; ++pass) {
        for (int a = 0; a < m.na; ++a) {
            int it = m.typ[a];
            double xa = pass == 0 ? m.xn[a] : m.xo[a], ya = pass == 0 ? m.yn[a] : m.yo[a],
                   za = pass == 0 ? m.zn[a] : m.zo[a];
            double xij = yd_wrap(xa - xj), yij = yd_wrap(ya - yj), zij = yd_wrap(za - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            double fx = xij * pr.FR, fy = yij * pr.FR, fz = zij * pr.FR;
            if (pass == 0) {
                nb[0] = nb[0] + pr.pot; nb[1] = nb[1] - fx; nb[2] = nb[2] - fy; nb[3] = nb[3] - fz;
            } else {
                ob[0] = ob[0] + pr.pot; ob[1] = ob[1] - fx; ob[2] = ob[2] - fy; ob[3] = ob[3] - fz;
                acc[0] += pr.potew;
                acc[1] += pr.potlj6;
                acc[2] += pr.potlj12;
            }
        }
    }
    if (AOS) {  // the chain kernel's launch copy of the per-atom arrays
        auto u = yd_g(reinterpret_cast<double4 *>(c.uf)) + j;
        double4 v = *u;
        v.x = v.x + nb[0] - ob[0];
        v.y = v.y + nb[1] - ob[1];
        v.z = v.z + nb[2] - ob[2];
        v.w = v.w + nb[3] - ob[3];
        *u = v;
        return;
    }
    yd_g(c.uu)[j] = yd_g(c.uu)[j] + nb[0] - ob[0];
    yd_g(c.Fx)[j] = yd_g(c.Fx)[j] + nb[1] - ob[1];
    yd_g(c.Fy)[j] = yd_g(c.Fy)[j] + nb[2] - ob[2];
    yd_g(c.Fz)[j] = yd_g(c.Fz)[j] + nb[3] - ob[3];
}

// Commit of the real-space arrays over every other atom (the candidate list
// overflowed); acc[0..2] gather the old partial sum's uu_ew, uu_lj6, uu_lj12.
template <bool AOS = false>
__device__ void yk_rs_commit_all(const YdChain &c, const YkMol &m, const double *sig6, const double *feps,
                                 double (&acc)[3], int rank, int size) {
    const int me = m.imol + 1;
    for (int j = rank; j < c.p.natom; j += size)
        if (yd_g(c.molnum)[j] != me) yk_rs_apply<AOS>(c, m, sig6, feps, j, acc);
}

// Real-space pass of a trial move in two stages.
//  1. Scan: every other atom j is tested against the molecule's old and new
//     sites; survivors are marked in an LDS bitmap and compacted into the
//     candidate list.  The workgroup kernel tests packed positions (below)
//     in FP32 against a bound grown by the packing error, the wave kernel
//     float4 positions with a margin of 1e-5 box units^2 on r^2; either way
//     no atom the FP64 test accepts is lost.
//  2. Exact pass over the compacted list: the FP64 pair terms of the new
//     sites (EwaldSumRealSpace.f90:338-498, partial mode) and the overlap
//     test, with every lane busy.  The same list then drives the commit of
//     the per-atom arrays (only atoms within the cut-off of an old or new
//     site change, MC.f90:1251-1275).  When the list overflows the caller
//     uses the full FP64 pass yk_rs_move_n and the full commit.

// Packed position of the workgroup kernel's scan: the fractional part of
// each coordinate (box units) in 10 bits, x | y << 10 | z << 20, one 32-bit
// word per atom (so the copy fits in LDS next to two chains' workgroup data).
// A packed coordinate stands for the centre of its 1/1024 bin, within
// 0.5/1024 of the position, so the packed position is within sqrt(3)/2048
// box units of the true one (modulo the box).
__device__ __forceinline__ unsigned int yk_pack_pos(double x, double y, double z) {
    auto bin = [](double v) {
        const int i = (int)((v - floor(v)) * 1024.0);
        return (unsigned int)(i < 0 ? 0 : i > 1023 ? 1023 : i);
    };
    return bin(x) | (bin(y) << 10) | (bin(z) << 20);
}
__device__ __forceinline__ float4 yk_unpack_pos(unsigned int v) {
    const float s = 1.0f / 1024.0f;
    return make_float4(((float)(v & 1023u) + 0.5f) * s, ((float)((v >> 10) & 1023u) + 0.5f) * s,
                       ((float)(v >> 20) + 0.5f) * s, 0.0f);
}
// r^2 bound of the scan on packed positions: the cut-off grown by the
// packing error (8.5e-4 > sqrt(3)/2048 box units), squared, plus 1e-5 box^2
// for FP32 rounding.
__device__ __forceinline__ float yk_scan_r2(const YdChain &c) {
    const float r = sqrtf((float)c.p.rmax2) + 8.5e-4f;
    return r * r + 1e-5f;
}

// FP32 distance test of one atom against the molecule's old and new sites
// (margin 1e-5 box^2 on r^2, far above FP32 rounding).  The image is
// d - rint(d), the nearest one: it equals the pair sums' single-wrap image
// for |d| < 1.5 and is never farther, so every atom the exact pass can
// count passes the test.
template <int NA>
__device__ __forceinline__ bool yk_near(const float (&xof)[NA], const float (&yof)[NA], const float (&zof)[NA],
                                        const float (&xnf)[NA], const float (&ynf)[NA], const float (&znf)[NA],
                                        float4 pj, float r2f) {
    bool near = false;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        float dx = xnf[a] - pj.x, dy = ynf[a] - pj.y, dz = znf[a] - pj.z;
        dx -= rintf(dx); dy -= rintf(dy); dz -= rintf(dz);
        near |= dx * dx + dy * dy + dz * dz < r2f;
        float ex = xof[a] - pj.x, ey = yof[a] - pj.y, ez = zof[a] - pj.z;
        ex -= rintf(ex); ey -= rintf(ey); ez -= rintf(ez);
        near |= ex * ex + ey * ey + ez * ez < r2f;
    }
    return near;
}

// Cell-list form of the scan: only the atoms of the cells that can hold an
// atom within the scan radius of an old or new site are tested.  Cells are
// cubes of width 1/nc box units, any width (cells narrower than the cut-off
// give a finer stencil).  The visited cells: per (x, y) column of the
// range the sites' extent +- rv spans, the contiguous z run whose cells are
// within rv of the sites' bounding box (rv = the scan radius plus the
// packing error and rounding margins), so every atom the full scan would
// mark is tested and the candidate bitmap is the full scan's.  An axis
// whose range spans nc cells or more is visited whole (each cell once).
// Wave 0 builds the list and its bucket offsets lane-parallel (cl_cell:
// cell ids, cl_start: running atom offsets); then all threads scan the
// cells' atoms as one flat range.  cl_* are the workgroup's LDS arrays; all
// threads call this.  Returns false (the same in every thread) when the
// list would exceed YK_MAXCL cells: the caller scans every atom instead.
// split_G > 1 (a grid chain): workgroup split_g tests only its share
// [total g / G, total (g + 1) / G) of the visited cells' atoms, at most
// YK_NBCAP of them; when the cells hold more than G YK_NBCAP atoms it
// returns false (the same in every workgroup) and the caller scans atom
// ranges instead.
// LIST (grid chain): instead of marking bm, the candidates go straight to
// list[0 .. *nlist) in the order of the visited atoms, with wsum
// (YK_SCANU ints per wave) as the placement's LDS.
template <int NA, bool SPLIT = false, bool LIST = false>
__device__ __attribute__((noinline)) bool yk_rs_scan_cells(const YdChain &c, const YkMol &m, const unsigned int *pos, unsigned int *bm, int *cl_cell,
                                 int *cl_start, int *ncl_out, int split_g = 0, int split_G = 1,
                                 int *list = nullptr, int *nlist = nullptr, int *wsum = nullptr) {
    const int nc = c.nc, tid = threadIdx.x;
    const unsigned int f0 = (unsigned int)m.first, na = (unsigned int)m.na;
    const double w = 1.0 / (double)nc;
    // the scan radius (yk_scan_r2) + the packing error of the tested
    // position (< 8.5e-4) + FP32 and FP64 rounding, box units
    const double rv = sqrt((double)yk_scan_r2(c)) + 1.2e-3, rv2 = rv * rv;
    double lo[3], hi[3];
    int st[3], ln[3];
    bool whole[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double *po = d == 0 ? m.xo : d == 1 ? m.yo : m.zo;
        const double *pn = d == 0 ? m.xn : d == 1 ? m.yn : m.zn;
        const double ref = po[0];
        lo[d] = ref; hi[d] = ref;
        for (int k = 1; k < 2 * NA; ++k) {  // the other sites' nearest images around site 0
            const double v = ref + yd_wrap((k < NA ? po[k] : pn[k - NA]) - ref);
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = v > hi[d] ? v : hi[d];
        }
        const int i0 = (int)floor((lo[d] - rv + 0.5) * nc), i1 = (int)floor((hi[d] + rv + 0.5) * nc);
        whole[d] = i1 - i0 + 1 >= nc;
        st[d] = whole[d] ? 0 : i0;
        ln[d] = whole[d] ? nc : i1 - i0 + 1;
    }
    // distance from the sites' bounding box to unwrapped cell i on axis d
    auto gap = [&](int d, int i) {
        if (whole[d]) return 0.0;
        const double a = (double)i * w - 0.5, b = a + w;
        const double g1 = a - hi[d], g2 = lo[d] - b;
        return g1 > 0.0 ? g1 : g2 > 0.0 ? g2 : 0.0;
    };
    auto wrapc = [&](int i) { i %= nc; return i < 0 ? i + nc : i; };
    if (tid < 64) {
        const int lane = tid, ncol = ln[0] * ln[1];
        int total = 0, over = 0;
        for (int col0 = 0; col0 < ncol; col0 += 64) {
            const int col = col0 + lane;
            int zs = 0, zl = 0, cx = 0, cy = 0;
            if (col < ncol) {
                const int ix = st[0] + col / ln[1], iy = st[1] + col % ln[1];
                cx = wrapc(ix); cy = wrapc(iy);
                const double gx = gap(0, ix), gy = gap(1, iy), rem = rv2 - gx * gx - gy * gy;
                if (rem > 0.0) {
                    if (whole[2]) {
                        zs = 0; zl = nc;
                    } else {
                        const double rz = sqrt(rem);
                        const int z0 = (int)floor((lo[2] - rz + 0.5) * nc), z1 = (int)floor((hi[2] + rz + 0.5) * nc);
                        zs = z0; zl = z1 - z0 + 1;
                        if (zl >= nc) { zs = 0; zl = nc; }
                    }
                }
            }
            int incl = zl;  // inclusive scan of the run lengths inside the wave
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            const int sum = __shfl(incl, 63, 64);
            if (total + sum > YK_MAXCL) { over = 1; break; }
            const int e0 = total + incl - zl;
            for (int k = 0; k < zl; ++k) cl_cell[e0 + k] = (cx * nc + cy) * nc + wrapc(zs + k);
            total += sum;
        }
        if (!over) {
            // the buckets' atom counts and their running offsets
            int carry = 0;
            if (lane == 0) cl_start[0] = 0;
            for (int i0 = 0; i0 < total; i0 += 64) {
                const int i = i0 + lane;
                const int cnt = i < total ? yd_g(c.cell_cnt)[cl_cell[i]] : 0;
                int incl = cnt;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const int v = __shfl_up(incl, off, 64);
                    if (lane >= off) incl += v;
                }
                if (i < total) cl_start[i + 1] = carry + incl;
                carry += __shfl(incl, 63, 64);
            }
        }
        if (lane == 0) *ncl_out = over ? -1 : total;
    }
    __syncthreads();
    const int ncl = *ncl_out;
    if (ncl < 0) return false;  // uniform: the caller scans every atom instead
    int total = cl_start[ncl], tb = 0;
    if (SPLIT && split_G > 1) {
        if (total > split_G * YK_NBCAP) return false;
        tb = (int)((long long)total * split_g / split_G);
        total = (int)((long long)total * (split_g + 1) / split_G);
    }
    const float r2f = yk_scan_r2(c);
    float xof[NA], yof[NA], zof[NA], xnf[NA], ynf[NA], znf[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        xof[a] = (float)m.xo[a]; yof[a] = (float)m.yo[a]; zof[a] = (float)m.zo[a];
        xnf[a] = (float)m.xn[a]; ynf[a] = (float)m.yn[a]; znf[a] = (float)m.zn[a];
    }
    const int bs = (int)blockDim.x;
    if (LIST) {
        // thread tid tests items base + u bs + tid (one load round trip for
        // all of them); wave ballots count each u's finds and a block pass
        // over the counts (in wsum: YK_SCANU ints per wave) places them in
        // ascending item order
        const int lane = tid & 63, w = tid >> 6, nw = bs >> 6;
        const unsigned long long below = (1ull << lane) - 1ull;
        int run = 0;  // candidates listed so far (the same in every thread)
        for (int base = tb; base < total; base += YK_SCANU * bs) {  // uniform trip count
            int jj[YK_SCANU];
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) {
                const int t = base + u * bs + tid;
                jj[u] = -1;
                if (t < total) {
                    int lo = 0, hi = ncl - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (cl_start[mid] <= t) lo = mid;
                        else hi = mid - 1;
                    }
                    jj[u] = yd_g(c.cell_atoms)[(size_t)cl_cell[lo] * c.ccap + (t - cl_start[lo])];
                }
            }
            unsigned int pj[YK_SCANU];
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) pj[u] = jj[u] >= 0 ? pos[jj[u]] : 0u;
            int pre[YK_SCANU];
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) {
                const int j = jj[u];
                const bool take = j >= 0 && (unsigned int)j - f0 >= na &&
                                  yk_near<NA>(xof, yof, zof, xnf, ynf, znf, yk_unpack_pos(pj[u]), r2f);
                if (!take) jj[u] = -1;
                const unsigned long long bal = __ballot(take);
                pre[u] = __popcll(bal & below);
                if (lane == 0) wsum[w * YK_SCANU + u] = __popcll(bal);
            }
            __syncthreads();
            int o = run;
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) {
                int all = 0, bw = 0;
                for (int v = 0; v < nw; ++v) {
                    const int x = wsum[v * YK_SCANU + u];
                    all += x;
                    bw += v < w ? x : 0;
                }
                if (jj[u] >= 0) list[o + bw + pre[u]] = jj[u];
                o += all;
            }
            run = o;
            __syncthreads();  // wsum is reused
        }
        if (tid == 0) *nlist = run;
        return true;
    }
    for (int t0 = tb + tid; t0 < total; t0 += YK_SCANU * bs) {
        int jj[YK_SCANU];  // atom indices first, then their positions, YK_SCANU loads at a time
#pragma unroll
        for (int u = 0; u < YK_SCANU; ++u) {
            const int t = t0 + u * bs;
            jj[u] = -1;
            if (t < total) {
                int lo = 0, hi = ncl - 1;  // the bucket holding flat index t
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (cl_start[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                jj[u] = yd_g(c.cell_atoms)[(size_t)cl_cell[lo] * c.ccap + (t - cl_start[lo])];
            }
        }
        unsigned int pj[YK_SCANU];
#pragma unroll
        for (int u = 0; u < YK_SCANU; ++u) pj[u] = jj[u] >= 0 ? pos[jj[u]] : 0u;
#pragma unroll
        for (int u = 0; u < YK_SCANU; ++u) {
            const int j = jj[u];
            if (j >= 0 && (unsigned int)j - f0 >= na &&
                yk_near<NA>(xof, yof, zof, xnf, ynf, znf, yk_unpack_pos(pj[u]), r2f))
                atomicOr(&bm[j >> 5], 1u << (j & 31));
        }
    }
    return true;
}

// pos: packed positions (yk_pack_pos); the molecule's own atoms, indices
// first .. first + na - 1, are skipped.
// RANGE: only atoms [a0, a1), marked in bm relative to word wb.
template <int NA, bool RANGE = false>
__device__ __forceinline__ void yk_rs_scan(const YdChain &c, const YkMol &m, const unsigned int *pos, unsigned int *bm, int a0 = 0,
                           int a1 = -1, int wb = 0) {
    if (!RANGE) { a0 = 0; wb = 0; }
    const int N = RANGE ? a1 : c.p.natom, me = m.imol + 1;
    if (pos) {
        const float r2f = yk_scan_r2(c);
        const unsigned int f0 = (unsigned int)m.first, na = (unsigned int)m.na;
        float xof[NA], yof[NA], zof[NA], xnf[NA], ynf[NA], znf[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            xof[a] = (float)m.xo[a]; yof[a] = (float)m.yo[a]; zof[a] = (float)m.zo[a];
            xnf[a] = (float)m.xn[a]; ynf[a] = (float)m.yn[a]; znf[a] = (float)m.zn[a];
        }
        // YK_SCANU positions per thread are loaded before any is tested, so
        // that many loads are in flight at once when pos is in HBM
        const int bs = (int)blockDim.x;
        for (int j0 = a0 + (int)threadIdx.x; j0 < N; j0 += YK_SCANU * bs) {
            unsigned int pj[YK_SCANU];
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) {
                const int j = j0 + u * bs;
                pj[u] = j < N ? pos[j] : 0u;
            }
#pragma unroll
            for (int u = 0; u < YK_SCANU; ++u) {
                const int j = j0 + u * bs;
                if (j < N && (unsigned int)j - f0 >= na &&
                    yk_near<NA>(xof, yof, zof, xnf, ynf, znf, yk_unpack_pos(pj[u]), r2f))
                    atomicOr(&bm[(j >> 5) - wb], 1u << (j & 31));
            }
        }
    } else {
        for (int j = a0 + (int)threadIdx.x; j < N; j += (int)blockDim.x) {
            double xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
            int jm = yd_g(c.molnum)[j];
            bool near = false;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                double xo = yd_wrap(m.xo[a] - xj), yo = yd_wrap(m.yo[a] - yj), zo = yd_wrap(m.zo[a] - zj);
                near |= xo * xo + yo * yo + zo * zo < c.p.rmax2;
                double xn = yd_wrap(m.xn[a] - xj), yn = yd_wrap(m.yn[a] - yj), zn = yd_wrap(m.zn[a] - zj);
                near |= xn * xn + yn * yn + zn * zn < c.p.rmax2;
            }
            if (near && jm != me) atomicOr(&bm[(j >> 5) - wb], 1u << (j & 31));
        }
    }
}

// Accumulators of the exact pass: the partial-sum layout (YK_NRS) followed
// by the old sites' uu_ew, uu_lj6, uu_lj12.
#define YK_NRX (YK_NRS + 3)

// Exact FP64 pass over the compacted candidate list.  For every candidate j
// it evaluates the molecule's new sites (the partial sum and the overlap
// test, EwaldSumRealSpace.f90:338-498) and its old sites, and stores j's
// back terms of both (uu, -F summed over the sites in index order, as the
// back arrays of the two partial sums hold them, MC.f90:1251-1275) as one
// difference new - old in c.rsb, so that an accepted move commits them
// without a second pair pass.
template <int NA>
__device__ void yk_rs_exact(const YdChain &c, const YkMol &m, const double *sig6, const double *feps,
                            const int *list, int count, double (&acc)[YK_NRX], int &ov, int rank, int size) {
    const int T = c.p.ntype;
    for (int i = rank; i < count; i += size) {
        int j = list[i];
        int jt = yd_g(c.type)[j];
        double qj = yd_g(c.q)[j], dhj = yd_g(c.dh)[j], xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
        double nb[4] = {0, 0, 0, 0}, ob[4] = {0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xij = yd_wrap(m.xn[a] - xj), yij = yd_wrap(m.yn[a] - yj), zij = yd_wrap(m.zn[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (m.dh[a] + dhj) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) ov = 1;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            yk_rs_acc(acc, a, pr, xij, yij, zij);
            nb[0] = nb[0] + pr.pot;
            nb[1] = nb[1] - xij * pr.FR;
            nb[2] = nb[2] - yij * pr.FR;
            nb[3] = nb[3] - zij * pr.FR;
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xij = yd_wrap(m.xo[a] - xj), yij = yd_wrap(m.yo[a] - yj), zij = yd_wrap(m.zo[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            ob[0] = ob[0] + pr.pot;
            ob[1] = ob[1] - xij * pr.FR;
            ob[2] = ob[2] - yij * pr.FR;
            ob[3] = ob[3] - zij * pr.FR;
            acc[YK_NRS + 0] += pr.potew;
            acc[YK_NRS + 1] += pr.potlj6;
            acc[YK_NRS + 2] += pr.potlj12;
        }
        auto b = yd_g(c.rsb) + (size_t)i * 4;
        b[0] = nb[0] - ob[0]; b[1] = nb[1] - ob[1]; b[2] = nb[2] - ob[2]; b[3] = nb[3] - ob[3];
    }
}

// Commit of the per-atom arrays from the stored back terms of the list.
// Four candidates per thread are loaded before the first store (the list
// holds distinct atoms, so no store of one candidate aliases another's load).
// AOS: on the chain kernel's launch copy (c.uf) of the per-atom arrays.
template <bool AOS = false>
__device__ void yk_rs_commit_list(const YdChain &c, const int *list, int count, int rank, int size) {
    constexpr int U = 4;
    for (int i0 = rank; i0 < count; i0 += U * size) {
        int jj[U];
        double b[U][4], v[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * size;
            jj[u] = i < count ? list[i] : -1;
            if (jj[u] < 0) continue;
            auto bi = yd_g(c.rsb) + (size_t)i * 4;
            const int j = jj[u];
            b[u][0] = bi[0]; b[u][1] = bi[1]; b[u][2] = bi[2]; b[u][3] = bi[3];
            if (AOS) {
                const double4 w = yd_g(reinterpret_cast<const double4 *>(c.uf))[j];
                v[u][0] = w.x; v[u][1] = w.y; v[u][2] = w.z; v[u][3] = w.w;
            } else {
                v[u][0] = yd_g(c.uu)[j]; v[u][1] = yd_g(c.Fx)[j]; v[u][2] = yd_g(c.Fy)[j]; v[u][3] = yd_g(c.Fz)[j];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = jj[u];
            if (j < 0) continue;
            if (AOS) {
                yd_g(reinterpret_cast<double4 *>(c.uf))[j] =
                    make_double4(v[u][0] + b[u][0], v[u][1] + b[u][1], v[u][2] + b[u][2], v[u][3] + b[u][3]);
                continue;
            }
            yd_g(c.uu)[j] = v[u][0] + b[u][0];
            yd_g(c.Fx)[j] = v[u][1] + b[u][1];
            yd_g(c.Fy)[j] = v[u][2] + b[u][2];
            yd_g(c.Fz)[j] = v[u][3] + b[u][3];
        }
    }
}

template <bool RANGE = false>
__device__ __forceinline__ void yk_rs_scan_n(const YdChain &c, const YkMol &m, const unsigned int *pos,
                                             unsigned int *bm, int a0 = 0, int a1 = -1, int wb = 0) {
    switch (m.na) {
        case 1: yk_rs_scan<1, RANGE>(c, m, pos, bm, a0, a1, wb); break;
        case 2: yk_rs_scan<2, RANGE>(c, m, pos, bm, a0, a1, wb); break;
        case 3: yk_rs_scan<3, RANGE>(c, m, pos, bm, a0, a1, wb); break;
        default: yk_rs_scan<4, RANGE>(c, m, pos, bm, a0, a1, wb); break;
    }
}
template <bool SPLIT = false>
__device__ __forceinline__ bool yk_rs_scan_cells_n(const YdChain &c, const YkMol &m, const unsigned int *pos,
                                                   unsigned int *bm, int *cl_cell, int *cl_start, int *ncl,
                                                   int split_g = 0, int split_G = 1) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: return yk_rs_scan_cells<1, SPLIT>(c, m, pos, bm, cl_cell, cl_start, ncl, split_g, split_G);
        case 2: return yk_rs_scan_cells<2, SPLIT>(c, m, pos, bm, cl_cell, cl_start, ncl, split_g, split_G);
        case 3: return yk_rs_scan_cells<3, SPLIT>(c, m, pos, bm, cl_cell, cl_start, ncl, split_g, split_G);
        default: return yk_rs_scan_cells<4, SPLIT>(c, m, pos, bm, cl_cell, cl_start, ncl, split_g, split_G);
    }
}

static_assert(YK_SCANU * YK_NW <= YK_BMWORDS, "the list scan's counts fit in the bitmap's LDS");
__device__ __forceinline__ bool yk_rs_scan_cells_list_n(const YdChain &c, const YkMol &m, const unsigned int *pos,
                                                        int *cl_cell, int *cl_start, int *ncl, int split_g, int split_G,
                                                        int *list, int *nlist, int *wsum) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: return yk_rs_scan_cells<1, true, true>(c, m, pos, nullptr, cl_cell, cl_start, ncl, split_g, split_G, list, nlist, wsum);
        case 2: return yk_rs_scan_cells<2, true, true>(c, m, pos, nullptr, cl_cell, cl_start, ncl, split_g, split_G, list, nlist, wsum);
        case 3: return yk_rs_scan_cells<3, true, true>(c, m, pos, nullptr, cl_cell, cl_start, ncl, split_g, split_G, list, nlist, wsum);
        default: return yk_rs_scan_cells<4, true, true>(c, m, pos, nullptr, cl_cell, cl_start, ncl, split_g, split_G, list, nlist, wsum);
    }
}

// Rebuilds the cell buckets from the current positions (all threads of the
// workgroup); *ok = 0 if a bucket overflowed.
__device__ void yk_cells_build(const YdChain &c, int *ok) {
    const int nc = c.nc, ncell = nc * nc * nc, N = c.p.natom;
    for (int i = threadIdx.x; i < ncell; i += (int)blockDim.x) yd_g(c.cell_cnt)[i] = 0;
    if (threadIdx.x == 0) *ok = 1;
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += (int)blockDim.x) {
        int id = (yd_cell1(yd_g(c.x)[i], nc) * nc + yd_cell1(yd_g(c.y)[i], nc)) * nc + yd_cell1(yd_g(c.z)[i], nc);
        int slot = atomicAdd(&yd_g(c.cell_cnt)[id], 1);
        yd_g(c.atom_cell)[i] = id;
        yd_g(c.atom_slot)[i] = slot;
        if (slot < c.ccap) yd_g(c.cell_atoms)[(size_t)id * c.ccap + slot] = i;
        else *ok = 0;
    }
    __syncthreads();
}

// Keeps the buckets current after an accepted move (thread 0): each moved
// atom that changed cell leaves its bucket (the bucket's last atom takes its
// slot) and joins the new one.  *ok = 0 on overflow.
__device__ void yk_cells_move(const YdChain &c, const YkMol &m, int *ok) {
    const int nc = c.nc, cap = c.ccap;
    int oldc[YD_MAXA];  // the sites' cells, loaded together (a site's update
                        // below writes no other site's cell)
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) oldc[a] = a < m.na ? yd_g(c.atom_cell)[m.first + a] : -1;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) {
        if (a >= m.na) break;
        const int g = m.first + a;
        const int id = (yd_cell1(m.xn[a], nc) * nc + yd_cell1(m.yn[a], nc)) * nc + yd_cell1(m.zn[a], nc);
        const int old = oldc[a];
        if (id == old) continue;
        const int s = yd_g(c.atom_slot)[g], last = yd_g(c.cell_cnt)[old] - 1;
        const int moved = yd_g(c.cell_atoms)[(size_t)old * cap + last];
        yd_g(c.cell_atoms)[(size_t)old * cap + s] = moved;
        yd_g(c.atom_slot)[moved] = s;
        yd_g(c.cell_cnt)[old] = last;
        const int ns = yd_g(c.cell_cnt)[id];
        if (ns >= cap) { *ok = 0; return; }
        yd_g(c.cell_atoms)[(size_t)id * cap + ns] = g;
        yd_g(c.atom_slot)[g] = ns;
        yd_g(c.atom_cell)[g] = id;
        yd_g(c.cell_cnt)[id] = ns + 1;
    }
}

// Ascending compaction of the candidate bitmap into list[] (deterministic:
// thread t owns a contiguous run of words, offsets from a block prefix sum).
// *count gets the number of candidates (may exceed YK_NBCAP: list partial).
// bm[k] holds the bits of atoms 32 (wb + k) .. 32 (wb + k) + 31, k < nwords
// (a grid chain's workgroup compacts the words of its own atom range).
__device__ void yk_compact(const unsigned int *bm, int nwords, int *list, int *count, int *wsum, int wb = 0) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int bs = blockDim.x, per = (nwords + bs - 1) / bs, k0 = tid * per;
    int cnt = 0;
    for (int k = 0; k < per; ++k)
        if (k0 + k < nwords) cnt += __popc(bm[k0 + k]);
    int incl = cnt;  // inclusive scan inside the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int i = 0; i < (bs >> 6); ++i) {
        if (i < w) base += wsum[i];
        total += wsum[i];
    }
    int pos = base + incl - cnt;
    for (int k = 0; k < per; ++k) {
        if (k0 + k >= nwords) break;
        unsigned int b = bm[k0 + k];
        while (b) {
            int bit = __ffs(b) - 1;
            b &= b - 1;
            if (pos < YK_NBCAP) list[pos] = ((wb + k0 + k) << 5) + bit;
            ++pos;
        }
    }
    if (tid == 0) *count = total;
}

__device__ __forceinline__ void yk_rs_exact_n(const YdChain &c, const YkMol &m, const double *sig6,
                                              const double *feps, const int *list, int count,
                                              double (&acc)[YK_NRX], int &ov, int rank, int size) {
    switch (m.na) {
        case 1: yk_rs_exact<1>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        case 2: yk_rs_exact<2>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        case 3: yk_rs_exact<3>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        default: yk_rs_exact<4>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
    }
}

// Dispatch on the molecule size so every per-site loop has a compile-time
// trip count (keeps the per-site accumulators in registers).
template <bool F32, bool DELTA>
__device__ __forceinline__ void yk_sweep_n(const YdChain &c, const YkMol &m, double (&acc)[YK_NACC], int rank,
                                           int size) {
    switch (m.na) {
        case 1: yk_sweep<F32, DELTA, 1>(c, m, acc, rank, size); break;
        case 2: yk_sweep<F32, DELTA, 2>(c, m, acc, rank, size); break;
        case 3: yk_sweep<F32, DELTA, 3>(c, m, acc, rank, size); break;
        default: yk_sweep<F32, DELTA, 4>(c, m, acc, rank, size); break;
    }
}
__device__ __forceinline__ void yk_rs_move_n(const YdChain &c, const YkMol &m, const double *sig6,
                                             const double *feps, double (&acc)[YK_NRS], int &ov, int *list,
                                             int *count, int rank, int size) {
    switch (m.na) {
        case 1: yk_rs_move<1>(c, m, sig6, feps, acc, ov, list, count, rank, size); break;
        case 2: yk_rs_move<2>(c, m, sig6, feps, acc, ov, list, count, rank, size); break;
        case 3: yk_rs_move<3>(c, m, sig6, feps, acc, ov, list, count, rank, size); break;
        default: yk_rs_move<4>(c, m, sig6, feps, acc, ov, list, count, rank, size); break;
    }
}
__device__ __forceinline__ void yk_rs_new_n(const YdChain &c, const YkMol &m, const double *sig6,
                                            const double *feps, double (&acc)[YK_NRS], int &ov) {
    switch (m.na) {
        case 1: yk_rs_new<1>(c, m, sig6, feps, acc, ov); break;
        case 2: yk_rs_new<2>(c, m, sig6, feps, acc, ov); break;
        case 3: yk_rs_new<3>(c, m, sig6, feps, acc, ov); break;
        default: yk_rs_new<4>(c, m, sig6, feps, acc, ov); break;
    }
}

// MonteCarloMove.f90 on thread 0 ------------------------------------------

__device__ __forceinline__ void yk_prod(double x1, double y1, double z1, double x2, double y2, double z2, double *o) {
    o[0] = y1 * z2 - z1 * y2;
    o[1] = z1 * x2 - x1 * z2;
    o[2] = x1 * y2 - y1 * x2;
}

// MonteCarloMove.f90:109-266 chooseDisplacement
__device__ void yk_displacement(const YdParams &p, YkMcm &m, YdRng &rng) {
    double fxo = m.fo[0], fyo = m.fo[1], fzo = m.fo[2], ftot = m.ftotold;
    double umax = p.lambda_f * ftot * p.dr, cold;
    if (umax > 1e-2) {
        double e = exp(-2.0 * umax);
        double t = (umax * (1.0 + e) - (1.0 - e)) / 2.0;
        cold = 3.0 * t / yd_p3(umax);
        double xksi = 0.0;
        for (int g = 0; xksi < 1e-10; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            xksi = rng.next();
        }
        double z = umax + log(xksi * t), u;
        if (z <= log(4.0)) {
            double sv = exp(z);
            u = pow(3.0 * sv, 1.0 / 3.0) - sv / 10.0;
        } else {
            u = z - log(z / 2.0);
        }
        for (int k = 1; k <= 10; ++k) {
            e = exp(-2.0 * u);
            double ch = (1.0 + e) / 2.0, sh = (1.0 - e) / 2.0;
            if (u < 1e-5) sh = u - u * u;
            double g = u * ch - sh;
            if (u < 1e-3) g = yd_p3(u) / 3.0;
            double h = u + log(g), h1 = u * sh / g;
            if (k < 10) u = u + (z - h) / h1;
        }
        double deltar = u / (p.lambda_f * ftot);
        xksi = 1.0;
        for (int g = 0; xksi > (1.0 - 1e-10) || xksi < 1e-10; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            xksi = rng.next();
        }
        double xcostet = 1.0 + log(xksi + (1.0 - xksi) * e) / u;
        if (u < 1e-7) xcostet = 1.0 - 2.0 * (1.0 - xksi);
        if (xcostet > 1.0 - 1e-14) xcostet = 1.0 - 1e-14;
        if (xcostet < -1.0 + 1e-14) xcostet = -1.0 + 1e-14;
        double xsintet = sqrt(1.0 - xcostet * xcostet);
        double w = -1.0, wv[3] = {0, 0, 0};
        for (int g = 0; w <= 1e-10; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            double v2 = 2.0, vx = 0, vy = 0, vz = 0;
            for (int g = 0; v2 > 1.0; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
                vx = 2.0 * rng.next() - 1.0;
                vy = 2.0 * rng.next() - 1.0;
                vz = 2.0 * rng.next() - 1.0;
                v2 = vx * vx + vy * vy + vz * vz;
            }
            yk_prod(vx, vy, vz, fxo, fyo, fzo, wv);
            w = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
        }
        m.d[0] = deltar * (xcostet * fxo / ftot + xsintet * wv[0] / w);
        m.d[1] = deltar * (xcostet * fyo / ftot + xsintet * wv[1] / w);
        m.d[2] = deltar * (xcostet * fzo / ftot + xsintet * wv[2] / w);
    } else {
        double v2 = 2.0, vx = 0, vy = 0, vz = 0;
        for (int g = 0; v2 > 1.0; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            vx = 2.0 * rng.next() - 1.0;
            vy = 2.0 * rng.next() - 1.0;
            vz = 2.0 * rng.next() - 1.0;
            v2 = vx * vx + vy * vy + vz * vz;
        }
        m.d[0] = p.dr * vx; m.d[1] = p.dr * vy; m.d[2] = p.dr * vz;
        umax = 0.0; cold = 1.0;
    }
    m.cold = cold;
    m.umaxold = umax;
}

// The draws of yk_displacement without its arithmetic: the same branches and
// rejection tests (none of which depends on the Newton solve), so that the
// rotation's draws can start where the displacement's end while another
// wave computes the displacement.
__device__ void yk_displacement_draws(const YdParams &p, YkMcm &m, YdRng &rng) {
    const double fxo = m.fo[0], fyo = m.fo[1], fzo = m.fo[2];
    const double umax = p.lambda_f * m.ftotold * p.dr;
    if (umax > 1e-2) {
        double xksi = 0.0;
        for (int g = 0; xksi < 1e-10; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            xksi = rng.next();
        }
        xksi = 1.0;
        for (int g = 0; xksi > (1.0 - 1e-10) || xksi < 1e-10; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            xksi = rng.next();
        }
        double w = -1.0, wv[3] = {0, 0, 0};
        for (int g = 0; w <= 1e-10; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            double v2 = 2.0, vx = 0, vy = 0, vz = 0;
            for (int g = 0; v2 > 1.0; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
                vx = 2.0 * rng.next() - 1.0;
                vy = 2.0 * rng.next() - 1.0;
                vz = 2.0 * rng.next() - 1.0;
                v2 = vx * vx + vy * vy + vz * vz;
            }
            yk_prod(vx, vy, vz, fxo, fyo, fzo, wv);
            w = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
        }
    } else {
        double v2 = 2.0;
        for (int g = 0; v2 > 1.0; ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            double vx = 2.0 * rng.next() - 1.0;
            double vy = 2.0 * rng.next() - 1.0;
            double vz = 2.0 * rng.next() - 1.0;
            v2 = vx * vx + vy * vy + vz * vz;
        }
    }
}

// MonteCarloMove.f90:269-444 chooseRotation (WAVE: called by a whole wave
// with identical arguments in every lane).
// MonteCarloMove.f90's Q_old of the torque-biased rotation: shi_sans_exp(
// vmax) / vmax, vmax = lambda_c |C_old| d_angle (1 when vmax <= 1e-2), as
// yk_rotation computes it; by a whole wave.
__device__ __forceinline__ double yk_qold(const YdParams &p, const YkMcm &m) {
    const double vmax = p.lambda_c * m.couplold * p.d_angle;
    return vmax > 1e-2 ? yd_shi_sans_exp_wave(vmax) / vmax : 1.0;
}
// QPRE: qold is computed elsewhere (another wave of the chain evaluates
// yk_qold at the same time) and left unset here.
template <bool WAVE, bool QPRE = false>
__device__ void yk_rotation(const YdParams &p, YkMcm &m, YdRng &rng) {
    double cx = m.co[0], cy = m.co[1], cz = m.co[2], cabs = m.couplold;
    double vmax = p.lambda_c * cabs * p.d_angle, qold, angle, vx, vy, vz;
    if (vmax > 1e-2) {
        double emax = exp(-2.0 * vmax), shmax = (1.0 - emax) / 2.0;
        double xksi = 1.0, proba = 0.0, v = 0.0, e = 1.0;
        for (int g = 0; xksi >= proba; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            v = 0.0;
            for (int g = 0; fabs(v) < 1e-10; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
                xksi = rng.next();
                double xs = xksi * shmax;
                v = vmax + log(xs + sqrt(xs * xs + emax));
                if (xs * xs < 1e-14 * emax) v = xs / sqrt(emax);
            }
            e = exp(-2.0 * v);
            proba = (1.0 - e) / v / (1.0 + e);
            if (v < 1e-5) proba = (2.0 * v - 2.0 * v * v) / v / (1.0 + e);
            xksi = rng.next();
        }
        double cosx = v / vmax;
        qold = QPRE ? 0.0 : (WAVE ? yd_shi_sans_exp_wave(vmax) : yd_shi_sans_exp(vmax)) / vmax;
        xksi = 1.0;
        for (int g = 0; xksi > (1.0 - 1e-10); ++g) {
            if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            xksi = rng.next();
        }
        angle = p.d_angle * (1.0 + log(xksi + (1.0 - xksi) * e) / v);
        if (v < 1e-7) angle = p.d_angle * (1.0 - 2.0 * (1.0 - xksi));
        double w = -1.0, wv[3] = {0, 0, 0};
        vx = vy = vz = 0;
        for (int g = 0; w <= 1e-10; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            double v2 = 2.0;
            for (int g = 0; v2 > 1.0 || v2 < 1e-10; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
                vx = 2.0 * rng.next() - 1.0;
                vy = 2.0 * rng.next() - 1.0;
                vz = 2.0 * rng.next() - 1.0;
                v2 = vx * vx + vy * vy + vz * vz;
            }
            yk_prod(vx, vy, vz, cx, cy, cz, wv);
            w = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
        }
        double sinx = sqrt(1.0 - cosx * cosx);
        vx = cx / cabs * cosx + wv[0] / w * sinx;
        vy = cy / cabs * cosx + wv[1] / w * sinx;
        vz = cz / cabs * cosx + wv[2] / w * sinx;
    } else {
        double v2 = 2.0;
        vx = vy = vz = 0;
        for (int g = 0; v2 > 1.0 || v2 < 1e-10; ++g) {
                if (g >= YK_TRIES || m.stall) { m.stall = 1; break; }
            vx = 2.0 * rng.next() - 1.0;
            vy = 2.0 * rng.next() - 1.0;
            vz = 2.0 * rng.next() - 1.0;
            v2 = vx * vx + vy * vy + vz * vz;
        }
        double v = sqrt(v2);
        vx = vx / v; vy = vy / v; vz = vz / v;
        angle = p.d_angle * (2.0 * rng.next() - 1.0);
        vmax = 0.0; qold = 1.0;
    }
    m.ca = cos(angle); m.sa = sin(angle); m.ct = vz; m.st = sqrt(1.0 - vz * vz);
    m.cphi = 1.0; m.sphi = 0.0;
    if (m.st > 0.0) { m.cphi = vx / m.st; m.sphi = vy / m.st; }
    m.v[0] = vx; m.v[1] = vy; m.v[2] = vz;
    m.vmaxold = vmax; m.angle = angle;
    if (!QPRE) m.qold = qold;
}

// MonteCarloMove.f90:56-106 init_new
// QPRE: qnew is computed elsewhere (yk_qnew) and left unset here.
template <bool WAVE, bool QPRE = false>
__device__ void yk_init_new(const YdParams &p, YkMcm &m) {
    m.ftotnew = sqrt(m.fn[0] * m.fn[0] + m.fn[1] * m.fn[1] + m.fn[2] * m.fn[2]);
    m.couplnew = sqrt(m.cn[0] * m.cn[0] + m.cn[1] * m.cn[1] + m.cn[2] * m.cn[2]);
    double umaxnew = p.lambda_f * m.ftotnew * p.dr, cnew;
    if (umaxnew > 1e-2) {
        double e = exp(-2.0 * umaxnew);
        double t = (umaxnew * (1.0 + e) - (1.0 - e)) / 2.0;
        cnew = 3.0 * t / yd_p3(umaxnew);
    } else {
        umaxnew = 0.0; cnew = 1.0;
    }
    m.umaxnew = umaxnew; m.cnew = cnew;
    double vmaxnew = p.lambda_c * m.couplnew * p.d_angle, qnew = 1.0;
    if (vmaxnew > 1e-2) {
        if (!QPRE) qnew = (WAVE ? yd_shi_sans_exp_wave(vmaxnew) : yd_shi_sans_exp(vmaxnew)) / vmaxnew;
    } else {
        vmaxnew = 0.0;
    }
    m.vmaxnew = vmaxnew;
    if (!QPRE) m.qnew = qnew;
}

// MonteCarloMove.f90:447-570 accept_or_decline
__device__ int yk_accept(const YdParams &p, const YkMcm &m, double dutot, YdRng &rng) {
    double exposu = -dutot;
    double exposf = -p.lambda_f * ((m.fo[0] + m.fn[0]) * m.d[0] + (m.fo[1] + m.fn[1]) * m.d[1] +
                                   (m.fo[2] + m.fn[2]) * m.d[2]);
    double exposfcorr = m.umaxold - m.umaxnew;
    double exposc = -p.lambda_c * m.angle *
                    ((m.co[0] + m.cn[0]) * m.v[0] + (m.co[1] + m.cn[1]) * m.v[1] + (m.co[2] + m.cn[2]) * m.v[2]);
    double exposccorr = m.vmaxold - m.vmaxnew;
    double factor = m.cold * m.qold / (m.cnew * m.qnew) * exp(exposu + exposf + exposfcorr + exposc + exposccorr);
    double xksi = rng.next();
    return xksi > factor ? 0 : 1;
}

// BiasedRandom.f90:80-109 BiasedRandom_choose: first i with
// beg(i) < rnd < end(i), else 1 (index 0 here).  With flat weights the
// interval bounds are i/nmol to rounding, so the interval is looked up next
// to floor(rnd*nmol) first; a binary search over the non-decreasing begins
// covers any other weights.
__device__ int yk_choose(const YdChain &c, double rnd) {
    const int n = c.p.nmol;
    int g = (int)(rnd * n);
    if (g > n - 1) g = n - 1;
    const int i0 = g > 0 ? g - 1 : 0, i1 = g + 1 < n ? g + 1 : n - 1;
    double b[3], e[3];  // the three intervals are loaded before the first test
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int i = i0 + u < i1 ? i0 + u : i1;
        b[u] = yd_g(c.freq_beg)[i];
        e[u] = yd_g(c.freq_end)[i];
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)
        if (i0 + u <= i1 && rnd > b[u] && rnd < e[u]) return i0 + u;
    // the intervals are disjoint: only the last one with beg < rnd can hold rnd
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (yd_g(c.freq_beg)[mid] < rnd) lo = mid;
        else hi = mid - 1;
    }
    if (rnd > yd_g(c.freq_beg)[lo] && rnd < yd_g(c.freq_end)[lo]) return lo;
    return 0;
}

// ------------------------------------------------------------------------
// Lean per-move passes of the wave kernel.  The same arithmetic as
// yk_sweep / yk_rs_exact, split so that fewer values are live at once: the
// Coulomb set and the LJ sets are swept in separate k loops, and the exact
// real-space pass visits the molecule's sites one at a time.  A wave then
// needs fewer registers, so more replicas share a SIMD.

// Coulomb set of a sweep (MC.f90:963-968 / 1136-1173, Coulomb parts):
// DELTA adds dU_C (acc[0]) and evaluates the forces against S + dS.
// The loads of a thread's next k (index, beta, S) are issued before the
// current k's arithmetic.
// RANGE: the sweep stops at kend instead of nk (a workgroup of a grid chain
// sweeps its own range of k); a separate instantiation, so that the
// one-workgroup kernels keep code specialised to the whole range.
template <bool F32, bool DELTA, int NA, bool RANGE = false>
__device__ void yl_sweep_coul(const YdChain &c, const YkMol &m, double (&acc)[YK_NACC], int rank, int size,
                              int kend = -1) {
    const int nk = RANGE ? kend : c.p.nk;
    const double *tabs = DELTA ? m.tab_n : m.tab_o;
    int kxyz_n = 0;
    double bt_n = 0.0;
    YkRaw8<F32> r_n;
    if (rank < nk) {
        kxyz_n = yd_g(c.kxyz)[rank];
        bt_n = yd_g(c.beta)[rank];
        yk_raw8_load<F32, true>(c, 0, rank, r_n);
    }
    for (int kk = rank; kk < nk; kk += size) {
        const int kxyz = kxyz_n, kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        const double bt = bt_n;
        const YkRaw8<F32> r = r_n;
        if (kk + size < nk) {
            kxyz_n = yd_g(c.kxyz)[kk + size];
            bt_n = yd_g(c.beta)[kk + size];
            yk_raw8_load<F32, true>(c, 0, kk + size, r_n);
        }
        double C[8];
        yk_raw8_get<F32, true>(r, C);
        double kn[NA][8];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (m.q[a] != 0.0) yd_kr_nz(tabs + a * 6 * YD_KP, m.q[a], kx, ky, kz, kn[a]);
        }
        if (DELTA) {
            double ns[8], os[8], ko[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) { ns[i] = 0.0; os[i] = 0.0; }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                if (m.q[a] == 0.0) continue;
                yd_kr_nz(m.tab_o + a * 6 * YD_KP, m.q[a], kx, ky, kz, ko);
#pragma unroll
                for (int i = 0; i < 8; ++i) { ns[i] = ns[i] + kn[a][i]; os[i] = os[i] + ko[i]; }
            }
            double dc[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) dc[i] = ns[i] - os[i];
            acc[0] += yd_du_k(C, dc, ky, kz) * bt;
#pragma unroll
            for (int i = 0; i < 8; ++i) C[i] = C[i] + dc[i];
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (m.q[a] == 0.0) continue;
            double f[3] = {0, 0, 0};
            yd_force_k(C, kn[a], kx, ky, kz, bt, f);
            acc[YK_FK(a, 0)] += f[0];
            acc[YK_FK(a, 1)] += f[1];
            acc[YK_FK(a, 2)] += f[2];
        }
    }
}

// LJ sets of a sweep: DELTA adds dU_LJ (acc[1]); forces of the LJ sites
// against every LJ set (+ its delta).  The LJ q-weighted sin/cos of a
// charged site is its Coulomb one times 1/q (RhoSquared.f90:173-176).
template <bool F32, bool DELTA, int NA, bool RANGE = false>
__device__ void yl_sweep_lj(const YdChain &c, const YkMol &m, double (&acc)[YK_NACC], int rank, int size,
                            int kend = -1) {
    const int nk = c.p.nk, T = c.p.ntype, ke = RANGE ? kend : nk;
    if (T == 0 || m.np == 0) return;
    const double *tabs = DELTA ? m.tab_n : m.tab_o;
    // the next k's index is loaded ahead; the S rows are loaded per set (a
    // prefetch of the next set's row costs registers this loop does not have)
    int kxyz_n = 0;
    if (rank < ke) kxyz_n = yd_g(c.kxyz)[rank];
    for (int kk = rank; kk < ke; kk += size) {
        const int kxyz = kxyz_n, kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        if (kk + size < ke) kxyz_n = yd_g(c.kxyz)[kk + size];
        double kl[NA][8];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            if (m.typ[a] <= 0) continue;
            double qa = m.q[a];
            if (qa != 0.0) {
                yd_kr_nz(tabs + a * 6 * YD_KP, qa, kx, ky, kz, kl[a]);
                double inv = 1.0 / qa;
#pragma unroll
                for (int i = 0; i < 8; ++i) kl[a][i] = kl[a][i] * inv;
            } else {
                yd_kr_nz(tabs + a * 6 * YD_KP, 1.0, kx, ky, kz, kl[a]);
            }
        }
        double dl[YK_MAXP][8];
        if (DELTA) {
            double nl[YK_MAXP][8], ol[YK_MAXP][8];
#pragma unroll
            for (int q = 0; q < YK_MAXP; ++q)
#pragma unroll
                for (int i = 0; i < 8; ++i) { nl[q][i] = 0.0; ol[q][i] = 0.0; }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                if (m.typ[a] <= 0) continue;
                double qa = m.q[a], lo[8];
                if (qa != 0.0) {
                    yd_kr_nz(m.tab_o + a * 6 * YD_KP, qa, kx, ky, kz, lo);
                    double inv = 1.0 / qa;
#pragma unroll
                    for (int i = 0; i < 8; ++i) lo[i] = lo[i] * inv;
                } else {
                    yd_kr_nz(m.tab_o + a * 6 * YD_KP, 1.0, kx, ky, kz, lo);
                }
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q == m.lslot[a]) {
#pragma unroll
                        for (int i = 0; i < 8; ++i) { nl[q][i] = nl[q][i] + kl[a][i]; ol[q][i] = ol[q][i] + lo[i]; }
                    }
            }
#pragma unroll
            for (int q = 0; q < YK_MAXP; ++q)
#pragma unroll
                for (int i = 0; i < 8; ++i) dl[q][i] = nl[q][i] - ol[q][i];
        }
        for (int t1 = 0; t1 < T; ++t1) {
            double C1[8], bl[YK_MAXP];
            yk_ld8<F32>(c, 1 + t1, kk, C1);
#pragma unroll
            for (int q = 0; q < YK_MAXP; ++q)
                bl[q] = q < m.np ? yd_g(c.beta_lj)[((size_t)t1 * T + (m.ptype[q] - 1)) * nk + kk] : 0.0;
            if (DELTA) {
                int s1 = -1;
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q < m.np && m.ptype[q] == t1 + 1) s1 = q;
                double tw[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    double d1 = 0.0;
#pragma unroll
                    for (int q = 0; q < YK_MAXP; ++q)
                        if (q == s1) d1 = dl[q][i];
                    tw[i] = C1[i] + C1[i] + d1;
                    C1[i] = C1[i] + d1;
                }
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q < m.np) acc[1] += yd_dulj_k(tw, dl[q], ky, kz) * bl[q];
            }
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const int ta = m.typ[a];
                if (ta <= 0) continue;
                double b = 0.0;
#pragma unroll
                for (int q = 0; q < YK_MAXP; ++q)
                    if (q == m.lslot[a]) b = bl[q];
                if (m.lslot[a] < 0) b = yd_g(c.beta_lj)[((size_t)(ta - 1) * T + t1) * nk + kk];
                double f[3] = {0, 0, 0};
                yd_force_k(C1, kl[a], kx, ky, kz, b, f);
                acc[YK_FK(a, 0)] += f[0];
                acc[YK_FK(a, 1)] += f[1];
                acc[YK_FK(a, 2)] += f[2];
            }
        }
    }
}

template <bool F32, bool DELTA, bool RANGE = false>
__device__ __forceinline__ void yl_sweep_n(const YdChain &c, const YkMol &m, double (&acc)[YK_NACC], int rank,
                                           int size, int kend = -1) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: yl_sweep_coul<F32, DELTA, 1, RANGE>(c, m, acc, rank, size, kend); yl_sweep_lj<F32, DELTA, 1, RANGE>(c, m, acc, rank, size, kend); break;
        case 2: yl_sweep_coul<F32, DELTA, 2, RANGE>(c, m, acc, rank, size, kend); yl_sweep_lj<F32, DELTA, 2, RANGE>(c, m, acc, rank, size, kend); break;
        case 3: yl_sweep_coul<F32, DELTA, 3, RANGE>(c, m, acc, rank, size, kend); yl_sweep_lj<F32, DELTA, 3, RANGE>(c, m, acc, rank, size, kend); break;
        default: yl_sweep_coul<F32, DELTA, 4, RANGE>(c, m, acc, rank, size, kend); yl_sweep_lj<F32, DELTA, 4, RANGE>(c, m, acc, rank, size, kend); break;
    }
}

// Exact real-space pass (same terms and stored back terms as yk_rs_exact):
// the new sites, whose per-site sums need compile-time accumulator indices,
// are unrolled; the old sites, which only feed site-independent sums, are
// visited one at a time.
template <int NA, bool AOS>
__device__ void yl_rs_exact(const YdChain &c, const YkMol &m, const double *sig6, const double *feps,
                            const int *list, int count, double (&acc)[YK_NRX], int &ov, int rank, int size) {
    const int T = c.p.ntype;
    // the next candidate's atom data are loaded before the current one's pairs
    int jt_n = 0;
    double q_n = 0.0, dh_n = 0.0, x_n = 0.0, y_n = 0.0, z_n = 0.0;
    auto load_cand = [&](int i) {
        const int j = list[i];
        if (AOS) {  // one 64-byte record of the launch copy
            auto r4 = yd_g(reinterpret_cast<const double4 *>(c.arec)) + 2 * (size_t)j;
            const double4 a = r4[0], b = r4[1];
            x_n = a.x; y_n = a.y; z_n = a.z; q_n = a.w; dh_n = b.x; jt_n = (int)b.y;
        } else {
            jt_n = yd_g(c.type)[j]; q_n = yd_g(c.q)[j]; dh_n = yd_g(c.dh)[j]; x_n = yd_g(c.x)[j]; y_n = yd_g(c.y)[j]; z_n = yd_g(c.z)[j];
        }
    };
    if (rank < count) load_cand(rank);
    for (int i = rank; i < count; i += size) {
        const int jt = jt_n;
        const double qj = q_n, dhj = dh_n, xj = x_n, yj = y_n, zj = z_n;
        if (i + size < count) load_cand(i + size);
        double nb[4] = {0, 0, 0, 0}, ob[4] = {0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xij = yd_wrap(m.xn[a] - xj), yij = yd_wrap(m.yn[a] - yj), zij = yd_wrap(m.zn[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (m.dh[a] + dhj) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) ov = 1;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            yk_rs_acc(acc, a, pr, xij, yij, zij);
            nb[0] = nb[0] + pr.pot;
            nb[1] = nb[1] - xij * pr.FR;
            nb[2] = nb[2] - yij * pr.FR;
            nb[3] = nb[3] - zij * pr.FR;
        }
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            int it = m.typ[a];
            double xij = yd_wrap(m.xo[a] - xj), yij = yd_wrap(m.yo[a] - yj), zij = yd_wrap(m.zo[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            ob[0] = ob[0] + pr.pot;
            ob[1] = ob[1] - xij * pr.FR;
            ob[2] = ob[2] - yij * pr.FR;
            ob[3] = ob[3] - zij * pr.FR;
            acc[YK_NRS + 0] += pr.potew;
            acc[YK_NRS + 1] += pr.potlj6;
            acc[YK_NRS + 2] += pr.potlj12;
        }
        auto b = yd_g(c.rsb) + (size_t)i * 4;
        b[0] = nb[0] - ob[0]; b[1] = nb[1] - ob[1]; b[2] = nb[2] - ob[2]; b[3] = nb[3] - ob[3];
    }
}

template <bool AOS = false>
__device__ __forceinline__ void yl_rs_exact_n(const YdChain &c, const YkMol &m, const double *sig6,
                                              const double *feps, const int *list, int count,
                                              double (&acc)[YK_NRX], int &ov, int rank, int size) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: yl_rs_exact<1, AOS>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        case 2: yl_rs_exact<2, AOS>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        case 3: yl_rs_exact<3, AOS>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
        default: yl_rs_exact<4, AOS>(c, m, sig6, feps, list, count, acc, ov, rank, size); break;
    }
}

// yd_block_sum over the accumulators a move uses: entries with !used(i)
// (uniform) are zero in every thread and are written as zero without the
// shuffles; the used ones are summed in yd_block_sum's order.
// Workgroup barrier that orders LDS accesses only: the waves' global stores
// still in flight are not waited for.  Used where no other wave reads those
// stores before the next full barrier (__syncthreads waits for every wave's
// stores), so a phase's last stores overlap the next phase.
__device__ __forceinline__ void yk_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// (its barriers are yk_lds_sync: the one-workgroup chain kernel calls it
// after its sweeps, which store nothing, and after the exact real-space
// pass, whose back terms are read back by the thread that stored them)
template <int NV, typename Used>
__device__ __forceinline__ void yk_block_sum_used(double (&v)[NV], double *red, double *out, Used used) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        if (!used(i)) continue;
        double a = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        v[i] = a;
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[w * NV + i] = used(i) ? v[i] : 0.0;
    }
    yk_lds_sync();
    if (threadIdx.x < NV) {
        double a = 0.0;
        for (int j = 0; j < nw; ++j) a += red[j * NV + threadIdx.x];
        out[threadIdx.x] = a;
    }
    yk_lds_sync();
}

// Phase timers of the chain kernel: thread 0 adds wall-clock ticks between
// workgroup barriers into st->phase[] (pick, tables, old sweep, proposal,
// real-space pass, new sweep, acceptance, commit, finalize).
#define YK_TICK(i)                                                  \
    do {                                                            \
        if (tid == 0) {                                             \
            unsigned long long t_ = wall_clock64();                 \
            sh.ph[i] += t_ - sh.t_last;                             \
            sh.t_last = t_;                                         \
        }                                                           \
    } while (0)

// ------------------------------------------------------------------------
// Scalar sections of a move, shared by the two chain kernels.  The
// workgroup kernel runs them on thread 0; the wave kernel (WAVE = true) runs
// them on all 64 lanes of the replica's wave with identical values in every
// lane (LDS stores of one value from all lanes; global stores from lane 0),
// which lets shi_sans_exp spread its trapezoid terms over the lanes.

// Molecule pick (runmc.f90:206, 232, 258) and the molecule's scalar data.
// The molecule of step mv (-1: an exchange step) and its draws.
__device__ __forceinline__ int yk_pick_imol(const YdChain &c, YdRng &rng, int mode, int mv) {
    if (mode == 1) {
        rng.cu = yd_g(c.cu)[mv];
        return yd_g(c.cimol)[mv];
    }
    if ((mv + c.mv_begin) % c.p.nmol == 0) (void)rng.next();  // pressure-step draw per cycle
    if (rng.next() < c.p.pxch) return -1;  // exchange step (runmc.f90:232-254): handed to the host
    return yk_choose(c, rng.next());
}

__device__ __forceinline__ void yk_sec_pick_as(const YdChain &c, YkMol &m, int imol) {
    m.imol = imol;
    if (imol < 0) return;
    m.first = yd_g(c.first)[imol];
    m.na = yd_g(c.last)[imol] - m.first + 1;
}

__device__ __forceinline__ void yk_sec_pick(const YdChain &c, YkMol &m, YdRng &rng, int mode, int mv) {
    yk_sec_pick_as(c, m, yk_pick_imol(c, rng, mode, mv));
}

// LJ types present in the molecule (after the per-site load).
__device__ __forceinline__ void yk_sec_slots(YkMol &m) {
    m.np = 0;
    for (int a = 0; a < m.na; ++a) {
        m.lslot[a] = -1;
        if (m.typ[a] > 0) {
            int sl = -1;
            for (int q = 0; q < m.np; ++q)
                if (m.ptype[q] == m.typ[a]) sl = q;
            if (sl < 0 && m.np < YK_MAXP) { sl = m.np; m.ptype[m.np++] = m.typ[a]; }
            m.lslot[a] = sl;
        }
    }
}

// Proposal (MC.f90:984-1067): old total force and torque from the old
// k-space sweep (out[YK_FK]) plus the per-atom real-space and external
// parts, then the force/torque-biased displacement and rotation.  Returns
// the stall flag.
template <bool WAVE>
__device__ void yk_sec_propose_pre(const YdChain &c, YkMol &m, YkScalar &t0, const double *out,
                                   const double (*fr_old)[3], const double *uu_old) {
    const YdParams &p = c.p;
    YkMcm &mcm = t0.mcm;
    int na = m.na;
    double fx[YD_MAXA], fy[YD_MAXA], fz[YD_MAXA];
    double U_old = 0.0;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) U_old = U_old + uu_old[a];
    t0.U_old = U_old;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        fx[a] = fr_old[a][0] + YD_FOUR_PI * out[YK_FK(a, 0)];
        fy[a] = fr_old[a][1] + YD_FOUR_PI * out[YK_FK(a, 1)];
        fz[a] = fr_old[a][2] + YD_FOUR_PI * out[YK_FK(a, 2)];
        fx[a] = fx[a] + p.m2kext * m.q[a] * t0.mu[0];
        fy[a] = fy[a] + p.m2kext * m.q[a] * t0.mu[1];
        fz[a] = fz[a] + p.m2kext * m.q[a] * t0.mu[2];
    }
    double mo[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        mo[0] = mo[0] + m.q[a] * m.xo[a];
        mo[1] = mo[1] + m.q[a] * m.yo[a];
        mo[2] = mo[2] + m.q[a] * m.zo[a];
    }
    t0.muo[0] = mo[0]; t0.muo[1] = mo[1]; t0.muo[2] = mo[2];
    double Fo[3] = {0, 0, 0}, To[3] = {0, 0, 0}, com[3] = {0, 0, 0}, tm = 0.0, tq[3];
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) { Fo[0] = Fo[0] + fx[a]; Fo[1] = Fo[1] + fy[a]; Fo[2] = Fo[2] + fz[a]; }
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        com[0] = com[0] + m.mass[a] * m.xo[a];
        com[1] = com[1] + m.mass[a] * m.yo[a];
        com[2] = com[2] + m.mass[a] * m.zo[a];
        tm = tm + m.mass[a];
    }
    com[0] = com[0] / tm; com[1] = com[1] / tm; com[2] = com[2] / tm;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        yk_prod(m.xo[a] - com[0], m.yo[a] - com[1], m.zo[a] - com[2], fx[a], fy[a], fz[a], tq);
        To[0] = To[0] + tq[0]; To[1] = To[1] + tq[1]; To[2] = To[2] + tq[2];
    }
    for (int d = 0; d < 3; ++d) { mcm.fo[d] = Fo[d]; mcm.co[d] = To[d]; t0.com[d] = com[d]; }
    mcm.ftotold = sqrt(Fo[0] * Fo[0] + Fo[1] * Fo[1] + Fo[2] * Fo[2]);
    mcm.couplold = sqrt(To[0] * To[0] + To[1] * To[1] + To[2] * To[2]);
    mcm.stall = 0;
}

// The new positions from the displacement and the rotation (MC.f90:1045-1067).
__device__ int yk_sec_propose_post(YkMol &m, YkScalar &t0) {
    YkMcm &mcm = t0.mcm;
    const int na = m.na;
    const double *com = t0.com;
    double cn[3];
    for (int d = 0; d < 3; ++d) cn[d] = com[d] + mcm.d[d] - round(com[d] + mcm.d[d]);
    if (na > 1) {
        double cphi = mcm.cphi, sphi = mcm.sphi, st = mcm.st, ct = mcm.ct, ca = mcm.ca, sa = mcm.sa;
        double ca1 = 1.0 - ca, R[9];
        R[0] = (cphi * cphi * (st * st) * ca1 + ca);
        R[1] = (sphi * cphi * (st * st) * ca1 - ct * sa);
        R[2] = st * (ct * ca1 * cphi + sa * sphi);
        R[3] = (cphi * sphi * (st * st) * ca1 + ct * sa);
        R[4] = (sphi * sphi * (st * st) * ca1 + ca);
        R[5] = st * (ct * ca1 * sphi - sa * cphi);
        R[6] = st * (ct * ca1 * cphi - sa * sphi);
        R[7] = st * (ct * ca1 * sphi + sa * cphi);
        R[8] = (st * st * ca + ct * ct);
#pragma unroll
        for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
            double xr = m.xo[a] - com[0], yr = m.yo[a] - com[1], zr = m.zo[a] - com[2];
            m.xn[a] = (R[0] * xr + R[1] * yr + R[2] * zr) + cn[0];
            m.yn[a] = (R[3] * xr + R[4] * yr + R[5] * zr) + cn[1];
            m.zn[a] = (R[6] * xr + R[7] * yr + R[8] * zr) + cn[2];
        }
    } else {
        m.xn[0] = cn[0]; m.yn[0] = cn[1]; m.zn[0] = cn[2];
    }
    return mcm.stall;
}

// The whole proposal on one wave (the wave kernel): forces and torques,
// displacement, rotation, new positions (MC.f90:984-1067).
template <bool WAVE>
__device__ int yk_sec_propose(const YdChain &c, YkMol &m, YkScalar &t0, const double *out,
                              const double (*fr_old)[3], const double *uu_old) {
    yk_sec_propose_pre<WAVE>(c, m, t0, out, fr_old, uu_old);
    yk_displacement(c.p, t0.mcm, t0.rng);
    yk_rotation<WAVE>(c.p, t0.mcm, t0.rng);
    return yk_sec_propose_post(m, t0);
}

// Acceptance (MC.f90:1176-1233): new forces and torque, dU_ext, the
// acceptance factor and the draw.  out holds the new k-space sweep's sums.
// Returns 1 if accepted; the move record goes to yd_g(c.rec)[mv] when present.
// The molecule's force and torque at its new position (MC.f90:1176-1212):
// the real-space per-site forces, the k-space ones (4 pi out), the external
// field's, and the torque about the new centre of mass.
__device__ __forceinline__ void yk_new_ft(const YdParams &p, const YkMol &m, const YkScalar &t0, const double *out,
                                          double (&Fn)[3], double (&Tn)[3], double (&mn)[3], double &du_ext) {
    const int na = m.na;
    double fx[YD_MAXA], fy[YD_MAXA], fz[YD_MAXA];
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        fx[a] = t0.frn[a][0] + YD_FOUR_PI * out[YK_FK(a, 0)];
        fy[a] = t0.frn[a][1] + YD_FOUR_PI * out[YK_FK(a, 1)];
        fz[a] = t0.frn[a][2] + YD_FOUR_PI * out[YK_FK(a, 2)];
    }
    double mun[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        mun[0] = mun[0] + m.q[a] * m.xn[a];
        mun[1] = mun[1] + m.q[a] * m.yn[a];
        mun[2] = mun[2] + m.q[a] * m.zn[a];
    }
    double dmu[3] = {mun[0] - t0.muo[0], mun[1] - t0.muo[1], mun[2] - t0.muo[2]};
    du_ext = p.kext * ((2 * t0.mu[0] + dmu[0]) * dmu[0] + (2 * t0.mu[1] + dmu[1]) * dmu[1] +
                       (2 * t0.mu[2] + dmu[2]) * dmu[2]);
    for (int d = 0; d < 3; ++d) mn[d] = t0.mu[d] + dmu[d];
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        fx[a] = fx[a] + p.m2kext * m.q[a] * mn[0];
        fy[a] = fy[a] + p.m2kext * m.q[a] * mn[1];
        fz[a] = fz[a] + p.m2kext * m.q[a] * mn[2];
    }
    double com[3] = {0, 0, 0}, tm = 0.0, tq[3];
    Fn[0] = Fn[1] = Fn[2] = 0.0;
    Tn[0] = Tn[1] = Tn[2] = 0.0;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) { Fn[0] = Fn[0] + fx[a]; Fn[1] = Fn[1] + fy[a]; Fn[2] = Fn[2] + fz[a]; }
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        com[0] = com[0] + m.mass[a] * m.xn[a];
        com[1] = com[1] + m.mass[a] * m.yn[a];
        com[2] = com[2] + m.mass[a] * m.zn[a];
        tm = tm + m.mass[a];
    }
    com[0] = com[0] / tm; com[1] = com[1] / tm; com[2] = com[2] / tm;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < na) {
        yk_prod(m.xn[a] - com[0], m.yn[a] - com[1], m.zn[a] - com[2], fx[a], fy[a], fz[a], tq);
        Tn[0] = Tn[0] + tq[0]; Tn[1] = Tn[1] + tq[1]; Tn[2] = Tn[2] + tq[2];
    }
}

// MonteCarloMove.f90's Q_new of the acceptance factor from the new torque,
// as yk_init_new computes it; by a whole wave.
__device__ __forceinline__ double yk_qnew(const YdParams &p, const double (&Tn)[3]) {
    const double couplnew = sqrt(Tn[0] * Tn[0] + Tn[1] * Tn[1] + Tn[2] * Tn[2]);
    const double vmax = p.lambda_c * couplnew * p.d_angle;
    return vmax > 1e-2 ? yd_shi_sans_exp_wave(vmax) / vmax : 1.0;
}

// MC.f90:1176-1233 up to the acceptance factor's inputs: the energy
// change (returned) and the new-side quantities of the force bias in
// t0.mcm.  QPRE: Q_new is computed elsewhere (another wave evaluates
// yk_qnew at the same time) and left unset.
template <bool WAVE, bool QPRE = false>
__device__ double yk_sec_accept_pre(const YdChain &c, const YkMol &m, YkScalar &t0, const double *out) {
    const YdParams &p = c.p;
    YkMcm &mcm = t0.mcm;
    double U_new = 0.0;
#pragma unroll
    for (int a = 0; a < YD_MAXA; ++a) if (a < m.na) U_new = U_new + t0.uun[a];
    double du_real = U_new - t0.U_old, duc = out[0], dulj = out[1];
    double Fn[3], Tn[3], mn[3], du_ext;
    yk_new_ft(p, m, t0, out, Fn, Tn, mn, du_ext);
    for (int d = 0; d < 3; ++d) { mcm.fn[d] = Fn[d]; mcm.cn[d] = Tn[d]; }
    double du_total = du_real + (duc + dulj) + du_ext;
    yk_init_new<WAVE, QPRE>(p, mcm);
    t0.du_real = du_real; t0.duc = duc; t0.dulj = dulj; t0.du_ext = du_ext;
    for (int d = 0; d < 3; ++d) t0.mu_new[d] = mn[d];
    return du_total;
}

// The acceptance test (one draw) and the move's record.
__device__ int yk_sec_accept_post(const YdChain &c, const YkMol &m, YkScalar &t0, double du_total, int mv) {
    const YkMcm &mcm = t0.mcm;
    int accd = yk_accept(c.p, mcm, du_total, t0.rng);
    if (c.rec && (threadIdx.x & 63) == 0) {
        yamc_move_rec r = {};
        r.imol = m.imol; r.accepted = accd; r.overlap = 0; r.imol2 = -1;
        r.du_real = t0.du_real; r.du_kc = t0.duc; r.du_klj = t0.dulj; r.du_ext = t0.du_ext; r.du_total = du_total;
        for (int d = 0; d < 3; ++d) {
            r.f_old[d] = mcm.fo[d]; r.t_old[d] = mcm.co[d]; r.f_new[d] = mcm.fn[d]; r.t_new[d] = mcm.cn[d];
            r.delta[d] = mcm.d[d];
        }
        r.angle = mcm.angle;
        yd_g(c.rec)[mv] = r;
    }
    return accd;
}

// MC.f90:1176-1233 on one wave.
template <bool WAVE>
__device__ int yk_sec_accept(const YdChain &c, const YkMol &m, YkScalar &t0, const double *out, int mv) {
    const double du_total = yk_sec_accept_pre<WAVE>(c, m, t0, out);
    return yk_sec_accept_post(c, m, t0, du_total, mv);
}

// Record of a move rejected before the acceptance test: overlap = 1 (hard
// core, MC.f90:1080-1084) or 2 (a proposal loop over its bound).
__device__ __forceinline__ void yk_sec_reject_rec(const YdChain &c, const YkMol &m, const YkMcm &mcm, int mv,
                                                  int overlap) {
    if (!c.rec) return;
    yamc_move_rec r = {};
    r.imol = m.imol; r.overlap = overlap; r.imol2 = -1;
    if (overlap == 1) {
        for (int d = 0; d < 3; ++d) { r.f_old[d] = mcm.fo[d]; r.t_old[d] = mcm.co[d]; r.delta[d] = mcm.d[d]; }
        r.angle = mcm.angle;
    }
    yd_g(c.rec)[mv] = r;
}

// Commit of the moved molecule's own per-atom arrays, positions and the
// energy components (MC.f90:1242-1318); single lane.
// cached: the carried energy sums are t0's copies (kept current here)
// instead of the state's.
__device__ __forceinline__ void yk_sec_finalize(const YdChain &c, const YkMol &m, YkScalar &t0, float4 *pos,
                                                bool aos = false, unsigned int *posq = nullptr, bool cached = false) {
    int f = m.first;
    YdState *st = c.st;
    // the state's sums are read before the per-atom stores, which the
    // compiler cannot tell apart from them
    const double e_kc = cached ? t0.e_kc : st->e_kc, e_klj = cached ? t0.e_klj : st->e_klj,
                 uu_ew = cached ? t0.uu_ew : st->uu_ew, uu_lj6 = cached ? t0.uu_lj6 : st->uu_lj6,
                 uu_lj12 = cached ? t0.uu_lj12 : st->uu_lj12;
    for (int a = 0; a < m.na; ++a) {
        yd_g(c.uu)[f + a] = t0.uun[a];
        yd_g(c.Fx)[f + a] = t0.frn[a][0];
        yd_g(c.Fy)[f + a] = t0.frn[a][1];
        yd_g(c.Fz)[f + a] = t0.frn[a][2];
        yd_g(c.x)[f + a] = m.xn[a];
        yd_g(c.y)[f + a] = m.yn[a];
        yd_g(c.z)[f + a] = m.zn[a];
        if (pos) pos[f + a] = make_float4((float)m.xn[a], (float)m.yn[a], (float)m.zn[a], __int_as_float(m.imol + 1));
        if (posq) posq[f + a] = yk_pack_pos(m.xn[a], m.yn[a], m.zn[a]);
        if (aos) {  // the chain kernel's launch copies
            yd_g(reinterpret_cast<double4 *>(c.uf))[f + a] = make_double4(t0.uun[a], t0.frn[a][0], t0.frn[a][1], t0.frn[a][2]);
            auto r4 = yd_g(reinterpret_cast<double4 *>(c.arec)) + 2 * (size_t)(f + a);
            r4[0] = make_double4(m.xn[a], m.yn[a], m.zn[a], m.q[a]);
        }
    }
    const double n_kc = e_kc + t0.duc, n_klj = e_klj + t0.dulj, n_ew = uu_ew + t0.cmp_new[0] - t0.cmp_old[0],
                 n_lj6 = uu_lj6 + t0.cmp_new[1] - t0.cmp_old[1], n_lj12 = uu_lj12 + t0.cmp_new[2] - t0.cmp_old[2];
    st->e_kc = n_kc;
    st->e_klj = n_klj;
    st->uu_ew = n_ew;
    st->uu_lj6 = n_lj6;
    st->uu_lj12 = n_lj12;
    if (cached) {
        t0.e_kc = n_kc; t0.e_klj = n_klj; t0.uu_ew = n_ew; t0.uu_lj6 = n_lj6; t0.uu_lj12 = n_lj12;
    }
    for (int d = 0; d < 3; ++d) st->mu[d] = t0.mu_new[d];
}

// ========================================================================
// mu = sum_j q_j r_j from scratch (EwaldSumExternal.f90:165-183) over the
// workgroup into t0.mu (thread 0).
__device__ __forceinline__ void yk_chain_mu(const YdChain &c, YkShared &sh, int tid) {
    double v[3] = {0, 0, 0};
    for (int i = tid; i < c.p.natom; i += (int)blockDim.x) {
        double q = yd_g(c.q)[i];
        v[0] += q * yd_g(c.x)[i]; v[1] += q * yd_g(c.y)[i]; v[2] += q * yd_g(c.z)[i];
    }
    yd_block_sum<3>(v, sh.red, sh.out);
    if (tid == 0) {
        sh.t0.mu[0] = sh.out[0]; sh.t0.mu[1] = sh.out[1]; sh.t0.mu[2] = sh.out[2];
    }
}

static_assert(offsetof(YkShared, bm) % sizeof(double) == 0, "the rebuild's LDS tables live in bm");
static_assert(sizeof(YkShared::red) >= sizeof(double) * 2 * (YK_KE_THREADS / 64), "red holds the energy sums");

// The periodic from-scratch rebuild of a chain inside its own launch
// (calc_total_energy, EwaldSumKSpace.f90:225-260, plus mu): what a
// launch-level rebuild (yk_atom_tables, yk_full_sums, yk_full_sums_reduce,
// yk_kenergy) followed by a new launch computes, with the same additions in
// the same order, so a chain rebuilt here equals the chain run in segments
// with rebuilds between them bit for bit.
// S(k): for each set and pass over KT tiles of blockDim consecutive k (one
// thread per k of each tile, the KT tiles' sums in registers), the set's
// atoms in list order, group by group (c.fs_ng equal groups, each group's
// sum added to the total in group order), the axis tables of each chunk of
// atoms built in LDS once per pass (the scan bitmap's space, free between
// moves) and read by every k of the pass.
// Here over the k vectors [kb, ke) (a grid chain's workgroup rebuilds its
// own range; the per-k arithmetic does not depend on the range).
template <int KT>
__device__ void yk_rebuild_sums(const YdChain &c, YkShared &sh, int tid, int kb, int ke, double *xtra,
                                int xtra_bytes) {
    const int bs = (int)blockDim.x;
    const int T = c.p.ntype, kmax = c.p.kmax, kp = kmax + 1, nkr = c.p.nk;  // nkr: the row length of S
    const int nk = ke;  // (the end of the range)
    // per atom of a chunk: its six axis tables and its 1/q factor
    const int stride = 6 * kp + 1;
    double *tb = reinterpret_cast<double *>(sh.bm);
    const int C1 = (int)(sizeof(sh.bm) / (sizeof(double) * (size_t)stride));
    int C = C1 + (xtra ? xtra_bytes / (int)(sizeof(double) * stride) : 0);
    if (3 * C > bs) C = bs / 3;  // one table item per thread
    auto tabs = [&](int a) { return a < C1 ? tb + (size_t)a * stride : xtra + (size_t)(a - C1) * stride; };
    for (int set = 0; set <= T; ++set) {
        const int l0 = yd_g(c.set_off)[set], l1 = yd_g(c.set_off)[set + 1], ng = c.fs_ng;
        const int gsize = (l1 - l0 + ng - 1) / ng;
        for (int k0 = kb; k0 < nk; k0 += KT * bs) {
            int kx[KT], ky[KT], kz[KT];
            double tot[KT][8], sg[KT][8];
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int kk = k0 + t * bs + tid;
                kx[t] = ky[t] = kz[t] = 0;
                if (kk < nk) {
                    const int kxyz = yd_g(c.kxyz)[kk];
                    kx[t] = kxyz & 255; ky[t] = (kxyz >> 8) & 255; kz[t] = (kxyz >> 16) & 255;
                }
#pragma unroll
                for (int m = 0; m < 8; ++m) { tot[t][m] = 0.0; sg[t][m] = 0.0; }
            }
            // this thread's table item of a chunk: atom tid / 3, axis tid % 3;
            // its coordinate and charge are loaded one chunk ahead
            const int ia = tid / 3, id = tid % 3;
            double r_n = 0.0, q_n = 0.0;
            auto load_item = [&](int base) {
                if (tid < 3 * C && base + ia < l1) {
                    const int at = yd_g(c.set_atoms)[base + ia];
                    r_n = id == 0 ? yd_g(c.x)[at] : id == 1 ? yd_g(c.y)[at] : yd_g(c.z)[at];
                    q_n = yd_g(c.q)[at];
                }
            };
            load_item(l0);
            int gend = l0 + gsize;  // end of the current group
            for (int base = l0; base < l1; base += C) {
                const int nch = l1 - base < C ? l1 - base : C;
                const double r = r_n, q = q_n;
                __syncthreads();  // the previous chunk's tables are no longer read
                if (tid < 3 * nch) {
                    // yd_kr_atom's factors for this set: the Coulomb set holds
                    // charged atoms only (q sin/cos); an LJ set takes (q
                    // sin/cos) (1/q) of a charged atom, else sin/cos.  The x
                    // table holds q cos, q sin: yd_kr_s's products q cx cy
                    // are (q cx) cy, so the values are the same to the bit.
                    double *ta = tabs(ia), *t = ta + 2 * id * kp;
                    yd_sincos_axis(r, kmax, t, t + kp);
                    if (id == 0) {
                        const double qk = set == 0 || q != 0.0 ? q : 1.0;
                        for (int k = 0; k <= kmax; ++k) {
                            t[k] = qk * t[k];
                            t[kp + k] = qk * t[kp + k];
                        }
                        ta[6 * kp] = set > 0 && q != 0.0 ? 1.0 / q : 0.0;
                    }
                }
                load_item(base + C);
                __syncthreads();
                for (int a = 0; a < nch; ++a) {
                    if (base + a == gend) {  // group boundary: the group's sum into the total
#pragma unroll
                        for (int t = 0; t < KT; ++t)
#pragma unroll
                            for (int m = 0; m < 8; ++m) {
                                tot[t][m] = tot[t][m] + sg[t][m];
                                sg[t][m] = 0.0;
                            }
                        gend += gsize;
                    }
                    const double *ta = tabs(a);
                    const double inv = ta[6 * kp];  // 0: no 1/q factor
#pragma unroll
                    for (int t = 0; t < KT; ++t) {
                        if (k0 + t * bs + tid >= nk) continue;
                        double kr[8];
                        yd_kr_pre(ta, kp, kx[t], ky[t], kz[t], kr);
                        if (inv != 0.0) {
#pragma unroll
                            for (int m = 0; m < 8; ++m) kr[m] = kr[m] * inv;
                        }
#pragma unroll
                        for (int m = 0; m < 8; ++m) sg[t][m] = sg[t][m] + kr[m];
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int kk = k0 + t * bs + tid;
                if (kk >= nk) continue;
                if (l1 > l0) {
#pragma unroll
                    for (int m = 0; m < 8; ++m) tot[t][m] = tot[t][m] + sg[t][m];  // the last non-empty group
                }
                // the slots yd_kr_s sets to zero at this k (ky = 0 or kz = 0)
                // were summed as they came; the sum of those zeros is +0
                if (kz[t] == 0) { tot[t][2] = 0.0; tot[t][3] = 0.0; tot[t][6] = 0.0; tot[t][7] = 0.0; }
                if (ky[t] == 0) { tot[t][4] = 0.0; tot[t][5] = 0.0; tot[t][6] = 0.0; tot[t][7] = 0.0; }
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const size_t idx = ((size_t)set * 8 + m) * nkr + kk;
                    if (c.p.precision == YAMC_FP32C) {
                        const float hi = (float)tot[t][m];
                        yd_g(c.Shi)[idx] = hi;
                        yd_g(c.Slo)[idx] = (float)(tot[t][m] - (double)hi);
                    } else {
                        yd_g(c.S)[idx] = tot[t][m];
                    }
                }
            }
        }
    }
}

// E_kC, E_kLJ from the rebuilt S(k): yk_kenergy's per-thread k strides and
// block sum of YK_KE_THREADS threads, each thread here standing in for
// YK_KE_THREADS / blockDim of them; into e[0..1] (LDS) and, with store, the
// chain state.
__device__ void yk_rebuild_energies(const YdChain &c, YkShared &sh, int tid, bool store, double *e) {
    const int bs = (int)blockDim.x, nk = c.p.nk;
    {
        const int J = YK_KE_THREADS / bs, nw = bs >> 6, lane = tid & 63, w = tid >> 6;
        for (int j = 0; j < J; ++j) {
            double v[2] = {0.0, 0.0};
            for (int kk = tid + bs * j; kk < nk; kk += YK_KE_THREADS) yk_kenergy_k(c, kk, v);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                double a = v[i];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
                if (lane == 0) sh.red[(w + nw * j) * 2 + i] = a;
            }
        }
        __syncthreads();
        if (tid < 2) {
            double a = 0.0;
            for (int vw = 0; vw < YK_KE_THREADS / 64; ++vw) a += sh.red[vw * 2 + tid];
            e[tid] = a;
            if (store) {
                if (tid == 0) yd_g(c.st)->e_kc = a;
                else yd_g(c.st)->e_klj = a;
            }
        }
        __syncthreads();
    }
}

// The whole rebuild of a one-workgroup chain: S(k) over all k, the
// energies, then mu and the cell buckets, as at the start of a launch.
template <int KT>
__device__ void yk_chain_rebuild(const YdChain &c, YkShared &sh, int tid, double *xtra, int xtra_bytes, double *e) {
    yk_rebuild_sums<KT>(c, sh, tid, 0, c.p.nk, xtra, xtra_bytes);
    __syncthreads();  // S(k) complete (workgroup-scope visibility of the stores)
    yk_rebuild_energies(c, sh, tid, true, e);
    yk_chain_mu(c, sh, tid);
    if (c.nc > 0) yk_cells_build(c, &sh.cell_ok);
    __syncthreads();
}

// ========================================================================
// The workgroup chain kernel: one workgroup per chain, every per-move sum
// spread over the workgroup.  mode 0: run_mc cycles (runmc.f90:204-268)
// with the Philox stream, nsteps = cycles; mode 1: nsteps moves with
// caller-given molecule and constant uniform per move.
// Both shapes are built for two waves per SIMD (the 256-thread form so that
// two chains share a CU when there are more chains than CUs).
template <bool F32, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
yk_chain(const YdChain *chains, const double *const *sig6s, const double *const *feps_s, int nsteps, int mode,
         int max_moves, int lds_atoms, int resync_every) {
    __shared__ YkShared sh;
    // packed positions when the launch provides LDS for them (8-byte aligned:
    // the periodic rebuild also keeps FP64 tables there)
    extern __shared__ __align__(16) unsigned int lds_pos[];
    const YdChain &c = chains[blockIdx.x];
    const double *sig6 = sig6s[blockIdx.x], *feps = feps_s[blockIdx.x];
    const int tid = threadIdx.x;
    const YdParams &p = c.p;
    int total = mode == 1 ? nsteps : nsteps * p.nmol;
    if (max_moves > 0 && total > max_moves) total = max_moves;
    const bool use_lds = lds_atoms >= p.natom;
    // packed copy of the positions for the real-space scan (yk_pack_pos): in
    // LDS when the launch gives room for it, else in the context's pos4 buffer
    unsigned int *const spos = use_lds ? lds_pos : reinterpret_cast<unsigned int *>(c.pos4);
    for (int i = tid; i < p.natom; i += (int)blockDim.x) {
        spos[i] = yk_pack_pos(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i]);
        // the launch copies: one record per atom for the exact pass, the
        // per-atom real-space arrays as (uu, F) quadruples
        auto r4 = yd_g(reinterpret_cast<double4 *>(c.arec)) + 2 * (size_t)i;
        r4[0] = make_double4(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i], yd_g(c.q)[i]);
        r4[1] = make_double4(yd_g(c.dh)[i], (double)yd_g(c.type)[i], 0.0, 0.0);
        yd_g(reinterpret_cast<double4 *>(c.uf))[i] = make_double4(yd_g(c.uu)[i], yd_g(c.Fx)[i], yd_g(c.Fy)[i], yd_g(c.Fz)[i]);
    }
    if (c.nc > 0) yk_cells_build(c, &sh.cell_ok);  // workgroup-uniform; ends with a barrier
    YkScalar &t0 = sh.t0;
    // mu = sum_j q_j r_j from scratch once per launch (EwaldSumExternal.f90:
    // 165-183); afterwards it is carried incrementally, mu + dmu on each
    // accepted move, which is the value movemc recomputes (MC.f90:916).
    yk_chain_mu(c, sh, tid);
    if (tid == 0) {
        for (int i = 0; i < 9; ++i) sh.ph[i] = 0;
        sh.t_last = wall_clock64();
        t0.acc_count = 0; t0.nmv = 0; t0.nov = 0; t0.prepicked = 0;
        t0.e_kc = yd_g(c.st)->e_kc; t0.e_klj = yd_g(c.st)->e_klj;
        t0.uu_ew = yd_g(c.st)->uu_ew; t0.uu_lj6 = yd_g(c.st)->uu_lj6; t0.uu_lj12 = yd_g(c.st)->uu_lj12;
        YdRng &rng = t0.rng;
        rng.key0 = yd_g(c.st)->key[0]; rng.key1 = yd_g(c.st)->key[1];
        rng.s0 = yd_g(c.st)->stream[0]; rng.s1 = yd_g(c.st)->stream[1];
        rng.ctr = yd_g(c.st)->ctr; rng.constant = mode == 1; rng.cu = 0.0;
        rng.reset_cache();
        rng.list = yd_g(c.st)->ulist; rng.nlist = yd_g(c.st)->nlist;
    }
    for (int mv = 0; mv < total; ++mv) {
        // the periodic FP64 rebuild between move resync_every k - 1 and
        // resync_every k (workgroup-uniform)
        if (resync_every > 0 && mv > 0 && mv % resync_every == 0) {
            __syncthreads();  // the last commit's and finalize's stores
            // the chunk tables also take the LDS of the packed positions,
            // which are packed again afterwards (the positions are unchanged)
            yk_chain_rebuild<2>(c, sh, tid, use_lds ? reinterpret_cast<double *>(lds_pos) : nullptr,
                                use_lds ? 4 * p.natom : 0, sh.out);  // ends with a barrier
            if (use_lds) {
                for (int i = tid; i < p.natom; i += (int)blockDim.x) spos[i] = yk_pack_pos(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i]);
                __syncthreads();
            }
            if (tid == 0) { t0.e_kc = yd_g(c.st)->e_kc; t0.e_klj = yd_g(c.st)->e_klj; }
        }
        if (tid == 0) {
            __builtin_amdgcn_s_setprio(3);
            if (t0.prepicked) {  // taken during the last move's acceptance
                yk_sec_pick_as(c, sh.m, t0.nimol);
                t0.rng = t0.rng_next;
                t0.prepicked = 0;
            } else {
                yk_sec_pick(c, sh.m, t0.rng, mode, mv);
            }
            sh.ov = 0;
            sh.nbcount = 0;
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
        if (sh.m.imol < 0) break;  // exchange step: the launch ends here and the host runs it
        if (tid < sh.m.na) {  // one lane per site loads the molecule
            YkMol &m = sh.m;
            int a = tid, g = m.first + a;
            m.q[a] = yd_g(c.q)[g]; m.typ[a] = yd_g(c.type)[g]; m.mass[a] = yd_g(c.mass)[g]; m.dh[a] = yd_g(c.dh)[g];
            m.xo[a] = yd_g(c.x)[g]; m.yo[a] = yd_g(c.y)[g]; m.zo[a] = yd_g(c.z)[g];
            const double4 u = yd_g(reinterpret_cast<const double4 *>(c.uf))[g];
            sh.uu_old[a] = u.x;
            sh.fr_old[a][0] = u.y; sh.fr_old[a][1] = u.z; sh.fr_old[a][2] = u.w;
        }
        __syncthreads();
        if (tid == 0) yk_sec_slots(sh.m);
        YK_TICK(0);
        // ---- old axis tables --------------------------------------------
        yk_mol_tables(c, sh.m, false, tid);
        __syncthreads();
        YK_TICK(1);
        // ---- old k-space forces (MC.f90:946-982) ------------------------
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, false>(c, sh.m, acc, tid, blockDim.x);
            const int na = __builtin_amdgcn_readfirstlane(sh.m.na);
            yk_block_sum_used(acc, sh.red, sh.out, [na](int i) { return i < 2 + 3 * na; });
        }
        YK_TICK(2);
        // ---- proposal (MC.f90:984-1067): forces and torques on wave 0,
        // then the displacement on wave 0 and the rotation on wave 1 at
        // once (the rotation's draws start after the displacement's, which
        // wave 1 replays without the arithmetic), then the new positions
        // The one- and two-wave sections of a move (proposal, acceptance,
        // pick, finalize) run at raised wave priority: on a CU shared with
        // another chain they are this chain's critical path, while the other
        // chain's wide phases can wait a few cycles.
        if (tid < 128) __builtin_amdgcn_s_setprio(3);
        if (tid < 64) {
            yk_sec_propose_pre<true>(c, sh.m, t0, sh.out, sh.fr_old, sh.uu_old);
            t0.mcm2 = t0.mcm;
            t0.rng2 = t0.rng;
        }
        __syncthreads();
        if (tid < 64) {
            yk_displacement(p, t0.mcm, t0.rng);
        } else if (tid < 128) {
            yk_displacement_draws(p, t0.mcm2, t0.rng2);
            yk_rotation<true, true>(p, t0.mcm2, t0.rng2);
        } else if (tid < 192) {  // Q_old of the rotation, off the rotation wave's path
            const double qo = yk_qold(p, t0.mcm);
            if (tid == 128) t0.qold_pre = qo;
        }
        __syncthreads();
        if (tid < 64) {
            YkMcm &a = t0.mcm;
            const YkMcm &b = t0.mcm2;
            a.ca = b.ca; a.sa = b.sa; a.ct = b.ct; a.st = b.st; a.cphi = b.cphi; a.sphi = b.sphi;
            a.v[0] = b.v[0]; a.v[1] = b.v[1]; a.v[2] = b.v[2];
            a.vmaxold = b.vmaxold; a.angle = b.angle; a.qold = t0.qold_pre;
            a.stall = b.stall;  // the replay stalls exactly where the displacement does
            t0.rng = t0.rng2;
            int stall = yk_sec_propose_post(sh.m, t0);
            if (tid == 0) sh.stall = stall;
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        YK_TICK(3);
        if (sh.stall) {  // proposal loop over its bound: rejected, no acceptance draw
            if (tid == 0) {
                ++t0.nmv;
                yk_sec_reject_rec(c, sh.m, t0.mcm, mv, 2);
            }
            __syncthreads();
            continue;
        }
        // ---- new real-space partial sum (MC.f90:1077-1090) --------------
        {
            double acc[YK_NRX];
#pragma unroll
            for (int i = 0; i < YK_NRX; ++i) acc[i] = 0.0;
            int ov = 0;
            yk_mol_tables(c, sh.m, true, tid);
            const int nwords = (p.natom + 31) >> 5;
            if (nwords <= YK_BMWORDS) {
                for (int i = tid; i < nwords; i += (int)blockDim.x) sh.bm[i] = 0u;
                __syncthreads();
                // two inlined copies, so that the LDS copy is read with LDS
                // instructions (a pointer that may be either is a flat one)
                bool scanned = false;
                if (use_lds) {
                    if (c.nc > 0 && sh.cell_ok)
                        scanned = yk_rs_scan_cells_n(c, sh.m, lds_pos, sh.bm, sh.cl_cell, sh.cl_start, &sh.ncl);
                    if (!scanned) yk_rs_scan_n(c, sh.m, lds_pos, sh.bm);
                } else {
                    if (c.nc > 0 && sh.cell_ok)
                        scanned = yk_rs_scan_cells_n(c, sh.m, spos, sh.bm, sh.cl_cell, sh.cl_start, &sh.ncl);
                    if (!scanned) yk_rs_scan_n(c, sh.m, spos, sh.bm);
                }
                __syncthreads();
                yk_compact(sh.bm, nwords, sh.nblist, &sh.nbcount, sh.wsum);
            } else if (tid == 0) {
                sh.nbcount = YK_NBCAP + 1;  // too many atoms for the bitmap: full passes
            }
            __syncthreads();
            if (sh.nbcount <= YK_NBCAP) {
                yl_rs_exact_n<true>(c, sh.m, sig6, feps, sh.nblist, sh.nbcount, acc, ov, tid, blockDim.x);
            } else {  // list overflow: full FP64 pass (the commit then visits every atom)
                int junk = 0;
                double a2[YK_NRS];
#pragma unroll
                for (int i = 0; i < YK_NRS; ++i) a2[i] = 0.0;
                yk_rs_move_n(c, sh.m, sig6, feps, a2, ov, sh.nblist, &junk, tid, blockDim.x);
#pragma unroll
                for (int i = 0; i < YK_NRS; ++i) acc[i] = a2[i];
            }
            if (ov) sh.ov = 1;
            const int na = __builtin_amdgcn_readfirstlane(sh.m.na);
            yk_block_sum_used(acc, sh.red, sh.out, [na](int i) {
                return i < na || (i >= YD_MAXA && i < YD_MAXA + 3 * na) || i >= 4 * YD_MAXA;
            });
            if (tid == 0) {
                for (int a = 0; a < sh.m.na; ++a) {
                    t0.uun[a] = sh.out[a];
                    t0.frn[a][0] = sh.out[YD_MAXA + 3 * a + 0];
                    t0.frn[a][1] = sh.out[YD_MAXA + 3 * a + 1];
                    t0.frn[a][2] = sh.out[YD_MAXA + 3 * a + 2];
                }
                for (int i = 0; i < 3; ++i) {
                    t0.cmp_new[i] = sh.out[4 * YD_MAXA + i];
                    t0.cmp_old[i] = sh.out[YK_NRS + i];
                }
            }
        }
        YK_TICK(4);
        if (sh.ov) {  // overlap: rejected, no acceptance draw (MC.f90:1080-1084)
            if (tid == 0) {
                ++t0.nmv; ++t0.nov;
                yk_sec_reject_rec(c, sh.m, t0.mcm, mv, 1);
            }
            __syncthreads();
            continue;
        }
        // ---- new k-space sweep: dU and new forces (MC.f90:1096-1173) ----
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, true>(c, sh.m, acc, tid, blockDim.x);
            const int na = __builtin_amdgcn_readfirstlane(sh.m.na);
            yk_block_sum_used(acc, sh.red, sh.out, [na](int i) { return i < 2 + 3 * na; });
        }
        YK_TICK(5);
        // ---- acceptance on wave 0 (MC.f90:1176-1233); wave 1 meanwhile
        // takes the next move's pick: the acceptance takes exactly one draw,
        // so the pick's draws start one draw after t0.rng2 (the state the
        // proposal left, which wave 0 does not write here) -------------------
        // (wave 2 meanwhile computes Q_new from the new torque, off wave 0's
        // path; wave 0 takes it after the barrier for the acceptance test)
        if (tid < 64) {
            __builtin_amdgcn_s_setprio(3);
            const double du = yk_sec_accept_pre<true, true>(c, sh.m, t0, sh.out);
            if (tid == 0) t0.du_total = du;
            __builtin_amdgcn_s_setprio(0);
        } else if (tid < 128) {
            if (mode == 0 && mv + 1 < total) {
                // in LDS, the same values from every lane of the wave
                t0.rng_next = t0.rng2;
                (void)t0.rng_next.next();
                const int nimol = yk_pick_imol(c, t0.rng_next, mode, mv + 1);
                t0.nimol = nimol;
                t0.prepicked = 1;
            }
        } else if (tid < 192) {
            __builtin_amdgcn_s_setprio(3);
            double Fn[3], Tn[3], mn[3], du_ext;
            yk_new_ft(p, sh.m, t0, sh.out, Fn, Tn, mn, du_ext);
            const double qn = yk_qnew(p, Tn);
            if (tid == 128) t0.qnew_pre = qn;
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
        if (tid < 64) {
            __builtin_amdgcn_s_setprio(3);
            t0.mcm.qnew = t0.qnew_pre;
            int accd = yk_sec_accept_post(c, sh.m, t0, t0.du_total, mv);
            if (tid == 0) {
                sh.accepted = accd;
                ++t0.nmv;
                t0.acc_count += accd;
            }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
        YK_TICK(6);
        if (!sh.accepted) continue;
        // ---- commit (MC.f90:1242-1318) ----------------------------------
        yk_commit<F32>(c, sh.m, tid, blockDim.x);
        if (sh.nbcount <= YK_NBCAP) {
            yk_rs_commit_list<true>(c, sh.nblist, sh.nbcount, tid, blockDim.x);
        } else {
            double acc[3] = {0, 0, 0};
            yk_rs_commit_all<true>(c, sh.m, sig6, feps, acc, tid, blockDim.x);
            yd_block_sum<3>(acc, sh.red, sh.out);
            if (tid == 0)
                for (int i = 0; i < 3; ++i) t0.cmp_old[i] = sh.out[i];
        }
        // The commit's and the finalize's global stores are not waited for
        // here: the next move reads them only after its pick barrier (a full
        // one), and a rebuild starts with one.
        yk_lds_sync();
        YK_TICK(7);
        if (tid == 0) {
            __builtin_amdgcn_s_setprio(3);
            yk_sec_finalize(c, sh.m, t0, nullptr, true, spos, true);
            if (c.nc > 0 && sh.cell_ok) yk_cells_move(c, sh.m, &sh.cell_ok);
            for (int d = 0; d < 3; ++d) t0.mu[d] = t0.mu_new[d];
            __builtin_amdgcn_s_setprio(0);
        }
        yk_lds_sync();
        YK_TICK(8);
    }
    // the launch copy of the per-atom real-space arrays back into the state
    __syncthreads();
    for (int i = tid; i < p.natom; i += (int)blockDim.x) {
        const double4 u = yd_g(reinterpret_cast<const double4 *>(c.uf))[i];
        yd_g(c.uu)[i] = u.x; yd_g(c.Fx)[i] = u.y; yd_g(c.Fy)[i] = u.z; yd_g(c.Fz)[i] = u.w;
    }
    if (tid == 0) {
        if (mode != 1) yd_g(c.st)->ctr = t0.rng.ctr;
        yd_g(c.st)->n_moves += t0.nmv;
        yd_g(c.st)->n_accepted += t0.acc_count;
        yd_g(c.st)->n_overlap += t0.nov;
        yd_g(c.st)->since_resync += t0.nmv;
        for (int i = 0; i < 9; ++i) yd_g(c.st)->phase[i] += sh.ph[i];
    }
}

#define YK_CHAIN_INST(F, B)                                                                                    \
    template __global__ void yk_chain<F, B>(const YdChain *, const double *const *, const double *const *, int, int, \
                                            int, int, int);
YK_CHAIN_INST(false, 256)
YK_CHAIN_INST(false, 512)
YK_CHAIN_INST(true, 256)
YK_CHAIN_INST(true, 512)

// ========================================================================
// Grid chain: one Markov chain spread over G co-resident workgroups (a
// cooperative launch), for systems too large for one CU to step quickly.
// Workgroup g owns a contiguous range of k vectors (its share of every
// sweep and of the S(k) commit: the same range every move, so its S(k)
// values are only ever read and written by it) and a contiguous range of
// atoms (its share of the real-space scan, exact pass and back-term commit).
// Every workgroup runs the scalar sections itself (pick, proposal,
// acceptance) from the same RNG state and the same reduced sums, so all of
// them take the same decisions; workgroup 0 alone writes the shared state
// (positions and per-atom arrays of the moved molecule, energies, records,
// cell buckets).  Two grid barriers per move:
//   B1 after the old sweep: the per-workgroup sums of the old k-space forces;
//   B2 after the real-space pass and the new sweep: their sums.
// The partial sums of the workgroups are added in a fixed order (a lane
// butterfly over the workgroups).
//
// Ordering of the shared data: a workgroup's writes of move m (its S(k)
// range, its atoms' back terms, workgroup 0's finalize) are released before
// it arrives at B1 of move m + 1 and acquired by every workgroup leaving
// it.  Reads that precede B1 in a move (the picked molecule's positions for
// the old sweep) can race only with workgroup 0's finalize of the previous
// move; every workgroup therefore keeps that molecule's new positions in
// its own LDS for one move and takes them from there.  The per-site
// real-space arrays of the picked molecule are read after B1.
#define YG_PART 64                 // doubles per workgroup per phase in the partials buffer
#define YG_SPIN (1u << 22)         // bound of a barrier wait, in polls (seconds)

struct YgShared {
    YkShared s;
    double out2[YK_NRX + 1 + YK_NACC];              // the reduced real-space and new-sweep sums
    double lx[YD_MAXA], ly[YD_MAXA], lz[YD_MAXA];  // the last finalized molecule's positions
    int last_imol;                                  // -1: none in the previous move
    int gen, sync_ok;
};

// Grid-wide barrier of the G workgroups of a chain: bar (64 bits) counts
// arrivals since the launch; barrier number n (gen) is complete when it
// reaches G (n + 1).  Thread 0 releases the workgroup's writes at agent
// scope (they reach the L2 of every XCD), arrives, polls the count and
// acquires.  The wait is bounded: past YG_SPIN polls it returns false, and
// the caller stops the chain (every workgroup then stops at the same barrier
// or the next one).
__device__ bool yg_sync(unsigned int *bar, unsigned int G, int &gen, int *ok_lds) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned long long *cnt = reinterpret_cast<unsigned long long *>(bar);
        const unsigned long long target = (unsigned long long)(gen + 1) * G;
        __threadfence();
        const unsigned long long a = atomicAdd(cnt, 1ull) + 1ull;
        if (a < target) {
            unsigned int polls = 0;
            while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++polls > YG_SPIN) { ok = 0; break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __threadfence();
        gen = gen + 1;
        *ok_lds = ok;
    }
    __syncthreads();
    return *ok_lds != 0;
}

// Sums of the workgroups' partial vectors part[g * YG_PART + i], i < n,
// over g < G into out[i], by the first nw waves of the workgroup: wave w
// takes the entries i = w, w + nw, ...  For each chunk of 64 workgroups lane
// l holds workgroup g0 + l's value (all loads of the chunk are issued before
// the first sum), the lanes are summed by an xor butterfly (every lane ends
// with the same sum, in the same order on every run) and the chunks in
// order.  n <= PER nw; the caller synchronizes before out[] is read.
template <int PER>
__device__ void yg_reduce(const double *part, int G, int n, int nw, double *out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w >= nw) return;
    for (int g0 = 0; g0 < G; g0 += 64) {
        const int g = g0 + lane;
        double v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = w + u * nw;
            v[u] = i < n && g < G ? __hip_atomic_load(part + (size_t)g * YG_PART + i, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = w + u * nw;
            if (i < n) {  // wave-uniform
                double t = v[u];
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
                if (lane == 0) out[i] = g0 == 0 ? t : out[i] + t;
            }
        }
    }
}
#define YG_N1 YK_NACC                    // entries of the first partials (old sweep)
#define YG_N2 (YK_NRX + 1 + YK_NACC)     // of the second (real space, overlap flag, new sweep)
static_assert(YG_N1 <= 5 * 3 && YG_N2 <= 10 * 4 && YG_N2 < YG_PART, "partials fit the reductions");

// Full FP64 pass over atoms [a0, a1) when the workgroup's candidate list
// overflowed: the new sites' partial sums and overlap test and the old
// sites' components, like the exact pass, without stored back terms (the
// commit then recomputes them over the same range).
template <int NA>
__device__ void yg_rs_full(const YdChain &c, const YkMol &m, const double *sig6, const double *feps, int a0,
                           int a1, double (&acc)[YK_NRX], int &ov) {
    const int T = c.p.ntype, me = m.imol + 1;
    for (int j = a0 + (int)threadIdx.x; j < a1; j += (int)blockDim.x) {
        if (yd_g(c.molnum)[j] == me) continue;
        const int jt = yd_g(c.type)[j];
        const double qj = yd_g(c.q)[j], dhj = yd_g(c.dh)[j], xj = yd_g(c.x)[j], yj = yd_g(c.y)[j], zj = yd_g(c.z)[j];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const int it = m.typ[a];
            double xij = yd_wrap(m.xn[a] - xj), yij = yd_wrap(m.yn[a] - yj), zij = yd_wrap(m.zn[a] - zj);
            double r2 = xij * xij + yij * yij + zij * zij;
            const double d12 = (m.dh[a] + dhj) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) ov = 1;
            if (r2 < c.p.rmax2) {
                const bool lj = it != 0 && jt != 0;
                const int off = lj ? (it - 1) * T + (jt - 1) : 0;
                YdPair pr;
                yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
                yk_rs_acc(acc, a, pr, xij, yij, zij);
            }
            xij = yd_wrap(m.xo[a] - xj); yij = yd_wrap(m.yo[a] - yj); zij = yd_wrap(m.zo[a] - zj);
            r2 = xij * xij + yij * yij + zij * zij;
            if (r2 < c.p.rmax2) {
                const bool lj = it != 0 && jt != 0;
                const int off = lj ? (it - 1) * T + (jt - 1) : 0;
                YdPair pr;
                yd_pair(c.p, r2, false, m.q[a], qj, lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
                acc[YK_NRS + 0] += pr.potew;
                acc[YK_NRS + 1] += pr.potlj6;
                acc[YK_NRS + 2] += pr.potlj12;
            }
        }
    }
}

__device__ __forceinline__ void yg_rs_full_n(const YdChain &c, const YkMol &m, const double *sig6,
                                             const double *feps, int a0, int a1, double (&acc)[YK_NRX], int &ov) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: yg_rs_full<1>(c, m, sig6, feps, a0, a1, acc, ov); break;
        case 2: yg_rs_full<2>(c, m, sig6, feps, a0, a1, acc, ov); break;
        case 3: yg_rs_full<3>(c, m, sig6, feps, a0, a1, acc, ov); break;
        default: yg_rs_full<4>(c, m, sig6, feps, a0, a1, acc, ov); break;
    }
}

// Per-workgroup chain descriptors of a grid chain from the context's one:
// workgroup g gets its own back-term buffer, workgroup 0 alone the move
// records; the barrier words start at zero.
__global__ void yk_gchain_prep(const YdChain *src, YdChain *dst, YgGroup *grp, int G, double *rsb, unsigned int *bar,
                               double *part, const double *sig6, const double *feps) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G) {
        YdChain ch = *src;
        ch.rsb = rsb + (size_t)g * YD_NBCAP * 4;
        if (g > 0) ch.rec = nullptr;
        dst[g] = ch;
        YgGroup r;
        r.g = g; r.G = G; r.rep = 0; r.bar = bar; r.part = part; r.sig6 = sig6; r.feps = feps;
        grp[g] = r;
    }
    if (g < 2) bar[g] = 0u;
}

// The same for a batch of grid chains: workgroup w runs workgroup grp[w].g
// of chain grp[w].rep (src holds the chains' current descriptors); the first
// workgroup of each chain clears its barrier words.
__global__ void yk_gchain_bprep(const YdChain *src, YdChain *dst, const YgGroup *grp, int nwg, double *rsb) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwg) {
        const YgGroup r = grp[w];
        YdChain ch = src[r.rep];
        ch.rsb = rsb + (size_t)w * YD_NBCAP * 4;
        if (r.g > 0) ch.rec = nullptr;
        dst[w] = ch;
        if (r.g == 0) { r.bar[0] = 0u; r.bar[1] = 0u; }
    }
}

template <bool F32, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
yk_gchain(const YdChain *chains, const YgGroup *groups, int nsteps, int mode, int max_moves, int lds_atoms,
          int resync_every) {
    static_assert(BS == 256, "the grid chain's reductions and proposal take four waves");
    __shared__ YgShared gs;
    extern __shared__ unsigned int lds_pos[];
    YkShared &sh = gs.s;
    const YgGroup gr = groups[blockIdx.x];
    const int G = gr.G, g = gr.g, tid = threadIdx.x;
    unsigned int *const bar = gr.bar;
    double *const part = gr.part;
    const double *const sig6 = gr.sig6, *const feps = gr.feps;
    const YdChain &c = chains[blockIdx.x];
    const YdParams &p = c.p;
    const bool lead = g == 0;
    int total = mode == 1 ? nsteps : nsteps * p.nmol;
    if (max_moves > 0 && total > max_moves) total = max_moves;
    // own ranges: k in [k0, k1), atoms in [a0, a1)
    const int k0 = (int)((long long)p.nk * g / G), k1 = (int)((long long)p.nk * (g + 1) / G);
    const int a0 = (int)((long long)p.natom * g / G), a1 = (int)((long long)p.natom * (g + 1) / G);
    double *part1 = part, *part2 = part + (size_t)G * YG_PART;
    const bool use_lds = lds_atoms >= p.natom;
    unsigned int *const spos = use_lds ? lds_pos : reinterpret_cast<unsigned int *>(c.pos4);
    if (use_lds) {
        for (int i = tid; i < p.natom; i += BS) spos[i] = yk_pack_pos(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i]);
    } else {
        for (int i = a0 + tid; i < a1; i += BS) spos[i] = yk_pack_pos(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i]);
    }
    for (int i = a0 + tid; i < a1; i += BS) {  // the launch copies of the own atoms
        auto r4 = yd_g(reinterpret_cast<double4 *>(c.arec)) + 2 * (size_t)i;
        r4[0] = make_double4(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i], yd_g(c.q)[i]);
        r4[1] = make_double4(yd_g(c.dh)[i], (double)yd_g(c.type)[i], 0.0, 0.0);
        yd_g(reinterpret_cast<double4 *>(c.uf))[i] = make_double4(yd_g(c.uu)[i], yd_g(c.Fx)[i], yd_g(c.Fy)[i], yd_g(c.Fz)[i]);
    }
    if (tid == 0) { gs.gen = 0; gs.last_imol = -1; }
    if (c.nc > 0 && lead) yk_cells_build(c, &sh.cell_ok);
    if (c.nc > 0 && lead && tid == 0) part1[YG_PART - 1] = sh.cell_ok;  // header word: cells usable
    YkScalar &t0 = sh.t0;
    yk_chain_mu(c, sh, tid);
    if (tid == 0) {
        for (int i = 0; i < 9; ++i) sh.ph[i] = 0;
        sh.t_last = wall_clock64();
        t0.acc_count = 0; t0.nmv = 0; t0.nov = 0; t0.prepicked = 0;
        t0.e_kc = yd_g(c.st)->e_kc; t0.e_klj = yd_g(c.st)->e_klj;
        t0.uu_ew = yd_g(c.st)->uu_ew; t0.uu_lj6 = yd_g(c.st)->uu_lj6; t0.uu_lj12 = yd_g(c.st)->uu_lj12;
        YdRng &rng = t0.rng;
        rng.key0 = yd_g(c.st)->key[0]; rng.key1 = yd_g(c.st)->key[1];
        rng.s0 = yd_g(c.st)->stream[0]; rng.s1 = yd_g(c.st)->stream[1];
        rng.ctr = yd_g(c.st)->ctr; rng.constant = mode == 1; rng.cu = 0.0;
        rng.reset_cache();
        rng.list = yd_g(c.st)->ulist; rng.nlist = yd_g(c.st)->nlist;
    }
    bool ok = yg_sync(bar, G, gs.gen, &gs.sync_ok);
    if (ok && c.nc > 0 && !lead && tid == 0)
        sh.cell_ok = (int)__hip_atomic_load(part1 + YG_PART - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int mv = 0; ok && mv < total; ++mv) {
        // The periodic FP64 rebuild between move resync_every k - 1 and
        // resync_every k (uniform over the chain's workgroups): after a grid
        // barrier (the last move's S(k) commits and the lead's finalize of
        // the positions), each workgroup rebuilds S(k) over its own k range
        // with the one-workgroup rebuild's per-k arithmetic, and the lead
        // the cell buckets; after a second barrier every workgroup sums the
        // energies over all k (the same sum in each, stored by the lead) and
        // recomputes mu, as a new launch would.
        if (resync_every > 0 && mv > 0 && mv % resync_every == 0) {
            if (!(ok = yg_sync(bar, G, gs.gen, &gs.sync_ok))) break;
            yk_rebuild_sums<1>(c, sh, tid, k0, k1, use_lds ? reinterpret_cast<double *>(lds_pos) : nullptr,
                               use_lds ? 4 * p.natom : 0);
            __syncthreads();
            if (use_lds) {  // the chunk tables took the packed positions' LDS
                for (int i = tid; i < p.natom; i += BS) spos[i] = yk_pack_pos(yd_g(c.x)[i], yd_g(c.y)[i], yd_g(c.z)[i]);
            }
            if (c.nc > 0 && lead) yk_cells_build(c, &sh.cell_ok);
            if (c.nc > 0 && lead && tid == 0) part1[YG_PART - 1] = sh.cell_ok;
            if (!(ok = yg_sync(bar, G, gs.gen, &gs.sync_ok))) break;
            if (c.nc > 0 && !lead && tid == 0)
                sh.cell_ok = (int)__hip_atomic_load(part1 + YG_PART - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            yk_rebuild_energies(c, sh, tid, lead, gs.out2);
            yk_chain_mu(c, sh, tid);
            if (tid == 0) { t0.e_kc = gs.out2[0]; t0.e_klj = gs.out2[1]; }
            __syncthreads();
        }
        if (tid == 0) {
            if (t0.prepicked) {  // taken during the last move's acceptance
                yk_sec_pick_as(c, sh.m, t0.nimol);
                t0.rng = t0.rng_next;
                t0.prepicked = 0;
            } else {
                yk_sec_pick(c, sh.m, t0.rng, mode, mv);
            }
            sh.ov = 0;
            sh.nbcount = 0;
        }
        __syncthreads();
        if (sh.m.imol < 0) break;  // exchange step: the launch ends here and the host runs it
        const bool cached = gs.last_imol == sh.m.imol;
        if (tid < sh.m.na) {
            YkMol &m = sh.m;
            const int a = tid, gi = m.first + a;
            m.q[a] = yd_g(c.q)[gi]; m.typ[a] = yd_g(c.type)[gi]; m.mass[a] = yd_g(c.mass)[gi]; m.dh[a] = yd_g(c.dh)[gi];
            if (cached) { m.xo[a] = gs.lx[a]; m.yo[a] = gs.ly[a]; m.zo[a] = gs.lz[a]; }
            else { m.xo[a] = yd_g(c.x)[gi]; m.yo[a] = yd_g(c.y)[gi]; m.zo[a] = yd_g(c.z)[gi]; }
        }
        __syncthreads();
        if (tid == 0) { yk_sec_slots(sh.m); gs.last_imol = -1; }
        YK_TICK(0);
        yk_mol_tables(c, sh.m, false, tid);
        __syncthreads();
        YK_TICK(1);
        // ---- old k-space forces over the own k range --------------------
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, false, true>(c, sh.m, acc, k0 + tid, BS, k1);
            yd_block_sum<YK_NACC>(acc, sh.red, sh.out);
            if (tid < YK_NACC) part1[(size_t)g * YG_PART + tid] = sh.out[tid];
        }
        if (!(ok = yg_sync(bar, G, gs.gen, &gs.sync_ok))) break;
        if (c.nc > 0 && !lead && tid == 0)
            sh.cell_ok = (int)__hip_atomic_load(part1 + YG_PART - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // waves 0-2 reduce; wave 3 meanwhile loads the per-site real-space
        // arrays (written before B1 by their owner)
        yg_reduce<5>(part1, G, YG_N1, 3, sh.out);
        if (tid >= 192 && tid - 192 < sh.m.na) {
            const int a = tid - 192;
            const double4 u = yd_g(reinterpret_cast<const double4 *>(c.uf))[sh.m.first + a];
            sh.uu_old[a] = u.x;
            sh.fr_old[a][0] = u.y; sh.fr_old[a][1] = u.z; sh.fr_old[a][2] = u.w;
        }
        __syncthreads();
        YK_TICK(2);
        // ---- proposal (every workgroup, the same result) ----------------
        if (tid < 64) {
            yk_sec_propose_pre<true>(c, sh.m, t0, sh.out, sh.fr_old, sh.uu_old);
            t0.mcm2 = t0.mcm;
            t0.rng2 = t0.rng;
        }
        __syncthreads();
        if (tid < 64) {
            yk_displacement(p, t0.mcm, t0.rng);
        } else if (tid < 128) {
            yk_displacement_draws(p, t0.mcm2, t0.rng2);
            yk_rotation<true, true>(p, t0.mcm2, t0.rng2);
        } else if (tid < 192) {  // Q_old of the rotation, off the rotation wave's path
            const double qo = yk_qold(p, t0.mcm);
            if (tid == 128) t0.qold_pre = qo;
        }
        __syncthreads();
        if (tid < 64) {
            YkMcm &a = t0.mcm;
            const YkMcm &b = t0.mcm2;
            a.ca = b.ca; a.sa = b.sa; a.ct = b.ct; a.st = b.st; a.cphi = b.cphi; a.sphi = b.sphi;
            a.v[0] = b.v[0]; a.v[1] = b.v[1]; a.v[2] = b.v[2];
            a.vmaxold = b.vmaxold; a.angle = b.angle; a.qold = t0.qold_pre;
            a.stall = b.stall;
            t0.rng = t0.rng2;
            int stall = yk_sec_propose_post(sh.m, t0);
            if (tid == 0) sh.stall = stall;
        }
        __syncthreads();
        YK_TICK(3);
        if (sh.stall) {
            if (tid == 0) {
                ++t0.nmv;
                yk_sec_reject_rec(c, sh.m, t0.mcm, mv, 2);
            }
            __syncthreads();
            continue;
        }
        // ---- real-space pass over the own atoms, then the new sweep over
        // the own k range; both partials go out before B2 -------------------
        {
            double acc[YK_NRX];
#pragma unroll
            for (int i = 0; i < YK_NRX; ++i) acc[i] = 0.0;
            int ov = 0;
            yk_mol_tables(c, sh.m, true, tid);
            // Candidates of the workgroup: with usable cells, its share of
            // the visited cells' atoms (at most YK_NBCAP, so the list cannot
            // overflow); else the atoms of its own range.  Either way each
            // atom is one workgroup's candidate at most.
            // With cells the scan writes the list itself, in the order of
            // the visited atoms; over the atom range it marks the bitmap's
            // words of the range (only they are cleared and compacted).
            bool listed = false;
            if (c.nc > 0 && sh.cell_ok)
                listed = yk_rs_scan_cells_list_n(c, sh.m, spos, sh.cl_cell, sh.cl_start, &sh.ncl, g, G, sh.nblist,
                                                 &sh.nbcount, reinterpret_cast<int *>(sh.bm));
            if (!listed) {
                const int w0 = a0 >> 5, w1 = (a1 + 31) >> 5;
                if (w1 - w0 <= YK_BMWORDS) {
                    for (int i = tid; i < w1 - w0; i += BS) sh.bm[i] = 0u;
                    __syncthreads();
                    yk_rs_scan_n<true>(c, sh.m, spos, sh.bm, a0, a1, w0);
                    __syncthreads();
                    yk_compact(sh.bm, w1 - w0, sh.nblist, &sh.nbcount, sh.wsum, w0);
                } else if (tid == 0) {
                    sh.nbcount = YK_NBCAP + 1;
                }
            }
            __syncthreads();
            if (sh.nbcount <= YK_NBCAP)
                yl_rs_exact_n<true>(c, sh.m, sig6, feps, sh.nblist, sh.nbcount, acc, ov, tid, BS);
            else
                yg_rs_full_n(c, sh.m, sig6, feps, a0, a1, acc, ov);
            // (LDS-only barriers: the back terms are read back by the thread
            // that stored them; the partials are released at the barrier)
            yk_block_sum_used(acc, sh.red, sh.out, [](int) { return true; });
            if (tid < YK_NRX) part2[(size_t)g * YG_PART + tid] = sh.out[tid];
            // overlap flag of the workgroup (any thread)
            if (ov) sh.ov = 1;
            yk_lds_sync();
            if (tid == 0) part2[(size_t)g * YG_PART + YK_NRX] = (double)sh.ov;
        }
        YK_TICK(4);
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, true, true>(c, sh.m, acc, k0 + tid, BS, k1);
            yd_block_sum<YK_NACC>(acc, sh.red, sh.out);
            if (tid < YK_NACC) part2[(size_t)g * YG_PART + YK_NRX + 1 + tid] = sh.out[tid];
        }
        YK_TICK(5);
        if (!(ok = yg_sync(bar, G, gs.gen, &gs.sync_ok))) break;
        yg_reduce<10>(part2, G, YG_N2, 4, gs.out2);
        __syncthreads();
        if (tid == 0) {
            const double *o = gs.out2;
            sh.ov = o[YK_NRX] != 0.0;
            for (int a = 0; a < sh.m.na; ++a) {
                t0.uun[a] = o[a];
                t0.frn[a][0] = o[YD_MAXA + 3 * a + 0];
                t0.frn[a][1] = o[YD_MAXA + 3 * a + 1];
                t0.frn[a][2] = o[YD_MAXA + 3 * a + 2];
            }
            for (int i = 0; i < 3; ++i) {
                t0.cmp_new[i] = o[4 * YD_MAXA + i];
                t0.cmp_old[i] = o[YK_NRS + i];
            }
        }
        __syncthreads();
        if (sh.ov) {
            if (tid == 0) {
                ++t0.nmv; ++t0.nov;
                yk_sec_reject_rec(c, sh.m, t0.mcm, mv, 1);
            }
            __syncthreads();
            continue;
        }
        // ---- acceptance (every workgroup, the same decision); wave 1
        // meanwhile takes the next move's pick, whose draws start one draw
        // after t0.rng2 (the acceptance takes exactly one), and wave 2
        // computes Q_new, which wave 0 takes after the barrier for the
        // acceptance test ---------------------------------------------------
        if (tid < 64) {
            const double du = yk_sec_accept_pre<true, true>(c, sh.m, t0, gs.out2 + YK_NRX + 1);
            if (tid == 0) t0.du_total = du;
        } else if (tid < 128) {
            if (mode == 0 && mv + 1 < total) {
                t0.rng_next = t0.rng2;
                (void)t0.rng_next.next();
                const int nimol = yk_pick_imol(c, t0.rng_next, mode, mv + 1);
                t0.nimol = nimol;
                t0.prepicked = 1;
            }
        } else if (tid < 192) {
            double Fn[3], Tn[3], mn[3], du_ext;
            yk_new_ft(p, sh.m, t0, gs.out2 + YK_NRX + 1, Fn, Tn, mn, du_ext);
            const double qn = yk_qnew(p, Tn);
            if (tid == 128) t0.qnew_pre = qn;
        }
        __syncthreads();
        if (tid < 64) {
            t0.mcm.qnew = t0.qnew_pre;
            int accd = yk_sec_accept_post(c, sh.m, t0, t0.du_total, mv);
            if (tid == 0) {
                sh.accepted = accd;
                ++t0.nmv;
                t0.acc_count += accd;
            }
        }
        __syncthreads();
        YK_TICK(6);
        if (!sh.accepted) continue;
        // ---- commit of the own k range and the own atoms' back terms ----
        yk_commit<F32, true>(c, sh.m, k0 + tid, BS, k1);
        if (sh.nbcount <= YK_NBCAP) {
            yk_rs_commit_list<true>(c, sh.nblist, sh.nbcount, tid, BS);
        } else {
            double acc[3] = {0, 0, 0};
            const int me = sh.m.imol + 1;
            for (int j = a0 + tid; j < a1; j += BS)
                if (yd_g(c.molnum)[j] != me) yk_rs_apply<true>(c, sh.m, sig6, feps, j, acc);
        }
        // The commit's and the finalize's global stores are not waited for
        // here: the next move reads them after its pick barrier (within the
        // workgroup) or after the first grid barrier, which starts with a
        // full workgroup barrier and the agent-scope release.
        yk_lds_sync();
        YK_TICK(7);
        if (tid == 0) {
            if (lead) {
                yk_sec_finalize(c, sh.m, t0, nullptr, true, use_lds ? nullptr : spos, true);
                if (c.nc > 0 && sh.cell_ok) {
                    yk_cells_move(c, sh.m, &sh.cell_ok);
                    part1[YG_PART - 1] = sh.cell_ok;
                }
            }
            for (int a = 0; a < sh.m.na; ++a) {
                if (use_lds) spos[sh.m.first + a] = yk_pack_pos(sh.m.xn[a], sh.m.yn[a], sh.m.zn[a]);
                gs.lx[a] = sh.m.xn[a]; gs.ly[a] = sh.m.yn[a]; gs.lz[a] = sh.m.zn[a];
            }
            gs.last_imol = sh.m.imol;
            for (int d = 0; d < 3; ++d) t0.mu[d] = t0.mu_new[d];
        }
        yk_lds_sync();
        YK_TICK(8);
    }
    // the launch copy of the per-atom arrays back into the state, own atoms
    // (after a barrier: workgroup 0's finalize wrote some of them)
    if (ok) ok = yg_sync(bar, G, gs.gen, &gs.sync_ok);
    if (ok) {
        for (int i = a0 + tid; i < a1; i += BS) {
            const double4 u = yd_g(reinterpret_cast<const double4 *>(c.uf))[i];
            yd_g(c.uu)[i] = u.x; yd_g(c.Fx)[i] = u.y; yd_g(c.Fy)[i] = u.z; yd_g(c.Fz)[i] = u.w;
        }
    }
    if (tid == 0 && lead) {
        if (mode != 1) yd_g(c.st)->ctr = t0.rng.ctr;
        yd_g(c.st)->n_moves += t0.nmv;
        yd_g(c.st)->n_accepted += t0.acc_count;
        yd_g(c.st)->n_overlap += t0.nov;
        yd_g(c.st)->since_resync += t0.nmv;
        for (int i = 0; i < 9; ++i) yd_g(c.st)->phase[i] += sh.ph[i];
        if (!ok) yd_g(c.st)->grid_fail = 1;
    }
}

#define YK_GCHAIN_INST(F, B) \
    template __global__ void yk_gchain<F, B>(const YdChain *, const YgGroup *, int, int, int, int, int);
YK_GCHAIN_INST(false, 256)
YK_GCHAIN_INST(true, 256)

// ========================================================================
// The wave chain kernel (batched replicas): one wavefront per replica,
// YW_WPB replicas per workgroup, no workgroup barriers.  Each wave runs
// run_mc cycles of its own chain exactly as the workgroup kernel does; its
// sums are wave-shuffle reductions, its real-space candidate list is built
// in ascending atom order from wave ballots, and several replicas share
// every SIMD so that one replica's latency is hidden by the others'.
#define YW_WPB 4

// Per-replica LDS of the wave kernel.
struct YwShared {
    YkMol m;
    YkScalar t0;
    double out[YK_NRX > YK_NACC ? YK_NRX : YK_NACC];
    double fr_old[YD_MAXA][3], uu_old[YD_MAXA];
};

// Orders the wave's LDS and global accesses before and after the call for
// every lane of the wave (the lanes of one wave run in lockstep; this keeps
// the compiler from moving accesses across and waits for them to land).
__device__ __forceinline__ void yw_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Sum of NV doubles over the wave; every lane gets the same sums (each
// butterfly step adds the same two operands in both partner lanes).
template <int NV>
__device__ __forceinline__ void yw_sum(double (&v)[NV]) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double a = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        v[i] = a;
    }
}

// Candidate scan of the wave kernel: atoms of other molecules within the
// cut-off (plus the FP32 margin) of an old or a new site, appended in
// ascending order through ballots.  Returns the count (the list holds the
// first YK_NBCAP).
template <int NA>
__device__ int yw_scan(const YdChain &c, const YkMol &m, int lane) {
    const int N = c.p.natom, me = m.imol + 1;
    const float r2f = (float)c.p.rmax2 + 1e-5f;
    float xof[NA], yof[NA], zof[NA], xnf[NA], ynf[NA], znf[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        xof[a] = (float)m.xo[a]; yof[a] = (float)m.yo[a]; zof[a] = (float)m.zo[a];
        xnf[a] = (float)m.xn[a]; ynf[a] = (float)m.yn[a]; znf[a] = (float)m.zn[a];
    }
    int cnt = 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int base = 0; base < N; base += 64) {
        int j = base + lane;
        bool near = false;
        if (j < N) {
            float4 pj = yd_g(c.pos4)[j];
            near = yk_near<NA>(xof, yof, zof, xnf, ynf, znf, pj, r2f) && __float_as_int(pj.w) != me;
        }
        unsigned long long mask = __ballot(near);
        if (near) {
            int pos = cnt + __popcll(mask & below);
            if (pos < YK_NBCAP) yd_g(c.nbl)[pos] = j;
        }
        cnt += __popcll(mask);
    }
    return cnt;
}

__device__ __forceinline__ int yw_scan_n(const YdChain &c, const YkMol &m, int lane) {
    switch (__builtin_amdgcn_readfirstlane(m.na)) {
        case 1: return yw_scan<1>(c, m, lane);
        case 2: return yw_scan<2>(c, m, lane);
        case 3: return yw_scan<3>(c, m, lane);
        default: return yw_scan<4>(c, m, lane);
    }
}

// OCC: waves per SIMD the register allocation is sized for (1: four
// replicas per CU; 2: eight, at the cost of some register spills).
template <bool F32, int OCC>
__global__ void __launch_bounds__(YW_WPB * 64) __attribute__((amdgpu_waves_per_eu(OCC, OCC)))
yk_wchain(const YdChain *chains, int nchains,
                                                         const double *const *sig6s, const double *const *feps_s,
                                                         int ncycles) {
    __shared__ YwShared shw[YW_WPB];
    // the replica index is uniform over the wave: keep it (and so the chain
    // descriptor's fields) in scalar registers
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ci = __builtin_amdgcn_readfirstlane(blockIdx.x * YW_WPB + wv);
    if (ci >= nchains) return;  // the whole wave leaves; nothing waits for it
    YwShared &sh = shw[wv];
    const YdChain &c = chains[ci];
    const double *sig6 = sig6s[ci], *feps = feps_s[ci];
    const YdParams &p = c.p;
    YkScalar &t0 = sh.t0;
    YkMol &m = sh.m;
    // FP32 copy of the positions for the scan, and mu from scratch
    // (EwaldSumExternal.f90:165-183), carried incrementally afterwards.
    {
        double v[3] = {0, 0, 0};
        for (int i = lane; i < p.natom; i += 64) {
            double x = yd_g(c.x)[i], y = yd_g(c.y)[i], z = yd_g(c.z)[i], q = yd_g(c.q)[i];
            yd_g(c.pos4)[i] = make_float4((float)x, (float)y, (float)z, __int_as_float(yd_g(c.molnum)[i]));
            v[0] += q * x; v[1] += q * y; v[2] += q * z;
        }
        yw_sum<3>(v);
        t0.mu[0] = v[0]; t0.mu[1] = v[1]; t0.mu[2] = v[2];
        t0.acc_count = 0; t0.nmv = 0; t0.nov = 0;
        YdRng &rng = t0.rng;
        rng.key0 = yd_g(c.st)->key[0]; rng.key1 = yd_g(c.st)->key[1];
        rng.s0 = yd_g(c.st)->stream[0]; rng.s1 = yd_g(c.st)->stream[1];
        rng.ctr = yd_g(c.st)->ctr; rng.constant = 0; rng.cu = 0.0;
        rng.reset_cache();
        rng.list = yd_g(c.st)->ulist; rng.nlist = yd_g(c.st)->nlist;
    }
    yw_sync();
    const int total = ncycles * p.nmol;
    for (int mv = 0; mv < total; ++mv) {
        yk_sec_pick(c, m, t0.rng, 0, mv);
        yw_sync();
        if (m.imol < 0) break;  // exchange step: the host takes over (launched only without them)
        if (lane < m.na) {
            int a = lane, g = m.first + a;
            m.q[a] = yd_g(c.q)[g]; m.typ[a] = yd_g(c.type)[g]; m.mass[a] = yd_g(c.mass)[g]; m.dh[a] = yd_g(c.dh)[g];
            m.xo[a] = yd_g(c.x)[g]; m.yo[a] = yd_g(c.y)[g]; m.zo[a] = yd_g(c.z)[g];
            sh.uu_old[a] = yd_g(c.uu)[g];
            sh.fr_old[a][0] = yd_g(c.Fx)[g]; sh.fr_old[a][1] = yd_g(c.Fy)[g]; sh.fr_old[a][2] = yd_g(c.Fz)[g];
        }
        yw_sync();
        yk_sec_slots(m);
        yk_mol_tables(c, m, false, lane);
        yw_sync();
        // ---- old k-space forces (MC.f90:946-982) ------------------------
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, false>(c, m, acc, lane, 64);
            yw_sum<YK_NACC>(acc);
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) sh.out[i] = acc[i];
        }
        yw_sync();
        // ---- proposal (MC.f90:984-1067) ---------------------------------
        int stall = yk_sec_propose<true>(c, m, t0, sh.out, sh.fr_old, sh.uu_old);
        yw_sync();
        if (stall) {
            ++t0.nmv;
            if (lane == 0) yk_sec_reject_rec(c, m, t0.mcm, mv, 2);
            yw_sync();
            continue;
        }
        // ---- new real-space partial sum (MC.f90:1077-1090) --------------
        yk_mol_tables(c, m, true, lane);
        int nb = yw_scan_n(c, m, lane);
        yw_sync();
        int ov = 0;
        {
            double acc[YK_NRX];
#pragma unroll
            for (int i = 0; i < YK_NRX; ++i) acc[i] = 0.0;
            if (nb <= YK_NBCAP) {
                yl_rs_exact_n(c, m, sig6, feps, c.nbl, nb, acc, ov, lane, 64);
            } else {
                int junk = 0;
                double a2[YK_NRS];
#pragma unroll
                for (int i = 0; i < YK_NRS; ++i) a2[i] = 0.0;
                yk_rs_move_n(c, m, sig6, feps, a2, ov, c.nbl, &junk, lane, 64);
#pragma unroll
                for (int i = 0; i < YK_NRS; ++i) acc[i] = a2[i];
            }
            yw_sum<YK_NRX>(acc);
            for (int a = 0; a < m.na; ++a) {
                t0.uun[a] = acc[a];
                t0.frn[a][0] = acc[YD_MAXA + 3 * a + 0];
                t0.frn[a][1] = acc[YD_MAXA + 3 * a + 1];
                t0.frn[a][2] = acc[YD_MAXA + 3 * a + 2];
            }
            for (int i = 0; i < 3; ++i) {
                t0.cmp_new[i] = acc[4 * YD_MAXA + i];
                t0.cmp_old[i] = acc[YK_NRS + i];
            }
        }
        ov = __any(ov);
        yw_sync();
        if (ov) {  // overlap: rejected, no acceptance draw (MC.f90:1080-1084)
            ++t0.nmv; ++t0.nov;
            if (lane == 0) yk_sec_reject_rec(c, m, t0.mcm, mv, 1);
            yw_sync();
            continue;
        }
        // ---- new k-space sweep: dU and new forces (MC.f90:1096-1173) ----
        {
            double acc[YK_NACC];
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
            yl_sweep_n<F32, true>(c, m, acc, lane, 64);
            yw_sum<YK_NACC>(acc);
#pragma unroll
            for (int i = 0; i < YK_NACC; ++i) sh.out[i] = acc[i];
        }
        yw_sync();
        // ---- acceptance (MC.f90:1176-1233) ------------------------------
        int accd = yk_sec_accept<true>(c, m, t0, sh.out, mv);
        ++t0.nmv;
        t0.acc_count += accd;
        yw_sync();
        if (!accd) continue;
        // ---- commit (MC.f90:1242-1318) ----------------------------------
        yk_commit<F32>(c, m, lane, 64);
        if (nb <= YK_NBCAP) {
            yk_rs_commit_list(c, c.nbl, nb, lane, 64);
        } else {
            double acc[3] = {0, 0, 0};
            yk_rs_commit_all(c, m, sig6, feps, acc, lane, 64);
            yw_sum<3>(acc);
            for (int i = 0; i < 3; ++i) t0.cmp_old[i] = acc[i];
        }
        yw_sync();
        if (lane == 0) yk_sec_finalize(c, m, t0, c.pos4);
        for (int d = 0; d < 3; ++d) t0.mu[d] = t0.mu_new[d];
        yw_sync();
    }
    if (lane == 0) {
        yd_g(c.st)->ctr = t0.rng.ctr;
        yd_g(c.st)->n_moves += t0.nmv;
        yd_g(c.st)->n_accepted += t0.acc_count;
        yd_g(c.st)->n_overlap += t0.nov;
        yd_g(c.st)->since_resync += t0.nmv;
    }
}

// Only the FP64 form of the wave kernel is built: the FP32C form made an
// illegal memory access on its first GPU launch and stays switched off until
// the cause is found (FP32C batches run the workgroup kernel).
template __global__ void yk_wchain<false, 1>(const YdChain *, int, const double *const *, const double *const *, int);
template __global__ void yk_wchain<false, 2>(const YdChain *, int, const double *const *, const double *const *, int);

// ========================================================================
// Single-call entry points (one workgroup): the per-function parity surface.

// calc_dU_coulomb + calc_dU_LJ for molecule imol moved to (x,y,z).
template <bool F32>
__global__ void __launch_bounds__(YK_BS) yk_kspace_delta(YdChain c, int imol, const double *x, const double *y,
                                                         const double *z, double *out) {
    __shared__ YkShared sh;
    if (threadIdx.x == 0) {
        yk_load_mol(c, sh.m, imol);
        for (int a = 0; a < sh.m.na; ++a) { sh.m.xn[a] = x[a]; sh.m.yn[a] = y[a]; sh.m.zn[a] = z[a]; }
    }
    __syncthreads();
    yk_mol_tables(c, sh.m, false, threadIdx.x);
    __syncthreads();
    yk_mol_tables(c, sh.m, true, threadIdx.x);
    __syncthreads();
    double acc[YK_NACC];
    for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
    yk_sweep_n<F32, true>(c, sh.m, acc, threadIdx.x, blockDim.x);
    yd_block_sum<YK_NACC>(acc, sh.red, sh.out);
    if (threadIdx.x == 0) { out[0] = sh.out[0]; out[1] = sh.out[1]; }
}
template __global__ void yk_kspace_delta<false>(YdChain, int, const double *, const double *, const double *,
                                                double *);
template __global__ void yk_kspace_delta<true>(YdChain, int, const double *, const double *, const double *,
                                               double *);

// ForceKSpace_coulomb + _LJ for n <= YD_MAXA atoms first..first+n-1 at (x,y,z).
template <bool F32>
__global__ void __launch_bounds__(YK_BS) yk_kspace_forces(YdChain c, int first, int n, const double *x,
                                                          const double *y, const double *z, double *out) {
    __shared__ YkShared sh;
    if (threadIdx.x == 0) {
        YkMol &m = sh.m;
        m.imol = -1; m.first = first; m.na = n;
        for (int a = 0; a < n; ++a) {
            m.q[a] = yd_g(c.q)[first + a];
            m.typ[a] = yd_g(c.type)[first + a];
            m.xo[a] = x[a]; m.yo[a] = y[a]; m.zo[a] = z[a];
        }
        yk_sec_slots(m);  // atoms beyond YK_MAXP distinct LJ types keep slot -1
    }
    __syncthreads();
    yk_mol_tables(c, sh.m, false, threadIdx.x);
    __syncthreads();
    double acc[YK_NACC];
    for (int i = 0; i < YK_NACC; ++i) acc[i] = 0.0;
    yk_sweep_n<F32, false>(c, sh.m, acc, threadIdx.x, blockDim.x);
    yd_block_sum<YK_NACC>(acc, sh.red, sh.out);
    if (threadIdx.x < 3 * n) {
        int a = threadIdx.x / 3, d = threadIdx.x % 3;
        out[threadIdx.x] = YD_FOUR_PI * sh.out[YK_FK(a, d)];
    }
}
template __global__ void yk_kspace_forces<false>(YdChain, int, int, const double *, const double *, const double *,
                                                 double *);
template __global__ void yk_kspace_forces<true>(YdChain, int, int, const double *, const double *, const double *,
                                                double *);

// Real-space partial sum of molecule imol at (x,y,z), with back arrays.
// out: [0..na) uu, [YD_MAXA + 3a + d] F, [16..18] comps, [19] overlap.
__global__ void __launch_bounds__(YK_BS) yk_rspace_partial(YdChain c, int imol, const double *x, const double *y,
                                                           const double *z, const double *sig6, const double *feps,
                                                           double *out, double *uub, double *fxb, double *fyb,
                                                           double *fzb) {
    __shared__ YkShared sh;
    if (threadIdx.x == 0) {
        yk_load_mol(c, sh.m, imol);
        for (int a = 0; a < sh.m.na; ++a) { sh.m.xn[a] = x[a]; sh.m.yn[a] = y[a]; sh.m.zn[a] = z[a]; }
        sh.ov = 0;
    }
    __syncthreads();
    double acc[YK_NRS];
    for (int i = 0; i < YK_NRS; ++i) acc[i] = 0.0;
    int ov = 0;
    yk_rs_new_n(c, sh.m, sig6, feps, acc, ov);
    if (ov) sh.ov = 1;
    yd_block_sum<YK_NRS>(acc, sh.red, sh.out);
    if (threadIdx.x < YK_NRS) out[threadIdx.x] = sh.out[threadIdx.x];
    if (threadIdx.x == 0) out[YK_NRS] = sh.ov;
    if (uub) {
        const int N = c.p.natom, T = c.p.ntype, me = imol + 1;
        const YkMol &m = sh.m;
        for (int j = threadIdx.x; j < N; j += (int)blockDim.x) {
            double b[4] = {0, 0, 0, 0};
            if (yd_g(c.molnum)[j] != me) {
                int jt = yd_g(c.type)[j];
                for (int a = 0; a < m.na; ++a) {
                    int it = m.typ[a];
                    double xij = yd_wrap(m.xn[a] - yd_g(c.x)[j]), yij = yd_wrap(m.yn[a] - yd_g(c.y)[j]),
                           zij = yd_wrap(m.zn[a] - yd_g(c.z)[j]);
                    double r2 = xij * xij + yij * yij + zij * zij;
                    if (r2 >= c.p.rmax2) continue;
                    bool lj = it != 0 && jt != 0;
                    int off = lj ? (it - 1) * T + (jt - 1) : 0;
                    YdPair pr;
                    yd_pair(c.p, r2, false, m.q[a], yd_g(c.q)[j], lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
                    b[0] = b[0] + pr.pot;
                    b[1] = b[1] - xij * pr.FR;
                    b[2] = b[2] - yij * pr.FR;
                    b[3] = b[3] - zij * pr.FR;
                }
            }
            uub[j] = b[0]; fxb[j] = b[1]; fyb[j] = b[2]; fzb[j] = b[3];
        }
    }
}

// ========================================================================
// Exchange steps (mc_xchange, MC.f90:1325-1954).

// Joint delta of the sums: for every set (Coulomb, then the present LJ
// types), new minus old, each summed over the atoms in order into nulified
// sums (RhoSquared_addAtom + RhoSquared_sub, MC.f90:1464-1488, 1602-1641).
// D[(1+T)][8][nk] FP64; sets of absent types are not written.
__global__ void __launch_bounds__(256) yk_xchg_delta(YdChain c, YxAtoms at, double *D) {
    __shared__ double tab[2][YX_MAXA * 6 * YD_KP];
    const int tid = threadIdx.x, n = at.n, nk = c.p.nk, T = c.p.ntype;
    for (int w = tid; w < 2 * 3 * n; w += (int)blockDim.x) {
        int s = w / (3 * n), r = w % (3 * n), a = r / 3, d = r % 3;
        double *t = tab[s] + a * 6 * YD_KP + 2 * d * YD_KP;
        yd_sincos_axis(at.pos[s][d][a], c.p.kmax, t, t + YD_KP);
    }
    __syncthreads();
    for (int kk = blockIdx.x * blockDim.x + tid; kk < nk; kk += gridDim.x * blockDim.x) {
        int kxyz = yd_g(c.kxyz)[kk], kx = kxyz & 255, ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        for (int set = 0; set <= T; ++set) {
            if (set > 0 && !((at.pmask >> (set - 1)) & 1u)) continue;
            double ns[8] = {0, 0, 0, 0, 0, 0, 0, 0}, os[8] = {0, 0, 0, 0, 0, 0, 0, 0}, krc[8], krl[8];
            for (int a = 0; a < n; ++a) {
                if (set == 0 ? at.q[a] == 0.0 : at.typ[a] != set) continue;
                yd_kr_atom(tab[1] + a * 6 * YD_KP, at.q[a], at.typ[a], kx, ky, kz, krc, krl);
                for (int i = 0; i < 8; ++i) ns[i] = ns[i] + (set == 0 ? krc[i] : krl[i]);
                yd_kr_atom(tab[0] + a * 6 * YD_KP, at.q[a], at.typ[a], kx, ky, kz, krc, krl);
                for (int i = 0; i < 8; ++i) os[i] = os[i] + (set == 0 ? krc[i] : krl[i]);
            }
            size_t base = (size_t)set * 8 * nk + kk;
            for (int i = 0; i < 8; ++i) D[base + (size_t)i * nk] = ns[i] - os[i];
        }
    }
}

// Sum of the four sign slots of (2C + D1) D2 (EwaldSumKSpace.f90:528-753
// slot order: ppp, pmp, ppm, pmm).
__device__ __forceinline__ double yx_slots(const double *tw, const double *d2, int ky, int kz) {
    double loc = tw[0] * d2[0] + tw[1] * d2[1];
    if (ky > 0) loc = loc + tw[4] * d2[4] + tw[5] * d2[5];
    if (kz > 0) loc = loc + tw[2] * d2[2] + tw[3] * d2[3];
    if (ky > 0 && kz > 0) loc = loc + tw[6] * d2[6] + tw[7] * d2[7];
    return loc;
}

// calc_dU_coulomb and calc_dU_LJ of the joint delta (EwaldSumKSpace.f90:
// 528-753) against the current sums (FP32C: hi + lo); one workgroup.
template <bool F32>
__global__ void __launch_bounds__(256) yk_xchg_du(YdChain c, const double *D, unsigned pmask, double *out) {
    __shared__ double red[4 * 2], res[2];
    const int nk = c.p.nk, T = c.p.ntype;
    double acc[2] = {0.0, 0.0};
    for (int kk = threadIdx.x; kk < nk; kk += (int)blockDim.x) {
        int kxyz = yd_g(c.kxyz)[kk], ky = (kxyz >> 8) & 255, kz = (kxyz >> 16) & 255;
        double C[8], d[8], tw[8], d2[8];
        for (int set = 0; set <= T; ++set) {
            const bool pres = set == 0 || ((pmask >> (set - 1)) & 1u);
            size_t base = (size_t)set * 8 * nk + kk;
            for (int i = 0; i < 8; ++i) {
                size_t idx = base + (size_t)i * nk;
                C[i] = F32 ? (double)yd_g(c.Shi)[idx] + (double)yd_g(c.Slo)[idx] : yd_g(c.S)[idx];
                d[i] = pres ? D[idx] : 0.0;
                tw[i] = C[i] + C[i] + d[i];
            }
            if (set == 0) {
                acc[0] += yx_slots(tw, d, ky, kz) * yd_g(c.beta)[kk];
                continue;
            }
            for (int t2 = 1; t2 <= T; ++t2) {
                if (!((pmask >> (t2 - 1)) & 1u)) continue;
                size_t b2 = (size_t)t2 * 8 * nk + kk;
                for (int i = 0; i < 8; ++i) d2[i] = D[b2 + (size_t)i * nk];
                acc[1] += yx_slots(tw, d2, ky, kz) * yd_g(c.beta_lj)[((size_t)(set - 1) * T + (t2 - 1)) * nk + kk];
            }
        }
    }
    yd_block_sum<2>(acc, red, res);
    if (threadIdx.x == 0) { out[0] = res[0]; out[1] = res[1]; }
}
template __global__ void yk_xchg_du<false>(YdChain, const double *, unsigned, double *);
template __global__ void yk_xchg_du<true>(YdChain, const double *, unsigned, double *);

// S += D for the Coulomb set and the present LJ sets (FP32C: hi/lo split of
// the FP64 sum, as the move commit does).
template <bool F32>
__global__ void yk_xchg_commit_sums(YdChain c, const double *D, unsigned pmask) {
    const int nk = c.p.nk, T = c.p.ntype;
    const size_t n = (size_t)(1 + T) * 8 * nk;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int set = (int)(i / ((size_t)8 * nk));
        if (set > 0 && !((pmask >> (set - 1)) & 1u)) continue;
        if (F32) {
            double w = (double)yd_g(c.Shi)[i] + (double)yd_g(c.Slo)[i] + D[i];
            float hi = (float)w;
            yd_g(c.Shi)[i] = hi;
            yd_g(c.Slo)[i] = (float)(w - (double)hi);
        } else {
            yd_g(c.S)[i] = yd_g(c.S)[i] + D[i];
        }
    }
}
template __global__ void yk_xchg_commit_sums<false>(YdChain, const double *, unsigned);
template __global__ void yk_xchg_commit_sums<true>(YdChain, const double *, unsigned);

// Molecule-molecule real-space term: EwaldSumRealSpace_calc of molecule 1
// with j restricted to molecule 2's atoms (MC.f90:1573-1574, 1784-1785),
// positions as they stand in c.x/y/z.  out: uu_ew, uu_lj6, uu_lj12; the
// early return on an overlap is kept (EwaldSumRealSpace.f90:409-433).
__global__ void yk_rspace_pair(YdChain c, int imol1, int imol2, const double *sig6, const double *feps,
                               double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int T = c.p.ntype, f1 = yd_g(c.first)[imol1], l1 = yd_g(c.last)[imol1], f2 = yd_g(c.first)[imol2], l2 = yd_g(c.last)[imol2];
    double e0 = 0.0, e6 = 0.0, e12 = 0.0;
    for (int i = f1; i <= l1; ++i) {
        const int it = yd_g(c.type)[i];
        for (int j = f2; j <= l2; ++j) {
            const int jt = yd_g(c.type)[j];
            double xij = yd_wrap(yd_g(c.x)[i] - yd_g(c.x)[j]), yij = yd_wrap(yd_g(c.y)[i] - yd_g(c.y)[j]), zij = yd_wrap(yd_g(c.z)[i] - yd_g(c.z)[j]);
            double r2 = xij * xij + yij * yij + zij * zij;
            double d12 = (yd_g(c.dh)[i] + yd_g(c.dh)[j]) / 2.0;
            if (it > 0 && jt > 0 && r2 < d12 * d12) {
                out[0] = e0; out[1] = e6; out[2] = e12;
                return;
            }
            if (r2 >= c.p.rmax2) continue;
            bool lj = it != 0 && jt != 0;
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            YdPair pr;
            yd_pair(c.p, r2, false, yd_g(c.q)[i], yd_g(c.q)[j], lj, yd_g(sig6)[off], yd_g(feps)[off], pr);
            e0 = e0 + pr.potew;
            if (lj) {
                e6 = e6 + pr.potlj6;
                e12 = e12 + pr.potlj12;
            }
        }
    }
    out[0] = e0; out[1] = e6; out[2] = e12;
}

// Per-atom real-space arrays after an accepted exchange (MC.f90:1786-1810):
// every atom takes new1 - old1 + new2 - old2 of the back arrays, then the
// two molecules' atoms take their new partial sums (mv[m][a] = uu, fx, fy, fz).
// With n2 = 0 (one molecule, bn2/bo2 unused) it is movemc's commit.
__global__ void yk_xchg_commit_rs(YdChain c, const double *bn1, const double *bo1, const double *bn2,
                                  const double *bo2, const double *mv, int f1, int n1, int f2, int n2) {
    const int N = c.p.natom;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
        if (j >= f1 && j < f1 + n1) {
            const double *v = mv + 4 * (j - f1);
            yd_g(c.uu)[j] = v[0]; yd_g(c.Fx)[j] = v[1]; yd_g(c.Fy)[j] = v[2]; yd_g(c.Fz)[j] = v[3];
        } else if (j >= f2 && j < f2 + n2) {
            const double *v = mv + 4 * (YD_MAXA + j - f2);
            yd_g(c.uu)[j] = v[0]; yd_g(c.Fx)[j] = v[1]; yd_g(c.Fy)[j] = v[2]; yd_g(c.Fz)[j] = v[3];
        } else if (n2 > 0) {
            yd_g(c.uu)[j] = yd_g(c.uu)[j] + bn1[j] - bo1[j] + bn2[j] - bo2[j];
            yd_g(c.Fx)[j] = yd_g(c.Fx)[j] + bn1[N + j] - bo1[N + j] + bn2[N + j] - bo2[N + j];
            yd_g(c.Fy)[j] = yd_g(c.Fy)[j] + bn1[2 * N + j] - bo1[2 * N + j] + bn2[2 * N + j] - bo2[2 * N + j];
            yd_g(c.Fz)[j] = yd_g(c.Fz)[j] + bn1[3 * N + j] - bo1[3 * N + j] + bn2[3 * N + j] - bo2[3 * N + j];
        } else {  // one molecule (movemc's commit, MC.f90:1257-1275)
            yd_g(c.uu)[j] = yd_g(c.uu)[j] + bn1[j] - bo1[j];
            yd_g(c.Fx)[j] = yd_g(c.Fx)[j] + bn1[N + j] - bo1[N + j];
            yd_g(c.Fy)[j] = yd_g(c.Fy)[j] + bn1[2 * N + j] - bo1[2 * N + j];
            yd_g(c.Fz)[j] = yd_g(c.Fz)[j] + bn1[3 * N + j] - bo1[3 * N + j];
        }
    }
}
