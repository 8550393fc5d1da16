// yamc_device.h -- device-side data layout and arithmetic of the Ewald
// Monte Carlo path (gfx950).  Included by the kernel translation units only.
//
// Layout in HBM (one replica):
//   atoms       SoA x, y, z, q, dh (= hard core / L), mass : double[natom]
//               molnum (1-based), type (0 = no LJ)          : int[natom]
//   molecules   first, last (0-based inclusive)            : int[nmol]
//   k grid      kxyz packed kx | ky<<8 | kz<<16            : int[nk]
//               beta                                       : double[nk]
//               beta_lj, full symmetric T x T rows         : double[T*T*nk]
//   S(k)        [(1+T) sets][8 components][nk]  (set 0 = Coulomb,
//               set t = LJ type t; components cos/sin of ppp, ppm, pmp, pmm)
//               FP64: double; FP32C: float hi + float lo two-sum pairs
//   real space  uu, Fx, Fy, Fz : double[natom]
// Every k-space sweep walks k fastest, so a wavefront reads 64 consecutive
// doubles (512 B) of one component array per load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/yamc_ewald.h"

#define YD_PI 3.14159265358979323846
#define YD_TWO_PI (2.0 * YD_PI)
#define YD_FOUR_PI (4.0 * YD_PI)
#define YD_MAXA YAMC_MAX_MOL_ATOM
// Capacity of the per-move real-space candidate list (atoms within the
// cut-off of an old or a new site of the moved molecule).
#define YD_NBCAP 2048
#define YD_KP (YAMC_MAX_KMAX + 1)
#define YD_MAXT YAMC_MAX_LJ_TYPE

struct YdParams {
    int natom, nmol, ntype, kmax, nk, precision;
    double L, alpha, rmax2, xclb, kext, m2kext, dr, d_angle, lambda_f, lambda_c;
    double alpha_over_sqrt_pi, two_alpha_over_sqrt_pi, inv_alpha;
    double pxch;  // xchange_probability (runmc.f90:61); 0 = no exchange steps
};

// Scalars of the Monte Carlo state kept on the device.
struct YdState {
    double e_kc, e_klj, uu_ew, uu_lj6, uu_lj12, uu_ew_intra, uu_lj6_intra, uu_lj12_intra;
    double mu[3];
    double tail_c, tail6, tail12;
    uint32_t key[2], stream[2];
    unsigned long long ctr;
    // replay mode (ulist != null): the uniforms are ulist[ctr++] of a
    // recorded stream of nlist values instead of Philox
    const double *ulist;
    unsigned long long nlist;
    long long n_moves, n_accepted, n_overlap, since_resync;
    unsigned long long phase[10];  // wall-clock ticks per chain-kernel phase (thread 0)
    int grid_fail;                 // a grid chain's barrier wait ran past its bound (the chain stopped)
    int grid_pad;
};

// The two molecules of an exchange step (mc_xchange, MC.f90:1325-1954):
// their atoms, molecule 1 first, at the old and the new places.
#define YX_MAXA (2 * YD_MAXA)
struct YxAtoms {
    int n;                      // atoms of both molecules
    unsigned pmask;             // bit t-1: LJ type t is present among them
    int typ[YX_MAXA];
    double q[YX_MAXA];
    double pos[2][3][YX_MAXA];  // [old, new][x, y, z][atom]
};

struct YdChain {
    YdParams p;
    double *x, *y, *z, *q, *dh, *mass;
    int *molnum, *type, *first, *last;
    int *kxyz;
    double *beta, *beta_lj;
    double *S;            // FP64 sums
    float *Shi, *Slo;     // FP32C sums
    double *uu, *Fx, *Fy, *Fz;
    double *rsb;          // per candidate of the current move: new - old back terms (YD_NBCAP x 4)
    int *nbl;             // candidate list of the current move (YD_NBCAP), wave kernel
    float4 *pos4;         // FP32 positions + molecule number bits, wave kernel's scan
    // Launch-local atom records of the workgroup chain kernel, built at the
    // start of each launch: arec (8 doubles per atom: x, y, z, q, dh, type,
    // unused, unused) feeds the exact real-space pass with one 64-byte record
    // per candidate; uf (uu, Fx, Fy, Fz per atom) carries the per-atom
    // real-space arrays through the launch and is written back at its end.
    double *arec, *uf;
    // atoms of each S(k) set in index order, for the from-scratch sums:
    // set_atoms[set_off[s] .. set_off[s + 1]) = the charged atoms (s = 0) or
    // the atoms of LJ type s; fixed at creation (no step changes a type or
    // a charge at an index)
    int *set_atoms, *set_off;
    double *freq_beg, *freq_end;  // move intervals (BiasedRandom.f90:41-75)
    YdState *st;
    yamc_move_rec *rec;   // optional per-move records
    const int *cimol;     // constant-uniform mode: molecule per move
    const double *cu;     // constant-uniform mode: uniform per move
    int mv_begin;         // step number of the launch's first step in its run (cycle starts: multiples of nmol)
    // From-scratch S(k): the number of equal atom groups per set whose
    // partial sums are added in order (fixed per context, so every rebuild
    // path rounds alike).
    int fs_ng;
    // Cell list of the workgroup chain kernel's real-space scan (nc = 0: none).
    // nc^3 cubic cells of width 1/nc >= the cut-off (box units); cell
    // (ix, iy, iz) = floor((r + 0.5) nc) mod nc per axis, id (ix nc + iy) nc + iz;
    // bucket c holds cell_cnt[c] <= cap atoms in cell_atoms[c cap ..].
    int nc, ccap;
    int *cell_cnt, *cell_atoms, *atom_cell, *atom_slot;
};

// One workgroup of a grid chain (one chain over G co-resident workgroups):
// its index g among the G, the chain's index in a batch, the chain's barrier
// words and partial-sum buffer, and the chain's LJ tables.
struct YgGroup {
    int g, G, rep;
    unsigned int *bar;
    double *part;
    const double *sig6, *feps;
};

// Global-memory view of a device pointer.  The pointers of a chain
// descriptor are read from memory, so the compiler only knows them as
// generic pointers and emits flat loads and stores for them; a flat access
// counts against the LDS wait counter as well as the vector-memory one, so
// every wait for an LDS read would also wait for the flat loads in flight
// (a load issued ahead of its use would be waited for at the next LDS
// access).  Accesses through yd_g(p) are global_load / global_store.
// (The host pass of hipcc also parses device code; there the qualifier is
// left out, since the host target has no such address space.)
#if defined(__HIP_DEVICE_COMPILE__)
#define YD_GLOBAL __attribute__((address_space(1)))
#else
#define YD_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ YD_GLOBAL T *yd_g(T *p) {
    return (YD_GLOBAL T *)p;
}

// Cell coordinate of one position component (box units).
__device__ __forceinline__ int yd_cell1(double x, int nc) {
    int i = (int)floor((x + 0.5) * (double)nc) % nc;
    return i < 0 ? i + nc : i;
}

// ---------------------------------------------------------------- math ----
// Functions.f90:116-142 erfc_Luc
__device__ __forceinline__ double yd_erfc_luc(double x, double x2, double ee2) {
    const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027,
                 A5 = 1.061405429, pi05 = 1.7724538509055;
    if (x < 0.455) return 1.0 - 2.0 / pi05 * x * (1.0 - x2 * (1.0 / 3.0 - x2 * (0.1 - x2 * (1.0 / 42.0 - x2 / 216.0))));
    if (x > 3.0) return (1.0 - 0.5 * (1.0 - 1.5 / x2) / x2) / (pi05 * x) * ee2;
    double T = 1.0 / (1.0 + P * x);
    return T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5)))) * ee2;
}

// 1/x and 1/sqrt(x) for positive normal x: the hardware estimate refined by
// two Newton-Raphson steps, each of which squares the relative error (the
// results are within a few ulp of the correctly rounded quotient or root).
__device__ __forceinline__ double yd_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}
__device__ __forceinline__ double yd_rsqrt(double x) {
    const double h = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-h * y, y, 1.5);
    return y * fma(-h * y, y, 1.5);
}

// erfc_Luc for the real-space pair sums, given ix = 1/x: the asymptotic
// branch's quotients become products with ix and the middle branch's
// reciprocal is yd_rcp (the values differ from erfc_Luc by rounding only).
__device__ __forceinline__ double yd_erfc_luc_ix(double x, double x2, double ix, double ee2) {
    const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027,
                 A5 = 1.061405429, pi05 = 1.7724538509055;
    if (x < 0.455) return 1.0 - 2.0 / pi05 * x * (1.0 - x2 * (1.0 / 3.0 - x2 * (0.1 - x2 * (1.0 / 42.0 - x2 / 216.0))));
    if (x > 3.0) {
        const double ix2 = ix * ix;
        return (1.0 - 0.5 * (1.0 - 1.5 * ix2) * ix2) * ix * (1.0 / pi05) * ee2;
    }
    double T = yd_rcp(1.0 + P * x);
    return T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5)))) * ee2;
}

// Functions.f90:145-171 erf_Luc
__device__ __forceinline__ double yd_erf_luc(double x, double x2, double ee2) {
    const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027,
                 A5 = 1.061405429, pi05 = 1.7724538509055;
    if (x < 0.455) return 2.0 / pi05 * x * (1.0 - x2 * (1.0 / 3.0 - x2 * (0.1 - x2 * (1.0 / 42.0 - x2 / 216.0))));
    if (x > 3.0) return 1.0 - (1.0 - 0.5 * (1.0 - 1.5 / x2) / x2) / (pi05 * x) * ee2;
    double T = 1.0 / (1.0 + P * x);
    return 1.0 - T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5)))) * ee2;
}

__device__ __forceinline__ double yd_p3(double x) { return x * x * x; }
__device__ __forceinline__ double yd_p9(double x) { double x2 = x * x, x4 = x2 * x2; return x4 * x4 * x; }

// Functions.f90:225-289 shi_sans_exp_f (Shi(x) e^-x)
__device__ double yd_shi_sans_exp(double x) {
    const double dx0 = 0.2;
    double x1, shi_f;
    if (x <= 1.0) {
        double x2 = x * x, fac = 1.0, xn = x;
        shi_f = x;
        for (int i = 1; i <= 4; ++i) {
            double deuxi = 2.0 * i;
            xn = xn * x2;
            fac = fac * deuxi * (deuxi + 1.0);
            shi_f = shi_f + xn / (deuxi + 1.0) / fac;
        }
        return exp(-x) * shi_f;
    } else if (x > 20.0) {
        double som = 1.0, term = 1.0;
        for (int i = 1; i <= 6; ++i) {
            term = term * i / x;
            som = som + term;
        }
        return 1.0 / 2.0 / x * som;
    } else if (x < 5.0) {
        x1 = 1.0; shi_f = 1.057250875;
    } else if (x < 10.0) {
        x1 = 5.0; shi_f = 20.093211826;
    } else if (x < 15.0) {
        x1 = 10.0; shi_f = 1246.11449;
    } else {
        x1 = 15.0; shi_f = 117477.926;
    }
    int n = (int)((x - x1) / dx0 + 2);
    double dx = (x - x1) / (n - 1);
    shi_f = shi_f + dx * sinh(x1) / x1 / 2.0 + dx * dx / 12.0 * (cosh(x1) / x1 - sinh(x1) / (x1 * x1));
    double t = x1, s = sinh(x1);
    for (int i = 2; i <= n; ++i) {
        t = x1 + (double)(i - 1) * dx;
        s = sinh(t);
        shi_f = shi_f + dx * s / t;
    }
    shi_f = shi_f - dx * s / t / 2.0 - dx * dx / 12.0 * (cosh(t) / t - s / (t * t));
    return exp(-x) * shi_f;
}

// Lane l's value of v for a wave-uniform l (two scalar reads, no LDS round
// trip; the same bits a shuffle returns).
__device__ __forceinline__ double yd_readlane(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)b, l);
    const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// shi_sans_exp_f evaluated by a whole wave (every lane passes the same x and
// gets the same result).  The trapezoid terms dx sinh(t_i)/t_i, i = 2..n
// (n <= 27), are computed one per lane and then added in index order, which
// is the scalar loop's summation order, so the result is the scalar
// function's to the bit.
__device__ double yd_shi_sans_exp_wave(double x) {
    if (x <= 1.0 || x > 20.0) return yd_shi_sans_exp(x);
    const double dx0 = 0.2;
    double x1, shi_f;
    if (x < 5.0) {
        x1 = 1.0; shi_f = 1.057250875;
    } else if (x < 10.0) {
        x1 = 5.0; shi_f = 20.093211826;
    } else if (x < 15.0) {
        x1 = 10.0; shi_f = 1246.11449;
    } else {
        x1 = 15.0; shi_f = 117477.926;
    }
    int n = (int)((x - x1) / dx0 + 2);
    double dx = (x - x1) / (n - 1);
    shi_f = shi_f + dx * sinh(x1) / x1 / 2.0 + dx * dx / 12.0 * (cosh(x1) / x1 - sinh(x1) / (x1 * x1));
    const int lane = threadIdx.x & 63;
    double t = x1 + (double)(lane + 1) * dx, s = sinh(t), term = dx * s / t;  // i = lane + 2
    for (int l = 0; l < n - 1; ++l) shi_f = shi_f + yd_readlane(term, l);
    t = yd_readlane(t, n - 2);
    s = yd_readlane(s, n - 2);
    shi_f = shi_f - dx * s / t / 2.0 - dx * dx / 12.0 * (cosh(t) / t - s / (t * t));
    return exp(-x) * shi_f;
}

// ---------------------------------------------------------------- RNG -----
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void yd_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                          uint32_t k1, uint32_t *o) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

// Uniform source of one chain (thread 0 only).  Either the Philox stream of
// 24-bit uniforms (random.f90:65-93 draws a default real, 0 excluded) or a
// constant per move (random.f90:80-82).
struct YdRng {
    uint32_t key0, key1, s0, s1;
    unsigned long long ctr;
    int constant;
    double cu;
    const double *list;  // replay mode: recorded uniforms list[0 .. nlist)
    unsigned long long nlist;
    // the last Philox block (the four words of counter block cblk - 1;
    // cblk = 0: none): consecutive draws take the block's words in turn
    // instead of evaluating the generator once per draw
    unsigned long long cblk;
    uint32_t blk[4];
    // every constructor of a stream calls this after setting the fields above
    __device__ __forceinline__ void reset_cache() { cblk = 0; }
    __device__ double next() {
        if (constant) return cu;
        if (list) {  // past the end: 0.5, and the counter shows the overrun to the host
            const unsigned long long n = ctr++;
            return n < nlist ? list[n] : 0.5;
        }
        for (;;) {
            unsigned long long n = ctr++;
            const unsigned long long b = (n >> 2) + 1;
            if (b != cblk) {
                yd_philox((uint32_t)(n >> 2), (uint32_t)(n >> 34), s0, s1, key0, key1, blk);
                cblk = b;
            }
            uint32_t k = blk[n & 3] >> 8;
            if (k != 0) return (double)k * (1.0 / 16777216.0);
        }
    }
};

// ----------------------------------------------------------- reductions ---
// Block-wide sum of NV doubles; result valid in out[] for every thread after
// the call.  red must hold (blockDim/64) * NV doubles, out NV doubles.
template <int NV>
__device__ __forceinline__ void yd_block_sum(double (&v)[NV], double *red, double *out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double a = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        v[i] = a;
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) red[w * NV + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double a = 0.0;
        for (int j = 0; j < nw; ++j) a += red[j * NV + threadIdx.x];
        out[threadIdx.x] = a;
    }
    __syncthreads();
}

// ------------------------------------------------------ k-space helpers ---
// SumSinCosKR.f90:181-223 SinCosXYZ_fill for one coordinate: tab[0..kmax]
// cos, tab[KP..] sin of 2 pi k x, angle-addition recurrence.
__device__ __forceinline__ void yd_sincos_axis(double x, int kmax, double *c, double *s) {
    double cx = cos(YD_TWO_PI * x), sx = sin(YD_TWO_PI * x);
    c[0] = 1.0; s[0] = 0.0;
    for (int k = 1; k <= kmax; ++k) {
        c[k] = c[k - 1] * cx - s[k - 1] * sx;
        s[k] = s[k - 1] * cx + c[k - 1] * sx;
    }
}

// SumSinCosKR.f90:297-455 SinCosKR_fill evaluated at one k.  tab holds the
// six axis tables of one atom (cx, sx, cy, sy, cz, sz) with stride YD_KP.
// The xy factor is rebuilt from the tables at each k, which is the value the
// reference recurrence carries along a (kx, ky) row.
// (ks: stride of the six axis tables in tab.)
__device__ __forceinline__ void yd_kr_s(const double *tab, int ks, double q, int kx, int ky, int kz, double *o) {
    double cxk = tab[kx], sxk = tab[ks + kx], cyk = tab[2 * ks + ky], syk = tab[3 * ks + ky];
    double czk = tab[4 * ks + kz], szk = tab[5 * ks + kz];
    double qcc = q * cxk * cyk, qss = q * sxk * syk, qsc = q * sxk * cyk, qcs = q * cxk * syk;
    double qcpy = qcc - qss, qspy = qsc + qcs, qcmy = qcc + qss, qsmy = qsc - qcs;
    double a = qcpy * czk, b = qspy * szk, c = qspy * czk, d = qcpy * szk;
    o[0] = a - b;
    o[1] = c + d;
    o[2] = kz > 0 ? a + b : 0.0;
    o[3] = kz > 0 ? c - d : 0.0;
    double e = qcmy * czk, f = qsmy * szk, g = qsmy * czk, h = qcmy * szk;
    o[4] = ky > 0 ? e - f : 0.0;
    o[5] = ky > 0 ? g + h : 0.0;
    o[6] = (ky > 0 && kz > 0) ? e + f : 0.0;
    o[7] = (ky > 0 && kz > 0) ? g - h : 0.0;
}
// yd_kr_s with q folded into the x table (tab[kx] = q cos, tab[ks + kx] =
// q sin: the same products, since yd_kr_s's q cx cy is (q cx) cy) and the
// slots of ky = 0 or kz = 0 left unzeroed (the caller zeroes them once).
__device__ __forceinline__ void yd_kr_pre(const double *tab, int ks, int kx, int ky, int kz, double *o) {
    double qcx = tab[kx], qsx = tab[ks + kx], cyk = tab[2 * ks + ky], syk = tab[3 * ks + ky];
    double czk = tab[4 * ks + kz], szk = tab[5 * ks + kz];
    double qcc = qcx * cyk, qss = qsx * syk, qsc = qsx * cyk, qcs = qcx * syk;
    double qcpy = qcc - qss, qspy = qsc + qcs, qcmy = qcc + qss, qsmy = qsc - qcs;
    double a = qcpy * czk, b = qspy * szk, c = qspy * czk, d = qcpy * szk;
    o[0] = a - b;
    o[1] = c + d;
    o[2] = a + b;
    o[3] = c - d;
    double e = qcmy * czk, f = qsmy * szk, g = qsmy * czk, h = qcmy * szk;
    o[4] = e - f;
    o[5] = g + h;
    o[6] = e + f;
    o[7] = g - h;
}
// yd_kr_s without the zeroing of the slots of ky = 0 or kz = 0: for callers
// that read those slots only under the same ky > 0 / kz > 0 conditions
// (the sweeps' dU and force terms), which then see the same values.
__device__ __forceinline__ void yd_kr_nz(const double *tab, double q, int kx, int ky, int kz, double *o) {
    const int ks = YD_KP;
    double cxk = tab[kx], sxk = tab[ks + kx], cyk = tab[2 * ks + ky], syk = tab[3 * ks + ky];
    double czk = tab[4 * ks + kz], szk = tab[5 * ks + kz];
    double qcc = q * cxk * cyk, qss = q * sxk * syk, qsc = q * sxk * cyk, qcs = q * cxk * syk;
    double qcpy = qcc - qss, qspy = qsc + qcs, qcmy = qcc + qss, qsmy = qsc - qcs;
    double a = qcpy * czk, b = qspy * szk, c = qspy * czk, d = qcpy * szk;
    o[0] = a - b;
    o[1] = c + d;
    o[2] = a + b;
    o[3] = c - d;
    double e = qcmy * czk, f = qsmy * szk, g = qsmy * czk, h = qcmy * szk;
    o[4] = e - f;
    o[5] = g + h;
    o[6] = e + f;
    o[7] = g - h;
}
__device__ __forceinline__ void yd_kr(const double *tab, double q, int kx, int ky, int kz, double *o) {
    yd_kr_s(tab, YD_KP, q, kx, ky, kz, o);
}

// Coulomb and LJ parts of one atom's sin/cos at k (RhoSquared.f90:160-183):
// the LJ part of a charged LJ atom is the Coulomb part times 1/q.
__device__ __forceinline__ void yd_kr_atom_s(const double *tab, int ks, double q, int typ, int kx, int ky, int kz,
                                             double *krc, double *krl) {
    if (q != 0.0) {
        yd_kr_s(tab, ks, q, kx, ky, kz, krc);
        if (typ > 0) {
            double inv = 1.0 / q;
#pragma unroll
            for (int i = 0; i < 8; ++i) krl[i] = krc[i] * inv;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) krc[i] = 0.0;
        if (typ > 0) yd_kr_s(tab, ks, 1.0, kx, ky, kz, krl);
    }
}
__device__ __forceinline__ void yd_kr_atom(const double *tab, double q, int typ, int kx, int ky, int kz,
                                           double *krc, double *krl) {
    yd_kr_atom_s(tab, YD_KP, q, typ, kx, ky, kz, krc, krl);
}

// One k of ForceKSpace_coulomb/_LJ (ForceKSpace.f90:116-158): C, S of the
// four sign slots against the atom's q sin/cos, k-vector component signs
// folded per slot, times kx, ky, kz and beta.
__device__ __forceinline__ void yd_force_k(const double *C, const double *kr, int kx, int ky, int kz, double bt,
                                           double *f) {
    double A = C[0] * kr[1] - C[1] * kr[0];
    double fx = A, fy = A, fz = A;
    if (ky > 0) {
        A = C[4] * kr[5] - C[5] * kr[4];
        fx = fx + A; fy = fy - A; fz = fz + A;
    }
    if (kz > 0) {
        A = C[2] * kr[3] - C[3] * kr[2];
        fx = fx + A; fy = fy + A; fz = fz - A;
    }
    if (ky > 0 && kz > 0) {
        A = C[6] * kr[7] - C[7] * kr[6];
        fx = fx + A; fy = fy - A; fz = fz - A;
    }
    f[0] += fx * (double)kx * bt;
    f[1] += fy * (double)ky * bt;
    f[2] += fz * (double)kz * bt;
}

// One k of calc_dU_coulomb (EwaldSumKSpace.f90:581-614) without beta.
__device__ __forceinline__ double yd_du_k(const double *C, const double *D, int ky, int kz) {
    double loc = (C[0] + C[0] + D[0]) * D[0] + (C[1] + C[1] + D[1]) * D[1];
    if (ky > 0) loc = loc + (C[4] + C[4] + D[4]) * D[4] + (C[5] + C[5] + D[5]) * D[5];
    if (kz > 0) loc = loc + (C[2] + C[2] + D[2]) * D[2] + (C[3] + C[3] + D[3]) * D[3];
    if (ky > 0 && kz > 0) loc = loc + (C[6] + C[6] + D[6]) * D[6] + (C[7] + C[7] + D[7]) * D[7];
    return loc;
}

// One k of calc_dU_LJ (EwaldSumKSpace.f90:725-741): tw = 2 C^t1 + dC^t1.
__device__ __forceinline__ double yd_dulj_k(const double *tw, const double *D2, int ky, int kz) {
    double loc = tw[0] * D2[0] + tw[1] * D2[1];
    if (ky > 0) loc = loc + tw[4] * D2[4] + tw[5] * D2[5];
    if (kz > 0) loc = loc + tw[2] * D2[2] + tw[3] * D2[3];
    if (ky > 0 && kz > 0) loc = loc + tw[6] * D2[6] + tw[7] * D2[7];
    return loc;
}

// --------------------------------------------------- real-space helpers ---
// RealSumLocal.f90:63-168 for one pair.
struct YdPair {
    double potew, potlj6, potlj12, pot, FR;
};

__device__ __forceinline__ void yd_pair(const YdParams &p, double r2, bool same_mol, double qi, double qj, bool lj,
                                        double sigma6, double four_eps, YdPair &o) {
    // 1/r and 1/r^2 once, from a refined reciprocal square root; the
    // divisions of the reference expressions become multiplications (the
    // values differ from the reference's by rounding only).
    double ir = yd_rsqrt(r2), r = r2 * ir, ir2 = ir * ir, a = p.alpha * r, a2 = a * a, e = exp(-a2);
    double erfc_a = same_mol ? -yd_erf_luc(a, a2, e) : yd_erfc_luc_ix(a, a2, ir * p.inv_alpha, e);
    double qq = p.xclb * qi * qj;
    o.potew = qq * erfc_a * ir;
    o.pot = o.potew;
    o.FR = qq * (erfc_a * ir + p.two_alpha_over_sqrt_pi * e) * ir2;
    o.potlj6 = 0.0;
    o.potlj12 = 0.0;
    if (lj) {
        double sor6 = sigma6 * (ir2 * ir2 * ir2);
        double C6 = 1.0 + a2 * (1.0 + a2 * 0.5);
        double a6 = a2 * a2 * a2;
        double C12 = C6 + a6 * (1.0 + a2 * (1.0 + a2 * 0.2) * 0.25) * (1.0 / 6.0);
        o.potlj6 = four_eps * sor6 * e * C6;
        o.potlj12 = four_eps * (sor6 * sor6) * e * C12;
        o.pot = o.pot + o.potlj12 - o.potlj6;
        o.FR = o.FR + four_eps * 6.0 * sor6 * e *
                          (2.0 * sor6 * (C12 + (a6 * a6) * (1.0 / 720.0)) - (C6 + a6 * (1.0 / 6.0))) * ir2;
    }
}

// Single-wrap minimum image (EwaldSumRealSpace.f90:394-401).
__device__ __forceinline__ double yd_wrap(double d) {
    if (d > 0.5) d = d - 1.0;
    if (d < -0.5) d = d + 1.0;
    return d;
}
