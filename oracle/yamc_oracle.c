/*
 * yamc_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
 *
 * Plain-C FP64 restatement of YAMC's Ewald Monte Carlo energy/force path.
 * Loop order, summation order and expression association follow the
 * reference Fortran so results agree to rounding; each routine cites the
 * reference file:line it restates.  Build with -ffp-contract=off so no FMA
 * is introduced where the Fortran evaluates separate multiply/add steps.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load this
 * library.  The product (libyamc_ewald.so) never links or calls it.
 */
#include "yamc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define YO_PI 3.14159265358979323846   /* constants.f90:5 */
#define YO_TWO_PI (2.0 * YO_PI)        /* constants.f90:6 */
#define YO_FOUR_PI (4.0 * YO_PI)       /* constants.f90:7 */

typedef struct {
    int kmax, nk;
    int32_t *kx, *ky, *kz, *ip, *iq;
    double *xk, *xk2;
} yo_grid;

struct yo_sys {
    int natom, nmol, ntype, kmax, nk;
    double L, alpha, rmax2, xclb, kext, minus_two_kext, dr, d_angle, lambda_f, lambda_c;
    double alpha_over_sqrt_pi, two_alpha_over_sqrt_pi;
    double *xx, *yy, *zz, *charge, *hard_core, *mass;
    int32_t *molnum, *type, *first, *last, *tail_counts;
    double *lj6, *lj12, *four_eps, *sigma6;
    yo_grid g;
    double *beta, *beta6, *beta12, *beta_lj; /* beta_lj: ntype*ntype*nk */
    double *S;                               /* (1+ntype)*8*nk */
    double *uu, *Fx, *Fy, *Fz;
    double uu_ew, uu_lj6, uu_lj12, uu_ew_intra, uu_lj6_intra, uu_lj12_intra;
    double e_kc, e_klj, mu[3], tail_c, tail6, tail12;
    /* BiasedRandom.f90:41-75 intervals of the move and exchange pickers
     * (runmc.f90:26-66), and xchange_probability */
    double *freq_beg, *freq_end, *xfreq_beg, *xfreq_end, pxch;
    /* rng */
    int rng_mode;
    uint32_t key[2], stream[2];
    uint64_t ctr;
    double const_u;
    double *ulist;   /* replay mode: the uniforms of a recorded stream */
    uint64_t nlist;
    /* scratch */
    double *mol_old, *mol_new, *tot_new, *krc, *krl;
    double *back_new, *back_old; /* 4*natom each */
    int max_mol_atom;
};

/* ------------------------------------------------------------------------ */
/* FourierGrid.f90:37-77 (FourierGrid_calcNalloc) and :118-189 (init)        */
static int grid_walk(int kmax, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq,
                     double *xk, double *xk2) {
    double xkmax2 = (double)kmax * (double)kmax;
    int kk = 0;
    for (int ix = 0; ix <= kmax; ++ix) {
        double xkx2 = (double)(ix * ix);
        int kymax = (int)sqrt(xkmax2 - xkx2);
        for (int iy = 0; iy <= kymax; ++iy) {
            double xky2 = (double)(iy * iy);
            int kzmax = (int)sqrt(xkmax2 - xkx2 - xky2);
            for (int iz = 0; iz <= kzmax; ++iz) {
                double k2 = xkx2 + xky2 + (double)(iz * iz);
                if (k2 < 0.1) continue;
                if (kx) {
                    kx[kk] = ix; ky[kk] = iy; kz[kk] = iz;
                    xk2[kk] = k2; xk[kk] = sqrt(k2);
                    if (iy == 0) iq[kk] = (iz == 0) ? 1 : 3;
                    else iq[kk] = (iz == 0) ? 2 : 4;
                    ip[kk] = 0;
                }
                ++kk;
            }
            if (kx && kk > 0) ip[kk - 1] = 1;
        }
        if (kx && kk > 0) ip[kk - 1] = 2;
    }
    if (kx && kk > 0) ip[kk - 1] = 3;
    return kk;
}

int yo_grid_nk(int kmax) { return grid_walk(kmax, 0, 0, 0, 0, 0, 0, 0); }

int yo_grid_fill(int kmax, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq) {
    int nk = yo_grid_nk(kmax);
    double *xk = malloc(sizeof(double) * (nk + 1)), *xk2 = malloc(sizeof(double) * (nk + 1));
    grid_walk(kmax, kx, ky, kz, ip, iq, xk, xk2);
    free(xk); free(xk2);
    return nk;
}

/* ------------------------------------------------------------------------ */
/* Functions.f90:116-142 erfc_Luc                                           */
double yo_erfc_luc(double x, double x2, double ee2) {
    const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741,
                 A4 = -1.453152027, A5 = 1.061405429, pi05 = 1.7724538509055;
    if (x < 0.455)
        return 1.0 - 2.0 / pi05 * x * (1.0 - x2 * (1.0 / 3.0 - x2 * (0.1 - x2 * (1.0 / 42.0 - x2 / 216.0))));
    if (x > 3.0) {
        double ax = (1.0 - 0.5 * (1.0 - 1.5 / x2) / x2) / (pi05 * x);
        return ax * ee2;
    }
    double T = 1.0 / (1.0 + P * x);
    double ax = T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5))));
    return ax * ee2;
}

/* Functions.f90:145-171 erf_Luc */
double yo_erf_luc(double x, double x2, double ee2) {
    const double P = 0.3275911, A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741,
                 A4 = -1.453152027, A5 = 1.061405429, pi05 = 1.7724538509055;
    if (x < 0.455)
        return 2.0 / pi05 * x * (1.0 - x2 * (1.0 / 3.0 - x2 * (0.1 - x2 * (1.0 / 42.0 - x2 / 216.0))));
    if (x > 3.0) {
        double ax = (1.0 - 0.5 * (1.0 - 1.5 / x2) / x2) / (pi05 * x);
        return 1.0 - ax * ee2;
    }
    double T = 1.0 / (1.0 + P * x);
    double ax = T * (A1 + T * (A2 + T * (A3 + T * (A4 + T * A5))));
    return 1.0 - ax * ee2;
}

static double p3(double x) { return x * x * x; }
static double p4(double x) { double x2 = x * x; return x2 * x2; }
static double p5(double x) { double x2 = x * x; return x2 * x2 * x; }
static double p9(double x) { double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4; return x8 * x; }
static double p10(double x) { double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4; return x8 * x2; }
static double p11(double x) { double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4; return x8 * x2 * x; }

/* Functions.f90:26-111 erfc_Luc_bet6_bet12 */
void yo_erfc_bet6_bet12(double x, double x2, double e2, double *erfk, double *bet6, double *bet12) {
    double pi = 4.0 * atan(1.0);
    double pi12 = sqrt(pi);
    double pi15 = pow(pi, 1.5);
    const double xmin = 0.5, xmax = 8.0;
    if (x <= xmin) {
        *erfk = yo_erfc_luc(x, x2, e2);
        *bet6 = pi15 / 3.0 * p3(x) * (pi12 * *erfk + (0.5 / x2 - 1.0) / x * e2);
        *bet12 = pi15 / 1080.0 * p9(x) *
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
        double som = dx * e2 / 2.0 - dx * dx / 12.0 * 2.0 * x * e2;
        double som6 = dx * e2 / p4(x) / 2.0 - dx * dx / 12.0 * (2.0 / p3(x) + 4.0 / p5(x)) * e2;
        double som12 = dx * e2 / p10(x) / 2.0 - dx * dx / 12.0 * (2.0 / p9(x) + 10.0 / p11(x)) * e2;
        double y = x, e = e2;
        for (int i = 2; i <= n; ++i) {
            y = x + (double)(i - 1) * dx;
            e = exp(-(y * y));
            som = som + dx * e;
            som6 = som6 + dx * e / p4(y);
            som12 = som12 + dx * e / p10(y);
        }
        som = som - dx * e / 2.0 + dx * dx / 12.0 * 2.0 * y * e;
        double y2 = y * y;
        som = som + e / 2.0 / y * (1.0 - 0.5 / y2 * (1.0 - 1.5 / y2 * (1.0 - 2.5 / y2)));
        *erfk = 2.0 / pi12 * som;
        som6 = som6 - dx * e / p4(y) / 2.0 + dx * dx / 12.0 * (2.0 / p3(y) + 4.0 / p5(y)) * e;
        som6 = som6 + e / 2.0 / p5(y) * (1.0 - 2.5 / y2 * (1.0 - 3.5 / y2 * (1.0 - 4.5 / y2)));
        *bet6 = pi15 * p3(x) / 2.0 * som6;
        som12 = som12 - dx * e / p10(y) / 2.0 + dx * dx / 12.0 * (2.0 / p9(y) + 10.0 / p11(y)) * e;
        som12 = som12 + e / 2.0 / p11(y) * (1.0 - 5.5 / y2 * (1.0 - 6.5 / y2 * (1.0 - 7.5 / y2)));
        *bet12 = pi15 * p9(x) / 120.0 * som12;
    }
}

/* Functions.f90:225-289 shi_sans_exp_f */
double yo_shi_sans_exp(double x) {
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

/* ------------------------------------------------------------------------ */
/* SumSinCosKR.f90:181-223 SinCosXYZ_fill: tab = cx,sx,cy,sy,cz,sz [kmax+1] */
static void sincos_xyz(double x, double y, double z, int kmax, double *tab) {
    double *cxk = tab, *sxk = tab + (kmax + 1), *cyk = tab + 2 * (kmax + 1), *syk = tab + 3 * (kmax + 1),
           *czk = tab + 4 * (kmax + 1), *szk = tab + 5 * (kmax + 1);
    double cx = cos(YO_TWO_PI * x), sx = sin(YO_TWO_PI * x);
    double cy = cos(YO_TWO_PI * y), sy = sin(YO_TWO_PI * y);
    double cz = cos(YO_TWO_PI * z), sz = sin(YO_TWO_PI * z);
    cxk[0] = 1.0; sxk[0] = 0.0; cyk[0] = 1.0; syk[0] = 0.0; czk[0] = 1.0; szk[0] = 0.0;
    for (int k = 1; k <= kmax; ++k) {
        cxk[k] = cxk[k - 1] * cx - sxk[k - 1] * sx;
        sxk[k] = sxk[k - 1] * cx + cxk[k - 1] * sx;
        cyk[k] = cyk[k - 1] * cy - syk[k - 1] * sy;
        syk[k] = syk[k - 1] * cy + cyk[k - 1] * sy;
        czk[k] = czk[k - 1] * cz - szk[k - 1] * sz;
        szk[k] = szk[k - 1] * cz + czk[k - 1] * sz;
    }
}

/* SumSinCosKR.f90:297-455 SinCosKR_fill; out = 8 arrays of nk */
static void sincos_kr_fill(const yo_grid *g, const double *tab, double charge, double *out) {
    int kmax = g->kmax, nk = g->nk;
    const double *cxk = tab, *sxk = tab + (kmax + 1), *cyk = tab + 2 * (kmax + 1), *syk = tab + 3 * (kmax + 1),
                 *czk = tab + 4 * (kmax + 1), *szk = tab + 5 * (kmax + 1);
    double *cos_ppp = out, *sin_ppp = out + nk, *cos_ppm = out + 2 * nk, *sin_ppm = out + 3 * nk,
           *cos_pmp = out + 4 * nk, *sin_pmp = out + 5 * nk, *cos_pmm = out + 6 * nk, *sin_pmm = out + 7 * nk;
    double qcpy = charge, qspy = 0.0, qcmy = charge, qsmy = 0.0;
    for (int kk = 0; kk < nk; ++kk) {
        int kz = g->kz[kk], ky = g->ky[kk];
        double q_cosxy_cosz = qcpy * czk[kz], q_sinxy_sinz = qspy * szk[kz];
        double q_sinxy_cosz = qspy * czk[kz], q_cosxy_sinz = qcpy * szk[kz];
        cos_ppp[kk] = q_cosxy_cosz - q_sinxy_sinz;
        sin_ppp[kk] = q_sinxy_cosz + q_cosxy_sinz;
        if (kz > 0) {
            cos_ppm[kk] = q_cosxy_cosz + q_sinxy_sinz;
            sin_ppm[kk] = q_sinxy_cosz - q_cosxy_sinz;
        } else {
            cos_ppm[kk] = 0.0; sin_ppm[kk] = 0.0;
        }
        double q_cosxmy_cosz = qcmy * czk[kz], q_sinxmy_sinz = qsmy * szk[kz];
        double q_sinxmy_cosz = qsmy * czk[kz], q_cosxmy_sinz = qcmy * szk[kz];
        if (ky > 0) {
            cos_pmp[kk] = q_cosxmy_cosz - q_sinxmy_sinz;
            sin_pmp[kk] = q_sinxmy_cosz + q_cosxmy_sinz;
        } else {
            cos_pmp[kk] = 0.0; sin_pmp[kk] = 0.0;
        }
        if (kz > 0 && ky > 0) {
            cos_pmm[kk] = q_cosxmy_cosz + q_sinxmy_sinz;
            sin_pmm[kk] = q_sinxmy_cosz - q_cosxmy_sinz;
        } else {
            cos_pmm[kk] = 0.0; sin_pmm[kk] = 0.0;
        }
        if (g->ip[kk] != 0 && kk < nk - 1) {
            int nx = g->kx[kk + 1], ny = g->ky[kk + 1];
            double q_cosx_cosy = charge * cxk[nx] * cyk[ny];
            double q_sinx_siny = charge * sxk[nx] * syk[ny];
            double q_sinx_cosy = charge * sxk[nx] * cyk[ny];
            double q_cosx_siny = charge * cxk[nx] * syk[ny];
            qcpy = q_cosx_cosy - q_sinx_siny;
            qspy = q_sinx_cosy + q_cosx_siny;
            qcmy = q_cosx_cosy + q_sinx_siny;
            qsmy = q_sinx_cosy - q_cosx_siny;
        }
    }
}

void yo_sincos_kr(const yo_sys *s, double x, double y, double z, double charge, double *out) {
    double tab[6 * 64];
    sincos_xyz(x, y, z, s->kmax, tab);
    sincos_kr_fill(&s->g, tab, charge, out);
}

/* RhoSquared.f90:147-185 RhoSquared_addAtom; sums = (1+T)*8*nk, present[T] */
static void add_atom(const yo_sys *s, double *sums, int *present, double x, double y, double z, double charge,
                     int lj_type, double *krc, double *krl) {
    int nk = s->nk, n8 = 8 * nk;
    double tab[6 * 64];
    if (charge != 0.0) {
        sincos_xyz(x, y, z, s->kmax, tab);
        sincos_kr_fill(&s->g, tab, charge, krc);
        for (int i = 0; i < n8; ++i) sums[i] = sums[i] + krc[i];
    } else {
        memset(krc, 0, sizeof(double) * n8);
    }
    if (lj_type > 0) {
        if (charge != 0.0) {
            double inv = 1.0 / charge;
            for (int i = 0; i < n8; ++i) krl[i] = krc[i] * inv;
        } else {
            sincos_xyz(x, y, z, s->kmax, tab);
            sincos_kr_fill(&s->g, tab, 1.0, krl);
        }
        double *dst = sums + (size_t)lj_type * n8;
        for (int i = 0; i < n8; ++i) dst[i] = dst[i] + krl[i];
        if (present) present[lj_type - 1] = 1;
    }
}

/* ------------------------------------------------------------------------ */
/* EwaldSumKSpace.f90:263-350 initBeta, :352-386 initBetaLJ                   */
static void init_beta(yo_sys *s) {
    double poa = YO_PI / s->alpha, poa2 = poa * poa;
    double a3 = p3(s->alpha), a9 = p9(s->alpha);
    for (int kk = 0; kk < s->nk; ++kk) {
        double b2 = poa2 * s->g.xk2[kk];
        double e = exp(-b2);
        s->beta[kk] = s->xclb * e / s->g.xk2[kk] / YO_TWO_PI;
        double b = poa * s->g.xk[kk], erfc_b, B6, B12;
        yo_erfc_bet6_bet12(b, b2, e, &erfc_b, &B6, &B12);
        s->beta6[kk] = a3 * B6;
        s->beta12[kk] = a9 * B12;
        if (s->g.kx[kk] > 0) {
            s->beta[kk] = 2.0 * s->beta[kk];
            s->beta6[kk] = 2.0 * s->beta6[kk];
            s->beta12[kk] = 2.0 * s->beta12[kk];
        }
    }
}

static void init_beta_lj(yo_sys *s) {
    int T = s->ntype, nk = s->nk;
    for (int t1 = 0; t1 < T; ++t1)
        for (int t2 = t1; t2 < T; ++t2) {
            double A6 = s->lj6[t1 * T + t2], A12 = s->lj12[t1 * T + t2];
            double *d = s->beta_lj + (size_t)(t1 * T + t2) * nk;
            double *d2 = s->beta_lj + (size_t)(t2 * T + t1) * nk; /* aliased row (:128-131) */
            for (int kk = 0; kk < nk; ++kk) {
                d[kk] = A12 * s->beta12[kk] - A6 * s->beta6[kk];
                d2[kk] = d[kk];
            }
        }
}

/* EwaldSumKSpace.f90:389-416 calc_coulomb_energy (with rho^2 of RhoSquared.f90:245-294) */
static double coulomb_energy(const yo_sys *s) {
    int nk = s->nk;
    const double *C = s->S;
    double U = 0.0;
    for (int kk = 0; kk < nk; ++kk) {
        double r2 = C[kk] * C[kk] + C[nk + kk] * C[nk + kk];
        int sy = s->g.ky[kk] > 0, sz = s->g.kz[kk] > 0;
        if (sy) r2 = r2 + C[4 * nk + kk] * C[4 * nk + kk] + C[5 * nk + kk] * C[5 * nk + kk];
        if (sz) r2 = r2 + C[2 * nk + kk] * C[2 * nk + kk] + C[3 * nk + kk] * C[3 * nk + kk];
        if (sy && sz) r2 = r2 + C[6 * nk + kk] * C[6 * nk + kk] + C[7 * nk + kk] * C[7 * nk + kk];
        U = U + s->beta[kk] * r2;
    }
    return U;
}

/* EwaldSumKSpace.f90:418-526 calc_LJ_energy */
static double lj_energy(const yo_sys *s) {
    int nk = s->nk, T = s->ntype, n8 = 8 * nk;
    double U = 0.0;
    for (int t1 = 0; t1 < T; ++t1) {
        const double *A = s->S + (size_t)(1 + t1) * n8;
        for (int t2 = t1; t2 < T; ++t2) {
            const double *B = s->S + (size_t)(1 + t2) * n8;
            const double *bl = s->beta_lj + (size_t)(t1 * T + t2) * nk;
            double sk = 0.0;
            for (int kk = 0; kk < nk; ++kk) {
                double loc = A[kk] * B[kk] + A[nk + kk] * B[nk + kk];
                if (s->g.ky[kk] > 0) loc = loc + A[4 * nk + kk] * B[4 * nk + kk] + A[5 * nk + kk] * B[5 * nk + kk];
                if (s->g.kz[kk] > 0) loc = loc + A[2 * nk + kk] * B[2 * nk + kk] + A[3 * nk + kk] * B[3 * nk + kk];
                if (s->g.ky[kk] > 0 && s->g.kz[kk] > 0)
                    loc = loc + A[6 * nk + kk] * B[6 * nk + kk] + A[7 * nk + kk] * B[7 * nk + kk];
                sk = sk + bl[kk] * loc;
            }
            if (t1 != t2) sk = sk + sk;
            U = U + sk;
        }
    }
    return U;
}

/* EwaldSumKSpace.f90:225-260 calc_total_energy */
static void calc_total_energy(yo_sys *s) {
    size_t n = (size_t)(1 + s->ntype) * 8 * s->nk;
    memset(s->S, 0, sizeof(double) * n);
    for (int i = 0; i < s->natom; ++i)
        add_atom(s, s->S, NULL, s->xx[i], s->yy[i], s->zz[i], s->charge[i], s->type[i], s->krc, s->krl);
    s->e_kc = coulomb_energy(s);
    s->e_klj = lj_energy(s);
}

/* EwaldSumKSpace.f90:197-223 EwaldSumKSpace_init */
void yo_kspace_init(yo_sys *s, double scale_beta) {
    if (scale_beta > 1e-9) {
        for (int kk = 0; kk < s->nk; ++kk) s->beta[kk] = s->beta[kk] * scale_beta;
    } else {
        init_beta(s);
        init_beta_lj(s);
    }
    calc_total_energy(s);
}

void yo_get_beta(const yo_sys *s, double *beta, double *beta6, double *beta12, double *beta_lj) {
    size_t nk = s->nk;
    if (beta) memcpy(beta, s->beta, sizeof(double) * nk);
    if (beta6) memcpy(beta6, s->beta6, sizeof(double) * nk);
    if (beta12) memcpy(beta12, s->beta12, sizeof(double) * nk);
    if (beta_lj) memcpy(beta_lj, s->beta_lj, sizeof(double) * nk * s->ntype * s->ntype);
}

void yo_get_sums(const yo_sys *s, double *sums) {
    memcpy(sums, s->S, sizeof(double) * (size_t)(1 + s->ntype) * 8 * s->nk);
}

void yo_kspace_energy(const yo_sys *s, double *e_c, double *e_lj) {
    *e_c = s->e_kc;
    *e_lj = s->e_klj;
}

/* EwaldSumKSpace.f90:528-620 calc_dU_coulomb; C = total sums, D = delta */
static double du_coulomb(const yo_sys *s, const double *C, const double *D) {
    int nk = s->nk;
    double total = 0.0;
    for (int kk = 0; kk < nk; ++kk) {
        double tC = C[kk] + C[kk], dC = D[kk], tS = C[nk + kk] + C[nk + kk], dS = D[nk + kk];
        double loc = (tC + dC) * dC + (tS + dS) * dS;
        if (s->g.ky[kk] > 0) {
            tC = C[4 * nk + kk] + C[4 * nk + kk]; dC = D[4 * nk + kk];
            tS = C[5 * nk + kk] + C[5 * nk + kk]; dS = D[5 * nk + kk];
            loc = loc + (tC + dC) * dC + (tS + dS) * dS;
        }
        if (s->g.kz[kk] > 0) {
            tC = C[2 * nk + kk] + C[2 * nk + kk]; dC = D[2 * nk + kk];
            tS = C[3 * nk + kk] + C[3 * nk + kk]; dS = D[3 * nk + kk];
            loc = loc + (tC + dC) * dC + (tS + dS) * dS;
        }
        if (s->g.ky[kk] > 0 && s->g.kz[kk] > 0) {
            tC = C[6 * nk + kk] + C[6 * nk + kk]; dC = D[6 * nk + kk];
            tS = C[7 * nk + kk] + C[7 * nk + kk]; dS = D[7 * nk + kk];
            loc = loc + (tC + dC) * dC + (tS + dS) * dS;
        }
        total = total + loc * s->beta[kk];
    }
    return total;
}

/* EwaldSumKSpace.f90:623-753 calc_dU_LJ; C/D point at set 0 (coulomb) */
static double du_lj_calc(const yo_sys *s, const double *C, const double *D, const int *present) {
    int nk = s->nk, T = s->ntype, n8 = 8 * nk;
    double *tw = malloc(sizeof(double) * n8);
    double total = 0.0;
    for (int t1 = 0; t1 < T; ++t1) {
        const double *C1 = C + (size_t)(1 + t1) * n8, *D1 = D + (size_t)(1 + t1) * n8;
        for (int i = 0; i < n8; ++i) tw[i] = C1[i] + C1[i] + D1[i];
        for (int t2 = 0; t2 < T; ++t2) {
            if (!present[t2]) continue;
            const double *D2 = D + (size_t)(1 + t2) * n8;
            const double *bl = s->beta_lj + (size_t)(t1 * T + t2) * nk;
            for (int kk = 0; kk < nk; ++kk) {
                double loc = tw[kk] * D2[kk] + tw[nk + kk] * D2[nk + kk];
                if (s->g.ky[kk] > 0) loc = loc + tw[4 * nk + kk] * D2[4 * nk + kk] + tw[5 * nk + kk] * D2[5 * nk + kk];
                if (s->g.kz[kk] > 0) loc = loc + tw[2 * nk + kk] * D2[2 * nk + kk] + tw[3 * nk + kk] * D2[3 * nk + kk];
                if (s->g.ky[kk] > 0 && s->g.kz[kk] > 0)
                    loc = loc + tw[6 * nk + kk] * D2[6 * nk + kk] + tw[7 * nk + kk] * D2[7 * nk + kk];
                total = total + loc * bl[kk];
            }
        }
    }
    free(tw);
    return total;
}

/* ForceKSpace.f90:59-167 (coulomb) and :171-302 (LJ), shared slot kernel */
static void force_sweep(const yo_sys *s, const double *bt, const double *C, const double *kr, double *f) {
    int nk = s->nk;
    double fxt = 0.0, fyt = 0.0, fzt = 0.0;
    for (int kk = 0; kk < nk; ++kk) {
        double A = C[kk] * kr[nk + kk] - C[nk + kk] * kr[kk];
        double fx = A, fy = A, fz = A;
        if (s->g.ky[kk] > 0) {
            A = C[4 * nk + kk] * kr[5 * nk + kk] - C[5 * nk + kk] * kr[4 * nk + kk];
            fx = fx + A; fy = fy - A; fz = fz + A;
        }
        if (s->g.kz[kk] > 0) {
            A = C[2 * nk + kk] * kr[3 * nk + kk] - C[3 * nk + kk] * kr[2 * nk + kk];
            fx = fx + A; fy = fy + A; fz = fz - A;
        }
        if (s->g.ky[kk] > 0 && s->g.kz[kk] > 0) {
            A = C[6 * nk + kk] * kr[7 * nk + kk] - C[7 * nk + kk] * kr[6 * nk + kk];
            fx = fx + A; fy = fy - A; fz = fz - A;
        }
        fx = fx * (double)s->g.kx[kk] * bt[kk];
        fy = fy * (double)s->g.ky[kk] * bt[kk];
        fz = fz * (double)s->g.kz[kk] * bt[kk];
        fxt = fxt + fx; fyt = fyt + fy; fzt = fzt + fz;
    }
    f[0] += fxt; f[1] += fyt; f[2] += fzt;
}

static void force_coulomb(const yo_sys *s, const double *S, const double *krc, double *f) {
    double t[3] = {0, 0, 0};
    force_sweep(s, s->beta, S, krc, t);
    f[0] = YO_FOUR_PI * t[0]; f[1] = YO_FOUR_PI * t[1]; f[2] = YO_FOUR_PI * t[2];
}

static void force_lj(const yo_sys *s, const double *S, int ityp, const double *krl, double *f) {
    int T = s->ntype, nk = s->nk, n8 = 8 * nk;
    double t[3] = {0, 0, 0};
    /* the running totals carry over the type loop (ForceKSpace.f90:221-295) */
    for (int typ = 0; typ < T; ++typ) {
        const double *bl = s->beta_lj + (size_t)((ityp - 1) * T + typ) * nk;
        const double *C = S + (size_t)(1 + typ) * n8;
        for (int kk = 0; kk < nk; ++kk) {
            double A = C[kk] * krl[nk + kk] - C[nk + kk] * krl[kk];
            double fx = A, fy = A, fz = A;
            if (s->g.ky[kk] > 0) {
                A = C[4 * nk + kk] * krl[5 * nk + kk] - C[5 * nk + kk] * krl[4 * nk + kk];
                fx = fx + A; fy = fy - A; fz = fz + A;
            }
            if (s->g.kz[kk] > 0) {
                A = C[2 * nk + kk] * krl[3 * nk + kk] - C[3 * nk + kk] * krl[2 * nk + kk];
                fx = fx + A; fy = fy + A; fz = fz - A;
            }
            if (s->g.ky[kk] > 0 && s->g.kz[kk] > 0) {
                A = C[6 * nk + kk] * krl[7 * nk + kk] - C[7 * nk + kk] * krl[6 * nk + kk];
                fx = fx + A; fy = fy - A; fz = fz - A;
            }
            fx = fx * (double)s->g.kx[kk] * bl[kk];
            fy = fy * (double)s->g.ky[kk] * bl[kk];
            fz = fz * (double)s->g.kz[kk] * bl[kk];
            t[0] = t[0] + fx; t[1] = t[1] + fy; t[2] = t[2] + fz;
        }
    }
    f[0] = YO_FOUR_PI * t[0]; f[1] = YO_FOUR_PI * t[1]; f[2] = YO_FOUR_PI * t[2];
}

void yo_kspace_forces(const yo_sys *s, int first, int n, const double *x, const double *y, const double *z,
                      double *fx, double *fy, double *fz) {
    size_t n8 = 8 * (size_t)s->nk;
    double *krc = malloc(sizeof(double) * n8), *krl = malloc(sizeof(double) * n8);
    double *tmp = calloc((size_t)(1 + s->ntype) * n8, sizeof(double));
    for (int i = 0; i < n; ++i) {
        int a = first + i;
        double fc[3], fl[3] = {0, 0, 0};
        add_atom(s, tmp, NULL, x[i], y[i], z[i], s->charge[a], s->type[a], krc, krl);
        force_coulomb(s, s->S, krc, fc);
        if (s->type[a] > 0) force_lj(s, s->S, s->type[a], krl, fl);
        fx[i] = fc[0] + fl[0];
        fy[i] = fc[1] + fl[1];
        fz[i] = fc[2] + fl[2];
    }
    free(krc); free(krl); free(tmp);
}

static void mol_sums(const yo_sys *s, int imol, const double *x, const double *y, const double *z, double *sums,
                     int *present, double *krc, double *krl) {
    size_t n = (size_t)(1 + s->ntype) * 8 * s->nk;
    memset(sums, 0, sizeof(double) * n);
    if (present) memset(present, 0, sizeof(int) * s->ntype);
    int f = s->first[imol], na = s->last[imol] - f + 1;
    size_t n8 = 8 * (size_t)s->nk;
    for (int i = 0; i < na; ++i)
        add_atom(s, sums, present, x[i], y[i], z[i], s->charge[f + i], s->type[f + i], krc + i * n8, krl + i * n8);
}

void yo_kspace_delta(const yo_sys *s, int imol, const double *x, const double *y, const double *z, double *du_c,
                     double *du_lj) {
    size_t n = (size_t)(1 + s->ntype) * 8 * s->nk;
    int f = s->first[imol], na = s->last[imol] - f + 1;
    size_t n8 = 8 * (size_t)s->nk;
    double *old = malloc(sizeof(double) * n), *nw = malloc(sizeof(double) * n);
    double *krc = malloc(sizeof(double) * n8 * na), *krl = malloc(sizeof(double) * n8 * na);
    int *pres = calloc(s->ntype + 1, sizeof(int));
    mol_sums(s, imol, s->xx + f, s->yy + f, s->zz + f, old, NULL, krc, krl);
    mol_sums(s, imol, x, y, z, nw, pres, krc, krl);
    for (size_t i = 0; i < n; ++i) nw[i] = nw[i] - old[i];
    *du_c = du_coulomb(s, s->S, nw);
    *du_lj = du_lj_calc(s, s->S, nw, pres);
    free(old); free(nw); free(krc); free(krl); free(pres);
}

/* ------------------------------------------------------------------------ */
/* RealSumLocal.f90:63-168 pair kernel                                       */
typedef struct { double potew, potlj6, potlj12, pot, FR; } yo_pair;

static void pair_terms(const yo_sys *s, double r2, int same_mol, double qi, double qj, int lj, double sigma6,
                       double four_eps, yo_pair *p) {
    double r = sqrt(r2), a = s->alpha * r, a2 = a * a, e = exp(-a2);
    double erfc_a = same_mol ? -yo_erf_luc(a, a2, e) : yo_erfc_luc(a, a2, e);
    p->potew = s->xclb * qi * qj * erfc_a / r;
    p->pot = p->potew;
    p->FR = s->xclb * qi * qj * (erfc_a / r + s->two_alpha_over_sqrt_pi * e) / r2;
    p->potlj6 = p->potlj12 = 0.0;
    if (lj) {
        double r6 = r2 * r2 * r2;
        double sor6 = sigma6 / r6;
        double C6 = 1.0 + a2 * (1.0 + a2 / 2.0);
        double a6 = a2 * a2 * a2;
        double C12 = C6 + a6 * (1.0 + a2 * (1.0 + a2 / 5.0) / 4.0) / 6.0;
        p->potlj6 = four_eps * sor6 * e * C6;
        p->potlj12 = four_eps * (sor6 * sor6) * e * C12;
        p->pot = p->pot + p->potlj12 - p->potlj6;
        p->FR = p->FR + four_eps * 6.0 * sor6 * e * (2.0 * sor6 * (C12 + (a6 * a6) / 720.0) - (C6 + a6 / 6.0)) / r2;
    }
}

/* EwaldSumRealSpace.f90:233-500; full when imol < 0.  comps[6] out. */
/* jlo >= 0 restricts j to jlo..jhi (the optional first/last arguments,
 * EwaldSumRealSpace.f90:338-339, used for molecule-molecule terms). */
static int rspace_calc_range(const yo_sys *s, int imol, const double *xn, const double *yn, const double *zn,
                             int jlo, int jhi, double *comps, double *uu, double *fx, double *fy, double *fz,
                             double *uub, double *fxb, double *fyb, double *fzb) {
    int full = imol < 0, N = s->natom, T = s->ntype;
    int first = full ? 0 : s->first[imol];
    int natom = full ? N : s->last[imol] - first + 1;
    int ntot_m = full ? N : natom;
    memset(uu, 0, sizeof(double) * ntot_m);
    memset(fx, 0, sizeof(double) * ntot_m);
    memset(fy, 0, sizeof(double) * ntot_m);
    memset(fz, 0, sizeof(double) * ntot_m);
    if (!full) {
        memset(uub, 0, sizeof(double) * N);
        memset(fxb, 0, sizeof(double) * N);
        memset(fyb, 0, sizeof(double) * N);
        memset(fzb, 0, sizeof(double) * N);
    } else {
        uub = uu; fxb = fx; fyb = fy; fzb = fz; /* aliasing, EwaldSumRealSpace.f90:92,99-101 */
    }
    for (int c = 0; c < 6; ++c) comps[c] = 0.0;
    for (int i = 0; i < natom; ++i) {
        int j0 = full ? i + 1 : 0, j1 = N - 1;
        if (jlo >= 0) { j0 = jlo; j1 = jhi; }
        int ii = full ? i : first + i;
        int im = s->molnum[ii], it = s->type[ii];
        double qi = s->charge[ii], dhi = s->hard_core[ii] / s->L;
        for (int j = j0; j <= j1; ++j) {
            int jm = s->molnum[j], jt = s->type[j];
            if (!full && im == jm) continue;
            double qj = s->charge[j], dhj = s->hard_core[j] / s->L;
            double d12 = (dhi + dhj) / 2.0;
            double xij = xn[i] - s->xx[j], yij = yn[i] - s->yy[j], zij = zn[i] - s->zz[j];
            if (xij > 0.5) xij = xij - 1.0;
            if (xij < -0.5) xij = xij + 1.0;
            if (yij > 0.5) yij = yij - 1.0;
            if (yij < -0.5) yij = yij + 1.0;
            if (zij > 0.5) zij = zij - 1.0;
            if (zij < -0.5) zij = zij + 1.0;
            double r2 = xij * xij + yij * yij + zij * zij;
            if (it > 0 && jt > 0 && im != jm && r2 < d12 * d12) return 1;
            if (r2 >= s->rmax2) continue;
            int lj = (it != 0 && jt != 0);
            int off = lj ? (it - 1) * T + (jt - 1) : 0;
            yo_pair p;
            pair_terms(s, r2, im == jm, qi, qj, lj, s->sigma6[off], s->four_eps[off], &p);
            comps[0] = comps[0] + p.potew;
            if (im == jm) comps[3] = comps[3] + p.potew;
            if (lj) {
                comps[1] = comps[1] + p.potlj6;
                comps[2] = comps[2] + p.potlj12;
                if (im == jm) {
                    comps[4] = comps[4] + p.potlj6;
                    comps[5] = comps[5] + p.potlj12;
                }
            }
            if (im != jm) {
                uu[i] = uu[i] + p.pot;
                uub[j] = uub[j] + p.pot;
                double f;
                f = xij * p.FR; fx[i] = fx[i] + f; fxb[j] = fxb[j] - f;
                f = yij * p.FR; fy[i] = fy[i] + f; fyb[j] = fyb[j] - f;
                f = zij * p.FR; fz[i] = fz[i] + f; fzb[j] = fzb[j] - f;
            }
        }
    }
    return 0;
}

static int rspace_calc(const yo_sys *s, int imol, const double *xn, const double *yn, const double *zn,
                       double *comps, double *uu, double *fx, double *fy, double *fz, double *uub, double *fxb,
                       double *fyb, double *fzb) {
    return rspace_calc_range(s, imol, xn, yn, zn, -1, -1, comps, uu, fx, fy, fz, uub, fxb, fyb, fzb);
}

int yo_rspace_full(yo_sys *s, double *comps, double *uu, double *fx, double *fy, double *fz) {
    return rspace_calc(s, -1, s->xx, s->yy, s->zz, comps, uu, fx, fy, fz, 0, 0, 0, 0);
}

int yo_rspace_partial(const yo_sys *s, int imol, const double *x, const double *y, const double *z, double *comps,
                      double *uu, double *fx, double *fy, double *fz, double *uub, double *fxb, double *fyb,
                      double *fzb) {
    return rspace_calc(s, imol, x, y, z, comps, uu, fx, fy, fz, uub, fxb, fyb, fzb);
}

/* ------------------------------------------------------------------------ */
/* EwaldSumTails.f90:18-38 and :45-127                                       */
void yo_tails(const yo_sys *s, double *self_c, double *lj6, double *lj12) {
    double S = 0.0;
    for (int i = 0; i < s->natom; ++i) S = S + s->charge[i] * s->charge[i];
    *self_c = -s->xclb * s->alpha_over_sqrt_pi * S;
    int T = s->ntype;
    double a3 = p3(s->alpha), a6 = a3 * a3, a9 = a3 * a6, a12 = a6 * a6, pspi = pow(YO_PI, 1.5);
    double sij6 = 0.0, sij12 = 0.0, sj6 = 0.0, sj12 = 0.0;
    for (int t1 = 0; t1 < T; ++t1)
        for (int t2 = t1; t2 < T; ++t2) {
            int n12 = s->tail_counts[t1] * s->tail_counts[t2];
            double l6 = s->lj6[t1 * T + t2] * (double)n12, l12 = s->lj12[t1 * T + t2] * (double)n12;
            if (t1 != t2) { l6 = l6 + l6; l12 = l12 + l12; }
            sij6 = sij6 + l6;
            sij12 = sij12 + l12;
        }
    for (int t = 0; t < T; ++t) {
        sj6 = sj6 + s->lj6[t * T + t] * (double)s->tail_counts[t];
        sj12 = sj12 + s->lj12[t * T + t] * (double)s->tail_counts[t];
    }
    *lj6 = pspi * a3 / 6.0 * sij6 - a6 / 12.0 * sj6;
    *lj12 = pspi * a9 / 1080.0 * sij12 - a12 / 1440.0 * sj12;
}

/* EwaldSumExternal.f90:165-183 calc_mu */
void yo_mu(const yo_sys *s, double *mu) {
    double mx = 0, my = 0, mz = 0;
    for (int i = 0; i < s->natom; ++i) {
        mx = mx + s->charge[i] * s->xx[i];
        my = my + s->charge[i] * s->yy[i];
        mz = mz + s->charge[i] * s->zz[i];
    }
    mu[0] = mx; mu[1] = my; mu[2] = mz;
}

/* MC.f90:681-698 mc_total_energy */
void yo_mc_energies(yo_sys *s, double *e) {
    double U_tail = s->tail_c + s->tail12 - s->tail6;
    double U_ext = s->kext * (s->mu[0] * s->mu[0] + s->mu[1] * s->mu[1] + s->mu[2] * s->mu[2]);
    double ew = s->uu_ew + s->e_kc + s->tail_c + U_ext;
    double lj = s->uu_lj12 - s->uu_lj6 + s->e_klj + s->tail12 - s->tail6;
    e[0] = ew + lj; e[1] = ew; e[2] = lj;
    e[3] = s->uu_ew; e[4] = s->uu_lj6; e[5] = s->uu_lj12;
    e[6] = s->e_kc; e[7] = s->e_klj; e[8] = U_tail; e[9] = U_ext;
}

/* MC.f90:477-585 mc_calc_total_sums */
int yo_mc_init(yo_sys *s, double *e) {
    double comps[6];
    int ov = rspace_calc(s, -1, s->xx, s->yy, s->zz, comps, s->uu, s->Fx, s->Fy, s->Fz, 0, 0, 0, 0);
    s->uu_ew = comps[0]; s->uu_lj6 = comps[1]; s->uu_lj12 = comps[2];
    s->uu_ew_intra = comps[3]; s->uu_lj6_intra = comps[4]; s->uu_lj12_intra = comps[5];
    yo_kspace_init(s, 0.0);
    yo_mu(s, s->mu);
    yo_tails(s, &s->tail_c, &s->tail6, &s->tail12);
    if (e) yo_mc_energies(s, e);
    return ov;
}

void yo_get_real_arrays(const yo_sys *s, double *uu, double *fx, double *fy, double *fz) {
    size_t n = s->natom;
    if (uu) memcpy(uu, s->uu, sizeof(double) * n);
    if (fx) memcpy(fx, s->Fx, sizeof(double) * n);
    if (fy) memcpy(fy, s->Fy, sizeof(double) * n);
    if (fz) memcpy(fz, s->Fz, sizeof(double) * n);
}

/* ------------------------------------------------------------------------ */
/* RNG.  Philox-4x32-10 (Salmon et al., SC'11), 24-bit uniforms in (0,1) as
 * random.f90:65-93 delivers them (a default-real random_number, 0 excluded). */
static uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t *hi) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

void yo_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void yo_rng_philox(yo_sys *s, uint64_t seed, uint64_t stream, uint64_t counter) {
    s->rng_mode = YO_RNG_PHILOX;
    s->key[0] = (uint32_t)seed; s->key[1] = (uint32_t)(seed >> 32);
    s->stream[0] = (uint32_t)stream; s->stream[1] = (uint32_t)(stream >> 32);
    s->ctr = counter;
}

uint64_t yo_rng_counter(const yo_sys *s) { return s->ctr; }

void yo_rng_const(yo_sys *s, double u) {
    s->rng_mode = YO_RNG_CONST;
    s->const_u = u;
}

/* Replay mode: rand() returns the recorded uniforms u[0..n) in order (the
 * counter is the position); past the end it returns 0.5 and the counter
 * keeps counting, so the caller sees counter > n. */
void yo_rng_list(yo_sys *s, const double *u, uint64_t n) {
    free(s->ulist);
    s->ulist = malloc(sizeof(double) * (n ? n : 1));
    memcpy(s->ulist, u, sizeof(double) * n);
    s->nlist = n;
    s->ctr = 0;
    s->rng_mode = YO_RNG_LIST;
}

double yo_rand(yo_sys *s) {
    if (s->rng_mode == YO_RNG_CONST) return s->const_u;
    if (s->rng_mode == YO_RNG_LIST) {
        uint64_t n = s->ctr++;
        return n < s->nlist ? s->ulist[n] : 0.5;
    }
    for (;;) {
        uint64_t n = s->ctr++;
        uint32_t c[4] = {(uint32_t)(n >> 2), (uint32_t)(n >> 34), s->stream[0], s->stream[1]}, o[4];
        yo_philox4x32_10(c, s->key, o);
        uint32_t k = o[n & 3] >> 8;
        if (k != 0) return (double)k * (1.0 / 16777216.0);
    }
}

/* ------------------------------------------------------------------------ */
/* geometry.f90:65-201 */
static void prod_vect(double x1, double y1, double z1, double x2, double y2, double z2, double *o) {
    o[0] = y1 * z2 - z1 * y2;
    o[1] = z1 * x2 - x1 * z2;
    o[2] = x1 * y2 - y1 * x2;
}

typedef struct { double cphi, sphi, ct, st, ca, sa; } yo_rot;

static void rot_matrix(const yo_rot *r, double *m) {
    double cphi = r->cphi, sphi = r->sphi, st = r->st, ct = r->ct, ca = r->ca, sa = r->sa, ca1 = 1.0 - ca;
    m[0] = (cphi * cphi * (st * st) * ca1 + ca);
    m[1] = (sphi * cphi * (st * st) * ca1 - ct * sa);
    m[2] = st * (ct * ca1 * cphi + sa * sphi);
    m[3] = (cphi * sphi * (st * st) * ca1 + ct * sa);
    m[4] = (sphi * sphi * (st * st) * ca1 + ca);
    m[5] = st * (ct * ca1 * sphi - sa * cphi);
    m[6] = st * (ct * ca1 * cphi - sa * sphi);
    m[7] = st * (ct * ca1 * sphi + sa * cphi);
    m[8] = (st * st * ca + ct * ct);
}

static void center_of_mass(const double *x, const double *y, const double *z, const double *m, int n, double *c) {
    double sx = 0, sy = 0, sz = 0, tm = 0;
    for (int i = 0; i < n; ++i) {
        sx = sx + m[i] * x[i]; sy = sy + m[i] * y[i]; sz = sz + m[i] * z[i];
        tm = tm + m[i];
    }
    c[0] = sx / tm; c[1] = sy / tm; c[2] = sz / tm;
}

/* MonteCarloMove.f90:10-28 */
/* Bound on every rejection-sampling loop (the reference loops until success;
 * a constant uniform per move can make some loops endless).  A move over the
 * bound is rejected without an acceptance draw and recorded with overlap = 2. */
#define YO_TRIES 65536

typedef struct {
    int stall;
    double fo[3], ftotold, co[3], couplold, umaxold, vmaxold, cold, qold;
    double v[3], d[3], angle;
    double fn[3], ftotnew, cn[3], couplnew, umaxnew, vmaxnew, cnew, qnew;
} yo_mcm;

/* MonteCarloMove.f90:34-54 */
static void mcm_init_old(yo_mcm *m, const double *f, const double *t) {
    for (int i = 0; i < 3; ++i) { m->fo[i] = f[i]; m->co[i] = t[i]; }
    m->ftotold = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    m->couplold = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

/* MonteCarloMove.f90:56-106 */
static void mcm_init_new(const yo_sys *s, yo_mcm *m, const double *f, const double *t) {
    for (int i = 0; i < 3; ++i) { m->fn[i] = f[i]; m->cn[i] = t[i]; }
    m->ftotnew = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    m->couplnew = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    double umaxnew = s->lambda_f * m->ftotnew * s->dr, cnew;
    if (umaxnew > 1e-2) {
        double e = exp(-2.0 * umaxnew);
        double tt = (umaxnew * (1.0 + e) - (1.0 - e)) / 2.0;
        cnew = 3.0 * tt / p3(umaxnew);
    } else {
        umaxnew = 0.0; cnew = 1.0;
    }
    m->umaxnew = umaxnew; m->cnew = cnew;
    double vmaxnew = s->lambda_c * m->couplnew * s->d_angle, qnew;
    if (vmaxnew > 1e-2) qnew = yo_shi_sans_exp(vmaxnew) / vmaxnew;
    else { vmaxnew = 0.0; qnew = 1.0; }
    m->vmaxnew = vmaxnew; m->qnew = qnew;
}

/* MonteCarloMove.f90:109-266 chooseDisplacement */
static void mcm_displacement(yo_sys *s, yo_mcm *m) {
    double fxo = m->fo[0], fyo = m->fo[1], fzo = m->fo[2], ftot = m->ftotold;
    double umax = s->lambda_f * ftot * s->dr, cold, dx, dy, dz;
    if (umax > 1e-2) {
        double e = exp(-2.0 * umax);
        double t = (umax * (1.0 + e) - (1.0 - e)) / 2.0;
        cold = 3.0 * t / p3(umax);
        double xksi = 0.0;
        for (int g = 0; xksi < 1e-10; ++g) {
            if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            xksi = yo_rand(s);
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
            if (u < 1e-3) g = p3(u) / 3.0;
            double h = u + log(g), h1 = u * sh / g;
            if (k < 10) u = u + (z - h) / h1;
        }
        double deltar = u / (s->lambda_f * ftot);
        xksi = 1.0;
        for (int g = 0; xksi > (1.0 - 1e-10) || xksi < 1e-10; ++g) {
            if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            xksi = yo_rand(s);
        }
        double xcostet = 1.0 + log(xksi + (1.0 - xksi) * e) / u;
        if (u < 1e-7) xcostet = 1.0 - 2.0 * (1.0 - xksi);
        if (xcostet > 1.0 - 1e-14) xcostet = 1.0 - 1e-14;
        if (xcostet < -1.0 + 1e-14) xcostet = -1.0 + 1e-14;
        double xsintet = sqrt(1.0 - xcostet * xcostet);
        double w = -1.0, wv[3];
        for (int g = 0; w <= 1e-10; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            double v2 = 2.0, vx = 0, vy = 0, vz = 0;
            for (int g = 0; v2 > 1.0; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
                vx = 2.0 * yo_rand(s) - 1.0;
                vy = 2.0 * yo_rand(s) - 1.0;
                vz = 2.0 * yo_rand(s) - 1.0;
                v2 = vx * vx + vy * vy + vz * vz;
            }
            prod_vect(vx, vy, vz, fxo, fyo, fzo, wv);
            w = sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
        }
        dx = deltar * (xcostet * fxo / ftot + xsintet * wv[0] / w);
        dy = deltar * (xcostet * fyo / ftot + xsintet * wv[1] / w);
        dz = deltar * (xcostet * fzo / ftot + xsintet * wv[2] / w);
    } else {
        double v2 = 2.0, vx = 0, vy = 0, vz = 0;
        for (int g = 0; v2 > 1.0; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            vx = 2.0 * yo_rand(s) - 1.0;
            vy = 2.0 * yo_rand(s) - 1.0;
            vz = 2.0 * yo_rand(s) - 1.0;
            v2 = vx * vx + vy * vy + vz * vz;
        }
        dx = s->dr * vx; dy = s->dr * vy; dz = s->dr * vz;
        umax = 0.0; cold = 1.0;
    }
    m->d[0] = dx; m->d[1] = dy; m->d[2] = dz;
    m->cold = cold; m->umaxold = umax;
}

/* MonteCarloMove.f90:269-444 chooseRotation */
static void mcm_rotation(yo_sys *s, yo_mcm *m, yo_rot *rot) {
    double cx = m->co[0], cy = m->co[1], cz = m->co[2], cabs = m->couplold;
    double vmax = s->lambda_c * cabs * s->d_angle, qold, angle, vx, vy, vz;
    if (vmax > 1e-2) {
        double emax = exp(-2.0 * vmax), shmax = (1.0 - emax) / 2.0;
        double xksi = 1.0, proba = 0.0, v = 0.0, e = 1.0;
        for (int g = 0; xksi >= proba; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            v = 0.0;
            for (int g = 0; fabs(v) < 1e-10; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
                xksi = yo_rand(s);
                double xs = xksi * shmax;
                v = vmax + log(xs + sqrt(xs * xs + emax));
                if (xs * xs < 1e-14 * emax) v = xs / sqrt(emax);
            }
            e = exp(-2.0 * v);
            proba = (1.0 - e) / v / (1.0 + e);
            if (v < 1e-5) proba = (2.0 * v - 2.0 * v * v) / v / (1.0 + e);
            xksi = yo_rand(s);
        }
        double cosx = v / vmax;
        qold = yo_shi_sans_exp(vmax) / vmax;
        xksi = 1.0;
        for (int g = 0; xksi > (1.0 - 1e-10); ++g) {
            if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            xksi = yo_rand(s);
        }
        angle = s->d_angle * (1.0 + log(xksi + (1.0 - xksi) * e) / v);
        if (v < 1e-7) angle = s->d_angle * (1.0 - 2.0 * (1.0 - xksi));
        double w = -1.0, wv[3];
        for (int g = 0; w <= 1e-10; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            double v2 = 2.0;
            vx = vy = vz = 0;
            for (int g = 0; v2 > 1.0 || v2 < 1e-10; ++g) {
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
                vx = 2.0 * yo_rand(s) - 1.0;
                vy = 2.0 * yo_rand(s) - 1.0;
                vz = 2.0 * yo_rand(s) - 1.0;
                v2 = vx * vx + vy * vy + vz * vz;
            }
            prod_vect(vx, vy, vz, cx, cy, cz, wv);
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
                if (g >= YO_TRIES || m->stall) { m->stall = 1; break; }
            vx = 2.0 * yo_rand(s) - 1.0;
            vy = 2.0 * yo_rand(s) - 1.0;
            vz = 2.0 * yo_rand(s) - 1.0;
            v2 = vx * vx + vy * vy + vz * vz;
        }
        double v = sqrt(v2);
        vx = vx / v; vy = vy / v; vz = vz / v;
        angle = s->d_angle * (2.0 * yo_rand(s) - 1.0);
        vmax = 0.0; qold = 1.0;
    }
    double ca = cos(angle), sa = sin(angle), ct = vz, st = sqrt(1.0 - ct * ct);
    /* MonteCarloMove.f90:425: phi is left unset when st == 0 */
    double cphi = 1.0, sphi = 0.0;
    if (st > 0.0) { cphi = vx / st; sphi = vy / st; }
    rot->cphi = cphi; rot->sphi = sphi; rot->ct = ct; rot->st = st; rot->ca = ca; rot->sa = sa;
    m->v[0] = vx; m->v[1] = vy; m->v[2] = vz;
    m->vmaxold = vmax; m->angle = angle; m->qold = qold;
}

/* MonteCarloMove.f90:447-570 accept_or_decline */
static int mcm_accept(yo_sys *s, const yo_mcm *m, double dutot) {
    double exposu = -dutot;
    double exposf = -s->lambda_f * ((m->fo[0] + m->fn[0]) * m->d[0] + (m->fo[1] + m->fn[1]) * m->d[1] +
                                    (m->fo[2] + m->fn[2]) * m->d[2]);
    double exposfcorr = m->umaxold - m->umaxnew;
    double exposc = -s->lambda_c * m->angle *
                    ((m->co[0] + m->cn[0]) * m->v[0] + (m->co[1] + m->cn[1]) * m->v[1] + (m->co[2] + m->cn[2]) * m->v[2]);
    double exposccorr = m->vmaxold - m->vmaxnew;
    double factor = m->cold * m->qold / (m->cnew * m->qnew) * exp(exposu + exposf + exposfcorr + exposc + exposccorr);
    double xksi = yo_rand(s);
    if (xksi > factor) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* MC.f90:831-1322 movemc */
int yo_movemc(yo_sys *s, int imol, yo_move_rec *rec) {
    int N = s->natom, T = s->ntype, nk = s->nk;
    size_t n8 = 8 * (size_t)nk, nS = (size_t)(1 + T) * n8;
    yo_move_rec r;
    memset(&r, 0, sizeof r);
    r.imol = imol;
    r.imol2 = -1;
    yo_mu(s, s->mu); /* MC.f90:916 */
    int f = s->first[imol], last = s->last[imol], na = last - f + 1;
    const double *xo = s->xx + f, *yo = s->yy + f, *zo = s->zz + f;
    double U_old = 0.0;
    for (int i = 0; i < na; ++i) U_old = U_old + s->uu[f + i];
    double fxo[8], fyo[8], fzo[8], fxn[8], fyn[8], fzn[8];
    for (int i = 0; i < na; ++i) { fxo[i] = s->Fx[f + i]; fyo[i] = s->Fy[f + i]; fzo[i] = s->Fz[f + i]; }
    /* step 3: old molecule sums and k-space forces (MC.f90:946-982) */
    memset(s->mol_old, 0, sizeof(double) * nS);
    for (int i = 0; i < na; ++i) {
        int a = f + i, it = s->type[a];
        double fc[3], fl[3] = {0, 0, 0};
        add_atom(s, s->mol_old, NULL, s->xx[a], s->yy[a], s->zz[a], s->charge[a], it, s->krc + i * n8,
                 s->krl + i * n8);
        force_coulomb(s, s->S, s->krc + i * n8, fc);
        if (it > 0) force_lj(s, s->S, it, s->krl + i * n8, fl);
        fxo[i] = fxo[i] + fc[0] + fl[0];
        fyo[i] = fyo[i] + fc[1] + fl[1];
        fzo[i] = fzo[i] + fc[2] + fl[2];
    }
    /* step 3A: external forces (MC.f90:988-997) */
    for (int i = 0; i < na; ++i) {
        double q = s->charge[f + i];
        fxo[i] = fxo[i] + s->minus_two_kext * q * s->mu[0];
        fyo[i] = fyo[i] + s->minus_two_kext * q * s->mu[1];
        fzo[i] = fzo[i] + s->minus_two_kext * q * s->mu[2];
    }
    double muo[3] = {0, 0, 0};
    for (int i = 0; i < na; ++i) {
        double q = s->charge[f + i];
        muo[0] = muo[0] + q * xo[i]; muo[1] = muo[1] + q * yo[i]; muo[2] = muo[2] + q * zo[i];
    }
    /* step 4: total force and torque (MC.f90:1002-1025) */
    double Fo[3] = {0, 0, 0}, To[3] = {0, 0, 0}, com[3], tq[3];
    for (int i = 0; i < na; ++i) { Fo[0] = Fo[0] + fxo[i]; Fo[1] = Fo[1] + fyo[i]; Fo[2] = Fo[2] + fzo[i]; }
    center_of_mass(xo, yo, zo, s->mass + f, na, com);
    for (int i = 0; i < na; ++i) {
        prod_vect(xo[i] - com[0], yo[i] - com[1], zo[i] - com[2], fxo[i], fyo[i], fzo[i], tq);
        To[0] = To[0] + tq[0]; To[1] = To[1] + tq[1]; To[2] = To[2] + tq[2];
    }
    /* step 5: proposal (MC.f90:1031-1067) */
    yo_mcm m;
    memset(&m, 0, sizeof m);
    mcm_init_old(&m, Fo, To);
    mcm_displacement(s, &m);
    yo_rot rot;
    mcm_rotation(s, &m, &rot);
    if (m.stall) {
        r.overlap = 2;
        if (rec) *rec = r;
        return 0;
    }
    double cn[3];
    for (int d = 0; d < 3; ++d) cn[d] = com[d] + m.d[d] - round(com[d] + m.d[d]);
    double xn[8], yn[8], zn[8];
    if (na > 1) {
        double R[9];
        rot_matrix(&rot, R);
        for (int i = 0; i < na; ++i) {
            double xr = xo[i] - com[0], yr = yo[i] - com[1], zr = zo[i] - com[2];
            xn[i] = (R[0] * xr + R[1] * yr + R[2] * zr) + cn[0];
            yn[i] = (R[3] * xr + R[4] * yr + R[5] * zr) + cn[1];
            zn[i] = (R[6] * xr + R[7] * yr + R[8] * zr) + cn[2];
        }
    } else {
        xn[0] = cn[0]; yn[0] = cn[1]; zn[0] = cn[2];
    }
    for (int d = 0; d < 3; ++d) { r.f_old[d] = Fo[d]; r.t_old[d] = To[d]; r.delta[d] = m.d[d]; }
    r.angle = m.angle;
    /* step 6: new real-space partial sum (MC.f90:1077-1084) */
    double cmp_new[6], uun[8];
    double *bn = s->back_new;
    if (rspace_calc(s, imol, xn, yn, zn, cmp_new, uun, fxn, fyn, fzn, bn, bn + N, bn + 2 * N, bn + 3 * N)) {
        r.overlap = 1;
        r.accepted = 0;
        if (rec) *rec = r;
        return 0;
    }
    double U_new = 0.0;
    for (int i = 0; i < na; ++i) U_new = U_new + uun[i];
    double du_real = U_new - U_old;
    /* steps 8-9: new molecule sums, delta, new total (MC.f90:1096-1137) */
    int pres[64];
    mol_sums(s, imol, xn, yn, zn, s->mol_new, pres, s->krc, s->krl);
    for (size_t i = 0; i < nS; ++i) s->mol_new[i] = s->mol_new[i] - s->mol_old[i];
    for (size_t i = 0; i < nS; ++i) s->tot_new[i] = s->S[i] + s->mol_new[i];
    double frn[8][3];
    for (int i = 0; i < na; ++i) { frn[i][0] = fxn[i]; frn[i][1] = fyn[i]; frn[i][2] = fzn[i]; }
    for (int i = 0; i < na; ++i) {
        int a = f + i, it = s->type[a];
        double fc[3], fl[3] = {0, 0, 0};
        force_coulomb(s, s->tot_new, s->krc + i * n8, fc);
        if (it > 0) force_lj(s, s->tot_new, it, s->krl + i * n8, fl);
        fxn[i] = frn[i][0] + fc[0] + fl[0];
        fyn[i] = frn[i][1] + fc[1] + fl[1];
        fzn[i] = frn[i][2] + fc[2] + fl[2];
    }
    /* steps 10-11 (MC.f90:1168-1198) */
    double du_kc = du_coulomb(s, s->S, s->mol_new);
    double du_klj = du_lj_calc(s, s->S, s->mol_new, pres);
    double du_f = du_kc + du_klj;
    double mun[3] = {0, 0, 0};
    for (int i = 0; i < na; ++i) {
        double q = s->charge[f + i];
        mun[0] = mun[0] + q * xn[i]; mun[1] = mun[1] + q * yn[i]; mun[2] = mun[2] + q * zn[i];
    }
    double dmu[3] = {mun[0] - muo[0], mun[1] - muo[1], mun[2] - muo[2]};
    double du_ext = s->kext * ((2 * s->mu[0] + dmu[0]) * dmu[0] + (2 * s->mu[1] + dmu[1]) * dmu[1] +
                               (2 * s->mu[2] + dmu[2]) * dmu[2]);
    double mu_new[3] = {s->mu[0] + dmu[0], s->mu[1] + dmu[1], s->mu[2] + dmu[2]};
    for (int i = 0; i < na; ++i) {
        double q = s->charge[f + i];
        fxn[i] = fxn[i] + s->minus_two_kext * q * mu_new[0];
        fyn[i] = fyn[i] + s->minus_two_kext * q * mu_new[1];
        fzn[i] = fzn[i] + s->minus_two_kext * q * mu_new[2];
    }
    /* step 12 (MC.f90:1202-1224) */
    double Fn[3] = {0, 0, 0}, Tn[3] = {0, 0, 0}, comn[3];
    for (int i = 0; i < na; ++i) { Fn[0] = Fn[0] + fxn[i]; Fn[1] = Fn[1] + fyn[i]; Fn[2] = Fn[2] + fzn[i]; }
    center_of_mass(xn, yn, zn, s->mass + f, na, comn);
    for (int i = 0; i < na; ++i) {
        prod_vect(xn[i] - comn[0], yn[i] - comn[1], zn[i] - comn[2], fxn[i], fyn[i], fzn[i], tq);
        Tn[0] = Tn[0] + tq[0]; Tn[1] = Tn[1] + tq[1]; Tn[2] = Tn[2] + tq[2];
    }
    /* step 13 (MC.f90:1228-1233) */
    double du_total = du_real + du_f + du_ext;
    mcm_init_new(s, &m, Fn, Tn);
    int acc = mcm_accept(s, &m, du_total);
    r.accepted = acc;
    r.du_real = du_real; r.du_kc = du_kc; r.du_klj = du_klj; r.du_ext = du_ext; r.du_total = du_total;
    for (int d = 0; d < 3; ++d) { r.f_new[d] = Fn[d]; r.t_new[d] = Tn[d]; }
    if (acc) {
        /* step 14 (MC.f90:1242-1320) */
        for (size_t i = 0; i < nS; ++i) s->S[i] = s->S[i] + s->mol_new[i];
        double cmp_old[6], uuo[8], fxo2[8], fyo2[8], fzo2[8];
        double *bo = s->back_old;
        rspace_calc(s, imol, xo, yo, zo, cmp_old, uuo, fxo2, fyo2, fzo2, bo, bo + N, bo + 2 * N, bo + 3 * N);
        for (int j = 0; j < N; ++j) {
            s->uu[j] = s->uu[j] + bn[j] - bo[j];
            s->Fx[j] = s->Fx[j] + bn[N + j] - bo[N + j];
            s->Fy[j] = s->Fy[j] + bn[2 * N + j] - bo[2 * N + j];
            s->Fz[j] = s->Fz[j] + bn[3 * N + j] - bo[3 * N + j];
        }
        for (int i = 0; i < na; ++i) {
            s->uu[f + i] = uun[i];
            s->Fx[f + i] = frn[i][0]; s->Fy[f + i] = frn[i][1]; s->Fz[f + i] = frn[i][2];
        }
        s->e_kc = s->e_kc + du_kc;
        s->e_klj = s->e_klj + du_klj;
        s->uu_ew = s->uu_ew + cmp_new[0] - cmp_old[0];
        s->uu_lj6 = s->uu_lj6 + cmp_new[1] - cmp_old[1];
        s->uu_lj12 = s->uu_lj12 + cmp_new[2] - cmp_old[2];
        for (int d = 0; d < 3; ++d) s->mu[d] = mu_new[d];
        for (int i = 0; i < na; ++i) { s->xx[f + i] = xn[i]; s->yy[f + i] = yn[i]; s->zz[f + i] = zn[i]; }
    }
    if (rec) *rec = r;
    return acc;
}

/* BiasedRandom.f90:41-75 BiasedRandom_init: cumulative intervals of
 * num/sum(num) (flat when the sum is 0); a zero entry gets end = -1 so that
 * it is never chosen. */
static void biased_init(int n, const double *num, double *beg, double *end) {
    double S = 0.0;
    for (int i = 0; i < n; ++i) S = S + num[i];
    for (int i = 0; i < n; ++i) {
        double prob = S == 0 ? 1.0 / (double)n : num[i] / S;
        beg[i] = i == 0 ? 0.0 : end[i - 1];
        end[i] = beg[i] + prob;
    }
    if (S > 0)
        for (int i = 0; i < n; ++i)
            if (num[i] == 0) end[i] = -1.0;
}

/* BiasedRandom.f90:80-109 BiasedRandom_choose: first interval with
 * beg < rnd < end, else the first one. */
static int biased_choose(yo_sys *s, const double *beg, const double *end) {
    double rnd = yo_rand(s);
    for (int i = 0; i < s->nmol; ++i)
        if (rnd > beg[i] && rnd < end[i]) return i;
    return 0;
}

static int choose_molecule(yo_sys *s) { return biased_choose(s, s->freq_beg, s->freq_end); }

/* runmc.f90:26-66 init_freq with per-molecule move and exchange weights
 * (read_freq_file columns 2 and 3); NULL = all ones / all zeros. */
int yo_set_frequencies(yo_sys *s, const double *mv_num, const double *xchange_num) {
    int n = s->nmol;
    double *mv = malloc(sizeof(double) * n), *xc = malloc(sizeof(double) * n);
    double smv = 0.0, sxc = 0.0;
    for (int i = 0; i < n; ++i) {
        mv[i] = mv_num ? mv_num[i] : 1.0;
        xc[i] = xchange_num ? xchange_num[i] : 0.0;
    }
    for (int i = 0; i < n; ++i) smv = smv + mv[i];
    for (int i = 0; i < n; ++i) sxc = sxc + xc[i];
    biased_init(n, mv, s->freq_beg, s->freq_end);
    biased_init(n, xc, s->xfreq_beg, s->xfreq_end);
    s->pxch = sxc / (smv + sxc);
    free(mv);
    free(xc);
    return 0;
}

/* MC.f90:1325-1954 mc_xchange: molecules i1 and i2 swap centres of mass
 * (each keeps its orientation); three partial real sums with both at the
 * new places (new1 over j not in i1, new2 over j not in i2, new12 over the
 * pair only), k-space dU of the two molecules' joint delta, external dU,
 * accepted iff rand() <= exp(-dU) (MonteCarloMove.f90:572-597); no
 * acceptance draw after an overlap of new1 or new2.  The acceptance dU_real
 * is sum(uu new) - sum(uu old) over both molecules, which counts the pair
 * term twice; the stored components subtract it once (MC.f90:1829-1859). */
int yo_xchange(yo_sys *s, int i1, int i2, yo_move_rec *rec) {
    const int N = s->natom, nk = s->nk;
    const size_t n8 = 8 * (size_t)nk, nS = (size_t)(1 + s->ntype) * n8;
    yo_move_rec r;
    memset(&r, 0, sizeof r);
    r.imol = i1;
    r.imol2 = i2;
    yo_mu(s, s->mu); /* MC.f90:1429 */
    const int f1 = s->first[i1], n1 = s->last[i1] - f1 + 1, f2 = s->first[i2], n2 = s->last[i2] - f2 + 1;
    double xo1[8], yo1[8], zo1[8], xo2[8], yo2[8], zo2[8];
    for (int a = 0; a < n1; ++a) { xo1[a] = s->xx[f1 + a]; yo1[a] = s->yy[f1 + a]; zo1[a] = s->zz[f1 + a]; }
    for (int a = 0; a < n2; ++a) { xo2[a] = s->xx[f2 + a]; yo2[a] = s->yy[f2 + a]; zo2[a] = s->zz[f2 + a]; }
    double u1 = 0.0, u2 = 0.0;
    for (int a = 0; a < n1; ++a) u1 = u1 + s->uu[f1 + a];
    for (int a = 0; a < n2; ++a) u2 = u2 + s->uu[f2 + a];
    const double U_old = u1 + u2;
    /* old joint sums, atoms of i1 then of i2 (MC.f90:1464-1488) */
    memset(s->mol_old, 0, sizeof(double) * nS);
    for (int a = 0; a < n1; ++a)
        add_atom(s, s->mol_old, NULL, xo1[a], yo1[a], zo1[a], s->charge[f1 + a], s->type[f1 + a], s->krc, s->krl);
    for (int a = 0; a < n2; ++a)
        add_atom(s, s->mol_old, NULL, xo2[a], yo2[a], zo2[a], s->charge[f2 + a], s->type[f2 + a], s->krc, s->krl);
    double muo1[3] = {0, 0, 0}, muo2[3] = {0, 0, 0};
    for (int a = 0; a < n1; ++a) {
        double q = s->charge[f1 + a];
        muo1[0] = muo1[0] + q * xo1[a]; muo1[1] = muo1[1] + q * yo1[a]; muo1[2] = muo1[2] + q * zo1[a];
    }
    for (int a = 0; a < n2; ++a) {
        double q = s->charge[f2 + a];
        muo2[0] = muo2[0] + q * xo2[a]; muo2[1] = muo2[1] + q * yo2[a]; muo2[2] = muo2[2] + q * zo2[a];
    }
    /* swap the centres of mass (MC.f90:1514-1531) */
    double c1[3], c2[3];
    center_of_mass(xo1, yo1, zo1, s->mass + f1, n1, c1);
    center_of_mass(xo2, yo2, zo2, s->mass + f2, n2, c2);
    double xn1[8], yn1[8], zn1[8], xn2[8], yn2[8], zn2[8];
    for (int a = 0; a < n1; ++a) {
        xn1[a] = xo1[a] - c1[0] + c2[0]; yn1[a] = yo1[a] - c1[1] + c2[1]; zn1[a] = zo1[a] - c1[2] + c2[2];
    }
    for (int a = 0; a < n2; ++a) {
        xn2[a] = xo2[a] - c2[0] + c1[0]; yn2[a] = yo2[a] - c2[1] + c1[1]; zn2[a] = zo2[a] - c2[2] + c1[2];
    }
    /* new partial sums with both molecules at their new places (MC.f90:1547-1584) */
    double *bn1 = malloc(sizeof(double) * 4 * N), *bn2 = malloc(sizeof(double) * 4 * N);
    double *bo1 = malloc(sizeof(double) * 4 * N), *bo2 = malloc(sizeof(double) * 4 * N);
    double *b12 = malloc(sizeof(double) * 4 * N);
    for (int a = 0; a < n1; ++a) { s->xx[f1 + a] = xn1[a]; s->yy[f1 + a] = yn1[a]; s->zz[f1 + a] = zn1[a]; }
    for (int a = 0; a < n2; ++a) { s->xx[f2 + a] = xn2[a]; s->yy[f2 + a] = yn2[a]; s->zz[f2 + a] = zn2[a]; }
    double cn1[6], cn2[6], cn12[6], uun1[8], uun2[8], fn1[3][8], fn2[3][8], t12[4][8];
    int ov1 = rspace_calc(s, i1, xn1, yn1, zn1, cn1, uun1, fn1[0], fn1[1], fn1[2], bn1, bn1 + N, bn1 + 2 * N,
                          bn1 + 3 * N);
    int ov2 = rspace_calc(s, i2, xn2, yn2, zn2, cn2, uun2, fn2[0], fn2[1], fn2[2], bn2, bn2 + N, bn2 + 2 * N,
                          bn2 + 3 * N);
    (void)rspace_calc_range(s, i1, xn1, yn1, zn1, f2, f2 + n2 - 1, cn12, t12[0], t12[1], t12[2], t12[3], b12,
                            b12 + N, b12 + 2 * N, b12 + 3 * N);
    for (int a = 0; a < n1; ++a) { s->xx[f1 + a] = xo1[a]; s->yy[f1 + a] = yo1[a]; s->zz[f1 + a] = zo1[a]; }
    for (int a = 0; a < n2; ++a) { s->xx[f2 + a] = xo2[a]; s->yy[f2 + a] = yo2[a]; s->zz[f2 + a] = zo2[a]; }
    int acc = 0;
    if (ov1 || ov2) {
        r.overlap = 1;
    } else {
        double v1 = 0.0, v2 = 0.0;
        for (int a = 0; a < n1; ++a) v1 = v1 + uun1[a];
        for (int a = 0; a < n2; ++a) v2 = v2 + uun2[a];
        const double du_real = (v1 + v2) - U_old;
        /* new joint sums and the delta (MC.f90:1602-1641) */
        int pres[64];
        memset(s->mol_new, 0, sizeof(double) * nS);
        memset(pres, 0, sizeof pres);
        for (int a = 0; a < n1; ++a)
            add_atom(s, s->mol_new, pres, xn1[a], yn1[a], zn1[a], s->charge[f1 + a], s->type[f1 + a], s->krc, s->krl);
        for (int a = 0; a < n2; ++a)
            add_atom(s, s->mol_new, pres, xn2[a], yn2[a], zn2[a], s->charge[f2 + a], s->type[f2 + a], s->krc, s->krl);
        for (size_t i = 0; i < nS; ++i) s->mol_new[i] = s->mol_new[i] - s->mol_old[i];
        const double du_kc = du_coulomb(s, s->S, s->mol_new);
        const double du_klj = du_lj_calc(s, s->S, s->mol_new, pres);
        const double du_f = du_kc + du_klj;
        /* external term (MC.f90:1692-1716) */
        double mun1[3] = {0, 0, 0}, mun2[3] = {0, 0, 0}, dmu[3], mu_new[3];
        for (int a = 0; a < n1; ++a) {
            double q = s->charge[f1 + a];
            mun1[0] = mun1[0] + q * xn1[a]; mun1[1] = mun1[1] + q * yn1[a]; mun1[2] = mun1[2] + q * zn1[a];
        }
        for (int a = 0; a < n2; ++a) {
            double q = s->charge[f2 + a];
            mun2[0] = mun2[0] + q * xn2[a]; mun2[1] = mun2[1] + q * yn2[a]; mun2[2] = mun2[2] + q * zn2[a];
        }
        for (int d = 0; d < 3; ++d) dmu[d] = (mun1[d] - muo1[d]) + (mun2[d] - muo2[d]);
        const double du_ext = s->kext * ((2 * s->mu[0] + dmu[0]) * dmu[0] + (2 * s->mu[1] + dmu[1]) * dmu[1] +
                                         (2 * s->mu[2] + dmu[2]) * dmu[2]);
        for (int d = 0; d < 3; ++d) mu_new[d] = s->mu[d] + dmu[d];
        const double du_total = du_real + du_f + du_ext;
        /* MonteCarloMove.f90:572-597 */
        const double factor = exp(-du_total);
        acc = !(yo_rand(s) > factor);
        r.du_real = du_real; r.du_kc = du_kc; r.du_klj = du_klj; r.du_ext = du_ext; r.du_total = du_total;
        for (int d = 0; d < 3; ++d) r.delta[d] = c2[d] - c1[d];
        if (acc) { /* MC.f90:1754-1872 */
            for (size_t i = 0; i < nS; ++i) s->S[i] = s->S[i] + s->mol_new[i];
            double co1[6], co2[6], co12[6], uuo[8], fo[3][8];
            rspace_calc(s, i1, xo1, yo1, zo1, co1, uuo, fo[0], fo[1], fo[2], bo1, bo1 + N, bo1 + 2 * N, bo1 + 3 * N);
            rspace_calc(s, i2, xo2, yo2, zo2, co2, uuo, fo[0], fo[1], fo[2], bo2, bo2 + N, bo2 + 2 * N, bo2 + 3 * N);
            (void)rspace_calc_range(s, i1, xo1, yo1, zo1, f2, f2 + n2 - 1, co12, t12[0], t12[1], t12[2], t12[3], b12,
                                    b12 + N, b12 + 2 * N, b12 + 3 * N);
            for (int j = 0; j < N; ++j) {
                s->uu[j] = s->uu[j] + bn1[j] - bo1[j] + bn2[j] - bo2[j];
                s->Fx[j] = s->Fx[j] + bn1[N + j] - bo1[N + j] + bn2[N + j] - bo2[N + j];
                s->Fy[j] = s->Fy[j] + bn1[2 * N + j] - bo1[2 * N + j] + bn2[2 * N + j] - bo2[2 * N + j];
                s->Fz[j] = s->Fz[j] + bn1[3 * N + j] - bo1[3 * N + j] + bn2[3 * N + j] - bo2[3 * N + j];
            }
            for (int a = 0; a < n1; ++a) {
                s->uu[f1 + a] = uun1[a];
                s->Fx[f1 + a] = fn1[0][a]; s->Fy[f1 + a] = fn1[1][a]; s->Fz[f1 + a] = fn1[2][a];
            }
            for (int a = 0; a < n2; ++a) {
                s->uu[f2 + a] = uun2[a];
                s->Fx[f2 + a] = fn2[0][a]; s->Fy[f2 + a] = fn2[1][a]; s->Fz[f2 + a] = fn2[2][a];
            }
            s->e_kc = s->e_kc + du_kc;
            s->e_klj = s->e_klj + du_klj;
            s->uu_ew = s->uu_ew + cn1[0] - co1[0] + cn2[0] - co2[0] - (cn12[0] - co12[0]);
            s->uu_lj6 = s->uu_lj6 + cn1[1] - co1[1] + cn2[1] - co2[1] - (cn12[1] - co12[1]);
            s->uu_lj12 = s->uu_lj12 + cn1[2] - co1[2] + cn2[2] - co2[2] - (cn12[2] - co12[2]);
            for (int d = 0; d < 3; ++d) s->mu[d] = mu_new[d];
            for (int a = 0; a < n1; ++a) { s->xx[f1 + a] = xn1[a]; s->yy[f1 + a] = yn1[a]; s->zz[f1 + a] = zn1[a]; }
            for (int a = 0; a < n2; ++a) { s->xx[f2 + a] = xn2[a]; s->yy[f2 + a] = yn2[a]; s->zz[f2 + a] = zn2[a]; }
        }
    }
    r.accepted = acc;
    free(bn1); free(bn2); free(bo1); free(bo2); free(b12);
    if (rec) *rec = r;
    return acc;
}

/* runmc.f90:223-266: one step of a move cycle: the exchange draw, then an
 * exchange of two distinct molecules picked with the exchange weights, or a
 * movemc of one picked with the move weights. */
static int run_step(yo_sys *s, yo_move_rec *rec) {
    double rnd = yo_rand(s);
    if (rnd < s->pxch) {
        int i1 = biased_choose(s, s->xfreq_beg, s->xfreq_end), i2 = i1;
        while (i2 == i1) i2 = biased_choose(s, s->xfreq_beg, s->xfreq_end);
        return yo_xchange(s, i1, i2, rec);
    }
    return yo_movemc(s, choose_molecule(s), rec);
}

/* runmc.f90:204-268 without volume steps: per cycle one NPT draw, then
 * nmol steps. */
int yo_run_cycles(yo_sys *s, int ncycles, yo_move_rec *recs, int64_t *n_accepted) {
    int64_t acc = 0, k = 0;
    for (int c = 0; c < ncycles; ++c) {
        (void)yo_rand(s); /* runmc.f90:206 */
        for (int j = 0; j < s->nmol; ++j) {
            yo_move_rec r;
            acc += run_step(s, &r);
            if (recs) recs[k] = r;
            ++k;
        }
    }
    if (n_accepted) *n_accepted = acc;
    return 0;
}

int yo_run_const(yo_sys *s, int nmoves, const int32_t *imol, const double *u, yo_move_rec *recs) {
    int acc = 0;
    for (int k = 0; k < nmoves; ++k) {
        yo_rng_const(s, u[k]);
        acc += yo_movemc(s, imol[k], recs ? &recs[k] : NULL);
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* MC.f90:701-825 mcvol, with scale_box (scale_box.f90:6-62),
 * Parameters_recalc at the new length (parameters.f90:355-413) and
 * LJTypes_scale_sigma (LJTypes.f90:238-252).  The box-unit parameters are
 * rescaled from the current ones: xclb, kext and dr scale as 1/L, the LJ
 * tables as sigma^6 ~ L^-6 and sigma^12 ~ L^-12 (four_eps unchanged); alpha,
 * rmax and kmax are constant in box units.  pressure is in kT/A^3.
 * rand() order: lambda, then the acceptance draw, always: the reference's
 * test `(proba > 1) .or. (rand() < proba)` (MC.f90:804) leaves it to the
 * compiler whether rand() runs when proba > 1 (Fortran does not fix the
 * evaluation of .or. operands); the reference as compiled here evaluates
 * both operands, and its own run_mc stream (tests/golden/runmc_*.npz)
 * consumes the draw after every volume step. */
typedef struct {
    double L, xclb, kext, minus_two_kext, dr;
    double *xx, *yy, *zz, *S, *uu, *Fx, *Fy, *Fz, *beta, *lj6, *lj12, *sigma6;
    double uu_ew, uu_lj6, uu_lj12, uu_ew_intra, uu_lj6_intra, uu_lj12_intra, e_kc, e_klj, mu[3], tail_c, tail6,
        tail12;
} yo_saved;

static double *dupd(const double *a, size_t n) {
    double *b = malloc(sizeof(double) * (n ? n : 1));
    if (n) memcpy(b, a, sizeof(double) * n);
    return b;
}

static void save_state(const yo_sys *s, yo_saved *v) {
    size_t N = s->natom, T2 = (size_t)s->ntype * s->ntype, nS = (size_t)(1 + s->ntype) * 8 * s->nk;
    v->L = s->L; v->xclb = s->xclb; v->kext = s->kext; v->minus_two_kext = s->minus_two_kext; v->dr = s->dr;
    v->xx = dupd(s->xx, N); v->yy = dupd(s->yy, N); v->zz = dupd(s->zz, N);
    v->S = dupd(s->S, nS); v->uu = dupd(s->uu, N); v->Fx = dupd(s->Fx, N); v->Fy = dupd(s->Fy, N);
    v->Fz = dupd(s->Fz, N); v->beta = dupd(s->beta, s->nk);
    v->lj6 = dupd(s->lj6, T2); v->lj12 = dupd(s->lj12, T2); v->sigma6 = dupd(s->sigma6, T2);
    v->uu_ew = s->uu_ew; v->uu_lj6 = s->uu_lj6; v->uu_lj12 = s->uu_lj12;
    v->uu_ew_intra = s->uu_ew_intra; v->uu_lj6_intra = s->uu_lj6_intra; v->uu_lj12_intra = s->uu_lj12_intra;
    v->e_kc = s->e_kc; v->e_klj = s->e_klj;
    for (int d = 0; d < 3; ++d) v->mu[d] = s->mu[d];
    v->tail_c = s->tail_c; v->tail6 = s->tail6; v->tail12 = s->tail12;
}

static void restore_free(yo_sys *s, yo_saved *v, int restore) {
    size_t N = s->natom, T2 = (size_t)s->ntype * s->ntype, nS = (size_t)(1 + s->ntype) * 8 * s->nk;
    if (restore) {
        s->L = v->L; s->xclb = v->xclb; s->kext = v->kext; s->minus_two_kext = v->minus_two_kext; s->dr = v->dr;
        memcpy(s->xx, v->xx, sizeof(double) * N); memcpy(s->yy, v->yy, sizeof(double) * N);
        memcpy(s->zz, v->zz, sizeof(double) * N); memcpy(s->S, v->S, sizeof(double) * nS);
        memcpy(s->uu, v->uu, sizeof(double) * N); memcpy(s->Fx, v->Fx, sizeof(double) * N);
        memcpy(s->Fy, v->Fy, sizeof(double) * N); memcpy(s->Fz, v->Fz, sizeof(double) * N);
        memcpy(s->beta, v->beta, sizeof(double) * s->nk);
        if (T2) {
            memcpy(s->lj6, v->lj6, sizeof(double) * T2); memcpy(s->lj12, v->lj12, sizeof(double) * T2);
            memcpy(s->sigma6, v->sigma6, sizeof(double) * T2);
        }
        init_beta_lj(s);
        s->uu_ew = v->uu_ew; s->uu_lj6 = v->uu_lj6; s->uu_lj12 = v->uu_lj12;
        s->uu_ew_intra = v->uu_ew_intra; s->uu_lj6_intra = v->uu_lj6_intra; s->uu_lj12_intra = v->uu_lj12_intra;
        s->e_kc = v->e_kc; s->e_klj = v->e_klj;
        for (int d = 0; d < 3; ++d) s->mu[d] = v->mu[d];
        s->tail_c = v->tail_c; s->tail6 = v->tail6; s->tail12 = v->tail12;
    }
    double *p[] = {v->xx, v->yy, v->zz, v->S, v->uu, v->Fx, v->Fy, v->Fz, v->beta, v->lj6, v->lj12, v->sigma6};
    for (size_t i = 0; i < sizeof p / sizeof p[0]; ++i) free(p[i]);
}

int yo_volume_move(yo_sys *s, double pressure, double log_max_volume_scaling, double *new_box_length,
                   int *overlap_out) {
    const double L_old = s->L, V_old = L_old * L_old * L_old;
    double lambda = 2.0 * yo_rand(s) - 1.0;
    double V_new = exp(log(V_old) + lambda * log_max_volume_scaling);
    double L_new = pow(V_new, 1.0 / 3.0);
    double e[10];
    yo_mc_energies(s, e);
    double U_old = e[0];
    yo_saved v;
    save_state(s, &v);
    /* scale_box: COMs stay, intramolecular offsets shrink or grow */
    double sc = L_old / L_new;
    for (int m = 0; m < s->nmol; ++m) {
        int f = s->first[m], n = s->last[m] - f + 1;
        if (n == 1) continue;
        double c[3];
        center_of_mass(s->xx + f, s->yy + f, s->zz + f, s->mass + f, n, c);
        for (int j = f; j < f + n; ++j) {
            s->xx[j] = (s->xx[j] - c[0]) * sc + c[0];
            s->yy[j] = (s->yy[j] - c[1]) * sc + c[1];
            s->zz[j] = (s->zz[j] - c[2]) * sc + c[2];
        }
    }
    /* Parameters_recalc(L_new) */
    double xclb_old = s->xclb, xclb_new = (xclb_old * L_old) / L_new;
    s->L = L_new;
    s->xclb = xclb_new;
    s->kext = s->kext * (xclb_new / xclb_old);
    s->minus_two_kext = -2 * s->kext;
    s->dr = (s->dr * L_old) / L_new;
    /* LJTypes_scale_sigma + initBetaLJ */
    double sc2 = sc * sc, sc6 = sc2 * sc2 * sc2;
    for (int t = 0; t < s->ntype * s->ntype; ++t) {
        s->sigma6[t] = s->sigma6[t] * sc6;
        s->lj6[t] = s->lj6[t] * sc6;
        s->lj12[t] = s->lj12[t] * sc6 * sc6;
    }
    init_beta_lj(s);
    /* mc_calc_total_sums(scale_beta = xclb_new / xclb_old) */
    double comps[6];
    int ov = rspace_calc(s, -1, s->xx, s->yy, s->zz, comps, s->uu, s->Fx, s->Fy, s->Fz, 0, 0, 0, 0);
    s->uu_ew = comps[0]; s->uu_lj6 = comps[1]; s->uu_lj12 = comps[2];
    s->uu_ew_intra = comps[3]; s->uu_lj6_intra = comps[4]; s->uu_lj12_intra = comps[5];
    yo_kspace_init(s, xclb_new / xclb_old);
    yo_mu(s, s->mu);
    yo_tails(s, &s->tail_c, &s->tail6, &s->tail12);
    double proba = -1.0;
    if (!ov) {
        yo_mc_energies(s, e);
        double dU = e[0] - U_old, dV = V_new - V_old;
        double arg = -dU - pressure * dV + (double)(s->nmol + 1) * log(V_new / V_old);
        proba = exp(arg);
    }
    const double xksi = yo_rand(s);
    int accepted = proba > 1.0 || xksi < proba;
    restore_free(s, &v, !accepted);
    if (new_box_length) *new_box_length = s->L;
    if (overlap_out) *overlap_out = ov;
    return accepted;
}

/* runmc.f90:204-268 with volume steps: per cycle one draw decides between
 * mcvol (rnd > pressure_step_multiplier, psm > 0) and nmol single moves. */
int yo_run_npt(yo_sys *s, int ncycles, double pressure_step_multiplier, double pressure,
               double log_max_volume_scaling, int64_t *n_accepted, int64_t *n_vol_steps, int64_t *n_vol_accepted) {
    int64_t acc = 0, nv = 0, nva = 0;
    for (int c = 0; c < ncycles; ++c) {
        double rnd = yo_rand(s) * (pressure_step_multiplier + 1);
        if (pressure_step_multiplier > 0 && rnd > pressure_step_multiplier) {
            ++nv;
            nva += yo_volume_move(s, pressure, log_max_volume_scaling, NULL, NULL);
        } else {
            for (int j = 0; j < s->nmol; ++j) acc += run_step(s, NULL);
        }
    }
    if (n_accepted) *n_accepted = acc;
    if (n_vol_steps) *n_vol_steps = nv;
    if (n_vol_accepted) *n_vol_accepted = nva;
    return 0;
}

double yo_box_length(const yo_sys *s) { return s->L; }

/* mc_calc_forces.f90:230-316, with the movemc real-space convention (j > i) */
void yo_forces_refconv(const yo_sys *s, double *fx, double *fy, double *fz) {
    double mu[3];
    yo_mu(s, mu);
    size_t n8 = 8 * (size_t)s->nk;
    double *krc = malloc(sizeof(double) * n8), *krl = malloc(sizeof(double) * n8);
    double *tmp = calloc((size_t)(1 + s->ntype) * n8, sizeof(double));
    for (int i = 0; i < s->natom; ++i) {
        int it = s->type[i];
        double fc[3], fl[3] = {0, 0, 0};
        add_atom(s, tmp, NULL, s->xx[i], s->yy[i], s->zz[i], s->charge[i], it, krc, krl);
        force_coulomb(s, s->S, krc, fc);
        if (it > 0) force_lj(s, s->S, it, krl, fl);
        fx[i] = s->Fx[i] + fc[0] + fl[0];
        fy[i] = s->Fy[i] + fc[1] + fl[1];
        fz[i] = s->Fz[i] + fc[2] + fl[2];
        fx[i] = fx[i] + s->minus_two_kext * s->charge[i] * mu[0];
        fy[i] = fy[i] + s->minus_two_kext * s->charge[i] * mu[1];
        fz[i] = fz[i] + s->minus_two_kext * s->charge[i] * mu[2];
    }
    free(krc); free(krl); free(tmp);
}

/* ------------------------------------------------------------------------ */
yo_sys *yo_create(const yo_desc *d) {
    if (d->kmax < 1 || d->kmax > 60 || d->natom < 1 || d->nmol < 1 || d->ntype < 0) return NULL;
    yo_sys *s = calloc(1, sizeof *s);
    int N = d->natom, T = d->ntype;
    s->natom = N; s->nmol = d->nmol; s->ntype = T; s->kmax = d->kmax;
    s->L = d->box_length; s->alpha = d->alpha; s->rmax2 = d->rmax2; s->xclb = d->xclb; s->kext = d->kext;
    s->minus_two_kext = -2 * d->kext; /* parameters.f90:404 */
    s->dr = d->dr; s->d_angle = d->d_angle; s->lambda_f = d->lambda_f; s->lambda_c = d->lambda_c;
    s->alpha_over_sqrt_pi = d->alpha / sqrt(YO_PI); /* parameters.f90:400-401 */
    s->two_alpha_over_sqrt_pi = 2.0 * s->alpha_over_sqrt_pi;
#define DUPD(dst, src, n) do { dst = malloc(sizeof(double) * (size_t)(n) + 8); memcpy(dst, src, sizeof(double) * (size_t)(n)); } while (0)
#define DUPI(dst, src, n) do { dst = malloc(sizeof(int32_t) * (size_t)(n) + 8); memcpy(dst, src, sizeof(int32_t) * (size_t)(n)); } while (0)
    DUPD(s->charge, d->charge, N); DUPD(s->hard_core, d->hard_core_angstr, N); DUPD(s->mass, d->mass, N);
    DUPI(s->molnum, d->molnum, N); DUPI(s->type, d->lj_type, N);
    DUPI(s->first, d->mol_first, d->nmol); DUPI(s->last, d->mol_last, d->nmol);
    int T2 = T * T > 0 ? T * T : 1;
    DUPD(s->lj6, d->lj6, T * T); DUPD(s->lj12, d->lj12, T * T);
    DUPD(s->four_eps, d->four_eps, T * T); DUPD(s->sigma6, d->sigma6, T * T);
    DUPI(s->tail_counts, d->tail_counts, T);
    (void)T2;
    s->xx = calloc(N, sizeof(double)); s->yy = calloc(N, sizeof(double)); s->zz = calloc(N, sizeof(double));
    s->uu = calloc(N, sizeof(double)); s->Fx = calloc(N, sizeof(double));
    s->Fy = calloc(N, sizeof(double)); s->Fz = calloc(N, sizeof(double));
    /* grid */
    int nk = yo_grid_nk(d->kmax);
    s->nk = nk; s->g.nk = nk; s->g.kmax = d->kmax;
    s->g.kx = malloc(sizeof(int32_t) * nk); s->g.ky = malloc(sizeof(int32_t) * nk);
    s->g.kz = malloc(sizeof(int32_t) * nk); s->g.ip = malloc(sizeof(int32_t) * nk);
    s->g.iq = malloc(sizeof(int32_t) * nk);
    s->g.xk = malloc(sizeof(double) * nk); s->g.xk2 = malloc(sizeof(double) * nk);
    grid_walk(d->kmax, s->g.kx, s->g.ky, s->g.kz, s->g.ip, s->g.iq, s->g.xk, s->g.xk2);
    s->beta = calloc(nk, sizeof(double)); s->beta6 = calloc(nk, sizeof(double));
    s->beta12 = calloc(nk, sizeof(double));
    s->beta_lj = calloc((size_t)(T * T > 0 ? T * T : 1) * nk, sizeof(double));
    size_t nS = (size_t)(1 + T) * 8 * nk;
    s->S = calloc(nS, sizeof(double));
    s->mol_old = calloc(nS, sizeof(double)); s->mol_new = calloc(nS, sizeof(double));
    s->tot_new = calloc(nS, sizeof(double));
    int mx = 1;
    for (int m = 0; m < d->nmol; ++m) {
        int n = d->mol_last[m] - d->mol_first[m] + 1;
        if (n > mx) mx = n;
    }
    if (mx > 8) { yo_destroy(s); return NULL; }
    s->max_mol_atom = mx;
    s->krc = calloc((size_t)mx * 8 * nk, sizeof(double));
    s->krl = calloc((size_t)mx * 8 * nk, sizeof(double));
    s->back_new = calloc(4 * (size_t)N, sizeof(double));
    s->back_old = calloc(4 * (size_t)N, sizeof(double));
    s->freq_beg = malloc(sizeof(double) * d->nmol); s->freq_end = malloc(sizeof(double) * d->nmol);
    s->xfreq_beg = malloc(sizeof(double) * d->nmol); s->xfreq_end = malloc(sizeof(double) * d->nmol);
    yo_set_frequencies(s, NULL, NULL); /* freq_file 'none' (runmc.f90:41-45) */
    yo_rng_philox(s, 0, 0, 0);
    return s;
}

void yo_destroy(yo_sys *s) {
    if (!s) return;
    void *p[] = {s->charge, s->hard_core, s->mass, s->molnum, s->type, s->first, s->last, s->lj6, s->lj12,
                 s->four_eps, s->sigma6, s->tail_counts, s->xx, s->yy, s->zz, s->uu, s->Fx, s->Fy, s->Fz,
                 s->g.kx, s->g.ky, s->g.kz, s->g.ip, s->g.iq, s->g.xk, s->g.xk2, s->beta, s->beta6,
                 s->beta12, s->beta_lj, s->S, s->mol_old, s->mol_new, s->tot_new, s->krc, s->krl,
                 s->back_new, s->back_old, s->freq_beg, s->freq_end, s->xfreq_beg, s->xfreq_end, s->ulist};
    for (size_t i = 0; i < sizeof p / sizeof p[0]; ++i) free(p[i]);
    free(s);
}

int yo_set_positions(yo_sys *s, const double *x, const double *y, const double *z) {
    memcpy(s->xx, x, sizeof(double) * s->natom);
    memcpy(s->yy, y, sizeof(double) * s->natom);
    memcpy(s->zz, z, sizeof(double) * s->natom);
    return 0;
}

int yo_get_positions(const yo_sys *s, double *x, double *y, double *z) {
    memcpy(x, s->xx, sizeof(double) * s->natom);
    memcpy(y, s->yy, sizeof(double) * s->natom);
    memcpy(z, s->zz, sizeof(double) * s->natom);
    return 0;
}

int yo_nk(const yo_sys *s) { return s->nk; }
