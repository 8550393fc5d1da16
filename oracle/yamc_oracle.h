/*
 * yamc_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the YAMC Ewald Monte Carlo energy/force path, used
 * as the parity checker for the HIP library and as the CPU baseline.  Only
 * tests, the smoke check and the benchmark's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * Every function follows one reference routine, cited as file:line into the
 * YAMC sources (flat directory).
 */
#ifndef YAMC_ORACLE_H
#define YAMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field meaning as yamc_system_desc in include/yamc_ewald.h (declared
 * separately so the oracle never depends on product headers). */
typedef struct yo_desc {
    int32_t natom, nmol, ntype, kmax;
    int32_t precision, pad;
    double box_length, alpha, rmax2, xclb, kext, dr, d_angle, lambda_f, lambda_c;
    const double *charge, *hard_core_angstr, *mass;
    const int32_t *molnum, *lj_type, *mol_first, *mol_last;
    const double *lj6, *lj12, *four_eps, *sigma6;
    const int32_t *tail_counts;
} yo_desc;

typedef struct yo_sys yo_sys;

/* Per-step record produced by yo_movemc (MC.f90:831-1322) or yo_xchange
 * (MC.f90:1325-1954; imol2 = second molecule, -1 for a single move). */
typedef struct yo_move_rec {
    int32_t imol, accepted, overlap, imol2;
    double du_real, du_kc, du_klj, du_ext, du_total;
    double f_old[3], t_old[3], f_new[3], t_new[3];
    double delta[3];
    double angle;
} yo_move_rec;

/* RNG modes: Philox-4x32-10 stream, one constant uniform per move
 * (random.f90:80-82 RND_CONSTANT, set per move by the caller), or the replay
 * of a recorded list of uniforms (the reference's own rand() stream). */
enum { YO_RNG_PHILOX = 0, YO_RNG_CONST = 1, YO_RNG_LIST = 2 };

yo_sys *yo_create(const yo_desc *d);
void yo_destroy(yo_sys *s);
int yo_set_positions(yo_sys *s, const double *x, const double *y, const double *z);
int yo_get_positions(const yo_sys *s, double *x, double *y, double *z);

/* FourierGrid.f90:37-189 */
int yo_grid_nk(int kmax);
int yo_grid_fill(int kmax, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq);
int yo_nk(const yo_sys *s);

/* Functions.f90 */
double yo_erfc_luc(double x, double x2, double ee2);
double yo_erf_luc(double x, double x2, double ee2);
void yo_erfc_bet6_bet12(double x, double x2, double e2, double *erfk, double *bet6, double *bet12);
double yo_shi_sans_exp(double x);

/* SumSinCosKR.f90:181-455 for one atom: out = 8*nk (cos_ppp, sin_ppp,
 * cos_ppm, sin_ppm, cos_pmp, sin_pmp, cos_pmm, sin_pmm). */
void yo_sincos_kr(const yo_sys *s, double x, double y, double z, double charge, double *out);

/* EwaldSumKSpace.f90:197-260: beta (unless scale_beta > 1e-9) + full sums. */
void yo_kspace_init(yo_sys *s, double scale_beta);
void yo_get_beta(const yo_sys *s, double *beta, double *beta6, double *beta12, double *beta_lj);
void yo_get_sums(const yo_sys *s, double *sums);
void yo_kspace_energy(const yo_sys *s, double *e_c, double *e_lj);
/* EwaldSumKSpace.f90:528-753 for molecule imol moved to (x,y,z). */
void yo_kspace_delta(const yo_sys *s, int imol, const double *x, const double *y, const double *z,
                     double *du_c, double *du_lj);
/* ForceKSpace.f90:59-302: k-space force on n atoms (indices first..first+n-1
 * for charge/type) at given positions, against the current total sums. */
void yo_kspace_forces(const yo_sys *s, int first, int n, const double *x, const double *y,
                      const double *z, double *fx, double *fy, double *fz);

/* EwaldSumRealSpace.f90:233-500.  Full: comps[6] = uu_ew, uu_lj6, uu_lj12
 * and the three _intra terms; uu/F over natom. */
int yo_rspace_full(yo_sys *s, double *comps, double *uu, double *fx, double *fy, double *fz);
/* Partial sum of molecule imol at (x,y,z); uu/F over the molecule's atoms,
 * *_back over all atoms. Returns overlap flag. */
int yo_rspace_partial(const yo_sys *s, int imol, const double *x, const double *y, const double *z,
                      double *comps, double *uu, double *fx, double *fy, double *fz,
                      double *uu_back, double *fxb, double *fyb, double *fzb);

/* EwaldSumTails.f90:18-127 */
void yo_tails(const yo_sys *s, double *self_c, double *lj6, double *lj12);
/* EwaldSumExternal.f90:165-183 */
void yo_mu(const yo_sys *s, double *mu);

/* MC.f90:477-585 + 681-698: full recompute; e[10] = U_total, U_total_ew,
 * U_total_lj, uu_ew, uu_lj6, uu_lj12, E_kC, E_kLJ, U_tail, U_ext. */
int yo_mc_init(yo_sys *s, double *e);
void yo_mc_energies(yo_sys *s, double *e);
void yo_get_real_arrays(const yo_sys *s, double *uu, double *fx, double *fy, double *fz);

/* RNG control */
void yo_rng_philox(yo_sys *s, uint64_t seed, uint64_t stream, uint64_t counter);
uint64_t yo_rng_counter(const yo_sys *s);
void yo_rng_const(yo_sys *s, double u);
void yo_rng_list(yo_sys *s, const double *u, uint64_t n);
double yo_rand(yo_sys *s);
void yo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* MC.f90:831-1322 (one trial move of molecule imol, 0-based). */
int yo_movemc(yo_sys *s, int imol, yo_move_rec *rec);
/* runmc.f90:26-66: per-molecule move and exchange weights (NULL = ones /
 * zeros); xchange_probability = sum(xchange) / (sum(move) + sum(xchange)). */
int yo_set_frequencies(yo_sys *s, const double *mv_num, const double *xchange_num);
/* MC.f90:1325-1954 mc_xchange of molecules i1 != i2 (0-based). */
int yo_xchange(yo_sys *s, int i1, int i2, yo_move_rec *rec);
/* runmc.f90:204-268 without volume steps: per cycle one NPT draw, then nmol
 * steps of (exchange draw, then mc_xchange or molecule pick + movemc).
 * recs may be NULL. */
int yo_run_cycles(yo_sys *s, int ncycles, yo_move_rec *recs, int64_t *n_accepted);
/* Same move loop with a caller-given molecule per move and a constant
 * uniform per move (for traces against the compiled reference). */
int yo_run_const(yo_sys *s, int nmoves, const int32_t *imol, const double *u, yo_move_rec *recs);

/* Reference-convention total force (mc_calc_forces.f90:230-316 with the
 * movemc real-space convention j>i, quirk Q1 avoided): real + k + external. */
void yo_forces_refconv(const yo_sys *s, double *fx, double *fy, double *fz);
/* MC.f90:701-825 mcvol (NPT volume step); pressure in kT/A^3.  Returns 1 if
 * accepted; *new_box_length = box length after the step. */
int yo_volume_move(yo_sys *s, double pressure, double log_max_volume_scaling, double *new_box_length,
                   int *overlap);
/* runmc.f90:204-268 with volume steps. */
int yo_run_npt(yo_sys *s, int ncycles, double pressure_step_multiplier, double pressure,
               double log_max_volume_scaling, int64_t *n_accepted, int64_t *n_vol_steps, int64_t *n_vol_accepted);
double yo_box_length(const yo_sys *s);

#ifdef __cplusplus
}
#endif
#endif
