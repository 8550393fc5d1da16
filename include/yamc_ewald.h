/*
 * yamc_ewald.h -- C ABI of the MI355X Ewald Monte Carlo library
 * (libyamc_ewald.so, HIP kernels for gfx950 inside).
 *
 * The ABI takes plain pointers and sizes only.  Coordinates are in box units
 * (box = [-0.5, 0.5)^3), energies in kT, charges in e, exactly as in the YAMC
 * Fortran modules it replaces.  Every entry point returns a status:
 *   0 = ok, 1 = ERROR_IO, 2 = ERROR_LIMITS, 3 = ERROR_PARAMETER,
 *   4 = ERROR_INITIALIZATION, 5 = ERROR_WRONG_FUNCTION (error.f90:47-51),
 *   6 = YAMC_ERROR_DEVICE (HIP runtime failure or no usable device).
 * A real-space overlap is not an error: it is reported through an int flag,
 * as the logical `overlap` result of EwaldSumRealSpace_calc.
 *
 * Threading: a context is used by one host thread at a time; every call is
 * ordered on the context's own HIP stream and returns after that stream has
 * finished the call's work.
 */
#ifndef YAMC_EWALD_H
#define YAMC_EWALD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    YAMC_OK = 0,
    YAMC_ERROR_IO = 1,
    YAMC_ERROR_LIMITS = 2,
    YAMC_ERROR_PARAMETER = 3,
    YAMC_ERROR_INITIALIZATION = 4,
    YAMC_ERROR_WRONG_FUNCTION = 5,
    YAMC_ERROR_DEVICE = 6
};

/* Arithmetic of the state sums S(k) and the per-move k-space sweep. */
enum {
    YAMC_FP64 = 0,  /* plain FP64 everywhere */
    YAMC_FP32C = 1  /* S(k) kept as FP32 (hi, lo) two-sum pairs; sweeps read hi
                       (LJ sets) or hi + lo (Coulomb set); FP64 arithmetic,
                       acceptance, constants and periodic resync */
};

/* Limits of this build. */
#define YAMC_MAX_MOL_ATOM 4   /* atoms per molecule (SPC/E: 3, ions: 1) */
#define YAMC_MAX_KMAX 40      /* largest integer |k| cutoff */
#define YAMC_MAX_LJ_TYPE 20   /* LJTypes.f90:20 MAX_LJ_TYPE */

/*
 * System description.  The fields are the values the reference modules read
 * from `parameters` (parameters.f90:355-413), `TAtomicData`
 * (AtomicData.f90:9-22), `TLJTypes` (LJTypes.f90:31-63), `TMoleculeTable`
 * (MoleculeTable.f90:9-19) and `TComposition` (composition.f90:10-17).
 * Arrays are copied at creation and never retained.
 */
typedef struct yamc_system_desc {
    int32_t natom;          /* TAtomicData%natom */
    int32_t nmol;           /* TMoleculeTable%nmol */
    int32_t ntype;          /* TLJTypes%NType */
    int32_t kmax;           /* parameters: kmax */
    int32_t precision;      /* YAMC_FP64 or YAMC_FP32C */
    int32_t pad;
    double box_length;      /* parameters: BoxLength (Angstrom) */
    double alpha;           /* parameters: alpha (1/box length) */
    double rmax2;           /* parameters: rmax2 (box units^2) */
    double xclb;            /* parameters: xclb (Bjerrum length / L) */
    double kext;            /* parameters: kext = 2 pi xclb / (2 eps_ext + 1) */
    double dr;              /* parameters: dr (max displacement, box units) */
    double d_angle;         /* parameters: d_angle (radians) */
    double lambda_f;        /* parameters: xlambda_f */
    double lambda_c;        /* parameters: xlambda_c */
    const double *charge;            /* [natom] TAtomicData%charge */
    const double *hard_core_angstr;  /* [natom] TAtomicData%hard_core_angstr */
    const double *mass;              /* [natom] TMolecule%mass per atom */
    const int32_t *molnum;           /* [natom] molnum_by_atomnum, 1-based */
    const int32_t *lj_type;          /* [natom] type_by_index, 0 = no LJ */
    const int32_t *mol_first;        /* [nmol] first_atom, 0-based */
    const int32_t *mol_last;         /* [nmol] last_atom, 0-based inclusive */
    const double *lj6;               /* [ntype*ntype] LJ6Tab, (t1-1)*ntype+(t2-1) */
    const double *lj12;              /* [ntype*ntype] LJ12Tab */
    const double *four_eps;          /* [ntype*ntype] four_epsilon_tab */
    const double *sigma6;            /* [ntype*ntype] sigma6_tab */
    const int32_t *tail_counts;      /* [ntype] comp%mol_numbers(t) (EwaldSumTails.f90:88) */
} yamc_system_desc;

typedef struct yamc_ctx yamc_ctx;

/* Per-step record: the quantities movemc computes (MC.f90:831-1322), or
 * for an exchange step (mc_xchange, MC.f90:1325-1954) the two molecules
 * (imol, imol2), the decision, overlap, the dU terms and delta = the
 * centre-of-mass shift of imol.  imol2 = -1 for a single-molecule move. */
typedef struct yamc_move_rec {
    int32_t imol, accepted, overlap, imol2;
    double du_real, du_kc, du_klj, du_ext, du_total;
    double f_old[3], t_old[3], f_new[3], t_new[3];
    double delta[3];
    double angle;
} yamc_move_rec;

/* ---- life cycle -------------------------------------------------------- */
/* Replaces EwaldSumKSpace_alloc (EwaldSumKSpace.f90:88-149), FourierGrid_init
 * (FourierGrid.f90:118-189), EwaldSumRealSpace_alloc_full/_partial
 * (EwaldSumRealSpace.f90:70-131) and mc_alloc (MC.f90:257-379). */
int yamc_ctx_create(const yamc_system_desc *desc, yamc_ctx **out);
/* Replaces EwaldSumKSpace_dealloc (:151-194), EwaldSumRealSpace_dealloc_*
 * (:187-230), clearmc (MC.f90:604-679). */
int yamc_ctx_destroy(yamc_ctx *ctx);
/* Device index used by the context, number of k vectors (TFourierGrid%nk). */
int yamc_ctx_info(const yamc_ctx *ctx, int32_t *device, int32_t *nk);

/* Copies TAtomicData%xx,yy,zz in/out (box units). */
int yamc_set_positions(yamc_ctx *ctx, const double *xx, const double *yy, const double *zz);
int yamc_get_positions(yamc_ctx *ctx, double *xx, double *yy, double *zz);

/* FourierGrid_init (FourierGrid.f90:118-189): kx, ky, kz, ip, iq of nk. */
int yamc_grid(yamc_ctx *ctx, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq);

/* ---- k space ----------------------------------------------------------- */
/* EwaldSumKSpace_init (EwaldSumKSpace.f90:197-223): beta, beta6, beta12,
 * beta_LJ unless scale_beta > 1e-9 (then beta *= scale_beta), then
 * calc_total_energy (:225-260).  Outputs energy_coulomb and energy_LJ. */
int yamc_kspace_init(yamc_ctx *ctx, double scale_beta, double *e_coulomb, double *e_lj);
/* initBeta / initBetaLJ results (:263-386); beta_lj is [ntype*ntype*nk]. */
int yamc_kspace_get_beta(yamc_ctx *ctx, double *beta, double *beta6, double *beta12, double *beta_lj);
/* TRhoSquared sums (RhoSquared.f90:9-45) as [(1+ntype)][8][nk]: set 0 is
 * sumsincos_coulomb, set t is sumsincos_LJ(t); the 8 arrays are cos_ppp,
 * sin_ppp, cos_ppm, sin_ppm, cos_pmp, sin_pmp, cos_pmm, sin_pmm. */
int yamc_kspace_get_sums(yamc_ctx *ctx, double *sums);
/* calc_dU_coulomb + calc_dU_LJ (EwaldSumKSpace.f90:528-753) for molecule
 * imol (0-based) moved to (x,y,z) [n_m atoms]; the state is not changed. */
int yamc_kspace_delta(yamc_ctx *ctx, int32_t imol, const double *x, const double *y, const double *z,
                      double *du_coulomb, double *du_lj);
/* ForceKSpace_coulomb + ForceKSpace_LJ (ForceKSpace.f90:59-302) for the n
 * atoms first..first+n-1 placed at (x,y,z), against the current sums. */
int yamc_kspace_forces(yamc_ctx *ctx, int32_t first, int32_t n, const double *x, const double *y,
                       const double *z, double *fx, double *fy, double *fz);

/* ---- real space -------------------------------------------------------- */
/* EwaldSumRealSpace_calc, full sum (EwaldSumRealSpace.f90:233-500, j > i).
 * comps[6] = uu_ew, uu_lj6, uu_lj12, uu_ew_intra, uu_lj6_intra, uu_lj12_intra;
 * uu, fx, fy, fz over natom (may be NULL). */
int yamc_rspace_full(yamc_ctx *ctx, double *comps, double *uu, double *fx, double *fy, double *fz,
                     int32_t *overlap);
/* yamc_rspace_full with no allocation at all: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  Its
 * scratch is a static device array of 24 MiB that each GPU holds until the
 * process ends; systems whose full sum needs more (above about 780,000
 * atoms) get ERROR_LIMITS.  Calls on one GPU share that array: calls that
 * overlap in time on one GPU (two host threads, two contexts) race.  Same
 * results as yamc_rspace_full. */
int yamc_rspace_full_noalloc(yamc_ctx *ctx, double *comps, double *uu, double *fx, double *fy, double *fz,
                             int32_t *overlap);
/* EwaldSumRealSpace_set_molecule + _calc, partial sum of molecule imol at
 * (x,y,z): uu/F over its n_m atoms, *_back over all natom atoms (may be
 * NULL), comps[6] as above. */
int yamc_rspace_partial(yamc_ctx *ctx, int32_t imol, const double *x, const double *y, const double *z,
                        double *comps, double *uu, double *fx, double *fy, double *fz, double *uu_back,
                        double *fx_back, double *fy_back, double *fz_back, int32_t *overlap);

/* ---- constant terms ---------------------------------------------------- */
/* ewald_sum_coulomb_tail and ewald_sum_lj_tails (EwaldSumTails.f90:18-127). */
int yamc_tails(yamc_ctx *ctx, double *self_coulomb, double *lj6_tail, double *lj12_tail);
/* EwaldSumExternal_calc_mu (EwaldSumExternal.f90:165-183) over all atoms. */
int yamc_external_mu(yamc_ctx *ctx, double *mu3);

/* ---- Monte Carlo ------------------------------------------------------- */
/* mc_calc_total_sums + mc_total_energy (MC.f90:477-585, 681-698) on the
 * device.  e[10] = U_total, U_total_ew, U_total_lj, uu_ew, uu_lj6, uu_lj12,
 * E_kC, E_kLJ, U_tail, U_ext. */
int yamc_mc_init(yamc_ctx *ctx, double *e, int32_t *overlap);
/* Energies of the current (incrementally updated) state, same layout. */
int yamc_mc_energies(yamc_ctx *ctx, double *e);
/* Per-atom real-space energy and force arrays of the state
 * (RealSpace_TotalSum%uu, Fx, Fy, Fz). */
int yamc_mc_real_arrays(yamc_ctx *ctx, double *uu, double *fx, double *fy, double *fz);
/* RNG stream: Philox-4x32-10 keyed by seed, counter (high 64 bits) = stream,
 * word position = counter; 24-bit uniforms in (0,1) as random.f90 gives. */
int yamc_mc_set_rng(yamc_ctx *ctx, uint64_t seed, uint64_t stream, uint64_t counter);
int yamc_mc_get_rng(yamc_ctx *ctx, uint64_t *counter);
/* Replay mode: every uniform the context draws from now on (rand(),
 * random.f90:66-93, in the consumption order of run_mc, movemc, mcvol and
 * mc_xchange) is the next of the n recorded values u[] (each in (0,1)), the
 * counter being the position; a run that draws past the end fails with
 * ERROR_LIMITS.  For replaying a stream recorded from the reference's own
 * random_number; yamc_mc_set_rng returns to Philox.  n = 0 clears the list. */
int yamc_mc_set_rng_list(yamc_ctx *ctx, const double *u, int64_t n);
/* run_mc cycles (runmc.f90:204-268, no volume steps): per cycle one
 * pressure-step draw, then nmol steps of (exchange draw, then molecule pick
 * + movemc on the device, or an exchange step when exchange weights are set,
 * see yamc_mc_set_frequencies).  recs (may be NULL) receives ncycles*nmol
 * records.
 * resync_every > 0 rebuilds S(k) from scratch in FP64 after that many moves
 * (FP32C form; with exchange steps set, at the end of the first whole cycle
 * after that many moves); 0 = never. */
int yamc_mc_run(yamc_ctx *ctx, int32_t ncycles, int32_t resync_every, yamc_move_rec *recs,
                int64_t *n_accepted);
/* movemc with a caller-given molecule and one constant uniform per move
 * (random.f90:80-82 RND_CONSTANT semantics), for traces against the
 * reference. */
int yamc_mc_run_const(yamc_ctx *ctx, int32_t nmoves, const int32_t *imol, const double *u,
                      yamc_move_rec *recs, int64_t *n_accepted);
/* Rebuilds S(k) and the k-space energies from the current positions
 * (calc_total_energy, EwaldSumKSpace.f90:225-260) in FP64. */
int yamc_resync(yamc_ctx *ctx);

/* ---- exchange steps ---------------------------------------------------- */
/* init_freq (runmc.f90:26-66) with the per-molecule weights of a frequency
 * file (read_freq_file columns 2 and 3): move_weights[nmol] (NULL = all 1),
 * xchange_weights[nmol] (NULL = all 0).  The move picker and the exchange
 * picker take intervals weight / sum (BiasedRandom.f90:41-75; a zero weight
 * is never picked) and a step is an exchange with probability
 * sum(xchange) / (sum(move) + sum(xchange)).  Nonzero exchange weights must
 * cover at least two molecules.  yamc_mc_run then runs exchange steps where
 * the reference would; batches refuse contexts with exchange weights. */
int yamc_mc_set_frequencies(yamc_ctx *ctx, const double *move_weights, const double *xchange_weights);
/* mc_xchange (MC.f90:1325-1954) of molecules imol1 != imol2 (0-based): the
 * two swap centres of mass, keeping their orientations; accepted with
 * min(1, exp(-dU)) (MonteCarloMove.f90:572-597), no acceptance draw after an
 * overlap.  u (may be NULL) is a constant uniform (RND_CONSTANT), NULL draws
 * from the context's stream.  rec (may be NULL) receives the step's record
 * (imol, imol2, decision, overlap, dU terms). */
int yamc_mc_xchange(yamc_ctx *ctx, int32_t imol1, int32_t imol2, const double *u, int32_t *accepted,
                    yamc_move_rec *rec);

/* ---- volume steps (NPT) ------------------------------------------------ */
/* mcvol (MC.f90:701-825): V_new = exp(ln V_old + (2u-1) ln max_volume_scaling),
 * molecules keep their centres of mass while their atoms' offsets scale
 * (scale_box.f90:6-62), the box-unit parameters, hard cores, LJ tables
 * (LJTypes_scale_sigma) and beta_LJ follow the new box, the whole state is
 * recomputed (mc_calc_total_sums with scale_beta = xclb_new / xclb_old) and
 * the step is accepted with min(1, exp(-dU - P dV + (nmol+1) ln(V_new/V_old)))
 * (the acceptance draw is taken after every step, overlap included: the
 * reference's `(proba > 1) .or. (rand() < proba)`, MC.f90:804, as compiled
 * evaluates both operands); a rejected
 * step restores the old state.  pressure_kT_A3 = P / (k_B T) in kT per
 * cubic Angstrom (pressure_angstr_kT, MC.f90:721).  u (may be NULL) is a
 * constant uniform for both draws (RND_CONSTANT); NULL draws them from the
 * context's stream.  box_length (may be NULL) receives the box after the step
 * (Angstrom).  Needs yamc_mc_init first. */
int yamc_mc_volume_move(yamc_ctx *ctx, double pressure_kT_A3, double log_max_volume_scaling, const double *u,
                        int32_t *accepted, double *box_length);
/* run_mc cycles with volume steps (runmc.f90:204-268): per cycle one draw
 * rnd = u (psm + 1); psm > 0 and rnd > psm gives a volume step, otherwise a
 * cycle of nmol moves (yamc_mc_run).  Counts of accepted moves, volume
 * steps and accepted volume steps are returned (each may be NULL). */
int yamc_mc_run_npt(yamc_ctx *ctx, int32_t ncycles, double pressure_step_multiplier, double pressure_kT_A3,
                    double log_max_volume_scaling, int64_t *n_accepted, int64_t *n_vol_steps,
                    int64_t *n_vol_accepted);
/* Current box length in Angstrom (atomic_data%BoxLength). */
int yamc_box_length(yamc_ctx *ctx, double *box_length);
/* The box change of mcvol without its acceptance test (MC.f90:760-777):
 * scale_box when scale_positions != 0 (molecules keep their centres of mass),
 * Parameters_recalc(box_length), LJTypes_scale_sigma, initBetaLJ, then
 * mc_calc_total_sums with beta scaled by xclb_new / xclb_old.  *overlap
 * (may be NULL) is the recompute's overlap flag. */
int yamc_set_box(yamc_ctx *ctx, double box_length, int32_t scale_positions, int32_t *overlap);

/* ---- host-driven moves ------------------------------------------------- */
/* For hosts that keep their own move loop over the per-function surface
 * (yamc_kspace_delta, yamc_rspace_partial, their own acceptance test): the
 * commit of an accepted movemc step (MC.f90:1242-1320) for molecule imol
 * moved from its current place to (x, y, z): S(k) += the molecule's delta
 * (RhoSquared_add), per-atom real-space arrays from the old and new partial
 * sums' back arrays, the real-space components, E_kC += du_c and
 * E_kLJ += du_lj (as yamc_kspace_delta returned them), positions and mu.
 * If the new place overlaps, nothing is committed and *overlap = 1. */
int yamc_mc_commit_move(yamc_ctx *ctx, int32_t imol, const double *x, const double *y, const double *z,
                        double du_c, double du_lj, int32_t *overlap);

/* ---- batched replicas -------------------------------------------------- */
/* Advances n independent contexts (uneven sizes, compositions and
 * temperatures allowed) by ncycles run_mc cycles each (runmc.f90:204-268)
 * in one persistent launch; the launch shape is the first context's
 * (yamc_set_block_size).  All contexts must live on the same device and
 * have run yamc_mc_init.
 * resync_every > 0: each replica's ncycles * nmol moves run in segments of
 * resync_every moves with a from-scratch FP64 rebuild of its S(k) between
 * them (all replicas' rebuilds in one batched pass), exactly the segments
 * and rebuilds of yamc_mc_run(ctx, ncycles, resync_every) alone; 0 = none.
 * Members set to G > 1 workgroups per chain (yamc_set_chain_groups) make a
 * batch of grid chains: all members' workgroups (one chain over G of them
 * each, G = 1 allowed for the others) in one cooperative launch, which
 * must fit on the device at once (else YAMC_ERROR_LIMITS); there each
 * member's workgroups rebuild its S(k) together inside the launch (each
 * its own k range), and a member with G > 1 advances exactly as
 * yamc_mc_run of that context alone (the members' workgroup counts when
 * the batch is staged fix its launch).
 * n_accepted[n] (may be NULL) receives each replica's accepted moves. */
int yamc_batch_run(yamc_ctx *const *ctxs, int32_t n, int32_t ncycles, int32_t resync_every, int64_t *n_accepted);
/* The same sweep captured once in a HIP graph: yamc_batch_create stages the
 * replicas and captures one sweep of ncycles cycles (its segments and
 * rebuilds included), yamc_batch_launch replays it (each replay continues
 * every replica's chain and RNG stream, and first re-reads every member's
 * current box parameters; a member whose state is not ready is refused),
 * yamc_batch_destroy frees the graph (a batch of grid chains is staged the
 * same way and launched without a graph).  The contexts must outlive the
 * batch; no reference counterpart (the reference runs one system serially). */
typedef struct yamc_batch yamc_batch;
int yamc_batch_create(yamc_ctx *const *ctxs, int32_t n, int32_t ncycles, int32_t resync_every, yamc_batch **out);
int yamc_batch_launch(yamc_batch *batch, int64_t *n_accepted);
int yamc_batch_destroy(yamc_batch *batch);

/* ---- forces ------------------------------------------------------------ */
/* Full-system forces.  f_grad = -dE_total/dr for all atoms (real + reciprocal
 * + LJ-Ewald + external, intramolecular real-space pairs included), from the
 * hand-written adjoint; f_ref = the reference force convention
 * (mc_calc_forces.f90:230-316 with the movemc real-space pairs j > i).
 * Either pointer triple may be NULL. */
int yamc_forces(yamc_ctx *ctx, double *gx, double *gy, double *gz, double *rx, double *ry, double *rz);

/* ---- analysis ----------------------------------------------------------- */
/* Centre-of-mass radial distribution counts of the current configuration
 * (mcrdf2.f90:262-336 count_distances, one frame, cnt4 column): molecule
 * COMs (mass-weighted, wrapped into the box), every ordered pair i != j with
 * the single minimum-image shift, r < rmax (rmax <= L/2), bin floor(r/dr) of
 * the (min, max) species pair.  counts[nspecies][nspecies][nbins] is
 * accumulated (+=), so successive frames add up; mol_species[nmol] is
 * 0-based, nspecies <= 8, nbins <= 4096. */
int yamc_rdf(yamc_ctx *ctx, const int32_t *mol_species, int32_t nspecies, double dr_angstr, double rmax_angstr,
             int32_t nbins, uint64_t *counts);

/* ---- per-procedure surface (the Fortran drop-in modules) ---------------- */
/* One call per procedure of the reference's EwaldSumKSpace, SumSinCosKR,
 * RhoSquared, FourierGrid, ForceKSpace, EwaldSumRealSpace, RealSumLocal and
 * EwaldSumTails modules, on the caller's host arrays: each call copies its
 * inputs to the device, computes there and copies its results back before it
 * returns.  fortran/shim/ holds modules with the reference's names, types and
 * procedure signatures that forward to these calls, so code written against
 * the reference's modules (MC.f90) builds and runs unchanged on the device.
 * Arrays are k-ordered as the workspace's grid (FourierGrid_init order); a
 * set of sums is 8 arrays cos_ppp, sin_ppp, cos_ppm, sin_ppm, cos_pmp,
 * sin_pmp, cos_pmm, sin_pmm (TSumSinCosKR / TSinCosKR fields, passed as 8
 * pointers); beta_LJ is ntype*ntype row pointers, row (t1-1)*ntype+(t2-1).
 * A workspace is used by one host thread at a time. */
typedef struct yamc_ws yamc_ws;
/* FourierGrid_calcNalloc + FourierGrid_init (FourierGrid.f90:37-189) for kmax
 * and the workspace's device buffers. */
int yamc_ws_create(int32_t kmax, yamc_ws **out);
int yamc_ws_destroy(yamc_ws *ws);
/* TFourierGrid fields nk, kx, ky, kz, ip, iq, xk, xk2 (any may be NULL). */
int yamc_ws_grid(const yamc_ws *ws, int32_t *nk, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq,
                 double *xk, double *xk2);
/* EwaldSumKSpace_initBeta (EwaldSumKSpace.f90:263-350): beta, beta6, beta12
 * [nk] for alpha (1/box length) and xclb. */
int yamc_op_beta(yamc_ws *ws, double alpha, double xclb, double *beta, double *beta6, double *beta12);
/* One row of EwaldSumKSpace_initBetaLJ (:352-386): out = lj12 beta12 - lj6 beta6. */
int yamc_op_beta_lj(yamc_ws *ws, double lj6, double lj12, const double *beta6, const double *beta12, double *out);
/* SinCosXYZ_fill (SumSinCosKR.f90:181-223): tab6 = cos_xkx, cos_yky, cos_zkz,
 * sin_xkx, sin_yky, sin_zkz, each [kmax+1]. */
int yamc_op_sincos_xyz(yamc_ws *ws, double x, double y, double z, double *const *tab6);
/* SinCosKR_fill (SumSinCosKR.f90:297-455) from the tables tab6 with charge q. */
int yamc_op_sincos_kr(yamc_ws *ws, const double *const *tab6, double q, double *const *kr8);
/* SinCosXYZ_fill + SinCosKR_fill of one atom at (x, y, z) with charge q. */
int yamc_op_sincos_kr_atom(yamc_ws *ws, double x, double y, double z, double q, double *const *kr8);
/* The sums RhoSquared_addAtoms (RhoSquared.f90:188-243) adds for n atoms:
 * sums[(1+ntype)*8] row pointers (set 0 = Coulomb, set t = LJ type t) are
 * overwritten with the atoms' sums; present[ntype] (may be NULL) = whether an
 * atom of type t is among them. */
int yamc_op_sums(yamc_ws *ws, int32_t n, const double *x, const double *y, const double *z, const double *q,
                 const int32_t *lj_type, int32_t ntype, double *const *sums, int32_t *present);
/* RhoSquared_calc_rho_squared (RhoSquared.f90:245-273). */
int yamc_op_rho_squared(yamc_ws *ws, const double *const *s8, double *rho2);
/* EwaldSumKSpace_calc_coulomb_energy (EwaldSumKSpace.f90:389-416) from
 * rho_squared [nk], and _calc_LJ_energy (:418-526). */
int yamc_op_energy_coulomb(yamc_ws *ws, const double *beta, const double *rho2, double *e);
int yamc_op_energy_lj(yamc_ws *ws, int32_t ntype, const double *const *beta_lj, const double *const *sums_lj,
                      double *e);
/* EwaldSumKSpace_calc_dU_coulomb / _calc_dU_LJ (EwaldSumKSpace.f90:528-753):
 * s8 / sums_lj the total sums, d8 / delta_lj the delta sums, present[ntype]
 * the delta's type_is_present. */
int yamc_op_du_coulomb(yamc_ws *ws, const double *beta, const double *const *s8, const double *const *d8, double *du);
int yamc_op_du_lj(yamc_ws *ws, int32_t ntype, const double *const *beta_lj, const double *const *sums_lj,
                  const double *const *delta_lj, const int32_t *present, double *du);
/* ForceKSpace_coulomb / ForceKSpace_LJ (ForceKSpace.f90:59-302) of the atom
 * whose sin/cos are kr8; ityp 1-based.  f3 = fx, fy, fz. */
int yamc_op_force_coulomb(yamc_ws *ws, const double *beta, const double *const *s8, const double *const *kr8,
                          double *f3);
int yamc_op_force_lj(yamc_ws *ws, int32_t ntype, int32_t ityp, const double *const *beta_lj,
                     const double *const *sums_lj, const double *const *kr8, double *f3);
/* Inputs of EwaldSumRealSpace_calc (EwaldSumRealSpace.f90:233-500). */
typedef struct yamc_rs_desc {
    int32_t natom;            /* TAtomicData%natom */
    int32_t ntype;            /* TLJTypes%NType */
    int32_t full;             /* 1: full sum (TEwaldSumRealSpace%full_sum), 0: partial */
    int32_t nmol_atom;        /* partial: the molecule's atoms */
    int32_t first_atom;       /* partial: its first atom, 0-based */
    int32_t j0, j1;           /* the optional first, last (0-based); -1: the defaults */
    int32_t stop_on_overlap;  /* 1 unless dont_stop_on_overlap is .TRUE. */
    int32_t pad;
    double box_length, alpha, rmax2, xclb;  /* parameters: BoxLength, alpha, rmax2, xclb */
    const double *xx, *yy, *zz, *charge, *hard_core_angstr;  /* [natom] */
    const int32_t *molnum, *lj_type;                         /* [natom] */
    const double *sigma2, *four_eps, *sigma6;                /* [ntype*ntype] */
    const double *xmol, *ymol, *zmol;                        /* partial: the molecule's positions */
} yamc_rs_desc;
/* EwaldSumRealSpace_calc: uu/F over the rows (full: natom, partial:
 * nmol_atom), the back arrays over natom (full sum: pass the same arrays as
 * uu/F, they are added to), comps[6] = uu_ew, uu_lj6, uu_lj12 and their
 * intramolecular parts; with stop_on_overlap the sums stop at the first
 * overlapping pair in the reference's loop order, as it returns there.
 * *overlap = the number of overlapping pairs the reference's loop visits (0
 * or 1 with stop_on_overlap), i.e. the 'O' marks it prints.  Any of the
 * array outputs may be NULL. */
int yamc_op_rspace(yamc_ws *ws, const yamc_rs_desc *d, double *uu, double *fx, double *fy, double *fz,
                   double *uu_back, double *fx_back, double *fy_back, double *fz_back, double *comps,
                   int32_t *overlap);
/* RealSumLocal (RealSumLocal.f90:63-168) for one pair at r2 > 0: out[14] =
 * potew, potlj6, potlj12, pot, FR, then the TRealSumLocal fields r, a, a2,
 * exp_minus_a2, erfc_a, C6, C12, a6, sigma_over_r_6.  lj = 0 leaves the LJ
 * terms out (potlj6 = potlj12 = 0). */
int yamc_op_pair(yamc_ws *ws, double alpha, double xclb, double r2, int32_t same_molecule, double qi, double qj,
                 int32_t lj, double sigma6, double four_eps, double *out);
/* ewald_sum_coulomb_tail + ewald_sum_lj_tails (EwaldSumTails.f90:18-127). */
int yamc_op_tails(yamc_ws *ws, int32_t natom, const double *charge, int32_t ntype, const double *lj6,
                  const double *lj12, const int32_t *counts, double alpha, double xclb, double *self_coulomb,
                  double *lj6_tail, double *lj12_tail);

/* ---- launch shape ------------------------------------------------------ */
/* Launch shape of the chain kernel for this context (and for a batch whose
 * first context it is): 256 or 512 = one workgroup of that size per chain;
 * 64 or 128 = one wavefront per chain, four chains per workgroup, registers
 * sized for one (64) or two (128) waves per SIMD, i.e. 4 or 8 chains in
 * flight per CU (FP64 run_mc cycles; FP32C contexts and constant-uniform
 * traces run these shapes as workgroups of 256).  No
 * reference counterpart (the reference is serial); the autotuner
 * (yamc_amd.tune) measures them. */
int yamc_set_block_size(yamc_ctx *ctx, int32_t block);
/* Workgroups per chain for this context's single-chain runs (yamc_mc_run,
 * yamc_mc_run_const, exchange and NPT runs): 1 = one workgroup (default);
 * G > 1 = one chain spread over G co-resident workgroups of 256 threads (a
 * cooperative launch) that split every k sweep and S(k) commit by k range
 * and the real-space pass by atom range, with two grid barriers per move
 * (the loops of EwaldSumKSpace.f90:528-753 and EwaldSumRealSpace.f90:
 * 367-496 split across CUs).  Same decisions and results as G = 1 up to the
 * summation order of the partial sums.  In a batch, the member's chain
 * takes its G workgroups too (yamc_batch_run).  No reference counterpart
 * (the reference is serial). */
int yamc_set_chain_groups(yamc_ctx *ctx, int32_t groups);
/* Cell list of the chain kernels' real-space scan: cubic cells of width at
 * least rc / div box units (div = 1, 2, 3 or 4; nc = floor(div / rc) cells
 * per axis), or none (div = 0: the scan visits every atom).  Fewer than 3
 * cells per axis means none.  A move then tests only the atoms of the cells
 * within the scan radius of its sites (a sphere-pruned stencil), which
 * finds exactly the candidates of the full scan, so results do not depend
 * on div.  Default at creation: 1 when that gives 4 or more cells per axis,
 * else 0.  The reference has no cell list (EwaldSumRealSpace.f90:367 loops
 * over all atoms); the autotuner (yamc_amd.tune) measures div. */
int yamc_set_cell_div(yamc_ctx *ctx, int32_t div);

/* ---- diagnostics ------------------------------------------------------- */
/* Time of the last yamc_mc_run / yamc_batch_run kernel launch in ms (HIP
 * events on the context stream). */
int yamc_last_kernel_ms(yamc_ctx *ctx, double *ms);
/* Cumulative wall-clock microseconds spent by the chain kernel in its nine
 * phases (pick, mu + tables, old k sweep, proposal, real-space new sum, new
 * k sweep, acceptance, commit, finalize); us[9]. */
int yamc_phase_times(yamc_ctx *ctx, double *us);
const char *yamc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
