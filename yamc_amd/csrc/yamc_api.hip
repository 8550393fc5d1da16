// yamc_api.hip -- C ABI of libyamc_ewald.so (declared in include/yamc_ewald.h).
//
// A context owns one replica's device state on its own HIP stream.  Host
// arrays passed in are copied to the device and never retained; every entry
// point synchronises its stream before returning, so outputs are ready.
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "yamc_device.h"

// kernels (yamc_kernels.hip)
__global__ void yk_beta(YdChain c, const double *xk2, double *beta6, double *beta12);
__global__ void yk_beta_lj(YdChain c, const double *lj6, const double *lj12, const double *beta6,
                           const double *beta12);
__global__ void yk_scale_beta(YdChain c, double s);
__global__ void yk_atom_tables(YdChain c, const double *x, const double *y, const double *z, double *tab);
__global__ void yk_full_sums(YdChain c, const double *tab, int ngroups, double *part);
__global__ void yk_full_sums_reduce(YdChain c, const double *part, int ngroups);
__global__ void yk_kenergy(YdChain c);

__global__ void yk_mu(YdChain c, double *mu_out);
__global__ void yk_tails(YdChain c, const double *lj6, const double *lj12, const int *counts);
__global__ void yk_scale_box(YdChain c, double sc);
__global__ void yk_rspace_full(YdChain c, const double *lj_sig6, const double *lj_4eps, int ngroups, double *part,
                               double *pat, int *ov);
__global__ void yk_rspace_full_atoms(YdChain c, const double *pat, int ngroups);
__global__ void yk_rspace_full_final(YdChain c, const double *part, int nblk, double *comps_out);
template <bool F32, int BS>
__global__ void yk_chain(const YdChain *chains, const double *const *sig6s, const double *const *feps_s, int nsteps,
                         int mode, int max_moves, int lds_atoms, int resync_every);
template <bool F32, int BS>
__global__ void yk_gchain(const YdChain *chains, const YgGroup *groups, int nsteps, int mode, int max_moves,
                          int lds_atoms, int resync_every);
__global__ void yk_gchain_prep(const YdChain *src, YdChain *dst, YgGroup *grp, int G, double *rsb, unsigned int *bar,
                               double *part, const double *sig6, const double *feps);
__global__ void yk_gchain_bprep(const YdChain *src, YdChain *dst, const YgGroup *grp, int nwg, double *rsb);
#define YG_PART 64  // yk_gchain: doubles per workgroup per phase of the partials buffer
template <bool F32, int OCC>
__global__ void yk_wchain(const YdChain *chains, int nchains, const double *const *sig6s, const double *const *feps_s,
                          int ncycles);  // FP64 instantiations only
#define YW_WPB 4  // replicas (waves) per workgroup of yk_wchain
template <bool F32>
__global__ void yk_kspace_delta(YdChain c, int imol, const double *x, const double *y, const double *z,
                                double *out);
template <bool F32>
__global__ void yk_kspace_forces(YdChain c, int first, int n, const double *x, const double *y, const double *z,
                                 double *out);
__global__ void yk_rspace_partial(YdChain c, int imol, const double *x, const double *y, const double *z,
                                  const double *sig6, const double *feps, double *out, double *uub, double *fxb,
                                  double *fyb, double *fzb);
__global__ void yk_xchg_delta(YdChain c, YxAtoms at, double *D);
template <bool F32>
__global__ void yk_xchg_du(YdChain c, const double *D, unsigned pmask, double *out);
template <bool F32>
__global__ void yk_xchg_commit_sums(YdChain c, const double *D, unsigned pmask);
__global__ void yk_rspace_pair(YdChain c, int imol1, int imol2, const double *sig6, const double *feps,
                               double *out);
__global__ void yk_xchg_commit_rs(YdChain c, const double *bn1, const double *bo1, const double *bn2,
                                  const double *bo2, const double *mv, int f1, int n1, int f2, int n2);
#define YK_NRS (4 * YD_MAXA + 3)  // yk_rspace_partial output: uu, F, 3 components; overlap flag after
int yk_launch_forces(const YdChain &c, const double *sig6, const double *feps, const double *tab, double *g,
                     double *r, hipStream_t s);  // 0, or -1 when its scratch allocation failed
void yk_launch_rdf(const YdChain &c, double *com, const int *species, int nspec, double dr, double rmax2, int nbins,
                   unsigned long long *hist, hipStream_t s);

#ifndef YK_BS
#define YK_BS 512
#endif

static thread_local char g_err[512] = "";

#include <stdarg.h>
static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return fail(YAMC_ERROR_DEVICE, "HIP error %s (line %d)", hipGetErrorString(e_), __LINE__); \
    } while (0)

struct yamc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    YdChain ch{};
    int natom = 0, nmol = 0, ntype = 0, nk = 0, kmax = 0, precision = 0;
    std::vector<int> kx, ky, kz, ip, iq;
    // device buffers owned by the context
    std::vector<void *> owned;
    double *d_lj6 = nullptr, *d_lj12 = nullptr, *d_sig6 = nullptr, *d_4eps = nullptr;
    int *d_counts = nullptr;
    double *d_xk2 = nullptr, *d_beta6 = nullptr, *d_beta12 = nullptr, *d_tab = nullptr;
    int *d_ov = nullptr;
    double *d_io = nullptr;  // small staging buffer for single-call results (64 doubles)
    double *d_mol = nullptr; // staged molecule coordinates (3 * YD_MAXA)
    YdChain *d_chain = nullptr;
    const double **d_sig6s = nullptr, **d_feps = nullptr;
    bool beta_ready = false, state_ready = false;
    int chain_bs = 256;  // launch shape of the chain kernel (yamc_set_block_size)
    int cell_div = 0;    // cells per cut-off of the scan's cell list (yamc_set_cell_div; 0: none)
    // grid chain (yamc_set_chain_groups): workgroups per chain and their
    // buffers (per-workgroup descriptors, back terms, partial sums, barrier)
    int groups = 1, gcap = 0;
    YdChain *d_gch = nullptr;
    YgGroup *d_ggrp = nullptr;
    double *d_grsb = nullptr, *d_gpart = nullptr;
    unsigned int *d_gbar = nullptr;
    // host copies of the box-dependent inputs, scaled to the current box
    std::vector<double> hard_core_angstr, lj6, lj12, sigma6;
    // device copies of the state kept across a volume step (allocated on first use)
    std::vector<void *> vol_save;
    // host copies of the per-atom and per-molecule inputs read by exchange steps
    std::vector<double> h_q, h_mass;
    std::vector<int> h_type, h_first, h_last;
    // exchange steps: move and exchange weights (runmc.f90:26-66) and device
    // scratch (joint delta, four back arrays, the molecules' new values)
    std::vector<double> xfreq_beg, xfreq_end;
    double pxch = 0.0;
    double *d_xd = nullptr, *d_xb = nullptr, *d_xmv = nullptr;
    // replay mode: the recorded uniforms, device copy and host copy
    double *d_ulist = nullptr;
    std::vector<double> h_ulist;
};

template <typename T>
static int dalloc(yamc_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, n * sizeof(T) + 64));
    HIPCHK(hipMemsetAsync(q, 0, n * sizeof(T) + 64, c->stream));
    c->owned.push_back(q);
    *p = (T *)q;
    return YAMC_OK;
}

#define TRY(x)                      \
    do {                            \
        int r_ = (x);               \
        if (r_ != YAMC_OK) return r_; \
    } while (0)

// Cell list of the chain kernels' scan: cells of width >= rc / div box units
// (nc = floor(div / rc) per axis, at most 40), none when div = 0 or nc < 3.
// Replaces the context's cell arrays; the buckets are built by each launch.
static int setup_cells(yamc_ctx *c, int div) {
    YdChain &h = c->ch;
    for (void *p : {(void *)h.cell_cnt, (void *)h.cell_atoms, (void *)h.atom_cell, (void *)h.atom_slot}) {
        if (!p) continue;
        for (size_t i = 0; i < c->owned.size(); ++i)
            if (c->owned[i] == p) { c->owned.erase(c->owned.begin() + (long)i); break; }
        HIPCHK(hipFree(p));
    }
    h.cell_cnt = h.cell_atoms = h.atom_cell = h.atom_slot = nullptr;
    h.nc = 0;
    h.ccap = 0;
    c->cell_div = 0;
    const double rc = sqrt(h.p.rmax2);
    int nc = div > 0 && rc > 0 ? (int)floor((double)div / (rc * (1.0 + 1e-9))) : 0;
    if (nc > 40) nc = 40;
    if (nc < 3) return YAMC_OK;
    const int N = h.p.natom, ncell = nc * nc * nc;
    const int cap = ((2 * N) / ncell + 64 + 63) / 64 * 64;
    TRY(dalloc(c, &h.cell_cnt, (size_t)ncell));
    TRY(dalloc(c, &h.cell_atoms, (size_t)ncell * cap));
    TRY(dalloc(c, &h.atom_cell, (size_t)N));
    TRY(dalloc(c, &h.atom_slot, (size_t)N));
    h.nc = nc;
    h.ccap = cap;
    c->cell_div = div;
    return YAMC_OK;
}

// Stream-ordered device scratch, freed (stream-ordered) on every exit path.
struct StreamScratch {
    hipStream_t s;
    std::vector<void *> held;
    explicit StreamScratch(hipStream_t st) : s(st) {}
    ~StreamScratch() {
        for (void *q : held) (void)hipFreeAsync(q, s);
    }
    template <typename T>
    int get(T **out, size_t n) {
        void *q = nullptr;
        HIPCHK(hipMallocAsync(&q, sizeof(T) * (n ? n : 1), s));
        held.push_back(q);
        *out = (T *)q;
        return YAMC_OK;
    }
};

// FourierGrid.f90:118-189: half-space k list, row flags ip, zero pattern iq.
static void build_grid(int kmax, std::vector<int> &kx, std::vector<int> &ky, std::vector<int> &kz,
                       std::vector<int> &ip, std::vector<int> &iq, std::vector<double> &xk2) {
    double xkmax2 = (double)kmax * kmax;
    for (int ix = 0; ix <= kmax; ++ix) {
        double xkx2 = (double)(ix * ix);
        int kymax = (int)sqrt(xkmax2 - xkx2);
        for (int iy = 0; iy <= kymax; ++iy) {
            double xky2 = (double)(iy * iy);
            int kzmax = (int)sqrt(xkmax2 - xkx2 - xky2);
            for (int iz = 0; iz <= kzmax; ++iz) {
                double k2 = xkx2 + xky2 + (double)(iz * iz);
                if (k2 < 0.1) continue;
                kx.push_back(ix); ky.push_back(iy); kz.push_back(iz); xk2.push_back(k2);
                iq.push_back(iy == 0 ? (iz == 0 ? 1 : 3) : (iz == 0 ? 2 : 4));
                ip.push_back(0);
            }
            if (!ip.empty()) ip.back() = 1;
        }
        if (!ip.empty()) ip.back() = 2;
    }
    if (!ip.empty()) ip.back() = 3;
}

// BiasedRandom.f90:41-75 BiasedRandom_init: cumulative intervals of
// num / sum(num) (flat when the sum is 0); a zero weight gets end = -1 so
// that its interval is never chosen.
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

static int sync(yamc_ctx *c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return YAMC_OK;
}

static int get_state(yamc_ctx *c, YdState *st);
static int put_state(yamc_ctx *c, const YdState *st);

extern "C" const char *yamc_last_error(void) { return g_err; }

// The error of a failing call in another source file of the library.
int yamc_internal_fail(int code, const char *msg) { return fail(code, "%s", msg); }

extern "C" int yamc_ctx_create(const yamc_system_desc *d, yamc_ctx **out) {
    if (!d || !out) return fail(YAMC_ERROR_PARAMETER, "yamc_ctx_create: null argument");
    *out = nullptr;
    if (d->natom < 1 || d->nmol < 1 || d->ntype < 0 || d->kmax < 1)
        return fail(YAMC_ERROR_PARAMETER, "yamc_ctx_create: bad sizes");
    if (d->kmax > YAMC_MAX_KMAX) return fail(YAMC_ERROR_LIMITS, "kmax %d exceeds YAMC_MAX_KMAX", d->kmax);
    if (d->ntype > YAMC_MAX_LJ_TYPE) return fail(YAMC_ERROR_LIMITS, "ntype %d exceeds MAX_LJ_TYPE", d->ntype);
    if (d->precision != YAMC_FP64 && d->precision != YAMC_FP32C)
        return fail(YAMC_ERROR_PARAMETER, "unknown precision %d", d->precision);
    for (int m = 0; m < d->nmol; ++m) {
        int f = d->mol_first[m], l = d->mol_last[m];
        if (f < 0 || l < f || l >= d->natom) return fail(YAMC_ERROR_PARAMETER, "bad molecule range %d", m);
        if (l - f + 1 > YAMC_MAX_MOL_ATOM)
            return fail(YAMC_ERROR_LIMITS, "molecule %d has more than YAMC_MAX_MOL_ATOM atoms", m);
        int seen[YAMC_MAX_MOL_ATOM], ns = 0;
        for (int a = f; a <= l; ++a) {
            if (d->molnum[a] != m + 1) return fail(YAMC_ERROR_PARAMETER, "molnum mismatch at atom %d", a);
            int t = d->lj_type[a];
            if (t < 0 || t > d->ntype) return fail(YAMC_ERROR_PARAMETER, "bad lj_type at atom %d", a);
            bool dup = false;
            for (int s = 0; s < ns; ++s) dup |= seen[s] == t;
            if (t > 0 && !dup) seen[ns++] = t;
        }
        if (ns > 2) return fail(YAMC_ERROR_LIMITS, "molecule %d has more than 2 LJ types", m);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(YAMC_ERROR_DEVICE, "no HIP device available");
    yamc_ctx *c = new yamc_ctx();
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    int N = d->natom, T = d->ntype, M = d->nmol;
    c->natom = N; c->nmol = M; c->ntype = T; c->kmax = d->kmax; c->precision = d->precision;
    std::vector<double> xk2;
    build_grid(d->kmax, c->kx, c->ky, c->kz, c->ip, c->iq, xk2);
    int nk = (int)c->kx.size();
    c->nk = nk;
    YdParams &p = c->ch.p;
    p.natom = N; p.nmol = M; p.ntype = T; p.kmax = d->kmax; p.nk = nk; p.precision = d->precision;
    p.L = d->box_length; p.alpha = d->alpha; p.rmax2 = d->rmax2; p.xclb = d->xclb; p.kext = d->kext;
    p.m2kext = -2 * d->kext;  // parameters.f90:404
    p.dr = d->dr; p.d_angle = d->d_angle; p.lambda_f = d->lambda_f; p.lambda_c = d->lambda_c;
    p.alpha_over_sqrt_pi = d->alpha / sqrt(YD_PI);  // parameters.f90:400-401
    p.two_alpha_over_sqrt_pi = 2.0 * p.alpha_over_sqrt_pi;
    p.inv_alpha = 1.0 / d->alpha;
    YdChain &h = c->ch;
    int TT = T * T > 0 ? T * T : 1;
    int rc = YAMC_OK;
    auto go = [&]() -> int {
        TRY(dalloc(c, &h.x, N)); TRY(dalloc(c, &h.y, N)); TRY(dalloc(c, &h.z, N));
        TRY(dalloc(c, &h.q, N)); TRY(dalloc(c, &h.dh, N)); TRY(dalloc(c, &h.mass, N));
        TRY(dalloc(c, &h.molnum, N)); TRY(dalloc(c, &h.type, N));
        TRY(dalloc(c, &h.first, M)); TRY(dalloc(c, &h.last, M));
        TRY(dalloc(c, &h.kxyz, nk)); TRY(dalloc(c, &h.beta, nk));
        TRY(dalloc(c, &h.beta_lj, (size_t)TT * nk));
        size_t nS = (size_t)(1 + T) * 8 * nk;
        if (d->precision == YAMC_FP32C) {
            TRY(dalloc(c, &h.Shi, nS)); TRY(dalloc(c, &h.Slo, nS));
        } else {
            TRY(dalloc(c, &h.S, nS));
        }
        TRY(dalloc(c, &h.uu, N)); TRY(dalloc(c, &h.Fx, N)); TRY(dalloc(c, &h.Fy, N)); TRY(dalloc(c, &h.Fz, N));
        TRY(dalloc(c, &h.rsb, (size_t)YD_NBCAP * 4));
        TRY(dalloc(c, &h.nbl, (size_t)YD_NBCAP)); TRY(dalloc(c, &h.pos4, (size_t)N));
        TRY(dalloc(c, &h.arec, (size_t)N * 8)); TRY(dalloc(c, &h.uf, (size_t)N * 4));
        std::vector<int> soff(T + 2, 0), satoms;  // per-set atom lists of the from-scratch S(k)
        for (int set = 0; set <= T; ++set) {
            soff[set] = (int)satoms.size();
            for (int i = 0; i < N; ++i)
                if (set == 0 ? d->charge[i] != 0.0 : d->lj_type[i] == set) satoms.push_back(i);
        }
        soff[T + 1] = (int)satoms.size();
        TRY(dalloc(c, &h.set_atoms, satoms.empty() ? (size_t)1 : satoms.size()));
        TRY(dalloc(c, &h.set_off, (size_t)(T + 2)));
        // on the context stream, after dalloc's clearing, and complete before
        // the host vectors go
        if (!satoms.empty())
            HIPCHK(hipMemcpyAsync(h.set_atoms, satoms.data(), sizeof(int) * satoms.size(), hipMemcpyHostToDevice,
                                  c->stream));
        HIPCHK(hipMemcpyAsync(h.set_off, soff.data(), sizeof(int) * (T + 2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        TRY(dalloc(c, &h.freq_end, M)); TRY(dalloc(c, &h.freq_beg, M)); TRY(dalloc(c, &h.st, 1));
        // cell list of the real-space scan: cells of width >= rc (box units;
        // rmax2 is constant in box units, so volume steps keep nc) when that
        // gives 4 or more per axis (yamc_set_cell_div changes it)
        {
            const double rc = sqrt(d->rmax2);
            const int nc = rc > 0 ? (int)floor(1.0 / (rc * (1.0 + 1e-9))) : 0;
            if (nc >= 4) TRY(setup_cells(c, 1));
        }
        TRY(dalloc(c, &c->d_lj6, TT)); TRY(dalloc(c, &c->d_lj12, TT));
        TRY(dalloc(c, &c->d_sig6, TT)); TRY(dalloc(c, &c->d_4eps, TT));
        TRY(dalloc(c, &c->d_counts, T > 0 ? T : 1));
        TRY(dalloc(c, &c->d_xk2, nk)); TRY(dalloc(c, &c->d_beta6, nk)); TRY(dalloc(c, &c->d_beta12, nk));
        TRY(dalloc(c, &c->d_tab, (size_t)N * 6 * YD_KP));
        {
            // atom groups of the from-scratch S(k): each set's atom list in
            // groups (of at least 64 atoms for the largest list), enough
            // (set, k, group) threads to fill the GPU (about 2^19), at most 64
            const int nsk = (1 + T) * nk;
            int ng = (1 << 19) / nsk;
            ng = ng < 1 ? 1 : ng > 64 ? 64 : ng;
            if (ng > N / 64) ng = N / 64 > 1 ? N / 64 : 1;
            h.fs_ng = ng;
        }
        TRY(dalloc(c, &c->d_ov, 1)); TRY(dalloc(c, &c->d_io, 64)); TRY(dalloc(c, &c->d_mol, 3 * YD_MAXA));
        TRY(dalloc(c, &c->d_chain, 1)); TRY(dalloc(c, &c->d_sig6s, 1)); TRY(dalloc(c, &c->d_feps, 1));
        // host staging
        c->hard_core_angstr.assign(d->hard_core_angstr, d->hard_core_angstr + N);
        c->h_q.assign(d->charge, d->charge + N);
        c->h_mass.assign(d->mass, d->mass + N);
        c->h_type.assign(d->lj_type, d->lj_type + N);
        c->h_first.assign(d->mol_first, d->mol_first + M);
        c->h_last.assign(d->mol_last, d->mol_last + M);
        if (T > 0) {
            c->lj6.assign(d->lj6, d->lj6 + T * T);
            c->lj12.assign(d->lj12, d->lj12 + T * T);
            c->sigma6.assign(d->sigma6, d->sigma6 + T * T);
        }
        std::vector<double> dh(N);
        for (int i = 0; i < N; ++i) dh[i] = d->hard_core_angstr[i] / d->box_length;  // EwaldSumRealSpace.f90:359,381
        std::vector<int> kxyz(nk);
        for (int k = 0; k < nk; ++k) kxyz[k] = c->kx[k] | (c->ky[k] << 8) | (c->kz[k] << 16);
        // BiasedRandom.f90:41-75 with mv_num(:) = 1 (runmc.f90:43)
        std::vector<double> fb(M), fe(M), ones(M, 1.0);
        biased_init(M, ones.data(), fb.data(), fe.data());
        c->xfreq_beg.assign(M, 0.0);
        c->xfreq_end.assign(M, 0.0);
        hipStream_t s = c->stream;
        HIPCHK(hipMemcpyAsync(h.q, d->charge, sizeof(double) * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.dh, dh.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.mass, d->mass, sizeof(double) * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.molnum, d->molnum, sizeof(int) * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.type, d->lj_type, sizeof(int) * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.first, d->mol_first, sizeof(int) * M, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.last, d->mol_last, sizeof(int) * M, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.kxyz, kxyz.data(), sizeof(int) * nk, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->d_xk2, xk2.data(), sizeof(double) * nk, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.freq_end, fe.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h.freq_beg, fb.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
        if (T > 0) {
            HIPCHK(hipMemcpyAsync(c->d_lj6, d->lj6, sizeof(double) * T * T, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(c->d_lj12, d->lj12, sizeof(double) * T * T, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(c->d_sig6, d->sigma6, sizeof(double) * T * T, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(c->d_4eps, d->four_eps, sizeof(double) * T * T, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(c->d_counts, d->tail_counts, sizeof(int) * T, hipMemcpyHostToDevice, s));
        }
        HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, s));
        const double *sp = c->d_sig6, *fp = c->d_4eps;
        HIPCHK(hipMemcpyAsync(c->d_sig6s, &sp, sizeof(sp), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->d_feps, &fp, sizeof(fp), hipMemcpyHostToDevice, s));
        return sync(c);
    };
    rc = go();
    if (rc != YAMC_OK) {
        yamc_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return YAMC_OK;
}

extern "C" int yamc_ctx_destroy(yamc_ctx *c) {
    if (!c) return YAMC_OK;
    // teardown: errors here have no one to report to
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->owned) (void)hipFree(p);
    for (void *p : c->vol_save) (void)hipFree(p);
    if (c->d_xd) (void)hipFree(c->d_xd);
    if (c->d_xb) (void)hipFree(c->d_xb);
    if (c->d_xmv) (void)hipFree(c->d_xmv);
    if (c->d_ulist) (void)hipFree(c->d_ulist);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return YAMC_OK;
}

extern "C" int yamc_ctx_info(const yamc_ctx *c, int32_t *device, int32_t *nk) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (device) *device = c->device;
    if (nk) *nk = c->nk;
    return YAMC_OK;
}

extern "C" int yamc_grid(yamc_ctx *c, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip, int32_t *iq) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    for (int k = 0; k < c->nk; ++k) {
        if (kx) kx[k] = c->kx[k];
        if (ky) ky[k] = c->ky[k];
        if (kz) kz[k] = c->kz[k];
        if (ip) ip[k] = c->ip[k];
        if (iq) iq[k] = c->iq[k];
    }
    return YAMC_OK;
}

extern "C" int yamc_set_positions(yamc_ctx *c, const double *x, const double *y, const double *z) {
    if (!c || !x || !y || !z) return fail(YAMC_ERROR_PARAMETER, "null argument");
    size_t n = sizeof(double) * c->natom;
    HIPCHK(hipMemcpyAsync(c->ch.x, x, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ch.y, y, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ch.z, z, n, hipMemcpyHostToDevice, c->stream));
    c->state_ready = false;
    return sync(c);
}

extern "C" int yamc_get_positions(yamc_ctx *c, double *x, double *y, double *z) {
    if (!c || !x || !y || !z) return fail(YAMC_ERROR_PARAMETER, "null argument");
    size_t n = sizeof(double) * c->natom;
    HIPCHK(hipMemcpyAsync(x, c->ch.x, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(y, c->ch.y, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(z, c->ch.z, n, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

// S(k) from scratch + k-space energies (EwaldSumKSpace.f90:225-260).
static int total_energy(yamc_ctx *c) {
    hipStream_t s = c->stream;
    int N = c->natom, nk = c->nk, T = c->ntype;
    yk_atom_tables<<<(N + 255) / 256, 256, 0, s>>>(c->ch, c->ch.x, c->ch.y, c->ch.z, c->d_tab);
    const int nsk = (1 + T) * nk;
    const int ng = c->ch.fs_ng;
    StreamScratch sc(s);
    double *part = nullptr;
    TRY(sc.get(&part, 8 * (size_t)nsk * ng));
    yk_full_sums<<<dim3((nsk + 127) / 128, ng), 128, 0, s>>>(c->ch, c->d_tab, ng, part);
    yk_full_sums_reduce<<<(8 * nsk + 255) / 256, 256, 0, s>>>(c->ch, part, ng);
    yk_kenergy<<<1, 1024, 0, s>>>(c->ch);
    HIPCHK(hipGetLastError());
    return YAMC_OK;
}

extern "C" int yamc_kspace_init(yamc_ctx *c, double scale_beta, double *e_c, double *e_lj) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    hipStream_t s = c->stream;
    int nk = c->nk, T = c->ntype;
    if (scale_beta > 1e-9) {
        if (!c->beta_ready) return fail(YAMC_ERROR_INITIALIZATION, "scale_beta before beta is initialised");
        yk_scale_beta<<<(nk + 255) / 256, 256, 0, s>>>(c->ch, scale_beta);
    } else {
        yk_beta<<<nk, 64, 0, s>>>(c->ch, c->d_xk2, c->d_beta6, c->d_beta12);
        if (T > 0) {
            long long n = (long long)T * T * nk;
            yk_beta_lj<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(c->ch, c->d_lj6, c->d_lj12, c->d_beta6, c->d_beta12);
        }
        c->beta_ready = true;
    }
    HIPCHK(hipGetLastError());
    TRY(total_energy(c));
    TRY(sync(c));
    if (e_c || e_lj) {
        YdState st;
        HIPCHK(hipMemcpy(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost));
        if (e_c) *e_c = st.e_kc;
        if (e_lj) *e_lj = st.e_klj;
    }
    return YAMC_OK;
}

extern "C" int yamc_kspace_get_beta(yamc_ctx *c, double *beta, double *beta6, double *beta12, double *beta_lj) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->beta_ready) return fail(YAMC_ERROR_INITIALIZATION, "beta not initialised");
    TRY(sync(c));  // the copies below run on the null stream; the context's stream is non-blocking
    size_t nk = c->nk;
    if (beta) HIPCHK(hipMemcpy(beta, c->ch.beta, sizeof(double) * nk, hipMemcpyDeviceToHost));
    if (beta6) HIPCHK(hipMemcpy(beta6, c->d_beta6, sizeof(double) * nk, hipMemcpyDeviceToHost));
    if (beta12) HIPCHK(hipMemcpy(beta12, c->d_beta12, sizeof(double) * nk, hipMemcpyDeviceToHost));
    if (beta_lj && c->ntype > 0)
        HIPCHK(hipMemcpy(beta_lj, c->ch.beta_lj, sizeof(double) * nk * c->ntype * c->ntype, hipMemcpyDeviceToHost));
    return YAMC_OK;
}

extern "C" int yamc_kspace_get_sums(yamc_ctx *c, double *sums) {
    if (!c || !sums) return fail(YAMC_ERROR_PARAMETER, "null argument");
    TRY(sync(c));  // the copies below run on the null stream; the context's stream is non-blocking
    size_t n = (size_t)(1 + c->ntype) * 8 * c->nk;
    if (c->precision == YAMC_FP32C) {
        std::vector<float> hi(n), lo(n);
        HIPCHK(hipMemcpy(hi.data(), c->ch.Shi, sizeof(float) * n, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lo.data(), c->ch.Slo, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) sums[i] = (double)hi[i] + (double)lo[i];
    } else {
        HIPCHK(hipMemcpy(sums, c->ch.S, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return YAMC_OK;
}

static int stage_mol(yamc_ctx *c, int n, const double *x, const double *y, const double *z) {
    double buf[3 * YD_MAXA];
    for (int a = 0; a < n; ++a) { buf[a] = x[a]; buf[YD_MAXA + a] = y[a]; buf[2 * YD_MAXA + a] = z[a]; }
    HIPCHK(hipMemcpyAsync(c->d_mol, buf, sizeof buf, hipMemcpyHostToDevice, c->stream));
    return YAMC_OK;
}

static int mol_size(yamc_ctx *c, int imol, int *na) {
    if (imol < 0 || imol >= c->nmol) return fail(YAMC_ERROR_PARAMETER, "molecule index %d out of range", imol);
    TRY(sync(c));  // the copies below run on the null stream; the context's stream is non-blocking
    int f, l;
    HIPCHK(hipMemcpy(&f, c->ch.first + imol, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&l, c->ch.last + imol, sizeof(int), hipMemcpyDeviceToHost));
    *na = l - f + 1;
    return YAMC_OK;
}

extern "C" int yamc_kspace_delta(yamc_ctx *c, int32_t imol, const double *x, const double *y, const double *z,
                                 double *du_c, double *du_lj) {
    if (!c || !x || !y || !z) return fail(YAMC_ERROR_PARAMETER, "null argument");
    if (!c->beta_ready) return fail(YAMC_ERROR_INITIALIZATION, "kspace not initialised");
    int na;
    TRY(mol_size(c, imol, &na));
    TRY(stage_mol(c, na, x, y, z));
    const double *dm = c->d_mol;
    if (c->precision == YAMC_FP32C)
        yk_kspace_delta<true><<<1, YK_BS, 0, c->stream>>>(c->ch, imol, dm, dm + YD_MAXA, dm + 2 * YD_MAXA, c->d_io);
    else
        yk_kspace_delta<false><<<1, YK_BS, 0, c->stream>>>(c->ch, imol, dm, dm + YD_MAXA, dm + 2 * YD_MAXA, c->d_io);
    HIPCHK(hipGetLastError());
    double o[2];
    HIPCHK(hipMemcpyAsync(o, c->d_io, sizeof o, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    if (du_c) *du_c = o[0];
    if (du_lj) *du_lj = o[1];
    return YAMC_OK;
}

extern "C" int yamc_kspace_forces(yamc_ctx *c, int32_t first, int32_t n, const double *x, const double *y,
                                  const double *z, double *fx, double *fy, double *fz) {
    if (!c || !x || !y || !z || !fx || !fy || !fz) return fail(YAMC_ERROR_PARAMETER, "null argument");
    if (!c->beta_ready) return fail(YAMC_ERROR_INITIALIZATION, "kspace not initialised");
    if (first < 0 || n < 0 || first + n > c->natom) return fail(YAMC_ERROR_PARAMETER, "atom range out of bounds");
    for (int b = 0; b < n; b += YD_MAXA) {
        int m = n - b < YD_MAXA ? n - b : YD_MAXA;
        TRY(stage_mol(c, m, x + b, y + b, z + b));
        const double *dm = c->d_mol;
        if (c->precision == YAMC_FP32C)
            yk_kspace_forces<true><<<1, YK_BS, 0, c->stream>>>(c->ch, first + b, m, dm, dm + YD_MAXA,
                                                                dm + 2 * YD_MAXA, c->d_io);
        else
            yk_kspace_forces<false><<<1, YK_BS, 0, c->stream>>>(c->ch, first + b, m, dm, dm + YD_MAXA,
                                                                 dm + 2 * YD_MAXA, c->d_io);
        HIPCHK(hipGetLastError());
        double o[3 * YD_MAXA];
        HIPCHK(hipMemcpyAsync(o, c->d_io, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, c->stream));
        TRY(sync(c));
        for (int a = 0; a < m; ++a) {
            fx[b + a] = o[3 * a];
            fy[b + a] = o[3 * a + 1];
            fz[b + a] = o[3 * a + 2];
        }
    }
    return YAMC_OK;
}

// Launch shape of the full real-space sum for N atoms: nb blocks of 256
// atoms, ng partner groups (of at least 64 atoms, about 2^19 (atom, group)
// threads, at most 64); its scratch is part[ng][nb][6] and pat[ng][4][N].
static void rspace_full_shape(int N, int *nb, int *ng) {
    *nb = (N + 255) / 256;
    int g = (1 << 19) / N;
    g = g < 1 ? 1 : g > 64 ? 64 : g;
    if (g > N / 64) g = N / 64 > 1 ? N / 64 : 1;
    *ng = g;
}

static size_t rspace_full_scratch_doubles(int N) {
    int nb, ng;
    rspace_full_shape(N, &nb, &ng);
    return 6 * (size_t)nb * ng + 4 * (size_t)N * ng;
}

// EwaldSumRealSpace_calc, full mode, with the caller's scratch (at least
// rspace_full_scratch_doubles(N) doubles): every value of it is written
// before it is read, so it needs no clearing.
static int rspace_full_run(yamc_ctx *c, double *scratch, double *comps, int32_t *overlap) {
    hipStream_t s = c->stream;
    const int N = c->natom;
    int nb, ng;
    rspace_full_shape(N, &nb, &ng);
    double *part = scratch, *pat = scratch + 6 * (size_t)nb * ng;
    HIPCHK(hipMemsetAsync(c->d_ov, 0, sizeof(int), s));
    yk_rspace_full<<<dim3(nb, ng), 256, 0, s>>>(c->ch, c->d_sig6, c->d_4eps, ng, part, pat, c->d_ov);
    yk_rspace_full_atoms<<<nb, 256, 0, s>>>(c->ch, pat, ng);
    yk_rspace_full_final<<<1, 256, 0, s>>>(c->ch, part, nb * ng, c->d_io);
    HIPCHK(hipGetLastError());
    double o[6];
    int ov = 0;
    HIPCHK(hipMemcpyAsync(o, c->d_io, sizeof o, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&ov, c->d_ov, sizeof ov, hipMemcpyDeviceToHost, s));
    TRY(sync(c));
    if (comps)
        for (int m = 0; m < 6; ++m) comps[m] = o[m];
    if (overlap) *overlap = ov;
    return YAMC_OK;
}

static int rspace_full(yamc_ctx *c, double *comps, int32_t *overlap) {
    StreamScratch sc(c->stream);
    double *scratch = nullptr;
    TRY(sc.get(&scratch, rspace_full_scratch_doubles(c->natom)));
    return rspace_full_run(c, scratch, comps, overlap);
}

extern "C" int yamc_rspace_full(yamc_ctx *c, double *comps, double *uu, double *fx, double *fy, double *fz,
                                int32_t *overlap) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    TRY(rspace_full(c, comps, overlap));
    c->state_ready = false;  // uu/F/components were overwritten; k-space and tails still need mc_init
    return yamc_mc_real_arrays(c, uu, fx, fy, fz);
}

// ---- the no-allocation variant of yamc_rspace_full ----------------------
// Its scratch (part + pat of the full real-space sum) is a file-scope device
// array: each GPU the library runs on holds one copy of it until the process
// ends.  The array's name is only a host-side stand-in; its device address
// on each GPU comes from hipGetSymbolAddress, taken once per device with
// that device current, and is the only handle host code passes on.
#define YAMC_NOALLOC_SCRATCH_DOUBLES (3u << 20)  // 24 MiB
#define YAMC_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[YAMC_NOALLOC_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[YAMC_NOALLOC_MAX_DEVICES] = {};

static int scratch_ptr(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= YAMC_NOALLOC_MAX_DEVICES) return fail(YAMC_ERROR_DEVICE, "device %d out of range", device);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        HIPCHK(hipGetDevice(&cur));
        if (cur != device) HIPCHK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(YAMC_ERROR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return YAMC_OK;
}

extern "C" int yamc_rspace_full_noalloc(yamc_ctx *c, double *comps, double *uu, double *fx, double *fy, double *fz,
                                        int32_t *overlap) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    const size_t need = rspace_full_scratch_doubles(c->natom);
    if (need > YAMC_NOALLOC_SCRATCH_DOUBLES)
        return fail(YAMC_ERROR_LIMITS,
                    "yamc_rspace_full_noalloc: %d atoms need %zu doubles of scratch, the static buffer holds %u "
                    "(use yamc_rspace_full)", c->natom, need, (unsigned)YAMC_NOALLOC_SCRATCH_DOUBLES);
    double *sp = nullptr;
    TRY(scratch_ptr(c->device, &sp));
    TRY(rspace_full_run(c, sp, comps, overlap));
    c->state_ready = false;  // as yamc_rspace_full
    return yamc_mc_real_arrays(c, uu, fx, fy, fz);
}

extern "C" int yamc_rspace_partial(yamc_ctx *c, int32_t imol, const double *x, const double *y, const double *z,
                                   double *comps, double *uu, double *fx, double *fy, double *fz, double *uu_back,
                                   double *fx_back, double *fy_back, double *fz_back, int32_t *overlap) {
    if (!c || !x || !y || !z) return fail(YAMC_ERROR_PARAMETER, "null argument");
    int na;
    TRY(mol_size(c, imol, &na));
    TRY(stage_mol(c, na, x, y, z));
    int N = c->natom;
    double *back = nullptr;
    bool want_back = uu_back || fx_back || fy_back || fz_back;
    StreamScratch sc(c->stream);
    if (want_back) TRY(sc.get(&back, 4 * (size_t)N));
    const double *dm = c->d_mol;
    yk_rspace_partial<<<1, YK_BS, 0, c->stream>>>(c->ch, imol, dm, dm + YD_MAXA, dm + 2 * YD_MAXA, c->d_sig6,
                                                  c->d_4eps, c->d_io, back, back ? back + N : nullptr,
                                                  back ? back + 2 * N : nullptr, back ? back + 3 * N : nullptr);
    HIPCHK(hipGetLastError());
    double o[4 * YD_MAXA + 4];
    HIPCHK(hipMemcpyAsync(o, c->d_io, sizeof o, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> hb;
    if (want_back) {
        hb.resize(4 * (size_t)N);
        HIPCHK(hipMemcpyAsync(hb.data(), back, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, c->stream));
    }
    TRY(sync(c));
    for (int a = 0; a < na; ++a) {
        if (uu) uu[a] = o[a];
        if (fx) fx[a] = o[YD_MAXA + 3 * a];
        if (fy) fy[a] = o[YD_MAXA + 3 * a + 1];
        if (fz) fz[a] = o[YD_MAXA + 3 * a + 2];
    }
    if (comps) {
        comps[0] = o[4 * YD_MAXA]; comps[1] = o[4 * YD_MAXA + 1]; comps[2] = o[4 * YD_MAXA + 2];
        comps[3] = comps[4] = comps[5] = 0.0;  // same-molecule pairs are skipped in the partial sum
    }
    if (overlap) *overlap = o[4 * YD_MAXA + 3] != 0.0;
    if (want_back) {
        if (uu_back) memcpy(uu_back, hb.data(), sizeof(double) * N);
        if (fx_back) memcpy(fx_back, hb.data() + N, sizeof(double) * N);
        if (fy_back) memcpy(fy_back, hb.data() + 2 * N, sizeof(double) * N);
        if (fz_back) memcpy(fz_back, hb.data() + 3 * N, sizeof(double) * N);
    }
    return YAMC_OK;
}

extern "C" int yamc_tails(yamc_ctx *c, double *self_c, double *lj6, double *lj12) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    yk_tails<<<1, 256, 0, c->stream>>>(c->ch, c->d_lj6, c->d_lj12, c->d_counts);
    HIPCHK(hipGetLastError());
    YdState st;
    HIPCHK(hipMemcpyAsync(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    if (self_c) *self_c = st.tail_c;
    if (lj6) *lj6 = st.tail6;
    if (lj12) *lj12 = st.tail12;
    return YAMC_OK;
}

extern "C" int yamc_external_mu(yamc_ctx *c, double *mu3) {
    if (!c || !mu3) return fail(YAMC_ERROR_PARAMETER, "null argument");
    yk_mu<<<1, 1024, 0, c->stream>>>(c->ch, c->d_io);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(mu3, c->d_io, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

// MC.f90:681-698 on the host from the device scalars.
static void energies_from_state(const YdState &st, double *e) {
    double U_tail = st.tail_c + st.tail12 - st.tail6;
    double U_ext = 0.0;  // filled by caller (needs kext)
    double ew = st.uu_ew + st.e_kc + st.tail_c;
    double lj = st.uu_lj12 - st.uu_lj6 + st.e_klj + st.tail12 - st.tail6;
    e[1] = ew; e[2] = lj; e[3] = st.uu_ew; e[4] = st.uu_lj6; e[5] = st.uu_lj12;
    e[6] = st.e_kc; e[7] = st.e_klj; e[8] = U_tail; e[9] = U_ext;
}

extern "C" int yamc_mc_energies(yamc_ctx *c, double *e) {
    if (!c || !e) return fail(YAMC_ERROR_PARAMETER, "null argument");
    YdState st;
    HIPCHK(hipMemcpyAsync(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    energies_from_state(st, e);
    double U_ext = c->ch.p.kext * (st.mu[0] * st.mu[0] + st.mu[1] * st.mu[1] + st.mu[2] * st.mu[2]);
    e[9] = U_ext;
    e[1] = e[1] + U_ext;  // U_total_ew = uu_ew + E_kC + U_tail_C + U_ext
    e[0] = e[1] + e[2];
    return YAMC_OK;
}

// mc_calc_total_sums (MC.f90:477-585): real space, k space (beta scaled by
// scale_beta when it is positive, else rebuilt), mu, tails.
static int calc_total_sums(yamc_ctx *c, double scale_beta, int32_t *overlap) {
    int32_t ov = 0;
    TRY(rspace_full(c, nullptr, &ov));
    TRY(yamc_kspace_init(c, scale_beta, nullptr, nullptr));
    double mu[3];
    TRY(yamc_external_mu(c, mu));
    TRY(yamc_tails(c, nullptr, nullptr, nullptr));
    c->state_ready = true;
    if (overlap) *overlap = ov;
    return YAMC_OK;
}

extern "C" int yamc_mc_init(yamc_ctx *c, double *e, int32_t *overlap) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    TRY(calc_total_sums(c, 0.0, overlap));
    if (e) TRY(yamc_mc_energies(c, e));
    return YAMC_OK;
}

extern "C" int yamc_mc_real_arrays(yamc_ctx *c, double *uu, double *fx, double *fy, double *fz) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    size_t n = sizeof(double) * c->natom;
    if (uu) HIPCHK(hipMemcpyAsync(uu, c->ch.uu, n, hipMemcpyDeviceToHost, c->stream));
    if (fx) HIPCHK(hipMemcpyAsync(fx, c->ch.Fx, n, hipMemcpyDeviceToHost, c->stream));
    if (fy) HIPCHK(hipMemcpyAsync(fy, c->ch.Fy, n, hipMemcpyDeviceToHost, c->stream));
    if (fz) HIPCHK(hipMemcpyAsync(fz, c->ch.Fz, n, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

extern "C" int yamc_mc_set_rng(yamc_ctx *c, uint64_t seed, uint64_t stream, uint64_t counter) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    YdState st;
    HIPCHK(hipMemcpyAsync(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    st.key[0] = (uint32_t)seed; st.key[1] = (uint32_t)(seed >> 32);
    st.stream[0] = (uint32_t)stream; st.stream[1] = (uint32_t)(stream >> 32);
    st.ctr = counter;
    st.ulist = nullptr;  // back to the Philox stream
    st.nlist = 0;
    HIPCHK(hipMemcpyAsync(c->ch.st, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

extern "C" int yamc_mc_set_rng_list(yamc_ctx *c, const double *u, int64_t n) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (n < 0 || (n > 0 && !u)) return fail(YAMC_ERROR_PARAMETER, "bad uniform list");
    for (int64_t k = 0; k < n; ++k)
        if (!(u[k] > 0.0 && u[k] < 1.0)) return fail(YAMC_ERROR_PARAMETER, "uniform %lld outside (0, 1)", (long long)k);
    YdState st;
    TRY(get_state(c, &st));
    if (c->d_ulist) {
        HIPCHK(hipFree(c->d_ulist));
        c->d_ulist = nullptr;
    }
    c->h_ulist.assign(u, u + n);
    if (n > 0) {
        HIPCHK(hipMalloc((void **)&c->d_ulist, sizeof(double) * n));
        HIPCHK(hipMemcpy(c->d_ulist, u, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    st.ulist = c->d_ulist;
    st.nlist = (unsigned long long)n;
    st.ctr = 0;
    return put_state(c, &st);
}

// After a run in replay mode: an error if the run drew past the list.
static int check_list(yamc_ctx *c) {
    YdState st;
    TRY(get_state(c, &st));
    if (st.ulist && st.ctr > st.nlist)
        return fail(YAMC_ERROR_LIMITS, "replayed uniform list exhausted (%llu drawn, %llu recorded)", st.ctr,
                    st.nlist);
    return YAMC_OK;
}

extern "C" int yamc_mc_get_rng(yamc_ctx *c, uint64_t *counter) {
    if (!c || !counter) return fail(YAMC_ERROR_PARAMETER, "null argument");
    YdState st;
    HIPCHK(hipMemcpyAsync(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    *counter = st.ctr;
    return YAMC_OK;
}

extern "C" int yamc_resync(yamc_ctx *c) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->beta_ready) return fail(YAMC_ERROR_INITIALIZATION, "kspace not initialised");
    TRY(total_energy(c));
    return sync(c);
}

template <bool F32, int BS>
static int launch_chain_t(yamc_ctx *c, const YdChain *d_chain, const double *const *sig6s,
                          const double *const *feps, int nchains, int nsteps, int mode, int max_natom,
                          int max_moves, int resync_every) {
    // The packed position copy of the real-space scan (4 bytes per atom) goes
    // into dynamic LDS when it fits next to the kernel's static LDS: within
    // half of a CU's LDS when two chains share a CU (more chains than CUs at
    // 256 threads), else within all of it; otherwise it stays in HBM.
    int lds_max = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    const void *fn = (const void *)yk_chain<F32, BS>;
    hipFuncAttributes fa;
    HIPCHK(hipFuncGetAttributes(&fa, fn));
    size_t want = (size_t)max_natom * 4;
    int lds_atoms = 0;
    size_t dyn = 0;
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    const bool two_per_cu = BS == 256 && nchains > ncu;
    const size_t room = two_per_cu ? (size_t)lds_max / 2 : (size_t)lds_max;
    if (want + fa.sharedSizeBytes + 1024 <= room) {
        lds_atoms = max_natom;
        dyn = want;
        HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    }
    yk_chain<F32, BS><<<nchains, BS, dyn, c->stream>>>(d_chain, sig6s, feps, nsteps, mode, max_moves, lds_atoms,
                                                       resync_every);
    HIPCHK(hipGetLastError());
    return YAMC_OK;
}

// One chain over c->groups co-resident workgroups (cooperative launch).
template <bool F32>
static int launch_gchain(yamc_ctx *c, const YdChain *d_chain, int nsteps, int mode, int max_natom, int max_moves) {
    const int G = c->groups;
    const void *fn = (const void *)yk_gchain<F32, 256>;
    int lds_max = 0, ncu = 0, coop = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device));
    if (!coop) return fail(YAMC_ERROR_DEVICE, "the device has no cooperative launch: use one workgroup per chain");
    hipFuncAttributes fa;
    HIPCHK(hipFuncGetAttributes(&fa, fn));
    // the packed positions go to LDS when they fit there (each workgroup
    // keeps a whole copy); else they stay in HBM (c.pos4)
    const size_t want = (size_t)max_natom * 4;
    const size_t room = G > ncu ? (size_t)lds_max / 2 : (size_t)lds_max;
    int lds_atoms = 0;
    size_t dyn = 0;
    if (want + fa.sharedSizeBytes + 1024 <= room) {
        lds_atoms = max_natom;
        dyn = want;
        HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    }
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, dyn));
    if (per_cu < 1 || G > per_cu * ncu)
        return fail(YAMC_ERROR_LIMITS, "%d workgroups per chain do not fit on the device at once (%d per CU, %d CUs)",
                    G, per_cu, ncu);
    yk_gchain_prep<<<(G + 63) / 64, 64, 0, c->stream>>>(d_chain, c->d_gch, c->d_ggrp, G, c->d_grsb, c->d_gbar,
                                                         c->d_gpart, c->d_sig6, c->d_4eps);
    HIPCHK(hipGetLastError());
    const YdChain *chains = c->d_gch;
    const YgGroup *grp = c->d_ggrp;
    int resync_every = 0;  // (a single grid chain is rebuilt between launches, over the whole GPU)
    void *args[] = {(void *)&chains, (void *)&grp, (void *)&nsteps, (void *)&mode, (void *)&max_moves,
                    (void *)&lds_atoms, (void *)&resync_every};
    HIPCHK(hipLaunchCooperativeKernel(fn, dim3(G), dim3(256), args, (unsigned int)dyn, c->stream));
    return YAMC_OK;
}

// The wave shapes (64, 128) run the wave kernel for FP64 run_mc cycles;
// FP32C chains, constant-uniform traces and runs with exchange steps (which
// launch step ranges, max_moves > 0) run them as workgroups of 256
// (the FP32C wave kernel is switched off: it made an illegal memory access
// on its first GPU launch and its cause is not yet found).
static int launch_chain(yamc_ctx *c, const YdChain *d_chain, const double *const *sig6s,
                        const double *const *feps, int nchains, int nsteps, int mode, int max_natom,
                        int max_moves = 0, int resync_every = 0) {
    const bool f32 = c->precision == YAMC_FP32C;
    if (c->groups > 1 && nchains == 1 && resync_every == 0)
        return f32 ? launch_gchain<true>(c, d_chain, nsteps, mode, max_natom, max_moves)
                   : launch_gchain<false>(c, d_chain, nsteps, mode, max_natom, max_moves);
    if ((c->chain_bs == 64 || c->chain_bs == 128) && mode == 0 && !f32 && max_moves == 0 &&
        resync_every == 0) {  // one wave per replica
        int nb = (nchains + YW_WPB - 1) / YW_WPB;
        if (c->chain_bs == 64) yk_wchain<false, 1><<<nb, YW_WPB * 64, 0, c->stream>>>(d_chain, nchains, sig6s, feps, nsteps);
        else yk_wchain<false, 2><<<nb, YW_WPB * 64, 0, c->stream>>>(d_chain, nchains, sig6s, feps, nsteps);
        HIPCHK(hipGetLastError());
        return YAMC_OK;
    }
    if (c->chain_bs == 512)
        return f32 ? launch_chain_t<true, 512>(c, d_chain, sig6s, feps, nchains, nsteps, mode, max_natom, max_moves,
                                               resync_every)
                   : launch_chain_t<false, 512>(c, d_chain, sig6s, feps, nchains, nsteps, mode, max_natom, max_moves,
                                                resync_every);
    return f32 ? launch_chain_t<true, 256>(c, d_chain, sig6s, feps, nchains, nsteps, mode, max_natom, max_moves,
                                           resync_every)
               : launch_chain_t<false, 256>(c, d_chain, sig6s, feps, nchains, nsteps, mode, max_natom, max_moves,
                                            resync_every);
}

extern "C" int yamc_set_block_size(yamc_ctx *c, int32_t block) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (block != 64 && block != 128 && block != 256 && block != 512)
        return fail(YAMC_ERROR_PARAMETER, "launch shape %d: must be 64 or 128 (one wave per replica), 256 or 512",
                    block);
    c->chain_bs = block;
    return YAMC_OK;
}

extern "C" int yamc_set_cell_div(yamc_ctx *c, int32_t div) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (div < 0 || div > 4) return fail(YAMC_ERROR_PARAMETER, "%d cells per cut-off: must be 0..4", div);
    HIPCHK(hipStreamSynchronize(c->stream));
    TRY(setup_cells(c, div));
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

extern "C" int yamc_set_chain_groups(yamc_ctx *c, int32_t groups) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (groups < 1 || groups > 256)
        return fail(YAMC_ERROR_PARAMETER, "%d workgroups per chain: must be 1..256", groups);
    if (groups > c->nk || groups > c->natom)
        return fail(YAMC_ERROR_PARAMETER, "%d workgroups per chain for %d k vectors and %d atoms", groups, c->nk,
                    c->natom);
    if (groups > c->gcap) {
        TRY(dalloc(c, &c->d_gch, (size_t)groups));
        TRY(dalloc(c, &c->d_ggrp, (size_t)groups));
        TRY(dalloc(c, &c->d_grsb, (size_t)groups * YD_NBCAP * 4));
        TRY(dalloc(c, &c->d_gpart, (size_t)2 * groups * YG_PART));
        TRY(dalloc(c, &c->d_gbar, 2));
        c->gcap = groups;
    }
    c->groups = groups;
    return YAMC_OK;
}

// A grid chain whose barrier wait ran past its bound stopped early.
static int check_grid(yamc_ctx *c, const YdState &st) {
    if (st.grid_fail) return fail(YAMC_ERROR_DEVICE, "grid chain stopped: a barrier wait ran past its bound");
    (void)c;
    return YAMC_OK;
}

static int run_with_exchanges(yamc_ctx *c, int cyc0, int ncyc, yamc_move_rec *d_rec);

extern "C" int yamc_mc_run(yamc_ctx *c, int32_t ncycles, int32_t resync_every, yamc_move_rec *recs,
                           int64_t *n_accepted) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    if (ncycles < 0) return fail(YAMC_ERROR_PARAMETER, "negative cycle count");
    YdState st0;
    HIPCHK(hipMemcpyAsync(&st0, c->ch.st, sizeof st0, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    // Without exchange steps the resyncs come every resync_every moves (below);
    // with them, at the end of the first whole cycle after that many moves.
    int seg = ncycles;
    if (resync_every > 0) {
        seg = (resync_every + c->nmol - 1) / c->nmol;
        if (seg < 1) seg = 1;
    }
    yamc_move_rec *d_rec = nullptr;
    size_t nrec = (size_t)ncycles * c->nmol;
    StreamScratch sc(c->stream);
    if (recs && nrec) TRY(sc.get(&d_rec, nrec));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (resync_every > 0 && c->pxch <= 0) {
        // segments of resync_every moves, each resumed where the last one
        // stopped inside its cycle (mv_begin), with a resync between them
        const long long total = (long long)ncycles * c->nmol;
        for (long long done = 0; done < total;) {
            const long long n = total - done < resync_every ? total - done : resync_every;
            YdChain ch = c->ch;
            ch.rec = d_rec ? d_rec + done : nullptr;
            ch.mv_begin = (int)(done % c->nmol);
            HIPCHK(hipMemcpyAsync(c->d_chain, &ch, sizeof ch, hipMemcpyHostToDevice, c->stream));
            const int nsteps = (int)((n + c->nmol - 1) / c->nmol) + 1;
            TRY(launch_chain(c, c->d_chain, c->d_sig6s, c->d_feps, 1, nsteps, 0, c->natom, (int)n));
            done += n;
            if (done < total) TRY(total_energy(c));
        }
        HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    } else {
        for (int done = 0; done < ncycles; done += seg) {
            int n = ncycles - done < seg ? ncycles - done : seg;
            if (c->pxch > 0) {
                TRY(run_with_exchanges(c, done, n, d_rec));
            } else {
                YdChain ch = c->ch;
                ch.rec = d_rec ? d_rec + (size_t)done * c->nmol : nullptr;
                HIPCHK(hipMemcpyAsync(c->d_chain, &ch, sizeof ch, hipMemcpyHostToDevice, c->stream));
                TRY(launch_chain(c, c->d_chain, c->d_sig6s, c->d_feps, 1, n, 0, c->natom));
            }
            if (resync_every > 0 && done + n < ncycles) TRY(total_energy(c));
        }
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    if (d_rec) {
        HIPCHK(hipMemcpyAsync(recs, d_rec, sizeof(yamc_move_rec) * nrec, hipMemcpyDeviceToHost, c->stream));
    }
    YdState st1;
    HIPCHK(hipMemcpyAsync(&st1, c->ch.st, sizeof st1, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    HIPCHK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (n_accepted) *n_accepted = st1.n_accepted - st0.n_accepted;
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    TRY(check_grid(c, st1));
    return check_list(c);
}

extern "C" int yamc_mc_run_const(yamc_ctx *c, int32_t nmoves, const int32_t *imol, const double *u,
                                 yamc_move_rec *recs, int64_t *n_accepted) {
    if (!c || !imol || !u) return fail(YAMC_ERROR_PARAMETER, "null argument");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    for (int k = 0; k < nmoves; ++k)
        if (imol[k] < 0 || imol[k] >= c->nmol) return fail(YAMC_ERROR_PARAMETER, "molecule index out of range at move %d", k);
    YdState st0;
    HIPCHK(hipMemcpyAsync(&st0, c->ch.st, sizeof st0, hipMemcpyDeviceToHost, c->stream));
    int *d_imol = nullptr;
    double *d_u = nullptr;
    yamc_move_rec *d_rec = nullptr;
    StreamScratch sc(c->stream);
    TRY(sc.get(&d_imol, (size_t)nmoves + 1));
    TRY(sc.get(&d_u, (size_t)nmoves + 1));
    HIPCHK(hipMemcpyAsync(d_imol, imol, sizeof(int) * nmoves, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_u, u, sizeof(double) * nmoves, hipMemcpyHostToDevice, c->stream));
    if (recs) TRY(sc.get(&d_rec, (size_t)nmoves + 1));
    YdChain ch = c->ch;
    ch.rec = d_rec; ch.cimol = d_imol; ch.cu = d_u;
    HIPCHK(hipMemcpyAsync(c->d_chain, &ch, sizeof ch, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    TRY(launch_chain(c, c->d_chain, c->d_sig6s, c->d_feps, 1, nmoves, 1, c->natom));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    if (recs) HIPCHK(hipMemcpyAsync(recs, d_rec, sizeof(yamc_move_rec) * nmoves, hipMemcpyDeviceToHost, c->stream));
    YdState st1;
    HIPCHK(hipMemcpyAsync(&st1, c->ch.st, sizeof st1, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    HIPCHK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (n_accepted) *n_accepted = st1.n_accepted - st0.n_accepted;
    return check_grid(c, st1);
}

// Staged batch: the chain descriptors of n contexts on the device.  A sweep
// of ncycles run_mc cycles per replica runs as one launch; with
// resync_every = R > 0 each chain rebuilds its S(k) from scratch in FP64
// after every R moves inside that launch (yk_chain_rebuild), with the same
// arithmetic as the launch-level rebuilds between the segments of
// yamc_mc_run(ctx, ncycles, R), so each replica's chain equals its single
// run.
struct yamc_batch {
    yamc_ctx *c0 = nullptr;
    int n = 0, ncycles = 0, max_natom = 0, resync_every = 0;
    std::vector<yamc_ctx *> ctxs;
    std::vector<YdChain> h_ch;  // n descriptors, rebuilt from the contexts before every launch
    YdChain *d_ch = nullptr;
    const double **d_s6 = nullptr, **d_fe = nullptr;
    long long *d_cnt = nullptr;  // accepted-move counters before and after a launch (2 x n)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // grid chains (a member set to more than one workgroup per chain): the
    // nwg workgroups of all members in one cooperative launch, with their
    // descriptors and group records, back terms, partial sums, barrier
    // words and the members' stop flags; launched without a graph
    bool grid = false;
    int nwg = 0;
    YdChain *d_gch = nullptr;
    YgGroup *d_grp = nullptr;
    double *d_grsb = nullptr, *d_gpart = nullptr;
    unsigned int *d_gbar = nullptr;
    int *d_fail = nullptr;
};

// Checks the members and writes their current descriptors (box parameters,
// buffers) into b->h_ch.
static int batch_descriptors(yamc_batch *b) {
    yamc_ctx *c0 = b->c0;
    b->h_ch.resize(b->n);
    for (int i = 0; i < b->n; ++i) {
        yamc_ctx *c = b->ctxs[i];
        if (!c) return fail(YAMC_ERROR_PARAMETER, "null context in batch");
        if (c->device != c0->device || c->precision != c0->precision)
            return fail(YAMC_ERROR_PARAMETER, "batch contexts differ in device or precision");
        if (!c->state_ready)
            return fail(YAMC_ERROR_INITIALIZATION, "batch member %d: yamc_mc_init has not run on its state", i);
        if (c->pxch > 0)
            return fail(YAMC_ERROR_PARAMETER, "batch member %d has exchange steps: run it with yamc_mc_run", i);
        if (!b->graph) b->max_natom = c->natom > b->max_natom ? c->natom : b->max_natom;
        b->h_ch[i] = c->ch;
        b->h_ch[i].rec = nullptr;
    }
    return YAMC_OK;
}

static int batch_stage(yamc_ctx *const *ctxs, int32_t n, int32_t ncycles, int32_t resync_every, yamc_batch *b) {
    if (!ctxs || n < 1) return fail(YAMC_ERROR_PARAMETER, "empty batch");
    if (ncycles < 0) return fail(YAMC_ERROR_PARAMETER, "negative cycle count");
    if (resync_every < 0) return fail(YAMC_ERROR_PARAMETER, "negative resync interval");
    yamc_ctx *c0 = ctxs[0];
    if (!c0) return fail(YAMC_ERROR_PARAMETER, "null context in batch");
    b->c0 = c0;
    b->n = n;
    b->ncycles = ncycles;
    b->resync_every = resync_every;
    b->ctxs.assign(ctxs, ctxs + n);
    TRY(batch_descriptors(b));
    std::vector<const double *> s6(n), fe(n);
    for (int i = 0; i < n; ++i) {
        HIPCHK(hipStreamSynchronize(ctxs[i]->stream));
        s6[i] = ctxs[i]->d_sig6;
        fe[i] = ctxs[i]->d_4eps;
    }
    HIPCHK(hipMalloc((void **)&b->d_ch, sizeof(YdChain) * n));
    HIPCHK(hipMalloc((void **)&b->d_s6, sizeof(void *) * n));
    HIPCHK(hipMalloc((void **)&b->d_fe, sizeof(void *) * n));
    HIPCHK(hipMalloc((void **)&b->d_cnt, sizeof(long long) * 2 * n));
    HIPCHK(hipMemcpy(b->d_ch, b->h_ch.data(), sizeof(YdChain) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_s6, s6.data(), sizeof(void *) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_fe, fe.data(), sizeof(void *) * n, hipMemcpyHostToDevice));
    for (int i = 0; i < n; ++i) b->grid = b->grid || ctxs[i]->groups > 1;
    if (b->grid) {
        int nwg = 0;
        size_t npart = 0;
        for (int i = 0; i < n; ++i) {
            nwg += ctxs[i]->groups;
            npart += (size_t)2 * ctxs[i]->groups * YG_PART;
        }
        b->nwg = nwg;
        HIPCHK(hipMalloc((void **)&b->d_gch, sizeof(YdChain) * nwg));
        HIPCHK(hipMalloc((void **)&b->d_grp, sizeof(YgGroup) * nwg));
        HIPCHK(hipMalloc((void **)&b->d_grsb, sizeof(double) * (size_t)nwg * YD_NBCAP * 4));
        HIPCHK(hipMalloc((void **)&b->d_gpart, sizeof(double) * npart));
        HIPCHK(hipMalloc((void **)&b->d_gbar, sizeof(unsigned int) * 2 * n));
        HIPCHK(hipMalloc((void **)&b->d_fail, sizeof(int) * n));
        std::vector<YgGroup> h(nwg);
        size_t off = 0;
        for (int i = 0, w = 0; i < n; ++i) {
            const int G = ctxs[i]->groups;
            for (int g = 0; g < G; ++g, ++w) {
                YgGroup &r = h[w];
                r.g = g; r.G = G; r.rep = i;
                r.bar = b->d_gbar + 2 * i;
                r.part = b->d_gpart + off;
                r.sig6 = s6[i]; r.feps = fe[i];
            }
            off += (size_t)2 * G * YG_PART;
        }
        HIPCHK(hipMemcpy(b->d_grp, h.data(), sizeof(YgGroup) * nwg, hipMemcpyHostToDevice));
    }
    return YAMC_OK;
}

static void batch_free(yamc_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_ch) (void)hipFree(b->d_ch);
    if (b->d_s6) (void)hipFree(b->d_s6);
    if (b->d_fe) (void)hipFree(b->d_fe);
    if (b->d_cnt) (void)hipFree(b->d_cnt);
    for (void *q : {(void *)b->d_gch, (void *)b->d_grp, (void *)b->d_grsb, (void *)b->d_gpart, (void *)b->d_gbar,
                    (void *)b->d_fail})
        if (q) (void)hipFree(q);
    delete b;
}

// One sweep of a batch of grid chains on stream s: the per-workgroup
// descriptors from the members' current ones, then one cooperative launch
// of all nwg workgroups (which must fit on the device at once).
template <bool F32>
static int launch_gbatch(yamc_batch *b, hipStream_t s) {
    yamc_ctx *c0 = b->c0;
    const void *fn = (const void *)yk_gchain<F32, 256>;
    int lds_max = 0, ncu = 0, coop = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c0->device));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c0->device));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c0->device));
    if (!coop) return fail(YAMC_ERROR_DEVICE, "the device has no cooperative launch: use one workgroup per chain");
    hipFuncAttributes fa;
    HIPCHK(hipFuncGetAttributes(&fa, fn));
    const size_t want = (size_t)b->max_natom * 4;
    const size_t room = b->nwg > ncu ? (size_t)lds_max / 2 : (size_t)lds_max;
    int lds_atoms = 0;
    size_t dyn = 0;
    if (want + fa.sharedSizeBytes + 1024 <= room) {
        lds_atoms = b->max_natom;
        dyn = want;
        HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    }
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, dyn));
    if (per_cu < 1 || b->nwg > per_cu * ncu)
        return fail(YAMC_ERROR_LIMITS, "%d workgroups of grid chains do not fit on the device at once (%d per CU, %d CUs)",
                    b->nwg, per_cu, ncu);
    yk_gchain_bprep<<<(b->nwg + 63) / 64, 64, 0, s>>>(b->d_ch, b->d_gch, b->d_grp, b->nwg, b->d_grsb);
    HIPCHK(hipGetLastError());
    const YdChain *chains = b->d_gch;
    const YgGroup *grp = b->d_grp;
    int nsteps = b->ncycles, mode = 0, max_moves = 0, resync_every = b->resync_every;
    void *args[] = {(void *)&chains, (void *)&grp, (void *)&nsteps, (void *)&mode, (void *)&max_moves,
                    (void *)&lds_atoms, (void *)&resync_every};
    HIPCHK(hipLaunchCooperativeKernel(fn, dim3(b->nwg), dim3(256), args, (unsigned int)dyn, s));
    return YAMC_OK;
}

// Enqueues one sweep of the batch on stream s: one kernel launch (so it can
// be captured), or for grid chains the descriptors' set-up and one
// cooperative launch.
static int batch_sweep(yamc_batch *b, hipStream_t s) {
    if (b->grid)
        return b->c0->precision == YAMC_FP32C ? launch_gbatch<true>(b, s) : launch_gbatch<false>(b, s);
    return launch_chain(b->c0, b->d_ch, b->d_s6, b->d_fe, b->n, b->ncycles, 0, b->max_natom, 0, b->resync_every);
}

// Copies every member's accepted-move counter into out[0..n) on the stream
// (one small kernel instead of a device-to-host copy per member).
// (fail: the members' grid-chain stop flags, or null)
__global__ void yk_batch_counts(const YdChain *chains, int n, long long *out, int *fail) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = chains[i].st->n_accepted;
        if (fail) fail[i] = chains[i].st->grid_fail;
    }
}

static int batch_counts_gather(const yamc_batch *b, int slot, hipStream_t s) {
    yk_batch_counts<<<(b->n + 255) / 256, 256, 0, s>>>(b->d_ch, b->n, b->d_cnt + (size_t)slot * b->n,
                                                       slot == 1 ? b->d_fail : nullptr);
    HIPCHK(hipGetLastError());
    return YAMC_OK;
}

// A member of a grid batch whose barrier wait ran past its bound.
static int batch_check_grid(const yamc_batch *b, hipStream_t s) {
    if (!b->grid) return YAMC_OK;
    yk_batch_counts<<<(b->n + 255) / 256, 256, 0, s>>>(b->d_ch, b->n, b->d_cnt + b->n, b->d_fail);
    HIPCHK(hipGetLastError());
    std::vector<int> h(b->n);
    HIPCHK(hipMemcpyAsync(h.data(), b->d_fail, sizeof(int) * b->n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < b->n; ++i)
        if (h[i]) return fail(YAMC_ERROR_DEVICE, "batch member %d: grid chain stopped, a barrier wait ran past its bound", i);
    return YAMC_OK;
}

// After gathers into slots 0 (before) and 1 (after) on stream s: accepted
// moves per member over the launch.
static int batch_counts_diff(const yamc_batch *b, hipStream_t s, int64_t *n_accepted) {
    std::vector<long long> h(2 * (size_t)b->n);
    HIPCHK(hipMemcpyAsync(h.data(), b->d_cnt, sizeof(long long) * 2 * b->n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < b->n; ++i) n_accepted[i] = h[(size_t)b->n + i] - h[i];
    return YAMC_OK;
}

// Before every launch: the members' current descriptors (a volume step or a
// box change since the last launch changes a member's box parameters) into
// the device array the launch reads; members must still have a ready state.
static int batch_refresh(yamc_batch *b, hipStream_t s) {
    TRY(batch_descriptors(b));
    for (int i = 0; i < b->n; ++i) HIPCHK(hipStreamSynchronize(b->ctxs[i]->stream));
    HIPCHK(hipMemcpyAsync(b->d_ch, b->h_ch.data(), sizeof(YdChain) * b->n, hipMemcpyHostToDevice, s));
    return YAMC_OK;
}

static int batch_go(yamc_batch *b, int64_t *n_accepted, bool graph) {
    yamc_ctx *c0 = b->c0;
    hipStream_t s = c0->stream;
    TRY(batch_refresh(b, s));
    if (n_accepted) TRY(batch_counts_gather(b, 0, s));
    HIPCHK(hipEventRecord(c0->ev0, s));
    if (graph && b->exec) HIPCHK(hipGraphLaunch(b->exec, s));
    else TRY(batch_sweep(b, s));
    HIPCHK(hipEventRecord(c0->ev1, s));
    if (n_accepted) TRY(batch_counts_gather(b, 1, s));
    TRY(sync(c0));
    HIPCHK(hipEventElapsedTime(&c0->last_ms, c0->ev0, c0->ev1));
    if (n_accepted) TRY(batch_counts_diff(b, s, n_accepted));
    TRY(batch_check_grid(b, s));
    return YAMC_OK;
}

extern "C" int yamc_batch_run(yamc_ctx *const *ctxs, int32_t n, int32_t ncycles, int32_t resync_every,
                              int64_t *n_accepted) {
    yamc_batch *b = new yamc_batch;
    int rc = batch_stage(ctxs, n, ncycles, resync_every, b);
    if (rc == YAMC_OK) rc = batch_go(b, n_accepted, false);
    batch_free(b);
    return rc;
}

extern "C" int yamc_batch_create(yamc_ctx *const *ctxs, int32_t n, int32_t ncycles, int32_t resync_every,
                                 yamc_batch **out) {
    if (!out) return fail(YAMC_ERROR_PARAMETER, "null output");
    *out = nullptr;
    if (ncycles < 1) return fail(YAMC_ERROR_PARAMETER, "ncycles must be positive");
    yamc_batch *b = new yamc_batch;
    int rc = batch_stage(ctxs, n, ncycles, resync_every, b);
    if (rc == YAMC_OK && !b->grid) {  // (a batch of grid chains is launched directly)
        yamc_ctx *c0 = b->c0;
        auto go = [&]() -> int {
            // The capture records kernel nodes only (no memset or copy nodes).
            HIPCHK(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
            int lrc = batch_sweep(b, c0->stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c0->stream, &g);
            if (lrc != YAMC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return lrc;
            }
            HIPCHK(e);
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            return YAMC_OK;
        };
        rc = go();
    }
    if (rc != YAMC_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return YAMC_OK;
}

extern "C" int yamc_batch_launch(yamc_batch *b, int64_t *n_accepted) {
    if (!b) return fail(YAMC_ERROR_PARAMETER, "null batch");
    return batch_go(b, n_accepted, true);
}

extern "C" int yamc_batch_destroy(yamc_batch *b) {
    if (b) batch_free(b);
    return YAMC_OK;
}

extern "C" int yamc_forces(yamc_ctx *c, double *gx, double *gy, double *gz, double *rx, double *ry, double *rz) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    int N = c->natom;
    hipStream_t s = c->stream;
    StreamScratch sc(s);
    double *dg = nullptr, *dr = nullptr;
    TRY(sc.get(&dg, 3 * (size_t)N));
    TRY(sc.get(&dr, 3 * (size_t)N));
    yk_atom_tables<<<(N + 255) / 256, 256, 0, s>>>(c->ch, c->ch.x, c->ch.y, c->ch.z, c->d_tab);
    HIPCHK(hipGetLastError());
    double mu[3];
    TRY(yamc_external_mu(c, mu));
    if (yk_launch_forces(c->ch, c->d_sig6, c->d_4eps, c->d_tab, dg, dr, s) != 0)
        return fail(YAMC_ERROR_DEVICE, "forces: scratch allocation failed");
    HIPCHK(hipGetLastError());
    std::vector<double> hg(3 * (size_t)N), hr(3 * (size_t)N);
    HIPCHK(hipMemcpyAsync(hg.data(), dg, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hr.data(), dr, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, s));
    TRY(sync(c));
    for (int i = 0; i < N; ++i) {
        if (gx) { gx[i] = hg[i]; gy[i] = hg[N + i]; gz[i] = hg[2 * N + i]; }
        if (rx) { rx[i] = hr[i]; ry[i] = hr[N + i]; rz[i] = hr[2 * N + i]; }
    }
    return YAMC_OK;
}

extern "C" int yamc_phase_times(yamc_ctx *c, double *us) {
    if (!c || !us) return fail(YAMC_ERROR_PARAMETER, "null argument");
    YdState st;
    HIPCHK(hipMemcpyAsync(&st, c->ch.st, sizeof st, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    for (int i = 0; i < 9; ++i) us[i] = khz > 0 ? (double)st.phase[i] * 1e3 / khz : 0.0;
    return YAMC_OK;
}

extern "C" int yamc_last_kernel_ms(yamc_ctx *c, double *ms) {
    if (!c || !ms) return fail(YAMC_ERROR_PARAMETER, "null argument");
    *ms = c->last_ms;
    return YAMC_OK;
}

extern "C" int yamc_rdf(yamc_ctx *c, const int32_t *mol_species, int32_t nspecies, double dr_angstr,
                        double rmax_angstr, int32_t nbins, uint64_t *counts) {
    if (!c || !mol_species || !counts) return fail(YAMC_ERROR_PARAMETER, "null argument");
    if (nspecies < 1 || nspecies > 8) return fail(YAMC_ERROR_PARAMETER, "nspecies %d outside 1..8", nspecies);
    if (nbins < 1 || nbins > 4096) return fail(YAMC_ERROR_PARAMETER, "nbins %d outside 1..4096", nbins);
    if (!(dr_angstr > 0.0)) return fail(YAMC_ERROR_PARAMETER, "dr must be positive");
    if (rmax_angstr > 0.5 * c->ch.p.L)
        return fail(YAMC_ERROR_PARAMETER, "rmax %g A exceeds half the box (%g A)", rmax_angstr, 0.5 * c->ch.p.L);
    for (int m = 0; m < c->nmol; ++m)
        if (mol_species[m] < 0 || mol_species[m] >= nspecies)
            return fail(YAMC_ERROR_PARAMETER, "species of molecule %d outside 0..%d", m, nspecies - 1);
    const size_t nh = (size_t)nspecies * nspecies * nbins;
    hipStream_t s = c->stream;
    double *com = nullptr;
    int *sp = nullptr;
    unsigned long long *hist = nullptr;
    StreamScratch sc(s);
    TRY(sc.get(&com, 3 * (size_t)c->nmol));
    TRY(sc.get(&sp, (size_t)c->nmol));
    TRY(sc.get(&hist, nh));
    HIPCHK(hipMemcpyAsync(sp, mol_species, sizeof(int) * c->nmol, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * nh, s));
    yk_launch_rdf(c->ch, com, sp, nspecies, dr_angstr, rmax_angstr * rmax_angstr, nbins, hist, s);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> h(nh);
    HIPCHK(hipMemcpyAsync(h.data(), hist, sizeof(unsigned long long) * nh, hipMemcpyDeviceToHost, s));
    TRY(sync(c));
    for (size_t k = 0; k < nh; ++k) counts[k] += h[k];
    return YAMC_OK;
}

// ---- volume steps ---------------------------------------------------------
static int get_state(yamc_ctx *c, YdState *st) {
    HIPCHK(hipMemcpyAsync(st, c->ch.st, sizeof *st, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

static int put_state(yamc_ctx *c, const YdState *st) {
    HIPCHK(hipMemcpyAsync(c->ch.st, st, sizeof *st, hipMemcpyHostToDevice, c->stream));
    return sync(c);
}
// Host Philox-4x32-10 on the context's stream (the device generator's word
// order: word n = lane n & 3 of block n >> 2; 24-bit uniforms, zero skipped).
static void host_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t *o) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        uint32_t n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

static double host_uniform(const yamc_ctx *c, const YdState &st, uint64_t *ctr) {
    if (st.ulist) {  // replay mode: the host copy of the recorded stream
        const uint64_t n = (*ctr)++;
        return n < c->h_ulist.size() ? c->h_ulist[n] : 0.5;
    }
    for (;;) {
        uint64_t n = (*ctr)++;
        uint32_t o[4];
        host_philox((uint32_t)(n >> 2), (uint32_t)(n >> 34), st.stream[0], st.stream[1], st.key[0], st.key[1], o);
        uint32_t k = o[n & 3] >> 8;
        if (k != 0) return (double)k * (1.0 / 16777216.0);
    }
}

// Draws one uniform from the context's stream (advancing it) or returns *u.
static int draw(yamc_ctx *c, const double *u, double *out) {
    if (u) {
        *out = *u;
        return YAMC_OK;
    }
    YdState st;
    TRY(get_state(c, &st));
    uint64_t ctr = st.ctr;
    *out = host_uniform(c, st, &ctr);
    st.ctr = ctr;
    return put_state(c, &st);
}

// Uploads the box-dependent inputs for box length L: hard cores in box units
// and the LJ tables (host copies, kept scaled to the current box), then
// rebuilds beta_LJ (EwaldSumKSpace_initBetaLJ, EwaldSumKSpace.f90:357-386).
static int upload_box_tables(yamc_ctx *c, double L) {
    const int N = c->natom, T = c->ntype;
    std::vector<double> dh(N);
    for (int i = 0; i < N; ++i) dh[i] = c->hard_core_angstr[i] / L;  // EwaldSumRealSpace.f90:359,381
    HIPCHK(hipMemcpyAsync(c->ch.dh, dh.data(), sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
    if (T > 0) {
        HIPCHK(hipMemcpyAsync(c->d_sig6, c->sigma6.data(), sizeof(double) * T * T, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_lj6, c->lj6.data(), sizeof(double) * T * T, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_lj12, c->lj12.data(), sizeof(double) * T * T, hipMemcpyHostToDevice, c->stream));
        long long n = (long long)T * T * c->nk;
        yk_beta_lj<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->ch, c->d_lj6, c->d_lj12, c->d_beta6,
                                                                        c->d_beta12);
        HIPCHK(hipGetLastError());
    }
    return sync(c);  // the host vectors are staging for the copies above
}

// Device buffers saved across a volume step: positions, S(k), per-atom
// real-space arrays, beta, chain state.
static int vol_buffers(yamc_ctx *c, std::vector<std::pair<void *, size_t>> &bufs) {
    const size_t N = c->natom, nS = (size_t)(1 + c->ntype) * 8 * c->nk;
    bufs = {{c->ch.x, N * 8}, {c->ch.y, N * 8}, {c->ch.z, N * 8}, {c->ch.uu, N * 8}, {c->ch.Fx, N * 8},
            {c->ch.Fy, N * 8}, {c->ch.Fz, N * 8}, {c->ch.beta, (size_t)c->nk * 8}, {c->ch.st, sizeof(YdState)}};
    if (c->precision == YAMC_FP32C) {
        bufs.push_back({c->ch.Shi, nS * 4});
        bufs.push_back({c->ch.Slo, nS * 4});
    } else {
        bufs.push_back({c->ch.S, nS * 8});
    }
    if (c->vol_save.empty())
        for (auto &b : bufs) {
            void *q = nullptr;
            HIPCHK(hipMalloc(&q, b.second));
            c->vol_save.push_back(q);
        }
    return YAMC_OK;
}

// The box change of a volume step: scale_box (optional) + Parameters_recalc
// (L_new) + LJTypes_scale_sigma + initBetaLJ (MC.f90:760-770: sigma scales
// by sc = L_old / L_new, so sigma^6 and LJ6 by sc^6, LJ12 by sc^12), then
// mc_calc_total_sums with beta scaled by xclb_new / xclb_old.
static int change_box(yamc_ctx *c, double L_new, bool scale_positions, int32_t *overlap) {
    YdParams &p = c->ch.p;
    const double L_old = p.L, sc = L_old / L_new;
    if (scale_positions) {
        yk_scale_box<<<(c->nmol + 255) / 256, 256, 0, c->stream>>>(c->ch, sc);
        HIPCHK(hipGetLastError());
    }
    const double xclb_new = (p.xclb * L_old) / L_new;
    p.L = L_new;
    p.kext = p.kext * (xclb_new / p.xclb);
    p.m2kext = -2 * p.kext;
    p.dr = (p.dr * L_old) / L_new;
    const double scale_beta = xclb_new / p.xclb;
    p.xclb = xclb_new;
    const double sc2 = sc * sc, sc6 = sc2 * sc2 * sc2;
    for (size_t t = 0; t < c->lj6.size(); ++t) {
        c->sigma6[t] = c->sigma6[t] * sc6;
        c->lj6[t] = c->lj6[t] * sc6;
        c->lj12[t] = c->lj12[t] * sc6 * sc6;
    }
    TRY(upload_box_tables(c, L_new));
    return calc_total_sums(c, scale_beta, overlap);
}

extern "C" int yamc_set_box(yamc_ctx *c, double box_length, int32_t scale_positions, int32_t *overlap) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    if (!(box_length > 0.0) || !std::isfinite(box_length))
        return fail(YAMC_ERROR_PARAMETER, "box length %g must be positive", box_length);
    TRY(sync(c));
    int32_t ov = 0;
    TRY(change_box(c, box_length, scale_positions != 0, &ov));
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    if (overlap) *overlap = ov;
    return YAMC_OK;
}

extern "C" int yamc_mc_volume_move(yamc_ctx *c, double pressure_kT_A3, double log_max_volume_scaling,
                                   const double *u, int32_t *accepted, double *box_length) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    TRY(sync(c));
    YdParams &p = c->ch.p;
    const double L_old = p.L, V_old = L_old * L_old * L_old;
    double r1;
    TRY(draw(c, u, &r1));
    const double lambda = 2.0 * r1 - 1.0;  // MC.f90:742-748
    const double V_new = exp(log(V_old) + lambda * log_max_volume_scaling);
    const double L_new = pow(V_new, 1.0 / 3.0);
    double e[10];
    TRY(yamc_mc_energies(c, e));
    const double U_old = e[0];
    // mc_save_state
    std::vector<std::pair<void *, size_t>> bufs;
    TRY(vol_buffers(c, bufs));
    for (size_t i = 0; i < bufs.size(); ++i)
        HIPCHK(hipMemcpyAsync(c->vol_save[i], bufs[i].first, bufs[i].second, hipMemcpyDeviceToDevice, c->stream));
    const YdParams p_old = p;
    const std::vector<double> lj6_old = c->lj6, lj12_old = c->lj12, sig6_old = c->sigma6;
    int32_t ov = 0;
    TRY(change_box(c, L_new, true, &ov));
    double proba = -1.0;
    if (!ov) {
        TRY(yamc_mc_energies(c, e));
        const double dU = e[0] - U_old, dV = V_new - V_old;
        const double arg = -dU - pressure_kT_A3 * dV + (double)(c->nmol + 1) * log(V_new / V_old);
        proba = exp(arg);
    }
    // MC.f90:804 `(proba > 1) .or. (rand() < proba)`: the reference as
    // compiled evaluates both operands, so the acceptance draw is taken after
    // every volume step (pinned by the reference's own run_mc stream)
    double r2;
    TRY(draw(c, u, &r2));
    const int acc = proba > 1.0 || r2 < proba;
    if (!acc) {  // mc_restore_state and the old box's parameters (MC.f90:806-812)
        YdState st, sv;
        TRY(get_state(c, &st));  // the generator keeps its position
        for (size_t i = 0; i < bufs.size(); ++i)
            HIPCHK(hipMemcpyAsync(bufs[i].first, c->vol_save[i], bufs[i].second, hipMemcpyDeviceToDevice, c->stream));
        TRY(get_state(c, &sv));
        sv.ctr = st.ctr;
        TRY(put_state(c, &sv));
        p = p_old;
        c->lj6 = lj6_old;
        c->lj12 = lj12_old;
        c->sigma6 = sig6_old;
        TRY(upload_box_tables(c, L_old));
    }
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    TRY(sync(c));
    if (accepted) *accepted = acc;
    if (box_length) *box_length = p.L;
    return YAMC_OK;
}

extern "C" int yamc_mc_run_npt(yamc_ctx *c, int32_t ncycles, double pressure_step_multiplier,
                               double pressure_kT_A3, double log_max_volume_scaling, int64_t *n_accepted,
                               int64_t *n_vol_steps, int64_t *n_vol_accepted) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    int64_t nacc = 0, nv = 0, nva = 0;
    for (int cyc = 0; cyc < ncycles; ++cyc) {
        // runmc.f90:206-208: one draw per cycle chooses a volume step or a
        // cycle of moves; the move kernel takes that draw itself, so it is
        // only looked at here
        YdState st;
        TRY(get_state(c, &st));
        uint64_t ctr = st.ctr;
        const double rnd = host_uniform(c, st, &ctr) * (pressure_step_multiplier + 1);
        if (pressure_step_multiplier > 0 && rnd > pressure_step_multiplier) {
            st.ctr = ctr;
            TRY(put_state(c, &st));
            int32_t a = 0;
            TRY(yamc_mc_volume_move(c, pressure_kT_A3, log_max_volume_scaling, nullptr, &a, nullptr));
            ++nv;
            nva += a;
        } else {
            int64_t n = 0;
            TRY(yamc_mc_run(c, 1, 0, nullptr, &n));
            nacc += n;
        }
    }
    if (n_accepted) *n_accepted = nacc;
    if (n_vol_steps) *n_vol_steps = nv;
    if (n_vol_accepted) *n_vol_accepted = nva;
    return check_list(c);
}

extern "C" int yamc_box_length(yamc_ctx *c, double *box_length) {
    if (!c || !box_length) return fail(YAMC_ERROR_PARAMETER, "null argument");
    *box_length = c->ch.p.L;
    return YAMC_OK;
}

// ---- exchange steps ---------------------------------------------------------
static int xchg_buffers(yamc_ctx *c) {
    if (c->d_xd) return YAMC_OK;
    const size_t nS = (size_t)(1 + c->ntype) * 8 * c->nk, N = c->natom;
    HIPCHK(hipMalloc((void **)&c->d_xd, sizeof(double) * nS));
    HIPCHK(hipMalloc((void **)&c->d_xb, sizeof(double) * 16 * N));
    HIPCHK(hipMalloc((void **)&c->d_xmv, sizeof(double) * 8 * YD_MAXA));
    return YAMC_OK;
}

static int write_pos(yamc_ctx *c, int f, int n, const double *x, const double *y, const double *z) {
    HIPCHK(hipMemcpyAsync(c->ch.x + f, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ch.y + f, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ch.z + f, z, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

// EwaldSumRealSpace_calc of molecule imol at (x,y,z), partial mode, with the
// other atoms where c.x/y/z hold them: o[a] uu, o[YD_MAXA + 3a + d] F,
// o[4 YD_MAXA + 0..2] uu_ew, uu_lj6, uu_lj12, o[YK_NRS] overlap; back
// arrays into back[4N] on the device.
static int partial_dev(yamc_ctx *c, int imol, const double *x, const double *y, const double *z, double *back,
                       double *o) {
    const int na = c->h_last[imol] - c->h_first[imol] + 1, N = c->natom;
    TRY(stage_mol(c, na, x, y, z));
    const double *dm = c->d_mol;
    yk_rspace_partial<<<1, YK_BS, 0, c->stream>>>(c->ch, imol, dm, dm + YD_MAXA, dm + 2 * YD_MAXA, c->d_sig6,
                                                  c->d_4eps, c->d_io, back, back + N, back + 2 * N, back + 3 * N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(o, c->d_io, sizeof(double) * (YK_NRS + 1), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

static int pair_term(yamc_ctx *c, int i1, int i2, double *o3) {
    yk_rspace_pair<<<1, 64, 0, c->stream>>>(c->ch, i1, i2, c->d_sig6, c->d_4eps, c->d_io);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(o3, c->d_io, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

// geometry.f90:170-201 center_of_mass
static void com(const double *x, const double *y, const double *z, const double *m, int n, double *o) {
    double sx = 0, sy = 0, sz = 0, tm = 0;
    for (int i = 0; i < n; ++i) {
        sx = sx + m[i] * x[i];
        sy = sy + m[i] * y[i];
        sz = sz + m[i] * z[i];
        tm = tm + m[i];
    }
    o[0] = sx / tm; o[1] = sy / tm; o[2] = sz / tm;
}

// mc_xchange (MC.f90:1325-1954) of molecules i1 != i2: they swap centres of
// mass, each keeping its orientation.  Real space: partial sums of each
// molecule with both at the new places and their pair term; k space: dU of
// the joint delta of both molecules' sums; external dU; accepted iff
// rand() <= exp(-dU) (MonteCarloMove.f90:572-597), no draw after an overlap.
// The acceptance dU_real = sum(uu new) - sum(uu old) over both molecules
// counts the pair term twice, as in the reference; the stored components
// subtract it once (MC.f90:1829-1859).
static int xchange_step(yamc_ctx *c, int i1, int i2, const double *u, yamc_move_rec *rec, int *accepted) {
    TRY(xchg_buffers(c));
    const int N = c->natom, f1 = c->h_first[i1], n1 = c->h_last[i1] - f1 + 1, f2 = c->h_first[i2],
              n2 = c->h_last[i2] - f2 + 1;
    double mu[3];
    TRY(yamc_external_mu(c, mu));  // from scratch (MC.f90:1429)
    double xo1[YD_MAXA], yo1[YD_MAXA], zo1[YD_MAXA], xo2[YD_MAXA], yo2[YD_MAXA], zo2[YD_MAXA], uo1[YD_MAXA],
        uo2[YD_MAXA];
    const std::pair<double *, int> reads[8] = {{xo1, 0}, {yo1, 1}, {zo1, 2}, {uo1, 3},
                                               {xo2, 0}, {yo2, 1}, {zo2, 2}, {uo2, 3}};
    const double *src[4] = {c->ch.x, c->ch.y, c->ch.z, c->ch.uu};
    for (int k = 0; k < 8; ++k) {
        const int f = k < 4 ? f1 : f2, n = k < 4 ? n1 : n2;
        HIPCHK(hipMemcpyAsync(reads[k].first, src[reads[k].second] + f, sizeof(double) * n, hipMemcpyDeviceToHost,
                              c->stream));
    }
    TRY(sync(c));
    double s1 = 0.0, s2 = 0.0;
    for (int a = 0; a < n1; ++a) s1 = s1 + uo1[a];
    for (int a = 0; a < n2; ++a) s2 = s2 + uo2[a];
    const double U_old = s1 + s2;
    const double *q1 = c->h_q.data() + f1, *q2 = c->h_q.data() + f2;
    double muo1[3] = {0, 0, 0}, muo2[3] = {0, 0, 0};
    for (int a = 0; a < n1; ++a) {
        muo1[0] = muo1[0] + q1[a] * xo1[a]; muo1[1] = muo1[1] + q1[a] * yo1[a]; muo1[2] = muo1[2] + q1[a] * zo1[a];
    }
    for (int a = 0; a < n2; ++a) {
        muo2[0] = muo2[0] + q2[a] * xo2[a]; muo2[1] = muo2[1] + q2[a] * yo2[a]; muo2[2] = muo2[2] + q2[a] * zo2[a];
    }
    double c1[3], c2[3];  // MC.f90:1514-1531
    com(xo1, yo1, zo1, c->h_mass.data() + f1, n1, c1);
    com(xo2, yo2, zo2, c->h_mass.data() + f2, n2, c2);
    double xn1[YD_MAXA], yn1[YD_MAXA], zn1[YD_MAXA], xn2[YD_MAXA], yn2[YD_MAXA], zn2[YD_MAXA];
    for (int a = 0; a < n1; ++a) {
        xn1[a] = xo1[a] - c1[0] + c2[0]; yn1[a] = yo1[a] - c1[1] + c2[1]; zn1[a] = zo1[a] - c1[2] + c2[2];
    }
    for (int a = 0; a < n2; ++a) {
        xn2[a] = xo2[a] - c2[0] + c1[0]; yn2[a] = yo2[a] - c2[1] + c1[1]; zn2[a] = zo2[a] - c2[2] + c1[2];
    }
    // new partial sums with both molecules at their new places (MC.f90:1547-1584)
    double *bn1 = c->d_xb, *bn2 = bn1 + 4 * (size_t)N, *bo1 = bn2 + 4 * (size_t)N, *bo2 = bo1 + 4 * (size_t)N;
    double on1[YK_NRS + 1], on2[YK_NRS + 1], cn12[3];
    TRY(write_pos(c, f1, n1, xn1, yn1, zn1));
    TRY(write_pos(c, f2, n2, xn2, yn2, zn2));
    TRY(partial_dev(c, i1, xn1, yn1, zn1, bn1, on1));
    TRY(partial_dev(c, i2, xn2, yn2, zn2, bn2, on2));
    TRY(pair_term(c, i1, i2, cn12));
    TRY(write_pos(c, f1, n1, xo1, yo1, zo1));
    TRY(write_pos(c, f2, n2, xo2, yo2, zo2));
    yamc_move_rec r = {};
    r.imol = i1;
    r.imol2 = i2;
    int acc = 0;
    YdState st;
    if (on1[YK_NRS] != 0.0 || on2[YK_NRS] != 0.0) {
        r.overlap = 1;
        TRY(get_state(c, &st));
        ++st.n_moves;
        ++st.n_overlap;
        TRY(put_state(c, &st));
    } else {
        double v1 = 0.0, v2 = 0.0;
        for (int a = 0; a < n1; ++a) v1 = v1 + on1[a];
        for (int a = 0; a < n2; ++a) v2 = v2 + on2[a];
        const double du_real = (v1 + v2) - U_old;
        // joint delta of the two molecules' sums and its dU (MC.f90:1602-1688)
        YxAtoms at = {};
        at.n = n1 + n2;
        for (int a = 0; a < at.n; ++a) {
            const bool m1 = a < n1;
            const int g = m1 ? f1 + a : f2 + a - n1, b = m1 ? a : a - n1;
            at.typ[a] = c->h_type[g];
            at.q[a] = c->h_q[g];
            if (at.typ[a] > 0) at.pmask |= 1u << (at.typ[a] - 1);
            at.pos[0][0][a] = m1 ? xo1[b] : xo2[b]; at.pos[0][1][a] = m1 ? yo1[b] : yo2[b];
            at.pos[0][2][a] = m1 ? zo1[b] : zo2[b];
            at.pos[1][0][a] = m1 ? xn1[b] : xn2[b]; at.pos[1][1][a] = m1 ? yn1[b] : yn2[b];
            at.pos[1][2][a] = m1 ? zn1[b] : zn2[b];
        }
        yk_xchg_delta<<<(c->nk + 255) / 256, 256, 0, c->stream>>>(c->ch, at, c->d_xd);
        if (c->precision == YAMC_FP32C) yk_xchg_du<true><<<1, 256, 0, c->stream>>>(c->ch, c->d_xd, at.pmask, c->d_io);
        else yk_xchg_du<false><<<1, 256, 0, c->stream>>>(c->ch, c->d_xd, at.pmask, c->d_io);
        HIPCHK(hipGetLastError());
        double du2[2];
        HIPCHK(hipMemcpyAsync(du2, c->d_io, sizeof du2, hipMemcpyDeviceToHost, c->stream));
        TRY(sync(c));
        const double du_kc = du2[0], du_klj = du2[1];
        // external term (MC.f90:1692-1716)
        double mun1[3] = {0, 0, 0}, mun2[3] = {0, 0, 0}, dmu[3], mu_new[3];
        for (int a = 0; a < n1; ++a) {
            mun1[0] = mun1[0] + q1[a] * xn1[a]; mun1[1] = mun1[1] + q1[a] * yn1[a]; mun1[2] = mun1[2] + q1[a] * zn1[a];
        }
        for (int a = 0; a < n2; ++a) {
            mun2[0] = mun2[0] + q2[a] * xn2[a]; mun2[1] = mun2[1] + q2[a] * yn2[a]; mun2[2] = mun2[2] + q2[a] * zn2[a];
        }
        for (int d = 0; d < 3; ++d) dmu[d] = (mun1[d] - muo1[d]) + (mun2[d] - muo2[d]);
        const double kext = c->ch.p.kext;
        const double du_ext =
            kext * ((2 * mu[0] + dmu[0]) * dmu[0] + (2 * mu[1] + dmu[1]) * dmu[1] + (2 * mu[2] + dmu[2]) * dmu[2]);
        for (int d = 0; d < 3; ++d) mu_new[d] = mu[d] + dmu[d];
        const double du_total = du_real + (du_kc + du_klj) + du_ext;
        double xksi;
        TRY(draw(c, u, &xksi));
        acc = !(xksi > exp(-du_total));
        r.du_real = du_real; r.du_kc = du_kc; r.du_klj = du_klj; r.du_ext = du_ext; r.du_total = du_total;
        for (int d = 0; d < 3; ++d) r.delta[d] = c2[d] - c1[d];
        double co12[3] = {0, 0, 0}, oo1[YK_NRS + 1], oo2[YK_NRS + 1];
        if (acc) {  // MC.f90:1754-1872
            if (c->precision == YAMC_FP32C)
                yk_xchg_commit_sums<true><<<(unsigned)((((size_t)(1 + c->ntype) * 8 * c->nk) + 255) / 256), 256, 0,
                                            c->stream>>>(c->ch, c->d_xd, at.pmask);
            else
                yk_xchg_commit_sums<false><<<(unsigned)((((size_t)(1 + c->ntype) * 8 * c->nk) + 255) / 256), 256, 0,
                                             c->stream>>>(c->ch, c->d_xd, at.pmask);
            HIPCHK(hipGetLastError());
            TRY(partial_dev(c, i1, xo1, yo1, zo1, bo1, oo1));
            TRY(partial_dev(c, i2, xo2, yo2, zo2, bo2, oo2));
            TRY(pair_term(c, i1, i2, co12));
            double mv[8 * YD_MAXA] = {};
            for (int a = 0; a < n1; ++a)
                for (int k = 0; k < 4; ++k) mv[4 * a + k] = k == 0 ? on1[a] : on1[YD_MAXA + 3 * a + k - 1];
            for (int a = 0; a < n2; ++a)
                for (int k = 0; k < 4; ++k) mv[4 * (YD_MAXA + a) + k] = k == 0 ? on2[a] : on2[YD_MAXA + 3 * a + k - 1];
            HIPCHK(hipMemcpyAsync(c->d_xmv, mv, sizeof mv, hipMemcpyHostToDevice, c->stream));
            yk_xchg_commit_rs<<<(N + 255) / 256, 256, 0, c->stream>>>(c->ch, bn1, bo1, bn2, bo2, c->d_xmv, f1, n1, f2,
                                                                      n2);
            HIPCHK(hipGetLastError());
            TRY(write_pos(c, f1, n1, xn1, yn1, zn1));
            TRY(write_pos(c, f2, n2, xn2, yn2, zn2));
        }
        TRY(get_state(c, &st));
        ++st.n_moves;
        if (acc) {
            const int E = 4 * YD_MAXA;
            ++st.n_accepted;
            st.e_kc = st.e_kc + du_kc;
            st.e_klj = st.e_klj + du_klj;
            st.uu_ew = st.uu_ew + on1[E] - oo1[E] + on2[E] - oo2[E] - (cn12[0] - co12[0]);
            st.uu_lj6 = st.uu_lj6 + on1[E + 1] - oo1[E + 1] + on2[E + 1] - oo2[E + 1] - (cn12[1] - co12[1]);
            st.uu_lj12 = st.uu_lj12 + on1[E + 2] - oo1[E + 2] + on2[E + 2] - oo2[E + 2] - (cn12[2] - co12[2]);
            for (int d = 0; d < 3; ++d) st.mu[d] = mu_new[d];
        }
        TRY(put_state(c, &st));
    }
    r.accepted = acc;
    if (rec) *rec = r;
    if (accepted) *accepted = acc;
    return YAMC_OK;
}

// BiasedRandom.f90:80-109 on the host, for the exchange picker.
static int host_choose(yamc_ctx *c, const std::vector<double> &beg, const std::vector<double> &end, int *out) {
    double rnd;
    TRY(draw(c, nullptr, &rnd));
    *out = 0;
    for (int i = 0; i < c->nmol; ++i)
        if (rnd > beg[i] && rnd < end[i]) {
            *out = i;
            break;
        }
    return YAMC_OK;
}

// Cycles cyc0 .. cyc0+ncyc-1 of a run with exchange steps (runmc.f90:223-266):
// the chain kernel runs the moves and ends its launch at an exchange step
// (its exchange draw taken); the host picks the two molecules from the
// exchange weights and runs the step, then the next launch continues.
static int run_with_exchanges(yamc_ctx *c, int cyc0, int ncyc, yamc_move_rec *d_rec) {
    const long long first = (long long)cyc0 * c->nmol, total = (long long)ncyc * c->nmol;
    long long done = 0;
    while (done < total) {
        YdState st0, st1;
        TRY(get_state(c, &st0));
        YdChain ch = c->ch;
        ch.rec = d_rec ? d_rec + first + done : nullptr;
        ch.mv_begin = (int)((first + done) % c->nmol);
        HIPCHK(hipMemcpyAsync(c->d_chain, &ch, sizeof ch, hipMemcpyHostToDevice, c->stream));
        const long long rem = total - done;
        const int nsteps = (int)((rem + c->nmol - 1) / c->nmol) + 1;
        TRY(launch_chain(c, c->d_chain, c->d_sig6s, c->d_feps, 1, nsteps, 0, c->natom, (int)rem));
        TRY(get_state(c, &st1));
        done += st1.n_moves - st0.n_moves;
        if (done >= total) break;
        int i1, i2;  // runmc.f90:236-247
        TRY(host_choose(c, c->xfreq_beg, c->xfreq_end, &i1));
        i2 = i1;
        for (int t = 0; i2 == i1; ++t) {
            if (t > 1000000) return fail(YAMC_ERROR_PARAMETER, "exchange weights pick one molecule only");
            TRY(host_choose(c, c->xfreq_beg, c->xfreq_end, &i2));
        }
        yamc_move_rec r;
        TRY(xchange_step(c, i1, i2, nullptr, &r, nullptr));
        if (d_rec) HIPCHK(hipMemcpyAsync(d_rec + first + done, &r, sizeof r, hipMemcpyHostToDevice, c->stream));
        TRY(sync(c));
        ++done;
    }
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

extern "C" int yamc_mc_set_frequencies(yamc_ctx *c, const double *move_weights, const double *xchange_weights) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    const int M = c->nmol;
    std::vector<double> mv(M), xc(M);
    double smv = 0.0, sxc = 0.0;
    int nx = 0;
    for (int i = 0; i < M; ++i) {
        mv[i] = move_weights ? move_weights[i] : 1.0;
        xc[i] = xchange_weights ? xchange_weights[i] : 0.0;
        if (!(mv[i] >= 0.0) || !(xc[i] >= 0.0) || !std::isfinite(mv[i]) || !std::isfinite(xc[i]))
            return fail(YAMC_ERROR_PARAMETER, "weights of molecule %d must be finite and >= 0", i);
        nx += xc[i] > 0.0;
    }
    for (int i = 0; i < M; ++i) smv = smv + mv[i];
    for (int i = 0; i < M; ++i) sxc = sxc + xc[i];
    if (smv + sxc <= 0.0) return fail(YAMC_ERROR_PARAMETER, "all weights are zero");
    if (sxc > 0.0 && nx < 2) return fail(YAMC_ERROR_PARAMETER, "exchange weights must cover at least two molecules");
    std::vector<double> fb(M), fe(M);
    biased_init(M, mv.data(), fb.data(), fe.data());
    c->xfreq_beg.assign(M, 0.0);
    c->xfreq_end.assign(M, 0.0);
    biased_init(M, xc.data(), c->xfreq_beg.data(), c->xfreq_end.data());
    c->pxch = sxc / (smv + sxc);  // runmc.f90:61
    c->ch.p.pxch = c->pxch;
    HIPCHK(hipMemcpyAsync(c->ch.freq_beg, fb.data(), sizeof(double) * M, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ch.freq_end, fe.data(), sizeof(double) * M, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_chain, &c->ch, sizeof(YdChain), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

extern "C" int yamc_mc_xchange(yamc_ctx *c, int32_t imol1, int32_t imol2, const double *u, int32_t *accepted,
                               yamc_move_rec *rec) {
    if (!c) return fail(YAMC_ERROR_PARAMETER, "null context");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    if (imol1 < 0 || imol1 >= c->nmol || imol2 < 0 || imol2 >= c->nmol || imol1 == imol2)
        return fail(YAMC_ERROR_PARAMETER, "exchange needs two distinct molecules in range (got %d, %d)", imol1, imol2);
    TRY(sync(c));
    int acc = 0;
    TRY(xchange_step(c, imol1, imol2, u, rec, &acc));
    if (accepted) *accepted = acc;
    return YAMC_OK;
}

// ---- host-driven moves ------------------------------------------------------
// movemc's commit (MC.f90:1242-1320) for molecule imol moved from its current
// place to (x, y, z): S(k) += the molecule's delta (FP32C: hi/lo re-split),
// per-atom real-space arrays from the back arrays of the old and new partial
// sums, the real-space components, E_kC += du_c, E_kLJ += du_lj (the values
// yamc_kspace_delta returned for the move), positions, mu.
extern "C" int yamc_mc_commit_move(yamc_ctx *c, int32_t imol, const double *x, const double *y, const double *z,
                                   double du_c, double du_lj, int32_t *overlap) {
    if (!c || !x || !y || !z) return fail(YAMC_ERROR_PARAMETER, "null argument");
    if (!c->state_ready) return fail(YAMC_ERROR_INITIALIZATION, "yamc_mc_init has not run on this state");
    if (imol < 0 || imol >= c->nmol) return fail(YAMC_ERROR_PARAMETER, "molecule index %d out of range", imol);
    TRY(sync(c));
    TRY(xchg_buffers(c));
    const int N = c->natom, f = c->h_first[imol], n = c->h_last[imol] - f + 1;
    double xo[YD_MAXA], yo[YD_MAXA], zo[YD_MAXA], xn[YD_MAXA], yn[YD_MAXA], zn[YD_MAXA];
    HIPCHK(hipMemcpyAsync(xo, c->ch.x + f, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(yo, c->ch.y + f, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(zo, c->ch.z + f, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    for (int a = 0; a < n; ++a) { xn[a] = x[a]; yn[a] = y[a]; zn[a] = z[a]; }
    double *bn = c->d_xb, *bo = bn + 4 * (size_t)N;
    double on[YK_NRS + 1], oo[YK_NRS + 1];
    TRY(partial_dev(c, imol, xn, yn, zn, bn, on));
    if (on[YK_NRS] != 0.0) {  // the new place overlaps: nothing is committed
        if (overlap) *overlap = 1;
        return YAMC_OK;
    }
    TRY(partial_dev(c, imol, xo, yo, zo, bo, oo));
    YxAtoms at = {};
    at.n = n;
    for (int a = 0; a < n; ++a) {
        at.typ[a] = c->h_type[f + a];
        at.q[a] = c->h_q[f + a];
        if (at.typ[a] > 0) at.pmask |= 1u << (at.typ[a] - 1);
        at.pos[0][0][a] = xo[a]; at.pos[0][1][a] = yo[a]; at.pos[0][2][a] = zo[a];
        at.pos[1][0][a] = xn[a]; at.pos[1][1][a] = yn[a]; at.pos[1][2][a] = zn[a];
    }
    yk_xchg_delta<<<(c->nk + 255) / 256, 256, 0, c->stream>>>(c->ch, at, c->d_xd);
    const unsigned nsb = (unsigned)((((size_t)(1 + c->ntype) * 8 * c->nk) + 255) / 256);
    if (c->precision == YAMC_FP32C) yk_xchg_commit_sums<true><<<nsb, 256, 0, c->stream>>>(c->ch, c->d_xd, at.pmask);
    else yk_xchg_commit_sums<false><<<nsb, 256, 0, c->stream>>>(c->ch, c->d_xd, at.pmask);
    HIPCHK(hipGetLastError());
    double mv[8 * YD_MAXA] = {};
    for (int a = 0; a < n; ++a)
        for (int k = 0; k < 4; ++k) mv[4 * a + k] = k == 0 ? on[a] : on[YD_MAXA + 3 * a + k - 1];
    HIPCHK(hipMemcpyAsync(c->d_xmv, mv, sizeof mv, hipMemcpyHostToDevice, c->stream));
    yk_xchg_commit_rs<<<(N + 255) / 256, 256, 0, c->stream>>>(c->ch, bn, bo, nullptr, nullptr, c->d_xmv, f, n, 0, 0);
    HIPCHK(hipGetLastError());
    TRY(write_pos(c, f, n, xn, yn, zn));
    double mu[3];
    TRY(yamc_external_mu(c, mu));
    YdState st;
    TRY(get_state(c, &st));
    const int E = 4 * YD_MAXA;
    st.e_kc = st.e_kc + du_c;
    st.e_klj = st.e_klj + du_lj;
    st.uu_ew = st.uu_ew + on[E] - oo[E];
    st.uu_lj6 = st.uu_lj6 + on[E + 1] - oo[E + 1];
    st.uu_lj12 = st.uu_lj12 + on[E + 2] - oo[E + 2];
    TRY(put_state(c, &st));
    if (overlap) *overlap = 0;
    return YAMC_OK;
}
