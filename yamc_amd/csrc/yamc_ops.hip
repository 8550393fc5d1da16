// yamc_ops.hip -- the per-procedure device operations of the C ABI
// (include/yamc_ewald.h, "per-procedure surface"): one call per reference
// procedure of EwaldSumKSpace, SumSinCosKR, RhoSquared, ForceKSpace,
// EwaldSumRealSpace, RealSumLocal and EwaldSumTails, on the caller's host
// arrays.  Each call gathers its inputs into one pinned host block, sends it
// with one copy, runs its kernels and brings its results back with one copy
// before it returns; a workspace (yamc_ws) holds a k grid and the buffers
// the calls reuse.  The Fortran modules of
// fortran/shim/ forward the reference's procedures to these calls.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "yamc_device.h"

int yamc_internal_fail(int code, const char *msg);  // yamc_api.hip: sets yamc_last_error
__global__ void yk_beta(YdChain c, const double *xk2, double *beta6, double *beta12);

static int ofail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return yamc_internal_fail(code, buf);
}

#define OCHK(expr)                                                                                         \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return ofail(YAMC_ERROR_DEVICE, "HIP error %s (line %d)", hipGetErrorString(e_), \
                                           __LINE__);                                                      \
    } while (0)
#define OTRY(x)                       \
    do {                              \
        int r_ = (x);                 \
        if (r_ != YAMC_OK) return r_; \
    } while (0)

// Device buffers of a workspace: slot -> (pointer, capacity in bytes).
enum { WB_L, WB_OUT, WB_N };

// A call's inputs are gathered into one pinned host block and sent with one
// copy; its results come back with one copy into a pinned landing block.
struct yamc_ws {
    int device = 0, kmax = 0, nk = 0;
    hipStream_t s = nullptr;
    std::vector<int> kx, ky, kz, ip, iq;
    std::vector<double> xk, xk2;
    int *d_kxyz = nullptr;
    double *d_xk2 = nullptr;
    void *buf[WB_N] = {};
    size_t cap[WB_N] = {};
    char *h_in = nullptr, *d_in = nullptr, *h_out = nullptr;  // staging (pinned), its device image, landing
    size_t h_in_cap = 0, d_in_cap = 0, h_out_cap = 0, in_used = 0;
};

// Room for n bytes in slot b (grown, never shrunk).
static int wbuf(yamc_ws *w, int b, size_t n, void **out) {
    if (w->cap[b] < n) {
        if (w->buf[b]) OCHK(hipFree(w->buf[b]));
        w->buf[b] = nullptr;
        w->cap[b] = 0;
        OCHK(hipMalloc(&w->buf[b], n));
        w->cap[b] = n;
    }
    *out = w->buf[b];
    return YAMC_OK;
}

// Pinned host block of at least n bytes keeping its first keep bytes.  Every
// call ends synchronised with its stream, so no copy still reads the block.
static int pinned(char **p, size_t *cap, size_t n, size_t keep) {
    if (*cap >= n) return YAMC_OK;
    size_t c = *cap ? *cap : ((size_t)1 << 16);
    while (c < n) c *= 2;
    char *q = nullptr;
    OCHK(hipHostMalloc((void **)&q, c));
    if (keep) memcpy(q, *p, keep);
    if (*p) OCHK(hipHostFree(*p));
    *p = q;
    *cap = c;
    return YAMC_OK;
}

static void stage_begin(yamc_ws *w) { w->in_used = 0; }

// Appends n bytes at h to the call's input block; *off = their offset.
static int stage_put(yamc_ws *w, const void *h, size_t n, size_t *off) {
    const size_t o = (w->in_used + 255) & ~(size_t)255;
    OTRY(pinned(&w->h_in, &w->h_in_cap, o + (n ? n : 1), w->in_used));
    if (n && h) memcpy(w->h_in + o, h, n);  // h NULL: room the caller fills
    w->in_used = o + (n ? n : 1);
    *off = o;
    return YAMC_OK;
}

// m arrays of nk doubles (host pointers) as one block [m][nk].
static int stage_rows(yamc_ws *w, const double *const *rows, int m, size_t *off) {
    const size_t row = sizeof(double) * w->nk, o = (w->in_used + 255) & ~(size_t)255;
    const size_t n = row * (m ? m : 1);
    OTRY(pinned(&w->h_in, &w->h_in_cap, o + n, w->in_used));
    for (int i = 0; i < m; ++i) {
        if (!rows[i]) return ofail(YAMC_ERROR_PARAMETER, "null array");
        memcpy(w->h_in + o + row * i, rows[i], row);
    }
    w->in_used = o + n;
    *off = o;
    return YAMC_OK;
}

// One copy of the staged inputs to the device.
static int stage_flush(yamc_ws *w) {
    if (w->d_in_cap < w->in_used) {
        if (w->d_in) OCHK(hipFree(w->d_in));
        w->d_in = nullptr;
        w->d_in_cap = 0;
        size_t c = (size_t)1 << 16;
        while (c < w->in_used) c *= 2;
        OCHK(hipMalloc((void **)&w->d_in, c));
        w->d_in_cap = c;
    }
    if (w->in_used) OCHK(hipMemcpyAsync(w->d_in, w->h_in, w->in_used, hipMemcpyHostToDevice, w->s));
    return YAMC_OK;
}

template <typename T>
static T *stage_dev(yamc_ws *w, size_t off) {
    return (T *)(w->d_in + off);
}

static int wsync(yamc_ws *w);

// n bytes of device memory d to the pinned landing block, synchronised.
static int fetch(yamc_ws *w, const void *d, size_t n, const char **h) {
    OTRY(pinned(&w->h_out, &w->h_out_cap, n ? n : 1, 0));
    if (n) OCHK(hipMemcpyAsync(w->h_out, d, n, hipMemcpyDeviceToHost, w->s));
    OTRY(wsync(w));
    *h = w->h_out;
    return YAMC_OK;
}

// Rows [m][nk] of a fetched block to the caller's arrays (NULL rows skipped).
static void scatter_rows(const yamc_ws *w, const char *h, double *const *rows, int m) {
    for (int i = 0; i < m; ++i)
        if (rows[i]) memcpy(rows[i], h + sizeof(double) * w->nk * i, sizeof(double) * w->nk);
}

static int wsync(yamc_ws *w) {
    OCHK(hipStreamSynchronize(w->s));
    return YAMC_OK;
}

// FourierGrid.f90:37-189 (FourierGrid_calcNalloc + FourierGrid_init):
// half-space k list in row order, ip = 1 at the end of a ky row, 2 at the
// end of a kx plane, 3 at the end; iq the zero pattern; xk = |k|, xk2 = k^2.
static void grid_walk(int kmax, yamc_ws *w) {
    const double xkmax2 = (double)kmax * kmax;
    for (int ix = 0; ix <= kmax; ++ix) {
        const double xkx2 = (double)(ix * ix);
        const int kymax = (int)sqrt(xkmax2 - xkx2);
        for (int iy = 0; iy <= kymax; ++iy) {
            const double xky2 = (double)(iy * iy);
            const int kzmax = (int)sqrt(xkmax2 - xkx2 - xky2);
            for (int iz = 0; iz <= kzmax; ++iz) {
                const double k2 = xkx2 + xky2 + (double)(iz * iz);
                if (k2 < 0.1) continue;
                w->kx.push_back(ix); w->ky.push_back(iy); w->kz.push_back(iz);
                w->xk2.push_back(k2); w->xk.push_back(sqrt(k2));
                w->iq.push_back(iy == 0 ? (iz == 0 ? 1 : 3) : (iz == 0 ? 2 : 4));
                w->ip.push_back(0);
            }
            if (!w->ip.empty()) w->ip.back() = 1;
        }
        if (!w->ip.empty()) w->ip.back() = 2;
    }
    if (!w->ip.empty()) w->ip.back() = 3;
}

extern "C" int yamc_ws_create(int32_t kmax, yamc_ws **out) {
    if (!out) return ofail(YAMC_ERROR_PARAMETER, "null output");
    *out = nullptr;
    if (kmax < 1 || kmax > YAMC_MAX_KMAX) return ofail(YAMC_ERROR_LIMITS, "kmax %d outside 1..%d", kmax, YAMC_MAX_KMAX);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ofail(YAMC_ERROR_DEVICE, "no HIP device available");
    yamc_ws *w = new yamc_ws();
    auto go = [&]() -> int {
        OCHK(hipGetDevice(&w->device));
        OCHK(hipStreamCreateWithFlags(&w->s, hipStreamNonBlocking));
        w->kmax = kmax;
        grid_walk(kmax, w);
        w->nk = (int)w->kx.size();
        std::vector<int> kxyz(w->nk);
        for (int k = 0; k < w->nk; ++k) kxyz[k] = w->kx[k] | (w->ky[k] << 8) | (w->kz[k] << 16);
        OCHK(hipMalloc((void **)&w->d_kxyz, sizeof(int) * w->nk));
        OCHK(hipMalloc((void **)&w->d_xk2, sizeof(double) * w->nk));
        OCHK(hipMemcpy(w->d_kxyz, kxyz.data(), sizeof(int) * w->nk, hipMemcpyHostToDevice));
        OCHK(hipMemcpy(w->d_xk2, w->xk2.data(), sizeof(double) * w->nk, hipMemcpyHostToDevice));
        return YAMC_OK;
    };
    int rc = go();
    if (rc != YAMC_OK) {
        yamc_ws_destroy(w);
        return rc;
    }
    *out = w;
    return YAMC_OK;
}

extern "C" int yamc_ws_destroy(yamc_ws *w) {
    if (!w) return YAMC_OK;
    if (w->s) (void)hipStreamSynchronize(w->s);
    for (int b = 0; b < WB_N; ++b)
        if (w->buf[b]) (void)hipFree(w->buf[b]);
    if (w->d_kxyz) (void)hipFree(w->d_kxyz);
    if (w->d_xk2) (void)hipFree(w->d_xk2);
    if (w->d_in) (void)hipFree(w->d_in);
    if (w->h_in) (void)hipHostFree(w->h_in);
    if (w->h_out) (void)hipHostFree(w->h_out);
    if (w->s) (void)hipStreamDestroy(w->s);
    delete w;
    return YAMC_OK;
}

extern "C" int yamc_ws_grid(const yamc_ws *w, int32_t *nk, int32_t *kx, int32_t *ky, int32_t *kz, int32_t *ip,
                            int32_t *iq, double *xk, double *xk2) {
    if (!w) return ofail(YAMC_ERROR_PARAMETER, "null workspace");
    if (nk) *nk = w->nk;
    for (int k = 0; k < w->nk; ++k) {
        if (kx) kx[k] = w->kx[k];
        if (ky) ky[k] = w->ky[k];
        if (kz) kz[k] = w->kz[k];
        if (ip) ip[k] = w->ip[k];
        if (iq) iq[k] = w->iq[k];
        if (xk) xk[k] = w->xk[k];
        if (xk2) xk2[k] = w->xk2[k];
    }
    return YAMC_OK;
}

// ---- k space ---------------------------------------------------------------

extern "C" int yamc_op_beta(yamc_ws *w, double alpha, double xclb, double *beta, double *beta6, double *beta12) {
    if (!w || !beta || !beta6 || !beta12) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (!(alpha > 0.0)) return ofail(YAMC_ERROR_PARAMETER, "alpha must be positive");
    YdChain c{};
    c.p.nk = w->nk;
    c.p.alpha = alpha;
    c.p.xclb = xclb;
    c.kxyz = w->d_kxyz;
    const size_t n = sizeof(double) * w->nk;
    void *po = nullptr;
    OTRY(wbuf(w, WB_OUT, 3 * n, &po));
    double *pb = (double *)po;
    c.beta = pb;
    yk_beta<<<w->nk, 64, 0, w->s>>>(c, w->d_xk2, pb + w->nk, pb + 2 * w->nk);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, 3 * n, &h));
    memcpy(beta, h, n);
    memcpy(beta6, h + n, n);
    memcpy(beta12, h + 2 * n, n);
    return YAMC_OK;
}

__global__ void yo_beta_lj(int nk, double lj6, double lj12, const double *b6, const double *b12, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) out[k] = lj12 * b12[k] - lj6 * b6[k];
}

extern "C" int yamc_op_beta_lj(yamc_ws *w, double lj6, double lj12, const double *beta6, const double *beta12,
                               double *out) {
    if (!w || !beta6 || !beta12 || !out) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    const size_t n = sizeof(double) * w->nk;
    size_t o6, o12;
    void *po = nullptr;
    stage_begin(w);
    OTRY(stage_put(w, beta6, n, &o6));
    OTRY(stage_put(w, beta12, n, &o12));
    OTRY(stage_flush(w));
    OTRY(wbuf(w, WB_OUT, n, &po));
    yo_beta_lj<<<(w->nk + 255) / 256, 256, 0, w->s>>>(w->nk, lj6, lj12, stage_dev<double>(w, o6),
                                                      stage_dev<double>(w, o12), (double *)po);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, n, &h));
    memcpy(out, h, n);
    return YAMC_OK;
}

// SinCosXYZ_fill (SumSinCosKR.f90:181-223) of one position: tab = cos x, sin
// x, cos y, sin y, cos z, sin z of 2 pi k r, k = 0..kmax, stride kp.
__global__ void yo_sincos_xyz(double x, double y, double z, int kmax, double *tab) {
    const int d = threadIdx.x;
    if (d >= 3) return;
    const int kp = kmax + 1;
    yd_sincos_axis(d == 0 ? x : d == 1 ? y : z, kmax, tab + 2 * d * kp, tab + (2 * d + 1) * kp);
}

// SinCosKR_fill (SumSinCosKR.f90:297-455) from the axis tables: kr[8][nk].
__global__ void yo_sincos_kr(int nk, const int *kxyz, const double *tab, int kp, double q, double *kr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const int v = kxyz[k];
    double o[8];
    yd_kr_s(tab, kp, q, v & 255, (v >> 8) & 255, (v >> 16) & 255, o);
#pragma unroll
    for (int m = 0; m < 8; ++m) kr[(size_t)m * nk + k] = o[m];
}

// tab6 order cos x, cos y, cos z, sin x, sin y, sin z; device order cx, sx,
// cy, sy, cz, sz
static const int kTabSlot[6] = {0, 2, 4, 1, 3, 5};

extern "C" int yamc_op_sincos_xyz(yamc_ws *w, double x, double y, double z, double *const *tab6) {
    if (!w || !tab6) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    for (int i = 0; i < 6; ++i)
        if (!tab6[i]) return ofail(YAMC_ERROR_PARAMETER, "null table");
    const int kp = w->kmax + 1;
    void *pt = nullptr;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 6 * kp, &pt));
    yo_sincos_xyz<<<1, 64, 0, w->s>>>(x, y, z, w->kmax, (double *)pt);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, pt, sizeof(double) * 6 * kp, &h));
    for (int i = 0; i < 6; ++i) memcpy(tab6[i], h + sizeof(double) * kTabSlot[i] * kp, sizeof(double) * kp);
    return YAMC_OK;
}

extern "C" int yamc_op_sincos_kr(yamc_ws *w, const double *const *tab6, double q, double *const *kr8) {
    if (!w || !tab6 || !kr8) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    const int kp = w->kmax + 1;
    stage_begin(w);
    size_t ot;
    OTRY(stage_put(w, nullptr, sizeof(double) * 6 * kp, &ot));
    for (int i = 0; i < 6; ++i) {
        if (!tab6[i]) return ofail(YAMC_ERROR_PARAMETER, "null table");
        memcpy(w->h_in + ot + sizeof(double) * kTabSlot[i] * kp, tab6[i], sizeof(double) * kp);
    }
    OTRY(stage_flush(w));
    void *pk = nullptr;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 8 * w->nk, &pk));
    yo_sincos_kr<<<(w->nk + 255) / 256, 256, 0, w->s>>>(w->nk, w->d_kxyz, stage_dev<double>(w, ot), kp, q,
                                                        (double *)pk);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, pk, sizeof(double) * 8 * w->nk, &h));
    scatter_rows(w, h, kr8, 8);
    return YAMC_OK;
}

extern "C" int yamc_op_sincos_kr_atom(yamc_ws *w, double x, double y, double z, double q, double *const *kr8) {
    if (!w || !kr8) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    const int kp = w->kmax + 1;
    void *pt = nullptr, *pk = nullptr;
    OTRY(wbuf(w, WB_L, sizeof(double) * 6 * kp, &pt));
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 8 * w->nk, &pk));
    yo_sincos_xyz<<<1, 64, 0, w->s>>>(x, y, z, w->kmax, (double *)pt);
    yo_sincos_kr<<<(w->nk + 255) / 256, 256, 0, w->s>>>(w->nk, w->d_kxyz, (double *)pt, kp, q, (double *)pk);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, pk, sizeof(double) * 8 * w->nk, &h));
    scatter_rows(w, h, kr8, 8);
    return YAMC_OK;
}

// RhoSquared_addAtoms' sums (RhoSquared.f90:160-243): thread per (set, k),
// the set's atoms in index order; set 0 = Coulomb (q sin/cos of charged
// atoms), set t = LJ type t (sin/cos, or (q sin/cos) (1/q) of a charged atom).
// The axis tables of every atom: tab[a][6][kp].
__global__ void yo_tables(int n, const double *x, const double *y, const double *z, int kp, double *tab) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double *t = tab + (size_t)a * 6 * kp;
    yd_sincos_axis(x[a], kp - 1, t, t + kp);
    yd_sincos_axis(y[a], kp - 1, t + 2 * kp, t + 3 * kp);
    yd_sincos_axis(z[a], kp - 1, t + 4 * kp, t + 5 * kp);
}

__global__ void yo_sums(int nk, const int *kxyz, const double *tab, const double *q, const int *typ, int nset,
                        const int *soff, const int *satoms, int kp, double *sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nset * nk) return;
    const int set = i / nk, k = i % nk, v = kxyz[k];
    const int kx = v & 255, ky = (v >> 8) & 255, kz = (v >> 16) & 255;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = soff[set]; t < soff[set + 1]; ++t) {
        const int a = satoms[t];
        double krc[8], krl[8];
        yd_kr_atom_s(tab + (size_t)a * 6 * kp, kp, q[a], typ[a], kx, ky, kz, krc, krl);
        const double *kr = set == 0 ? krc : krl;
#pragma unroll
        for (int m = 0; m < 8; ++m) s[m] = s[m] + kr[m];
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) sums[((size_t)set * 8 + m) * nk + k] = s[m];
}

extern "C" int yamc_op_sums(yamc_ws *w, int32_t n, const double *x, const double *y, const double *z,
                            const double *q, const int32_t *lj_type, int32_t ntype, double *const *sums,
                            int32_t *present) {
    if (!w || (n > 0 && (!x || !y || !z || !q || !lj_type)) || !sums) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (n < 0 || ntype < 0 || ntype > YAMC_MAX_LJ_TYPE) return ofail(YAMC_ERROR_PARAMETER, "bad sizes");
    const int nset = 1 + ntype;
    std::vector<int> soff(nset + 1, 0), satoms;
    for (int set = 0; set < nset; ++set) {
        soff[set] = (int)satoms.size();
        for (int a = 0; a < n; ++a) {
            if (lj_type[a] < 0 || lj_type[a] > ntype) return ofail(YAMC_ERROR_PARAMETER, "bad LJ type at atom %d", a);
            if (set == 0 ? q[a] != 0.0 : lj_type[a] == set) satoms.push_back(a);
        }
    }
    soff[nset] = (int)satoms.size();
    if (present)
        for (int t = 1; t <= ntype; ++t) present[t - 1] = soff[t + 1] > soff[t];
    size_t ox, oy, oz, oq, ot, ooff, oat;
    stage_begin(w);
    OTRY(stage_put(w, x, sizeof(double) * n, &ox));
    OTRY(stage_put(w, y, sizeof(double) * n, &oy));
    OTRY(stage_put(w, z, sizeof(double) * n, &oz));
    OTRY(stage_put(w, q, sizeof(double) * n, &oq));
    OTRY(stage_put(w, lj_type, sizeof(int32_t) * n, &ot));
    OTRY(stage_put(w, soff.data(), sizeof(int) * soff.size(), &ooff));
    OTRY(stage_put(w, satoms.data(), sizeof(int) * satoms.size(), &oat));
    OTRY(stage_flush(w));
    void *po = nullptr, *ptab = nullptr;
    const int kp = w->kmax + 1;
    OTRY(wbuf(w, WB_L, sizeof(double) * 6 * (size_t)kp * (n ? n : 1), &ptab));
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 8 * (size_t)nset * w->nk, &po));
    if (n)
        yo_tables<<<(n + 255) / 256, 256, 0, w->s>>>(n, stage_dev<double>(w, ox), stage_dev<double>(w, oy),
                                                     stage_dev<double>(w, oz), kp, (double *)ptab);
    yo_sums<<<(nset * w->nk + 127) / 128, 128, 0, w->s>>>(w->nk, w->d_kxyz, (double *)ptab, stage_dev<double>(w, oq),
                                                          stage_dev<int>(w, ot), nset, stage_dev<int>(w, ooff),
                                                          stage_dev<int>(w, oat), kp, (double *)po);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, sizeof(double) * 8 * (size_t)nset * w->nk, &h));
    scatter_rows(w, h, sums, 8 * nset);
    return YAMC_OK;
}

// Per-k terms reduced per block (deterministic): a block of 256 threads over
// k writes the sums of its NV values to part[block][NV]; the host adds the
// block sums in block order.
template <int NV>
__device__ __forceinline__ void yo_block_out(double (&v)[NV], double *part) {
    __shared__ double red[4 * NV], out[NV];
    yd_block_sum<NV>(v, red, out);
    if (threadIdx.x < NV) part[(size_t)blockIdx.x * NV + threadIdx.x] = out[threadIdx.x];
}

template <int NV>
static int reduce_blocks(yamc_ws *w, const double *dpart, int nblk, double *out) {
    const char *hc;
    OTRY(fetch(w, dpart, sizeof(double) * (size_t)nblk * NV, &hc));
    const double *h = (const double *)hc;
    for (int i = 0; i < NV; ++i) {
        double a = 0.0;
        for (int b = 0; b < nblk; ++b) a = a + h[(size_t)b * NV + i];
        out[i] = a;
    }
    return YAMC_OK;
}

// RhoSquared_calc_rho_squared (RhoSquared.f90:245-273).
__global__ void yo_rho2(int nk, const int *kxyz, const double *S, double *rho2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const int v = kxyz[k], ky = (v >> 8) & 255, kz = (v >> 16) & 255;
    double C[8];
    for (int m = 0; m < 8; ++m) C[m] = S[(size_t)m * nk + k];
    double r = C[0] * C[0] + C[1] * C[1];
    if (ky > 0) r = r + C[4] * C[4] + C[5] * C[5];
    if (kz > 0) r = r + C[2] * C[2] + C[3] * C[3];
    if (ky > 0 && kz > 0) r = r + C[6] * C[6] + C[7] * C[7];
    rho2[k] = r;
}

extern "C" int yamc_op_rho_squared(yamc_ws *w, const double *const *s8, double *rho2) {
    if (!w || !s8 || !rho2) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    size_t oS;
    stage_begin(w);
    OTRY(stage_rows(w, s8, 8, &oS));
    OTRY(stage_flush(w));
    void *po = nullptr;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * w->nk, &po));
    yo_rho2<<<(w->nk + 255) / 256, 256, 0, w->s>>>(w->nk, w->d_kxyz, stage_dev<double>(w, oS), (double *)po);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, sizeof(double) * w->nk, &h));
    memcpy(rho2, h, sizeof(double) * w->nk);
    return YAMC_OK;
}

// calc_coulomb_energy (EwaldSumKSpace.f90:389-416): sum_k beta rho^2, from
// the rho_squared the caller holds (RhoSquared_calc_rho_squared).
__global__ void yo_e_coulomb(int nk, const double *beta, const double *rho2, double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {k < nk ? beta[k] * rho2[k] : 0.0};
    yo_block_out<1>(v, part);
}

extern "C" int yamc_op_energy_coulomb(yamc_ws *w, const double *beta, const double *rho2, double *e) {
    if (!w || !beta || !rho2 || !e) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    size_t ob, orr;
    stage_begin(w);
    OTRY(stage_put(w, beta, sizeof(double) * w->nk, &ob));
    OTRY(stage_put(w, rho2, sizeof(double) * w->nk, &orr));
    OTRY(stage_flush(w));
    void *pp = nullptr;
    const int nblk = (w->nk + 255) / 256;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * nblk, &pp));
    yo_e_coulomb<<<nblk, 256, 0, w->s>>>(w->nk, stage_dev<double>(w, ob), stage_dev<double>(w, orr), (double *)pp);
    OCHK(hipGetLastError());
    return reduce_blocks<1>(w, (const double *)pp, nblk, e);
}

// calc_LJ_energy (EwaldSumKSpace.f90:418-526): sum over t1 <= t2 of
// (2 - delta) beta_LJ^{t1 t2} sum_slots (C1 C2 + S1 S2).
__global__ void yo_e_lj(int nk, const int *kxyz, int T, const double *BL, const double *S, double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (k < nk) {
        const int t = kxyz[k], ky = (t >> 8) & 255, kz = (t >> 16) & 255;
        for (int t1 = 0; t1 < T; ++t1) {
            double A[8];
            for (int m = 0; m < 8; ++m) A[m] = S[((size_t)t1 * 8 + m) * nk + k];
            for (int t2 = t1; t2 < T; ++t2) {
                double B[8];
                for (int m = 0; m < 8; ++m) B[m] = S[((size_t)t2 * 8 + m) * nk + k];
                double loc = A[0] * B[0] + A[1] * B[1];
                if (ky > 0) loc = loc + A[4] * B[4] + A[5] * B[5];
                if (kz > 0) loc = loc + A[2] * B[2] + A[3] * B[3];
                if (ky > 0 && kz > 0) loc = loc + A[6] * B[6] + A[7] * B[7];
                const double term = BL[((size_t)t1 * T + t2) * nk + k] * loc;
                v[0] += t1 != t2 ? term + term : term;
            }
        }
    }
    yo_block_out<1>(v, part);
}

extern "C" int yamc_op_energy_lj(yamc_ws *w, int32_t ntype, const double *const *beta_lj, const double *const *sums_lj,
                                 double *e) {
    if (!w || !e || (ntype > 0 && (!beta_lj || !sums_lj))) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (ntype < 0 || ntype > YAMC_MAX_LJ_TYPE) return ofail(YAMC_ERROR_PARAMETER, "bad ntype");
    if (ntype == 0) {
        *e = 0.0;
        return YAMC_OK;
    }
    size_t oB, oS;
    stage_begin(w);
    OTRY(stage_rows(w, beta_lj, ntype * ntype, &oB));
    OTRY(stage_rows(w, sums_lj, 8 * ntype, &oS));
    OTRY(stage_flush(w));
    void *pp = nullptr;
    const int nblk = (w->nk + 255) / 256;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * nblk, &pp));
    yo_e_lj<<<nblk, 256, 0, w->s>>>(w->nk, w->d_kxyz, ntype, stage_dev<double>(w, oB), stage_dev<double>(w, oS),
                                    (double *)pp);
    OCHK(hipGetLastError());
    return reduce_blocks<1>(w, (const double *)pp, nblk, e);
}

// calc_dU_coulomb (EwaldSumKSpace.f90:528-620).
__global__ void yo_du_coulomb(int nk, const int *kxyz, const double *beta, const double *S, const double *D,
                              double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (k < nk) {
        const int t = kxyz[k], ky = (t >> 8) & 255, kz = (t >> 16) & 255;
        double C[8], E[8];
        for (int m = 0; m < 8; ++m) {
            C[m] = S[(size_t)m * nk + k];
            E[m] = D[(size_t)m * nk + k];
        }
        v[0] = beta[k] * yd_du_k(C, E, ky, kz);
    }
    yo_block_out<1>(v, part);
}

extern "C" int yamc_op_du_coulomb(yamc_ws *w, const double *beta, const double *const *s8, const double *const *d8,
                                  double *du) {
    if (!w || !beta || !s8 || !d8 || !du) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    size_t ob, oS, oD;
    stage_begin(w);
    OTRY(stage_put(w, beta, sizeof(double) * w->nk, &ob));
    OTRY(stage_rows(w, s8, 8, &oS));
    OTRY(stage_rows(w, d8, 8, &oD));
    OTRY(stage_flush(w));
    void *pp = nullptr;
    const int nblk = (w->nk + 255) / 256;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * nblk, &pp));
    yo_du_coulomb<<<nblk, 256, 0, w->s>>>(w->nk, w->d_kxyz, stage_dev<double>(w, ob), stage_dev<double>(w, oS),
                                          stage_dev<double>(w, oD), (double *)pp);
    OCHK(hipGetLastError());
    return reduce_blocks<1>(w, (const double *)pp, nblk, du);
}

// calc_dU_LJ (EwaldSumKSpace.f90:623-753): for every t1 (all types) and every
// t2 whose delta is present, beta_LJ^{t1 t2} sum_slots (2 C^t1 + dC^t1) dC^t2.
__global__ void yo_du_lj(int nk, const int *kxyz, int T, unsigned pmask, const double *BL, const double *S,
                         const double *D, double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (k < nk) {
        const int t = kxyz[k], ky = (t >> 8) & 255, kz = (t >> 16) & 255;
        for (int t1 = 0; t1 < T; ++t1) {
            double tw[8];
            for (int m = 0; m < 8; ++m) {
                const double c1 = S[((size_t)t1 * 8 + m) * nk + k], d1 = D[((size_t)t1 * 8 + m) * nk + k];
                tw[m] = c1 + c1 + d1;
            }
            for (int t2 = 0; t2 < T; ++t2) {
                if (!((pmask >> t2) & 1u)) continue;
                double D2[8];
                for (int m = 0; m < 8; ++m) D2[m] = D[((size_t)t2 * 8 + m) * nk + k];
                v[0] += BL[((size_t)t1 * T + t2) * nk + k] * yd_dulj_k(tw, D2, ky, kz);
            }
        }
    }
    yo_block_out<1>(v, part);
}

extern "C" int yamc_op_du_lj(yamc_ws *w, int32_t ntype, const double *const *beta_lj, const double *const *sums_lj,
                             const double *const *delta_lj, const int32_t *present, double *du) {
    if (!w || !du || (ntype > 0 && (!beta_lj || !sums_lj || !delta_lj || !present)))
        return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (ntype < 0 || ntype > YAMC_MAX_LJ_TYPE) return ofail(YAMC_ERROR_PARAMETER, "bad ntype");
    unsigned pmask = 0;
    for (int t = 0; t < ntype; ++t)
        if (present[t]) pmask |= 1u << t;
    if (ntype == 0 || pmask == 0) {
        *du = 0.0;
        return YAMC_OK;
    }
    size_t oB, oS, oD;
    stage_begin(w);
    OTRY(stage_rows(w, beta_lj, ntype * ntype, &oB));
    OTRY(stage_rows(w, sums_lj, 8 * ntype, &oS));
    OTRY(stage_rows(w, delta_lj, 8 * ntype, &oD));
    OTRY(stage_flush(w));
    void *pp = nullptr;
    const int nblk = (w->nk + 255) / 256;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * nblk, &pp));
    yo_du_lj<<<nblk, 256, 0, w->s>>>(w->nk, w->d_kxyz, ntype, pmask, stage_dev<double>(w, oB),
                                     stage_dev<double>(w, oS), stage_dev<double>(w, oD), (double *)pp);
    OCHK(hipGetLastError());
    return reduce_blocks<1>(w, (const double *)pp, nblk, du);
}

// ForceKSpace_coulomb / ForceKSpace_LJ (ForceKSpace.f90:59-302): 4 pi
// sum_k k beta sum_slots +-(C kr_sin - S kr_cos) over the nset sets given,
// each with its own beta row.
__global__ void yo_force(int nk, const int *kxyz, int nset, const double *B, const double *S, const double *kr,
                         double *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (k < nk) {
        const int t = kxyz[k], kx = t & 255, ky = (t >> 8) & 255, kz = (t >> 16) & 255;
        double R[8];
        for (int m = 0; m < 8; ++m) R[m] = kr[(size_t)m * nk + k];
        for (int s = 0; s < nset; ++s) {
            double C[8];
            for (int m = 0; m < 8; ++m) C[m] = S[((size_t)s * 8 + m) * nk + k];
            yd_force_k(C, R, kx, ky, kz, B[(size_t)s * nk + k], v);
        }
    }
    yo_block_out<3>(v, part);
}

static int force_sets(yamc_ws *w, int nset, const double *const *brows, const double *const *srows,
                      const double *const *kr8, double *f3) {
    size_t oB, oS, oK;
    stage_begin(w);
    OTRY(stage_rows(w, brows, nset, &oB));
    OTRY(stage_rows(w, srows, 8 * nset, &oS));
    OTRY(stage_rows(w, kr8, 8, &oK));
    OTRY(stage_flush(w));
    void *pp = nullptr;
    const int nblk = (w->nk + 255) / 256;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 3 * nblk, &pp));
    yo_force<<<nblk, 256, 0, w->s>>>(w->nk, w->d_kxyz, nset, stage_dev<double>(w, oB), stage_dev<double>(w, oS),
                                     stage_dev<double>(w, oK), (double *)pp);
    OCHK(hipGetLastError());
    double f[3];
    OTRY(reduce_blocks<3>(w, (const double *)pp, nblk, f));
    for (int d = 0; d < 3; ++d) f3[d] = YD_FOUR_PI * f[d];
    return YAMC_OK;
}

extern "C" int yamc_op_force_coulomb(yamc_ws *w, const double *beta, const double *const *s8,
                                     const double *const *kr8, double *f3) {
    if (!w || !beta || !s8 || !kr8 || !f3) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    const double *b[1] = {beta};
    return force_sets(w, 1, b, s8, kr8, f3);
}

extern "C" int yamc_op_force_lj(yamc_ws *w, int32_t ntype, int32_t ityp, const double *const *beta_lj,
                                const double *const *sums_lj, const double *const *kr8, double *f3) {
    if (!w || !beta_lj || !sums_lj || !kr8 || !f3) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (ntype < 1 || ntype > YAMC_MAX_LJ_TYPE || ityp < 1 || ityp > ntype)
        return ofail(YAMC_ERROR_PARAMETER, "LJ type %d outside 1..%d", ityp, ntype);
    std::vector<const double *> rows(ntype);
    for (int t = 0; t < ntype; ++t) rows[t] = beta_lj[(size_t)(ityp - 1) * ntype + t];  // beta_LJ(ityp, t)
    return force_sets(w, ntype, rows.data(), sums_lj, kr8, f3);
}

// ---- real space --------------------------------------------------------------

struct YoRs {
    int natom, ntype, full, nrow, first_atom, j0, j1, stop;
    double L;
    YdParams p;
    const double *x, *y, *z, *q, *dh, *s2, *fe, *s6, *xm, *ym, *zm;
    const int *mol, *typ;
};

// Row atom i (full: atom i; partial: molecule atom first_atom + i) and its
// position.
__device__ __forceinline__ int yo_row_atom(const YoRs &r, int i) { return r.full ? i : r.first_atom + i; }

// The pair (row i, atom j) of EwaldSumRealSpace_calc (EwaldSumRealSpace.f90:
// 338-498): returns 0 when the loop skips it (outside the j range, or the
// same molecule in a partial sum), else 1 with the overlap test's result and,
// inside the cut-off, the pair terms.
__device__ __forceinline__ int yo_rs_pair(const YoRs &r, int i, int j, bool &ov, bool &inside, double &xij,
                                          double &yij, double &zij, YdPair &pr) {
    const int ii = yo_row_atom(r, i);
    const int j0 = r.j0 >= 0 ? r.j0 : (r.full ? i + 1 : 0);
    const int j1 = r.j1 >= 0 ? r.j1 : r.natom - 1;
    if (j < j0 || j > j1) return 0;
    const int imol = r.mol[ii], jmol = r.mol[j];
    if (!r.full && imol == jmol) return 0;
    const int it = r.typ[ii], jt = r.typ[j];
    xij = yd_wrap(r.xm[i] - r.x[j]);
    yij = yd_wrap(r.ym[i] - r.y[j]);
    zij = yd_wrap(r.zm[i] - r.z[j]);
    const double r2 = xij * xij + yij * yij + zij * zij;
    const double dhs = (r.dh[ii] / r.L + r.dh[j] / r.L) / 2;
    ov = it > 0 && jt > 0 && imol != jmol && r2 < dhs * dhs;
    inside = r2 < r.p.rmax2;
    if (inside) {
        const bool lj = it > 0 && jt > 0;
        const int off = lj ? (it - 1) * r.ntype + (jt - 1) : 0;
        yd_pair(r.p, r2, imol == jmol, r.q[ii], r.q[j], lj, lj ? r.s6[off] : 0.0, lj ? r.fe[off] : 0.0, pr);
    }
    return 1;
}

// Pass 1 (the reference returns at the first overlap in its loop order):
// the smallest linear index i * natom + j of an overlapping pair.
__global__ void yo_rs_first_overlap(YoRs r, unsigned long long *key) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.nrow || j >= r.natom) return;
    bool ov = false, in = false;
    double x, y, z;
    YdPair pr;
    YoRs rr = r;
    rr.p.rmax2 = -1.0;  // the overlap test only
    if (yo_rs_pair(rr, i, j, ov, in, x, y, z, pr) && ov) {
        atomicMin(key, (unsigned long long)i * (unsigned long long)r.natom + (unsigned long long)j);
        atomicAdd(key + 1, 1ull);
    }
}

// Pass 2: rows.  Thread per row i and column group g: the group's pairs
// before the cut-off key (the first overlap when the sum stops there, kc[0]
// = ~0 when none); partial uu(i), F(i) to pat[g][4][nrow] and the six
// energy components of the block to part[g][block][6].
__global__ void __launch_bounds__(256) yo_rs_rows(YoRs r, int ngroups, const unsigned long long *kc, double *pat,
                                                  double *part) {
    const unsigned long long key = r.stop ? kc[0] : ~0ull;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int jc = (r.natom + ngroups - 1) / ngroups, jb = (int)blockIdx.y * jc;
    const int je = jb + jc < r.natom ? jb + jc : r.natom;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < r.nrow) {
        double uu = 0, fx = 0, fy = 0, fz = 0;
        for (int j = jb; j < je; ++j) {
            if ((unsigned long long)i * r.natom + j >= key) break;
            bool ov, in;
            double x, y, z;
            YdPair pr;
            if (!yo_rs_pair(r, i, j, ov, in, x, y, z, pr) || !in) continue;
            const bool same = r.mol[yo_row_atom(r, i)] == r.mol[j];
            v[0] += pr.potew; v[1] += pr.potlj6; v[2] += pr.potlj12;
            if (same) {
                v[3] += pr.potew; v[4] += pr.potlj6; v[5] += pr.potlj12;
            } else {
                uu += pr.pot;
                fx += x * pr.FR; fy += y * pr.FR; fz += z * pr.FR;
            }
        }
        double *pa = pat + (size_t)blockIdx.y * 4 * r.nrow + i;
        pa[0] = uu;
        pa[(size_t)r.nrow] = fx;
        pa[2 * (size_t)r.nrow] = fy;
        pa[3 * (size_t)r.nrow] = fz;
    }
    __shared__ double red[4 * 6], out[6];
    yd_block_sum<6>(v, red, out);
    if (threadIdx.x < 6) part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = out[threadIdx.x];
}

// Pass 3: columns (the back arrays): thread per atom j and row group g.
__global__ void __launch_bounds__(256) yo_rs_cols(YoRs r, int ngroups, const unsigned long long *kc,
                                                  double *pcol) {
    const unsigned long long key = r.stop ? kc[0] : ~0ull;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= r.natom) return;
    const int ic = (r.nrow + ngroups - 1) / ngroups, ib = (int)blockIdx.y * ic;
    const int ie = ib + ic < r.nrow ? ib + ic : r.nrow;
    double uu = 0, fx = 0, fy = 0, fz = 0;
    for (int i = ib; i < ie; ++i) {
        if ((unsigned long long)i * r.natom + j >= key) break;
        bool ov, in;
        double x, y, z;
        YdPair pr;
        if (!yo_rs_pair(r, i, j, ov, in, x, y, z, pr) || !in) continue;
        if (r.mol[yo_row_atom(r, i)] == r.mol[j]) continue;
        uu += pr.pot;
        fx -= x * pr.FR; fy -= y * pr.FR; fz -= z * pr.FR;
    }
    double *pc = pcol + (size_t)blockIdx.y * 4 * r.natom + j;
    pc[0] = uu;
    pc[(size_t)r.natom] = fx;
    pc[2 * (size_t)r.natom] = fy;
    pc[3 * (size_t)r.natom] = fz;
}

// Group partials [ng][n] summed per element in group order (the order a
// host loop over the groups would add them in): out[t] = 0 + g0 + g1 + ...
__global__ void yo_group_sum(const double *part, int ng, size_t n, double *out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double a = 0.0;
    for (int g = 0; g < ng; ++g) a = a + part[(size_t)g * n + t];
    out[t] = a;
}

extern "C" int yamc_op_rspace(yamc_ws *w, const yamc_rs_desc *d, double *uu, double *fx, double *fy, double *fz,
                              double *uu_back, double *fx_back, double *fy_back, double *fz_back, double *comps,
                              int32_t *overlap) {
    if (!w || !d || !comps || !overlap) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    const int N = d->natom;
    if (N < 1 || d->ntype < 0 || d->ntype > YAMC_MAX_LJ_TYPE) return ofail(YAMC_ERROR_PARAMETER, "bad sizes");
    const int nrow = d->full ? N : d->nmol_atom;
    if (nrow < 1 || (!d->full && (d->first_atom < 0 || d->first_atom + nrow > N)))
        return ofail(YAMC_ERROR_PARAMETER, "bad molecule range");
    if (d->j1 >= N || d->j0 >= N || d->j0 < -1 || d->j1 < -1)
        return ofail(YAMC_ERROR_PARAMETER, "j range outside the atoms");
    if (nrow > 65535) return ofail(YAMC_ERROR_LIMITS, "%d rows: this per-procedure sum takes at most 65535", nrow);
    YoRs r{};
    r.natom = N; r.ntype = d->ntype; r.full = d->full != 0; r.nrow = nrow; r.first_atom = d->first_atom;
    r.j0 = d->j0; r.j1 = d->j1; r.stop = d->stop_on_overlap != 0; r.L = d->box_length;
    r.p.alpha = d->alpha; r.p.inv_alpha = 1.0 / d->alpha; r.p.rmax2 = d->rmax2; r.p.xclb = d->xclb;
    r.p.alpha_over_sqrt_pi = d->alpha / sqrt(YD_PI);  // parameters.f90:400-401
    r.p.two_alpha_over_sqrt_pi = 2.0 * r.p.alpha_over_sqrt_pi;
    const int TT = d->ntype * d->ntype > 0 ? d->ntype * d->ntype : 1;
    if (!d->xx || !d->yy || !d->zz || !d->charge || !d->hard_core_angstr || !d->molnum || !d->lj_type ||
        (d->ntype && (!d->sigma2 || !d->four_eps || !d->sigma6)))
        return ofail(YAMC_ERROR_PARAMETER, "null atom array");
    if (!d->full && (!d->xmol || !d->ymol || !d->zmol)) return ofail(YAMC_ERROR_PARAMETER, "null molecule positions");
    for (int a = 0; a < N; ++a)  // the pair tables are indexed by type
        if (d->lj_type[a] < 0 || d->lj_type[a] > d->ntype)
            return ofail(YAMC_ERROR_PARAMETER, "LJ type %d of atom %d outside 0..%d", d->lj_type[a], a, d->ntype);
    const double zero3[3] = {0.0, 0.0, 0.0};
    const unsigned long long kc0[2] = {~0ull, 0ull};  // first overlap key, number of overlapping pairs
    size_t ox, oy, oz, oq, oh, omol, otyp, os2, ofe, os6, om, okc;
    stage_begin(w);
    OTRY(stage_put(w, d->xx, sizeof(double) * N, &ox));
    OTRY(stage_put(w, d->yy, sizeof(double) * N, &oy));
    OTRY(stage_put(w, d->zz, sizeof(double) * N, &oz));
    OTRY(stage_put(w, d->charge, sizeof(double) * N, &oq));
    OTRY(stage_put(w, d->hard_core_angstr, sizeof(double) * N, &oh));
    OTRY(stage_put(w, d->molnum, sizeof(int32_t) * N, &omol));
    OTRY(stage_put(w, d->lj_type, sizeof(int32_t) * N, &otyp));
    OTRY(stage_put(w, d->ntype ? d->sigma2 : zero3, sizeof(double) * (d->ntype ? TT : 1), &os2));
    OTRY(stage_put(w, d->ntype ? d->four_eps : zero3, sizeof(double) * (d->ntype ? TT : 1), &ofe));
    OTRY(stage_put(w, d->ntype ? d->sigma6 : zero3, sizeof(double) * (d->ntype ? TT : 1), &os6));
    om = ox;
    if (!d->full) {
        OTRY(stage_put(w, nullptr, sizeof(double) * 3 * nrow, &om));
        double *m = (double *)(w->h_in + om);
        for (int a = 0; a < nrow; ++a) { m[a] = d->xmol[a]; m[nrow + a] = d->ymol[a]; m[2 * nrow + a] = d->zmol[a]; }
    }
    OTRY(stage_put(w, kc0, sizeof kc0, &okc));
    OTRY(stage_flush(w));
    r.x = stage_dev<double>(w, ox); r.y = stage_dev<double>(w, oy); r.z = stage_dev<double>(w, oz);
    r.q = stage_dev<double>(w, oq); r.dh = stage_dev<double>(w, oh);
    r.s2 = stage_dev<double>(w, os2); r.fe = stage_dev<double>(w, ofe); r.s6 = stage_dev<double>(w, os6);
    r.mol = stage_dev<int>(w, omol); r.typ = stage_dev<int>(w, otyp);
    if (d->full) {
        r.xm = r.x; r.ym = r.y; r.zm = r.z;
    } else {
        r.xm = stage_dev<double>(w, om); r.ym = r.xm + nrow; r.zm = r.xm + 2 * nrow;
    }
    unsigned long long *dkc = stage_dev<unsigned long long>(w, okc);
    // pass 1: the first overlap in the reference's loop order
    yo_rs_first_overlap<<<dim3((N + 255) / 256, nrow), 256, 0, w->s>>>(r, dkc);
    OCHK(hipGetLastError());
    // pass 2: rows, about 2^18 (row, group) threads, partials pat[ngr][4][nrow];
    // pass 3: columns, about 2^18 (atom, group) threads, col[ngc][4][N]; the
    // groups are summed on the device.  Output block: the two key words,
    // part[ngr][nbr][6], row sums [4][nrow], column sums [4][N].
    int ngr = (1 << 18) / nrow;
    ngr = ngr < 1 ? 1 : ngr > 256 ? 256 : ngr;
    if (ngr > N) ngr = N;
    const int nbr = (nrow + 255) / 256;
    int ngc = (1 << 18) / N;
    ngc = ngc < 1 ? 1 : ngc > 64 ? 64 : ngc;
    if (ngc > nrow) ngc = nrow;
    const size_t npat = 4 * (size_t)nrow * ngr, npart = 6 * (size_t)nbr * ngr, ncol = 4 * (size_t)N * ngc;
    const size_t nrowsum = 4 * (size_t)nrow, ncolsum = 4 * (size_t)N;
    void *pgrp = nullptr, *pout = nullptr;
    OTRY(wbuf(w, WB_L, sizeof(double) * (npat + ncol), &pgrp));
    OTRY(wbuf(w, WB_OUT, sizeof(double) * (2 + npart + nrowsum + ncolsum), &pout));
    double *dpat = (double *)pgrp, *dcol = dpat + npat;
    double *dpart = (double *)pout + 2, *drow = dpart + npart, *dcolsum = drow + nrowsum;
    yo_rs_rows<<<dim3(nbr, ngr), 256, 0, w->s>>>(r, ngr, dkc, dpat, dpart);
    OCHK(hipGetLastError());
    yo_rs_cols<<<dim3((N + 255) / 256, ngc), 256, 0, w->s>>>(r, ngc, dkc, dcol);
    OCHK(hipGetLastError());
    yo_group_sum<<<(unsigned)((nrowsum + 255) / 256), 256, 0, w->s>>>(dpat, ngr, nrowsum, drow);
    yo_group_sum<<<(unsigned)((ncolsum + 255) / 256), 256, 0, w->s>>>(dcol, ngc, ncolsum, dcolsum);
    OCHK(hipGetLastError());
    OCHK(hipMemcpyAsync(pout, dkc, sizeof kc0, hipMemcpyDeviceToDevice, w->s));
    const char *hc;
    OTRY(fetch(w, pout, sizeof(double) * (2 + npart + nrowsum + ncolsum), &hc));
    unsigned long long kc[2];
    memcpy(kc, hc, sizeof kc);
    const double *hpart = (const double *)hc + 2, *row = hpart + npart, *col = row + nrowsum;
    const bool ov = kc[0] != ~0ull;
    for (int m = 0; m < 6; ++m) {
        double a = 0.0;
        for (size_t b = 0; b < (size_t)nbr * ngr; ++b) a = a + hpart[b * 6 + m];
        comps[m] = a;
    }
    double *ro[4] = {uu, fx, fy, fz}, *co[4] = {uu_back, fx_back, fy_back, fz_back};
    for (int q = 0; q < 4; ++q) {
        if (ro[q])
            for (int i = 0; i < nrow; ++i) ro[q][i] = row[(size_t)q * nrow + i];
        if (co[q]) {
            if (d->full && co[q] == ro[q]) {  // full sum: the back arrays are the arrays themselves
                for (int j = 0; j < N; ++j) co[q][j] = co[q][j] + col[(size_t)q * N + j];
            } else {
                for (int j = 0; j < N; ++j) co[q][j] = col[(size_t)q * N + j];
            }
        }
    }
    *overlap = !ov ? 0 : r.stop ? 1 : kc[1] > 0x7fffffffull ? 0x7fffffff : (int32_t)kc[1];
    return YAMC_OK;
}

// RealSumLocal_init_electrostatics + _calc_electrostatics (+ _init_LJ +
// _calc_LJ when lj) for one pair (RealSumLocal.f90:63-168): out = potew,
// potlj6, potlj12, pot, FR and the TRealSumLocal fields r, a, a2,
// exp_minus_a2, erfc_a, C6, C12, a6, sigma_over_r_6 as yd_pair forms them.
__global__ void yo_pair(YdParams p, double r2, int same, double qi, double qj, int lj, double sigma6, double fe,
                        double *out) {
    if (threadIdx.x != 0) return;
    YdPair pr;
    yd_pair(p, r2, same != 0, qi, qj, lj != 0, sigma6, fe, pr);
    out[0] = pr.potew; out[1] = pr.potlj6; out[2] = pr.potlj12; out[3] = pr.pot; out[4] = pr.FR;
    const double ir = yd_rsqrt(r2), r = r2 * ir, ir2 = ir * ir, a = p.alpha * r, a2 = a * a, e = exp(-a2);
    const double C6 = 1.0 + a2 * (1.0 + a2 * 0.5), a6 = a2 * a2 * a2;
    out[5] = r; out[6] = a; out[7] = a2; out[8] = e;
    out[9] = same ? -yd_erf_luc(a, a2, e) : yd_erfc_luc_ix(a, a2, ir * p.inv_alpha, e);
    out[10] = C6;
    out[11] = C6 + a6 * (1.0 + a2 * (1.0 + a2 * 0.2) * 0.25) * (1.0 / 6.0);
    out[12] = a6;
    out[13] = sigma6 * (ir2 * ir2 * ir2);
}

extern "C" int yamc_op_pair(yamc_ws *w, double alpha, double xclb, double r2, int32_t same_molecule, double qi,
                            double qj, int32_t lj, double sigma6, double four_eps, double *out) {
    if (!w || !out) return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (!(r2 > 0.0)) return ofail(YAMC_ERROR_PARAMETER, "r2 = %g: a pair needs r2 > 0", r2);
    YdParams p{};
    p.alpha = alpha; p.inv_alpha = 1.0 / alpha; p.xclb = xclb;
    p.alpha_over_sqrt_pi = alpha / sqrt(YD_PI);
    p.two_alpha_over_sqrt_pi = 2.0 * p.alpha_over_sqrt_pi;
    void *po = nullptr;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 14, &po));
    yo_pair<<<1, 64, 0, w->s>>>(p, r2, same_molecule, qi, qj, lj, sigma6, four_eps, (double *)po);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, sizeof(double) * 14, &h));
    memcpy(out, h, sizeof(double) * 14);
    return YAMC_OK;
}

// ---- constant terms ------------------------------------------------------------

// ewald_sum_coulomb_tail + ewald_sum_lj_tails (EwaldSumTails.f90:18-127),
// one thread in the reference's summation order.
__global__ void yo_tails(int n, const double *q, int T, const double *lj6, const double *lj12, const int *counts,
                         double alpha, double xclb, double *out) {
    if (threadIdx.x != 0) return;
    double S = 0.0;
    for (int i = 0; i < n; ++i) S = S + q[i] * q[i];
    out[0] = -xclb * (alpha / sqrt(YD_PI)) * S;
    const double a3 = yd_p3(alpha), a6 = a3 * a3, a9 = a3 * a6, a12 = a6 * a6, pspi = pow(YD_PI, 1.5);
    double sij6 = 0.0, sij12 = 0.0, sj6 = 0.0, sj12 = 0.0;
    for (int t1 = 0; t1 < T; ++t1)
        for (int t2 = t1; t2 < T; ++t2) {
            const int n12 = counts[t1] * counts[t2];
            double l6 = lj6[t1 * T + t2] * (double)n12, l12 = lj12[t1 * T + t2] * (double)n12;
            if (t1 != t2) { l6 = l6 + l6; l12 = l12 + l12; }
            sij6 = sij6 + l6;
            sij12 = sij12 + l12;
        }
    for (int t = 0; t < T; ++t) {
        sj6 = sj6 + lj6[t * T + t] * (double)counts[t];
        sj12 = sj12 + lj12[t * T + t] * (double)counts[t];
    }
    out[1] = pspi * a3 / 6.0 * sij6 - a6 / 12.0 * sj6;
    out[2] = pspi * a9 / 1080.0 * sij12 - a12 / 1440.0 * sj12;
}

extern "C" int yamc_op_tails(yamc_ws *w, int32_t natom, const double *charge, int32_t ntype, const double *lj6,
                             const double *lj12, const int32_t *counts, double alpha, double xclb, double *self_c,
                             double *lj6_tail, double *lj12_tail) {
    if (!w || (natom > 0 && !charge) || (ntype > 0 && (!lj6 || !lj12 || !counts)))
        return ofail(YAMC_ERROR_PARAMETER, "null argument");
    if (natom < 0 || ntype < 0 || ntype > YAMC_MAX_LJ_TYPE) return ofail(YAMC_ERROR_PARAMETER, "bad sizes");
    const int TT = ntype * ntype > 0 ? ntype * ntype : 1;
    const double zero = 0.0;
    const int izero = 0;
    size_t oq, o6, o12, oc;
    stage_begin(w);
    OTRY(stage_put(w, natom ? charge : &zero, sizeof(double) * (natom ? natom : 1), &oq));
    OTRY(stage_put(w, ntype ? lj6 : &zero, sizeof(double) * (ntype ? TT : 1), &o6));
    OTRY(stage_put(w, ntype ? lj12 : &zero, sizeof(double) * (ntype ? TT : 1), &o12));
    OTRY(stage_put(w, ntype ? counts : &izero, sizeof(int32_t) * (ntype ? ntype : 1), &oc));
    OTRY(stage_flush(w));
    void *po = nullptr;
    OTRY(wbuf(w, WB_OUT, sizeof(double) * 3, &po));
    yo_tails<<<1, 64, 0, w->s>>>(natom, stage_dev<double>(w, oq), ntype, stage_dev<double>(w, o6),
                                 stage_dev<double>(w, o12), stage_dev<int>(w, oc), alpha, xclb, (double *)po);
    OCHK(hipGetLastError());
    const char *h;
    OTRY(fetch(w, po, sizeof(double) * 3, &h));
    double o[3];
    memcpy(o, h, sizeof o);
    if (self_c) *self_c = o[0];
    if (lj6_tail) *lj6_tail = o[1];
    if (lj12_tail) *lj12_tail = o[2];
    return YAMC_OK;
}
