// yamc_rdf.hip -- centre-of-mass radial distribution counts on the device
// (mcrdf2.f90:262-336 count_distances, one frame per call).
//
// Molecule COMs are mass-weighted means of their atoms (box units), wrapped
// into the box; every ordered pair (i, j), i != j, of molecules is tested
// with the single minimum-image shift of mcrdf2.f90:276-283 (valid for
// rmax <= L/2, the Nbox = 0 case), and counted in bin floor(r / dr) of the
// (min(species), max(species)) histogram (mcrdf2.f90:316-320, cnt4).
#include "yamc_device.h"

__global__ void yk_com(YdChain c, double *com) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.p.nmol) return;
    int f = c.first[m], l = c.last[m];
    // explicit round-to-nearest products and sums (no contraction), so the
    // bin of every pair is reproducible by any IEEE FP64 evaluation
    double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0;
    for (int a = f; a <= l; ++a) {
        double w = c.mass[a];
        sx = __dadd_rn(sx, __dmul_rn(w, c.x[a]));
        sy = __dadd_rn(sy, __dmul_rn(w, c.y[a]));
        sz = __dadd_rn(sz, __dmul_rn(w, c.z[a]));
        sm = __dadd_rn(sm, w);
    }
    double x = sx / sm, y = sy / sm, z = sz / sm;
    com[3 * m + 0] = x - round(x);
    com[3 * m + 1] = y - round(y);
    com[3 * m + 2] = z - round(z);
}

// Tile of 16 x 16 molecule pairs per 256-thread block; per-block LDS
// histogram flushed with 64-bit atomics.
#define YR_MAXBIN 4096
__global__ void __launch_bounds__(256) yk_rdf(int nmol, const double *com, const int *species, int nspec,
                                              double L, double dr, double rmax2, int nbins,
                                              unsigned long long *hist) {
    extern __shared__ unsigned int h[];
    const int nh = nspec * nspec * nbins;
    for (int k = threadIdx.x; k < nh; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    int i = blockIdx.y * 16 + (threadIdx.x >> 4), j = blockIdx.x * 16 + (threadIdx.x & 15);
    if (i < nmol && j < nmol && i != j) {
        double x1 = com[3 * i], y1 = com[3 * i + 1], z1 = com[3 * i + 2];
        double x2 = com[3 * j], y2 = com[3 * j + 1], z2 = com[3 * j + 2];
        if (x1 - x2 > 0.5) x2 = x2 + 1.0;
        if (x1 - x2 < -0.5) x2 = x2 - 1.0;
        if (y1 - y2 > 0.5) y2 = y2 + 1.0;
        if (y1 - y2 < -0.5) y2 = y2 - 1.0;
        if (z1 - z2 > 0.5) z2 = z2 + 1.0;
        if (z1 - z2 < -0.5) z2 = z2 - 1.0;
        double dx = __dmul_rn(x1 - x2, L), dy = __dmul_rn(y1 - y2, L), dz = __dmul_rn(z1 - z2, L);
        double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
        if (!(r2 > rmax2 - 1e-10)) {
            int b = (int)(sqrt(r2) / dr);
            int si = species[i], sj = species[j];
            int lo = si < sj ? si : sj, hi = si < sj ? sj : si;
            if (b < nbins) atomicAdd(&h[(lo * nspec + hi) * nbins + b], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nh; k += blockDim.x)
        if (h[k]) atomicAdd(&hist[k], (unsigned long long)h[k]);
}

void yk_launch_rdf(const YdChain &c, double *com, const int *species, int nspec, double dr_box_units_times_L,
                   double rmax2, int nbins, unsigned long long *hist, hipStream_t s) {
    int nmol = c.p.nmol;
    yk_com<<<(nmol + 255) / 256, 256, 0, s>>>(c, com);
    dim3 grid((nmol + 15) / 16, (nmol + 15) / 16);
    size_t lds = sizeof(unsigned int) * (size_t)nspec * nspec * nbins;
    yk_rdf<<<grid, 256, lds, s>>>(nmol, com, species, nspec, c.p.L, dr_box_units_times_L, rmax2, nbins, hist);
}
