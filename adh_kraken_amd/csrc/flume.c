/*
 * flume.c -- synthetic rectangular flume meshes for the SW2 workloads.
 *
 * Node numbering, element order and the alternating diagonals restate
 * create_rectangular_grid (src/tools/create_rectangular_grid.c:4-445):
 * node (i, j) = i*npy + j at (xmin + i*dx, ymin + j*dy); for each cell (i, j)
 * in i-major, j-minor order, even j gives triangles (n, n+npy, n+npy+1) and
 * (n, n+npy+1, n+1), odd j gives (n, n+npy, n+1) and (n+npy, n+npy+1, n+1),
 * all counter-clockwise.  The reference generator stops there (flat bed from
 * a polynomial, no boundary elements); this file adds what the workloads need:
 * the bed z = -slope*(x - xmin) + seeded Gaussian bumps, per-element Manning n,
 * an inflow-flux boundary along x = xmin and a tailwater-depth boundary along
 * x = xmax, both as 1-D elements ordered with the domain on their left so that
 * the normal of selem1d_load (selem_1d.c:53-54) points outward.
 *
 * Random numbers: SplitMix64, u = (x >> 11) * 2^-53; per bump the draws are
 * cx, cy, amplitude, sigma in that order.
 */
#include <math.h>
#include <stdint.h>

#include "../../include/adhk.h"

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int adhk_flume_counts(const adhk_flume_spec *f, int *nnodes, int *nelems2d, int *nelems1d)
{
    if (!f || f->npx < 2 || f->npy < 2) return -1;
    if (nnodes) *nnodes = f->npx * f->npy;
    if (nelems2d) *nelems2d = 2 * (f->npx - 1) * (f->npy - 1);
    if (nelems1d) *nelems1d = 2 * (f->npy - 1);
    return 0;
}

int adhk_flume_build(const adhk_flume_spec *f, double *x, double *y, double *z, int *tri, double *mann, int *seg,
                     int *seg_kind, double *seg_val)
{
    if (!f || f->npx < 2 || f->npy < 2) return -1;
    const int npx = f->npx, npy = f->npy;
    const double dx = (f->xmax - f->xmin) / (double)(npx - 1);
    const double dy = (f->ymax - f->ymin) / (double)(npy - 1);
    /* bumps */
    enum { MAXB = 4096 };
    double bcx[MAXB], bcy[MAXB], ba[MAXB], bs[MAXB];
    int nb = f->nbumps < MAXB ? f->nbumps : MAXB;
    uint64_t st = f->seed;
    for (int k = 0; k < nb; k++) {
        bcx[k] = f->xmin + (f->xmax - f->xmin) * unif(&st);
        bcy[k] = f->ymin + (f->ymax - f->ymin) * unif(&st);
        ba[k] = f->bump_amp_max * unif(&st);
        bs[k] = f->bump_sigma_min + (f->bump_sigma_max - f->bump_sigma_min) * unif(&st);
    }
    int np = 0;
    for (int i = 0; i < npx; i++)
        for (int j = 0; j < npy; j++) {
            double xx = f->xmin + i * dx, yy = f->ymin + j * dy;
            double zz = -f->slope * (xx - f->xmin);
            for (int k = 0; k < nb; k++) {
                double rx = xx - bcx[k], ry = yy - bcy[k];
                zz += ba[k] * exp(-(rx * rx + ry * ry) / (2.0 * bs[k] * bs[k]));
            }
            x[np] = xx;
            y[np] = yy;
            z[np] = zz;
            np++;
        }
    int k = 0;
    for (int i = 0; i < npx - 1; i++)
        for (int j = 0; j < npy - 1; j++) {
            int n = i * npy + j;
            if (j % 2 == 0) {
                tri[3 * k] = n;
                tri[3 * k + 1] = n + npy;
                tri[3 * k + 2] = n + npy + 1;
                k++;
                tri[3 * k] = n;
                tri[3 * k + 1] = n + npy + 1;
                tri[3 * k + 2] = n + 1;
                k++;
            } else {
                tri[3 * k] = n;
                tri[3 * k + 1] = n + npy;
                tri[3 * k + 2] = n + 1;
                k++;
                tri[3 * k] = n + npy;
                tri[3 * k + 1] = n + npy + 1;
                tri[3 * k + 2] = n + 1;
                k++;
            }
        }
    if (mann)
        for (int e = 0; e < k; e++) mann[e] = f->manning_n;
    /* boundary segments: inflow at x = xmin walked from j+1 to j, tailwater at x = xmax walked from j to j+1 */
    const double width = f->ymax - f->ymin;
    int s = 0;
    for (int j = 0; j < npy - 1; j++) {
        seg[2 * s] = j + 1;
        seg[2 * s + 1] = j;
        seg_kind[s] = ADHK_BC_FLUX;
        seg_val[2 * s] = -f->inflow_q / width;
        seg_val[2 * s + 1] = 0.0;
        s++;
    }
    const int last = (npx - 1) * npy;
    for (int j = 0; j < npy - 1; j++) {
        seg[2 * s] = last + j;
        seg[2 * s + 1] = last + j + 1;
        seg_kind[s] = ADHK_BC_DEPTH;
        seg_val[2 * s] = f->tailwater;
        seg_val[2 * s + 1] = 0.0;
        s++;
    }
    return 0;
}

int adhk_flume_initial_state(const adhk_flume_spec *f, int nnodes, const double *z, double *sol)
{
    if (!f || !z || !sol) return -1;
    const double wse = f->tailwater - f->slope * (f->xmax - f->xmin);
    for (int n = 0; n < nnodes; n++) {
        sol[3 * n] = f->ic_flat_wse ? wse - z[n] : f->tailwater;
        sol[3 * n + 1] = 0.0;
        sol[3 * n + 2] = 0.0;
    }
    return 0;
}
