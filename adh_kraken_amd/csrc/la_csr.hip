/*
 * la_csr.hip -- the reference's plain-pointer linear-algebra surface
 * (src/la/la.h) with its exact names and signatures, over host CSR arrays,
 * and a variant of its BiCGStab solve on device arrays that allocates nothing.
 *
 * Single rank: the split-CSR off-diagonal block must be absent
 * (indptr_off_diag == NULL), as it is when the reference runs on one process
 * (slin_sys_init_ptrs.c:26).  prep_umfpack/solve_umfpack keep their names but
 * build and apply a block-Jacobi preconditioner (block size 3 by default, the
 * SW2 node block; 1 for point Jacobi) instead of the UMFPACK LU
 * (bcgstab_solver.c:371-424).
 *
 * State: the reference keeps its BiCGStab vectors and factors in module statics
 * (bcgstab_solver.c:5-22).  Here each host thread owns its own workspace (device
 * buffers, stream, preconditioner), created on the thread's first call on the device
 * current then, grown on demand and released by free_bcgstab() on that thread.
 * Calls are thread-compatible: threads never share a workspace.
 *
 * BiCGStab (bcgstab_solver.c:65-335) runs as three fused kernels per iteration,
 * each a block-row pass of the CSR matrix or of the vectors, with the scalars
 * (alpha, omega, rho, beta, the residual norm and the stop test) kept on the device
 * and formed by one-workgroup finalizers that sum per-workgroup partials in a fixed
 * order.  The host launches the iterations in chunks and reads back one "done" flag
 * per chunk; the iterates never leave the device.
 *
 * Errors: no call aborts.  Argument and HIP errors leave the outputs untouched,
 * record the message (adhk_last_error) and a status (adhk_la_last_status: 0 or -1);
 * the int-returning calls also return -1.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "adhk_internal.h"

#define TPB 256
#define NPART_MAX 1024 /* partial sums per reduction: at most this many workgroups in a vector pass */
#define LA_CHUNK 8     /* BiCGStab iterations launched between read-backs of the "done" flag */
#define D_SMALL 2.220446049250313080847e-16
#define LA_TOL 1e-11  /* min_conv_tol, conv_tol (bcgstab_solver.c:20-21) */
#define LA_MAX_IT 100 /* max_it (:22) */

/* device scalars of one solve */
struct la_sc {
    double rho, rhop, alpha, omega, beta, rnorm, bnorm, tol;
    int it, max_it, done, pad;
};

/* device operands of one solve: CSR (n rows, nb = n / bs block rows), the block inverses and the vectors */
struct la_view {
    int n, nb, bs, nwg;
    const int *ip, *ci;
    const double *v, *minv, *scale;
    double *b, *x, *x0, *r, *q, *p, *Ap, *s, *As, *part;
    la_sc *sc;
};

/* ---------------------------------------------------------------------------------------------- kernels */

__device__ int kc_find(const int *ci, int lo, int hi, int t) /* binary_search_part.c:1 */
{
    hi -= 1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (ci[mid] == t) return mid;
        if (ci[mid] < t) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

__global__ void kc_spmv(int n, const int *ip, const int *ci, const double *v, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = ip[i]; j < ip[i + 1]; j++) s = __fma_rn(v[j], x[ci[j]], s);
    y[i] = s;
}

/* row maxima -> scale vector, b *= s, x /= s   (scale_linear_system.c:31-101) */
__global__ void kc_scale1(int n, const int *ip, const double *v, double *scale, double *b, double *x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int j = ip[i]; j < ip[i + 1]; j++) {
        double t = fabs(v[j]);
        if (t > mx) mx = t;
    }
    double s = (mx < D_SMALL) ? 1.0 / sqrt(D_SMALL) : 1.0 / sqrt(fabs(mx));
    scale[i] = s;
    b[i] *= s;
    x[i] /= s;
}
__global__ void kc_scale2(int n, const int *ip, const int *ci, double *v, const double *scale)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = ip[i]; j < ip[i + 1]; j++) v[j] *= scale[i] * scale[ci[j]];
}

/* block-Jacobi inverse of block row k (bs 3: the 3x3 node block; bs 1: the diagonal entry) */
__global__ void kc_bj_prep(int nb, int bs, const int *ip, const int *ci, const double *v, double *minv)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    if (bs == 1) {
        int idx = kc_find(ci, ip[k], ip[k + 1], k);
        double a = idx >= 0 ? v[idx] : 0.0;
        minv[k] = (a != 0.0) ? 1.0 / a : 1.0;
        return;
    }
    double a[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            int row = 3 * k + r;
            int idx = kc_find(ci, ip[row], ip[row + 1], 3 * k + c);
            a[3 * r + c] = idx >= 0 ? v[idx] : 0.0;
        }
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double *m = minv + 9 * (size_t)k;
    if (fabs(det) > 1e-300) {
        double id = 1.0 / det;
        m[0] = c00 * id;
        m[1] = (a[2] * a[7] - a[1] * a[8]) * id;
        m[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        m[3] = c01 * id;
        m[4] = (a[0] * a[8] - a[2] * a[6]) * id;
        m[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        m[6] = c02 * id;
        m[7] = (a[1] * a[6] - a[0] * a[7]) * id;
        m[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    } else {
        for (int i = 0; i < 9; i++) m[i] = 0.0;
        for (int i = 0; i < 3; i++) m[4 * i] = (a[4 * i] != 0.0) ? 1.0 / a[4 * i] : 1.0;
    }
}

__global__ void kc_bj_apply(int nb, int bs, const double *minv, const double *b, double *x)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    if (bs == 1) {
        x[k] = minv[k] * b[k];
        return;
    }
    const double *m = minv + 9 * (size_t)k;
    double b0 = b[3 * k], b1 = b[3 * k + 1], b2 = b[3 * k + 2];
    x[3 * k] = m[0] * b0 + m[1] * b1 + m[2] * b2;
    x[3 * k + 1] = m[3] * b0 + m[4] * b1 + m[5] * b2;
    x[3 * k + 2] = m[6] * b0 + m[7] * b1 + m[8] * b2;
}

__global__ void kc_unscale(int n, double *x, const double *x0, const double *s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (x[i] + x0[i]) * s[i]; /* unscale_linear_system, scale_linear_system.c:137-146 */
}

/* y = M^-1 (A col)[rows of block row k], the columns' values given by col(j) */
template <class F>
__device__ __forceinline__ void blk_op(const la_view &V, int k, F col, double *y)
{
    double t[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < V.bs; a++) {
        const int row = V.bs * k + a;
        double s = 0.0;
        for (int j = V.ip[row]; j < V.ip[row + 1]; j++) s = __fma_rn(V.v[j], col(V.ci[j]), s);
        t[a] = s;
    }
    if (V.bs == 1) {
        y[0] = V.minv[k] * t[0];
        return;
    }
    const double *m = V.minv + 9 * (size_t)k;
    for (int a = 0; a < 3; a++) y[a] = __fma_rn(m[3 * a + 2], t[2], __fma_rn(m[3 * a + 1], t[1], m[3 * a] * t[0]));
}

/* sums (op 0) or maxima (op 1) of up to 4 per-thread values over the workgroup -> part[c * NPART_MAX + wg] */
__device__ __forceinline__ void wg_partials(double *part, int nval, const double *val, const int *op)
{
    __shared__ double sh[4][TPB];
    for (int c = 0; c < nval; c++) sh[c][threadIdx.x] = val[c];
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < nval; c++)
                sh[c][threadIdx.x] = op[c] ? fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + o])
                                           : sh[c][threadIdx.x] + sh[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < nval; c++) part[c * NPART_MAX + blockIdx.x] = sh[c][0];
}

/* start (bcgstab_solver.c:128-176): x0 <- x, b <- M^-1 b, x <- M^-1 b; partial max|M^-1 b| (slot 2) */
__global__ __launch_bounds__(TPB) void kc_bicg_init1(la_view V)
{
    double bmax = 0.0;
    for (int k = blockIdx.x * TPB + threadIdx.x; k < V.nb; k += gridDim.x * TPB) {
        double bb[3], y[3];
        for (int a = 0; a < V.bs; a++) {
            const int i = V.bs * k + a;
            V.x0[i] = V.x[i];
            bb[a] = V.b[i];
        }
        if (V.bs == 1) {
            y[0] = V.minv[k] * bb[0];
        } else {
            const double *m = V.minv + 9 * (size_t)k;
            for (int a = 0; a < 3; a++)
                y[a] = __fma_rn(m[3 * a + 2], bb[2], __fma_rn(m[3 * a + 1], bb[1], m[3 * a] * bb[0]));
        }
        for (int a = 0; a < V.bs; a++) {
            const int i = V.bs * k + a;
            V.b[i] = y[a];
            V.x[i] = y[a];
            bmax = fmax(bmax, fabs(y[a]));
        }
    }
    const double val[1] = {bmax};
    const int op[1] = {1};
    wg_partials(V.part + 2 * NPART_MAX, 1, val, op);
}

/* r = M^-1 b - M^-1 A x, q = r, p = r (the first direction); partials max|r| (slot 0), r.q (slot 1) */
__global__ __launch_bounds__(TPB) void kc_bicg_init2(la_view V)
{
    double rmax = 0.0, rq = 0.0;
    for (int k = blockIdx.x * TPB + threadIdx.x; k < V.nb; k += gridDim.x * TPB) {
        double y[3];
        blk_op(V, k, [&](int j) { return V.x[j]; }, y);
        for (int a = 0; a < V.bs; a++) {
            const int i = V.bs * k + a;
            const double r = V.b[i] + (-1.0) * y[a];
            V.r[i] = r;
            V.q[i] = r;
            V.p[i] = r;
            rmax = fmax(rmax, fabs(r));
            rq = __fma_rn(r, r, rq);
        }
    }
    const double val[2] = {rmax, rq};
    const int op[2] = {1, 0};
    wg_partials(V.part, 2, val, op);
}

/* Ap = M^-1 A p; partial q.Ap (slot 0) */
__global__ __launch_bounds__(TPB) void kc_bicg_a(la_view V)
{
    if (V.sc->done) return;
    double qa = 0.0;
    for (int k = blockIdx.x * TPB + threadIdx.x; k < V.nb; k += gridDim.x * TPB) {
        double y[3];
        blk_op(V, k, [&](int j) { return V.p[j]; }, y);
        for (int a = 0; a < V.bs; a++) {
            const int i = V.bs * k + a;
            V.Ap[i] = y[a];
            qa = __fma_rn(V.q[i], y[a], qa);
        }
    }
    const double val[1] = {qa};
    const int op[1] = {0};
    wg_partials(V.part, 1, val, op);
}

/* s = r - alpha Ap, As = M^-1 A s (s of the gathered columns formed on the fly); partials As.s, As.As, As.q */
__global__ __launch_bounds__(TPB) void kc_bicg_s(la_view V)
{
    if (V.sc->done) return;
    const double alpha = V.sc->alpha;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int k = blockIdx.x * TPB + threadIdx.x; k < V.nb; k += gridDim.x * TPB) {
        double y[3];
        blk_op(V, k, [&](int j) { return __fma_rn(-alpha, V.Ap[j], V.r[j]); }, y);
        for (int a = 0; a < V.bs; a++) {
            const int i = V.bs * k + a;
            const double s = __fma_rn(-alpha, V.Ap[i], V.r[i]);
            V.s[i] = s;
            V.As[i] = y[a];
            d0 = __fma_rn(y[a], s, d0);
            d1 = __fma_rn(y[a], y[a], d1);
            d2 = __fma_rn(y[a], V.q[i], d2);
        }
    }
    const double val[3] = {d0, d1, d2};
    const int op[3] = {0, 0, 0};
    wg_partials(V.part, 3, val, op);
}

/* x = x + alpha p + omega s, r = s - omega As, p = (p - omega Ap) beta + r; partial max|r| */
__global__ __launch_bounds__(TPB) void kc_bicg_u(la_view V)
{
    if (V.sc->done) return;
    const double alpha = V.sc->alpha, omega = V.sc->omega, beta = V.sc->beta;
    double rmax = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < V.n; i += gridDim.x * TPB) {
        const double s = V.s[i];
        V.x[i] = __fma_rn(omega, s, __fma_rn(alpha, V.p[i], V.x[i]));
        const double r = __fma_rn(-omega, V.As[i], s);
        V.r[i] = r;
        V.p[i] = __fma_rn(__fma_rn(-omega, V.Ap[i], V.p[i]), beta, r);
        rmax = fmax(rmax, fabs(r));
    }
    const double val[1] = {rmax};
    const int op[1] = {1};
    wg_partials(V.part, 1, val, op);
}

enum { FIN_INIT, FIN_A, FIN_S, FIN_U };

/* one workgroup: the partials of the preceding pass, summed (or maxed) in index order, into the scalars
 * (bcgstab_solver.c:178-321) */
__global__ __launch_bounds__(TPB) void kc_bicg_fin(la_view V, int which)
{
    la_sc *sc = V.sc;
    if (which != FIN_INIT && sc->done) return;
    __shared__ double sh[4][TPB];
    const int nval = which == FIN_INIT ? 3 : which == FIN_S ? 3 : 1;
    int op[3] = {0, 0, 0};
    if (which == FIN_INIT) op[0] = op[2] = 1;
    if (which == FIN_U) op[0] = 1;
    for (int c = 0; c < nval; c++) {
        double a = 0.0;
        for (int w = threadIdx.x; w < V.nwg; w += TPB) {
            const double t = V.part[c * NPART_MAX + w];
            a = op[c] ? fmax(a, t) : a + t;
        }
        sh[c][threadIdx.x] = a;
    }
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < nval; c++)
                sh[c][threadIdx.x] = op[c] ? fmax(sh[c][threadIdx.x], sh[c][threadIdx.x + o])
                                           : sh[c][threadIdx.x] + sh[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (which == FIN_INIT) {
        sc->rnorm = sh[0][0];
        sc->rho = sh[1][0];
        sc->bnorm = sh[2][0];
        sc->rhop = sc->alpha = sc->omega = 1.0;
        sc->it = 0;
        sc->done = !(sc->rnorm > (sc->tol + sc->bnorm * sc->tol) && sc->it < sc->max_it);
    } else if (which == FIN_A) {
        sc->alpha = sc->rho / sh[0][0];
    } else if (which == FIN_S) {
        sc->omega = sh[0][0] / sh[1][0];
        sc->rhop = sc->rho;
        sc->rho = -1.0 * sc->omega * sh[2][0];
        sc->beta = (sc->rho * sc->alpha) / (sc->rhop * sc->omega); /* the next iteration's beta (:213) */
    } else {
        sc->rnorm = sh[0][0];
        sc->it += 1;
        sc->done = !(sc->rnorm > (sc->tol + sc->bnorm * sc->tol) && sc->it < sc->max_it);
    }
}

__global__ void kc_sc_init(la_sc *sc, double tol, int max_it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->rho = sc->rhop = sc->alpha = sc->omega = sc->beta = sc->rnorm = sc->bnorm = 0.0;
        sc->tol = tol;
        sc->it = sc->done = sc->pad = 0;
        sc->max_it = max_it;
    }
}

/* ------------------------------------------------------------------------------------- host helpers */

static int nblk(int n) { return (n + TPB - 1) / TPB; }

#define LCK(x)                                                                                                    \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("la: %s: %s", #x, hipGetErrorString(e_));                                              \
            return -1;                                                                                            \
        }                                                                                                         \
    } while (0)

static __thread int tl_status, tl_its;

extern "C" int adhk_la_last_status(void) { return tl_status; }
extern "C" int adhk_la_last_iterations(void) { return tl_its; }

/* The solve on prepared device operands: x (in: x0 of the scaled system; out: the unscaled solution
 * (x + x0) * scale), b (in: the scaled right-hand side; out: M^-1 b, as the reference leaves it,
 * bcgstab_solver.c:135-137).  V.minv must hold the block inverses. */
static int bicg_core(la_view V, hipStream_t st, double tol, int max_it, int *its)
{
    V.nwg = nblk(V.nb) < NPART_MAX ? (nblk(V.nb) < 1 ? 1 : nblk(V.nb)) : NPART_MAX;
    const dim3 g(V.nwg), t(TPB);
    hipLaunchKernelGGL(kc_sc_init, dim3(1), dim3(1), 0, st, V.sc, tol, max_it);
    hipLaunchKernelGGL(kc_bicg_init1, g, t, 0, st, V);
    hipLaunchKernelGGL(kc_bicg_init2, g, t, 0, st, V);
    hipLaunchKernelGGL(kc_bicg_fin, dim3(1), t, 0, st, V, (int)FIN_INIT);
    LCK(hipGetLastError());
    la_sc h;
    for (int chunk = 0; chunk <= max_it / LA_CHUNK + 1; chunk++) {
        LCK(hipMemcpyAsync(&h, V.sc, sizeof h, hipMemcpyDeviceToHost, st));
        LCK(hipStreamSynchronize(st));
        if (h.done) break;
        for (int k = 0; k < LA_CHUNK; k++) {
            hipLaunchKernelGGL(kc_bicg_a, g, t, 0, st, V);
            hipLaunchKernelGGL(kc_bicg_fin, dim3(1), t, 0, st, V, (int)FIN_A);
            hipLaunchKernelGGL(kc_bicg_s, g, t, 0, st, V);
            hipLaunchKernelGGL(kc_bicg_fin, dim3(1), t, 0, st, V, (int)FIN_S);
            hipLaunchKernelGGL(kc_bicg_u, g, t, 0, st, V);
            hipLaunchKernelGGL(kc_bicg_fin, dim3(1), t, 0, st, V, (int)FIN_U);
        }
        LCK(hipGetLastError());
    }
    hipLaunchKernelGGL(kc_unscale, dim3(nblk(V.n)), t, 0, st, V.n, V.x, V.x0, V.scale);
    LCK(hipGetLastError());
    LCK(hipMemcpyAsync(&h, V.sc, sizeof h, hipMemcpyDeviceToHost, st));
    LCK(hipStreamSynchronize(st));
    if (its) *its = h.it;
    return 0;
}

/* --------------------------------------------------------------------- per-thread workspace (la.h) */

typedef struct {
    hipStream_t st;
    int cap_n, cap_nnz, n, nnz, prep_n, prep_bs;
    int *ip, *ci;
    double *v, *minv, *b, *x, *x0, *r, *q, *p, *Ap, *s, *As, *scale, *tmp, *part;
    la_sc *sc;
} la_ws;

static __thread la_ws *tl_ws;
static int g_bs = 3;

extern "C" void adhk_la_set_block_size(int bs) { g_bs = (bs == 1) ? 1 : 3; }

static void ws_release(la_ws *w)
{
    void *ptrs[] = {w->ip, w->ci, w->v, w->minv, w->b, w->x, w->x0, w->r, w->q, w->p,
                    w->Ap, w->s, w->As, w->scale, w->tmp, w->part, w->sc};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
        if (ptrs[i]) (void)hipFree(ptrs[i]);
    if (w->st) (void)hipStreamDestroy(w->st);
    free(w);
}

/* this thread's workspace with room for n rows and nnz entries; the preconditioner survives growth */
static la_ws *ws_get(int n, int nnz)
{
    la_ws *w = tl_ws;
    if (!w) {
        w = (la_ws *)calloc(1, sizeof *w);
        if (!w) {
            adhk_set_error("la: out of host memory");
            return NULL;
        }
        if (hipStreamCreate(&w->st) != hipSuccess || hipMalloc((void **)&w->part, 4 * NPART_MAX * sizeof(double)) ||
            hipMalloc((void **)&w->sc, sizeof(la_sc))) {
            adhk_set_error("la: workspace creation failed");
            ws_release(w);
            return NULL;
        }
        tl_ws = w;
    }
    if (n > w->cap_n) {
        double **vs[] = {&w->b, &w->x, &w->x0, &w->r, &w->q, &w->p, &w->Ap, &w->s, &w->As, &w->scale, &w->tmp};
        for (size_t i = 0; i < sizeof vs / sizeof vs[0]; i++) {
            if (*vs[i]) (void)hipFree(*vs[i]);
            *vs[i] = NULL;
            if (hipMalloc((void **)vs[i], (size_t)n * sizeof(double)) != hipSuccess) goto oom;
        }
        if (w->ip) (void)hipFree(w->ip);
        w->ip = NULL;
        if (hipMalloc((void **)&w->ip, ((size_t)n + 1) * sizeof(int)) != hipSuccess) goto oom;
        double *mnew = NULL;
        if (hipMalloc((void **)&mnew, (size_t)n * 3 * sizeof(double)) != hipSuccess) goto oom;
        if (w->minv) {
            /* keep the prepared block inverses: a later solve of the prepared size still finds them */
            if (w->prep_n > 0 && hipMemcpyAsync(mnew, w->minv, (size_t)w->prep_n * 3 * sizeof(double),
                                                hipMemcpyDeviceToDevice, w->st) != hipSuccess) {
                (void)hipFree(mnew);
                goto oom;
            }
            (void)hipStreamSynchronize(w->st);
            (void)hipFree(w->minv);
        }
        w->minv = mnew;
        w->cap_n = n;
    }
    if (nnz > w->cap_nnz) {
        if (w->ci) (void)hipFree(w->ci);
        if (w->v) (void)hipFree(w->v);
        w->ci = NULL;
        w->v = NULL;
        if (hipMalloc((void **)&w->ci, (size_t)nnz * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&w->v, (size_t)nnz * sizeof(double)) != hipSuccess) {
            w->cap_nnz = 0;
            goto oom;
        }
        w->cap_nnz = nnz;
    }
    return w;
oom:
    adhk_set_error("la: device allocation failed for %d rows", n);
    ws_release(w);
    tl_ws = NULL;
    return NULL;
}

static la_view ws_view(la_ws *w, int n, int bs)
{
    la_view V;
    memset(&V, 0, sizeof V);
    V.n = n;
    V.bs = bs;
    V.nb = n / bs;
    V.ip = w->ip;
    V.ci = w->ci;
    V.v = w->v;
    V.minv = w->minv;
    V.scale = w->scale;
    V.b = w->b;
    V.x = w->x;
    V.x0 = w->x0;
    V.r = w->r;
    V.q = w->q;
    V.p = w->p;
    V.Ap = w->Ap;
    V.s = w->s;
    V.As = w->As;
    V.part = w->part;
    V.sc = w->sc;
    return V;
}

static int upload_csr(la_ws *w, const int *indptr, const int *cols, const double *vals, int n)
{
    const int nnz = indptr[n];
    w->n = n;
    w->nnz = nnz;
    LCK(hipMemcpyAsync(w->ip, indptr, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice, w->st));
    if (nnz > 0) {
        LCK(hipMemcpyAsync(w->ci, cols, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, w->st));
        LCK(hipMemcpyAsync(w->v, vals, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, w->st));
    }
    return 0;
}
static int up(la_ws *w, double *d, const double *h, int n)
{
    LCK(hipMemcpyAsync(d, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, w->st));
    return 0;
}
static int down(la_ws *w, double *h, const double *d, int n)
{
    LCK(hipMemcpyAsync(h, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, w->st));
    LCK(hipStreamSynchronize(w->st));
    return 0;
}

static int check_single_rank(const int *indptr_off_diag, int nghost)
{
    if (indptr_off_diag != NULL || nghost != 0) {
        adhk_set_error("la: split-CSR off-diagonal blocks (multi-rank) are not supported on this path");
        return -1;
    }
    return 0;
}
static int check_csr(const int *indptr, const int *cols, const double *vals, int n)
{
    if (n <= 0 || !indptr || indptr[0] != 0 || indptr[n] < 0 || (indptr[n] > 0 && (!cols || !vals))) {
        adhk_set_error("la: bad CSR arguments (n %d)", n);
        return -1;
    }
    return 0;
}
static void done_status(int rc) { tl_status = rc < 0 ? -1 : 0; }

/* ------------------------------------------------------------------------------- la.h entry points */

static int spmv_impl(double *Ax, int *indptr, int *cols, double *vals, int *indptr_off_diag, double *x, int nrows,
                     int nghost)
{
    if (check_single_rank(indptr_off_diag, nghost) || check_csr(indptr, cols, vals, nrows) || !Ax || !x) {
        if (!Ax || !x) adhk_set_error("split_CSR_mat_vec_mult: NULL vector");
        return -1;
    }
    la_ws *w = ws_get(nrows, indptr[nrows] > 0 ? indptr[nrows] : 1);
    if (!w || upload_csr(w, indptr, cols, vals, nrows) || up(w, w->x, x, nrows)) return -1;
    hipLaunchKernelGGL(kc_spmv, dim3(nblk(nrows)), dim3(TPB), 0, w->st, nrows, w->ip, w->ci, w->v, w->x, w->b);
    LCK(hipGetLastError());
    return down(w, Ax, w->b, nrows);
}

extern "C" void split_CSR_mat_vec_mult(double *Ax, int *indptr_diag, int *cols_diag, double *vals_diag,
                                       int *indptr_off_diag, int *cols_off_diag, double *vals_off_diag, double *x,
                                       int nrows, int *ghosts, int nghost)
{
    (void)cols_off_diag;
    (void)vals_off_diag;
    (void)ghosts;
    done_status(spmv_impl(Ax, indptr_diag, cols_diag, vals_diag, indptr_off_diag, x, nrows, nghost));
}

static int scale_impl(int *indptr, int *cols, double *vals, int *indptr_off_diag, double *b, double *x,
                      double *scale_vect, int local_size, int size, int nghost)
{
    if (check_single_rank(indptr_off_diag, nghost) || check_csr(indptr, cols, vals, local_size)) return -1;
    if (size != local_size) {
        adhk_set_error("scale_linear_system: size != local_size needs ghost dofs (multi-rank)");
        return -1;
    }
    if (!b || !x || !scale_vect) {
        adhk_set_error("scale_linear_system: NULL vector");
        return -1;
    }
    const int n = local_size;
    la_ws *w = ws_get(n, indptr[n] > 0 ? indptr[n] : 1);
    if (!w || upload_csr(w, indptr, cols, vals, n) || up(w, w->b, b, n) || up(w, w->x, x, n)) return -1;
    hipLaunchKernelGGL(kc_scale1, dim3(nblk(n)), dim3(TPB), 0, w->st, n, w->ip, w->v, w->scale, w->b, w->x);
    hipLaunchKernelGGL(kc_scale2, dim3(nblk(n)), dim3(TPB), 0, w->st, n, w->ip, w->ci, w->v, w->scale);
    LCK(hipGetLastError());
    if (down(w, b, w->b, n) || down(w, x, w->x, n) || down(w, scale_vect, w->scale, n)) return -1;
    LCK(hipMemcpyAsync(vals, w->v, (size_t)w->nnz * sizeof(double), hipMemcpyDeviceToHost, w->st));
    LCK(hipStreamSynchronize(w->st));
    return 0;
}

extern "C" void scale_linear_system(int *indptr_diag, int *cols_diag, double *vals_diag, int *indptr_off_diag,
                                    int *cols_off_diag, double *vals_off_diag, double *b, double *x,
                                    double *scale_vect, int local_size, int size, int rank, int *ghosts, int nghost)
{
    (void)cols_off_diag;
    (void)vals_off_diag;
    (void)rank;
    (void)ghosts;
    done_status(scale_impl(indptr_diag, cols_diag, vals_diag, indptr_off_diag, b, x, scale_vect, local_size, size,
                           nghost));
}

static int unscale_impl(double *x, double *x0, double *scale_vect, int n)
{
    if (n <= 0 || !x || !x0 || !scale_vect) {
        adhk_set_error("unscale_linear_system: bad arguments");
        return -1;
    }
    la_ws *w = ws_get(n, 1);
    if (!w || up(w, w->x, x, n) || up(w, w->x0, x0, n) || up(w, w->scale, scale_vect, n)) return -1;
    hipLaunchKernelGGL(kc_unscale, dim3(nblk(n)), dim3(TPB), 0, w->st, n, w->x, w->x0, w->scale);
    LCK(hipGetLastError());
    return down(w, x, w->x, n);
}

extern "C" void unscale_linear_system(double *x, double *x0, double *scale_vect, int local_size)
{
    done_status(unscale_impl(x, x0, scale_vect, local_size));
}

extern "C" int prep_umfpack(int *indptr_diag, int *cols_diag, double *vals_diag, int nrow)
{
    int rc = -1;
    if (!check_csr(indptr_diag, cols_diag, vals_diag, nrow)) {
        la_ws *w = ws_get(nrow, indptr_diag[nrow] > 0 ? indptr_diag[nrow] : 1);
        if (w && !upload_csr(w, indptr_diag, cols_diag, vals_diag, nrow)) {
            const int bs = (g_bs == 3 && nrow % 3 == 0) ? 3 : 1;
            const int nb = nrow / bs;
            hipLaunchKernelGGL(kc_bj_prep, dim3(nblk(nb)), dim3(TPB), 0, w->st, nb, bs, w->ip, w->ci, w->v, w->minv);
            if (hipGetLastError() == hipSuccess && hipStreamSynchronize(w->st) == hipSuccess) {
                w->prep_n = nrow;
                w->prep_bs = bs;
                rc = 0;
            } else {
                adhk_set_error("prep_umfpack: device error");
            }
        }
    }
    done_status(rc);
    return rc;
}

extern "C" int solve_umfpack(double *x, int *indptr_diag, int *cols_diag, double *vals_diag, double *b, int nrow)
{
    (void)indptr_diag;
    (void)cols_diag;
    (void)vals_diag;
    int rc = -1;
    la_ws *w = tl_ws;
    if (!w || w->prep_n != nrow || nrow <= 0 || !x || !b) {
        adhk_set_error("solve_umfpack: no preconditioner prepared for %d rows on this thread (prep_umfpack)", nrow);
    } else if (!up(w, w->b, b, nrow)) {
        const int nb = nrow / w->prep_bs;
        hipLaunchKernelGGL(kc_bj_apply, dim3(nblk(nb)), dim3(TPB), 0, w->st, nb, w->prep_bs, w->minv, w->b, w->x);
        if (hipGetLastError() == hipSuccess && !down(w, x, w->x, nrow)) rc = 0;
    }
    done_status(rc);
    return rc;
}

extern "C" void umfpack_clear(void)
{
    if (tl_ws) tl_ws->prep_n = 0;
    done_status(0);
}

extern "C" void free_bcgstab(void)
{
    if (tl_ws) ws_release(tl_ws);
    tl_ws = NULL;
    done_status(0);
}

extern "C" void Screen_print_CSR(int *indptr, int *indices, double *data, int n)
{
    for (int i = 0; i < n; i++)
        for (int j = indptr[i]; j < indptr[i + 1]; j++) printf("(%d,%d) %.17g\n", i, indices[j], data[j]);
}

/* solve_linear_sys_bcgstab, bcgstab_solver.c:65-335: left-preconditioned BiCGStab of the scaled system
 * with the block-Jacobi preconditioner of prep_umfpack, initial guess M^-1 b, stop when
 * |r|_inf <= 1e-11 + 1e-11 |M^-1 b|_inf or after 100 iterations; x (in: x0) <- (x + x0) * scale_vect;
 * b <- M^-1 b.  Returns 0, or -1 on an error (outputs untouched). */
extern "C" int solve_linear_sys_bcgstab(double *x, int *indptr_diag, int *cols_diag, double *vals_diag,
                                        int *indptr_off_diag, int *cols_off_diag, double *vals_off_diag, double *b,
                                        double *scale_vect, int local_size, int size, int rank, int *ghosts,
                                        int nghost)
{
    (void)cols_off_diag;
    (void)vals_off_diag;
    (void)rank;
    (void)ghosts;
    (void)size;
    const int n = local_size;
    int rc = -1;
    la_ws *w = tl_ws;
    if (check_single_rank(indptr_off_diag, nghost) || check_csr(indptr_diag, cols_diag, vals_diag, n)) {
    } else if (!x || !b || !scale_vect) {
        adhk_set_error("solve_linear_sys_bcgstab: NULL vector");
    } else if (!w || w->prep_n != n) {
        adhk_set_error("solve_linear_sys_bcgstab: no preconditioner prepared for %d rows on this thread "
                       "(prep_umfpack)", n);
    } else if ((w = ws_get(n, indptr_diag[n] > 0 ? indptr_diag[n] : 1)) != NULL &&
               !upload_csr(w, indptr_diag, cols_diag, vals_diag, n) && !up(w, w->x, x, n) && !up(w, w->b, b, n) &&
               !up(w, w->scale, scale_vect, n)) {
        int its = 0;
        if (!bicg_core(ws_view(w, n, w->prep_bs), w->st, LA_TOL, LA_MAX_IT, &its) && !down(w, b, w->b, n) &&
            !down(w, x, w->x, n)) {
            tl_its = its;
            rc = 0;
        }
    }
    done_status(rc);
    return rc;
}

/* ------------------------------------------- vector helpers of the struct-level surface (ref_shim.c) */

__global__ void kc_add(int n, double *y, const double *x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i]; /* increment_function, newton_tools.c:261 */
}

/* get_residual_norms, newton_tools.c:303-373: per workgroup the max |r| and max |dsol| with their first
 * dof, the sum of r^2 over the dofs not in bc_mask (nor outside include_dof), and a NaN count */
__global__ __launch_bounds__(TPB) void kc_norms(int n, const double *r, const double *d, const int *bc,
                                                const int *inc, double *part, int *ipart)
{
    __shared__ double sv[3][TPB];
    __shared__ int si[3][TPB];
    double rmax = 0.0, dmax = 0.0, l2 = 0.0;
    int ir = 0x7fffffff, id = 0x7fffffff, nan = 0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
        const int use = !(inc && inc[i] != 1) && !(bc && bc[i] == 1);
        const double a = fabs(r[i]), b = fabs(d[i]);
        if (r[i] != r[i]) nan++;
        if (use && b > dmax) { dmax = b; id = i; }
        if (use && a > rmax) { rmax = a; ir = i; }
        if (use) l2 = __fma_rn(r[i], r[i], l2);
    }
    sv[0][threadIdx.x] = rmax; si[0][threadIdx.x] = ir;
    sv[1][threadIdx.x] = dmax; si[1][threadIdx.x] = id;
    sv[2][threadIdx.x] = l2; si[2][threadIdx.x] = nan;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            for (int c = 0; c < 2; c++) { /* the larger value; on a tie the first dof, as the sequential scan */
                const double a = sv[c][threadIdx.x], b = sv[c][threadIdx.x + o];
                const int ia = si[c][threadIdx.x], ib = si[c][threadIdx.x + o];
                if (b > a || (b == a && ib < ia)) { sv[c][threadIdx.x] = b; si[c][threadIdx.x] = ib; }
            }
            sv[2][threadIdx.x] += sv[2][threadIdx.x + o];
            si[2][threadIdx.x] += si[2][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++) {
            part[c * NPART_MAX + blockIdx.x] = sv[c][0];
            ipart[c * NPART_MAX + blockIdx.x] = si[c][0];
        }
}

extern "C" int adhk_la_vec_add(double *y, const double *x, int n)
{
    if (n <= 0 || !y || !x) {
        adhk_set_error("increment: bad arguments");
        return -1;
    }
    la_ws *w = ws_get(n, 1);
    if (!w || up(w, w->x, y, n) || up(w, w->s, x, n)) return -1;
    hipLaunchKernelGGL(kc_add, dim3(nblk(n)), dim3(TPB), 0, w->st, n, w->x, w->s);
    LCK(hipGetLastError());
    return down(w, y, w->x, n);
}

extern "C" int adhk_la_residual_norms(int n, const double *residual, const double *dsol, const int *bc_mask,
                                      const int *include_dof, double *rmax, double *l2sum, double *dmax, int *irmax,
                                      int *idmax, int *nnan)
{
    if (n <= 0 || !residual || !dsol) {
        adhk_set_error("get_residual_norms: bad arguments");
        return -1;
    }
    /* room for the vectors, the two integer masks (in the int buffer of the CSR columns) and the integer
     * partials (in the scratch vector tmp) */
    const int nws = n > 2 * NPART_MAX ? n : 2 * NPART_MAX;
    la_ws *w = ws_get(nws, 2 * n);
    if (!w || up(w, w->r, residual, n) || up(w, w->s, dsol, n)) return -1;
    int *dbc = NULL, *dinc = NULL, *ipart = (int *)w->tmp;
    if (bc_mask) {
        dbc = w->ci;
        LCK(hipMemcpyAsync(dbc, bc_mask, (size_t)n * sizeof(int), hipMemcpyHostToDevice, w->st));
    }
    if (include_dof) {
        dinc = w->ci + n;
        LCK(hipMemcpyAsync(dinc, include_dof, (size_t)n * sizeof(int), hipMemcpyHostToDevice, w->st));
    }
    const int g = nblk(n) < NPART_MAX ? nblk(n) : NPART_MAX;
    hipLaunchKernelGGL(kc_norms, dim3(g), dim3(TPB), 0, w->st, n, w->r, w->s, dbc, dinc, w->part, ipart);
    LCK(hipGetLastError());
    double hp[3 * NPART_MAX];
    int hi[3 * NPART_MAX];
    LCK(hipMemcpyAsync(hp, w->part, 3 * NPART_MAX * sizeof(double), hipMemcpyDeviceToHost, w->st));
    LCK(hipMemcpyAsync(hi, ipart, 3 * NPART_MAX * sizeof(int), hipMemcpyDeviceToHost, w->st));
    LCK(hipStreamSynchronize(w->st));
    double rm = 0.0, dm = 0.0, s2 = 0.0;
    int ir = 0, id = 0, nn = 0;
    for (int b = 0; b < g; b++) {
        if (hp[b] > rm || (hp[b] == rm && hp[b] > 0.0 && hi[b] < ir)) { rm = hp[b]; ir = hi[b]; }
        if (hp[NPART_MAX + b] > dm || (hp[NPART_MAX + b] == dm && dm > 0.0 && hi[NPART_MAX + b] < id)) {
            dm = hp[NPART_MAX + b];
            id = hi[NPART_MAX + b];
        }
        s2 += hp[2 * NPART_MAX + b];
        nn += hi[2 * NPART_MAX + b];
    }
    *rmax = rm;
    *dmax = dm;
    *l2sum = s2;
    *irmax = rm > 0.0 ? ir : 0;
    *idmax = dm > 0.0 ? id : 0;
    *nnan = nn;
    return 0;
}

/* ---------------------------------------------------- BiCGStab on device arrays without allocation */

/* Scratch of adhk_bcgstab_device_noalloc, one copy per GPU (the array's host-side name is not a device
 * address: host code uses the per-device pointer from hipGetSymbolAddress).  Layout for n rows: the
 * vectors x0 r q p Ap s As (n each), the block inverses (3n), the partial sums (4 * NPART_MAX) and the
 * scalars. */
#define NOALLOC_DOUBLES ((size_t)10 * ADHK_NOALLOC_MAX_N + 4 * NPART_MAX + 16)
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

#define ADHK_MAX_DEVICES 64
static double *g_scratch[ADHK_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

/* this device's copy of scratch, looked up once per device with that device current */
static double *scratch_of_current_device(void)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= ADHK_MAX_DEVICES) {
        adhk_set_error("adhk_bcgstab_device_noalloc: no current device");
        return NULL;
    }
    pthread_mutex_lock(&g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = NULL;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            adhk_set_error("adhk_bcgstab_device_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        else
            g_scratch[dev] = (double *)ptr;
    }
    double *p = g_scratch[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    return p;
}

extern "C" int adhk_bcgstab_device_noalloc(double *x, const int *indptr, const int *cols, const double *vals,
                                           double *b, const double *scale_vect, int nrow, int bs, void *stream,
                                           int *its)
{
    if (nrow <= 0 || nrow > ADHK_NOALLOC_MAX_N) {
        adhk_set_error("adhk_bcgstab_device_noalloc: %d rows (1..%d supported: the scratch is a fixed %zu MB)",
                       nrow, ADHK_NOALLOC_MAX_N, NOALLOC_DOUBLES * sizeof(double) >> 20);
        return -1;
    }
    if (bs != 1 && bs != 3) bs = 3;
    if (bs == 3 && nrow % 3 != 0) bs = 1;
    if (!x || !indptr || !cols || !vals || !b || !scale_vect) {
        adhk_set_error("adhk_bcgstab_device_noalloc: NULL argument");
        return -1;
    }
    double *scr = scratch_of_current_device();
    if (!scr) return -1;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)nrow;
    la_view V;
    memset(&V, 0, sizeof V);
    V.n = nrow;
    V.bs = bs;
    V.nb = nrow / bs;
    V.ip = indptr;
    V.ci = cols;
    V.v = vals;
    V.scale = scale_vect;
    V.b = b;
    V.x = x;
    V.x0 = scr;
    V.r = scr + n;
    V.q = scr + 2 * n;
    V.p = scr + 3 * n;
    V.Ap = scr + 4 * n;
    V.s = scr + 5 * n;
    V.As = scr + 6 * n;
    double *minv = scr + 7 * n; /* 3n */
    V.minv = minv;
    V.part = scr + 10 * n;
    V.sc = (la_sc *)(scr + 10 * n + 4 * NPART_MAX);
    hipLaunchKernelGGL(kc_bj_prep, dim3(nblk(V.nb)), dim3(TPB), 0, st, V.nb, bs, indptr, cols, vals, minv);
    LCK(hipGetLastError());
    return bicg_core(V, st, LA_TOL, LA_MAX_IT, its);
}
