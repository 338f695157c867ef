/*
 * la_csr.hip -- the reference's plain-pointer linear-algebra surface
 * (src/la/la.h) with its exact names and signatures, over host CSR arrays.
 * Each call uploads its operands, computes on the GPU and downloads the
 * results, so a caller of the reference's la/ can link this instead.
 *
 * Single rank: the split-CSR off-diagonal block must be absent
 * (indptr_off_diag == NULL), as it is when the reference runs on one process
 * (slin_sys_init_ptrs.c:26).  prep_umfpack/solve_umfpack keep their names but
 * build and apply a block-Jacobi preconditioner (block size 3 by default, the
 * SW2 node block; 1 for point Jacobi) instead of the UMFPACK LU
 * (bcgstab_solver.c:371-424).  Module state mirrors the reference's statics
 * (bcgstab_solver.c:5-22): one workspace per process, not thread-safe.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "adhk_internal.h"

#define TPB 256
#define NPART 1024
#define D_SMALL 2.220446049250313080847e-16

typedef struct {
    hipStream_t st;
    int n, nnz, cap_n, cap_nnz, bs, prep_n, prep_bs;
    int *indptr, *cols;
    double *vals, *minv;
    double *b, *x, *r, *p, *Ap, *Mp, *q, *s, *As, *Ms, *x0, *scale, *tmp;
    double *part, *hpart;
} la_ws;

static la_ws W;
static int g_bs = 3;

extern "C" void adhk_la_set_block_size(int bs) { g_bs = (bs == 1) ? 1 : 3; }

static void die(const char *what, hipError_t e)
{
    fprintf(stderr, "adhk la: %s: %s\n", what, hipGetErrorString(e));
    adhk_set_error("la: %s: %s", what, hipGetErrorString(e));
    abort(); /* the reference aborts on solver errors (tl_error -> exit) */
}
#define HCK(x)                                                                                                    \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) die(#x, e_);                                                                        \
    } while (0)

static void ws_free(void)
{
    void *ptrs[] = {W.indptr, W.cols, W.vals, W.minv, W.b, W.x, W.r, W.p, W.Ap, W.Mp, W.q,
                    W.s, W.As, W.Ms, W.x0, W.scale, W.tmp, W.part};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
        if (ptrs[i]) hipFree(ptrs[i]);
    if (W.hpart) hipHostFree(W.hpart);
    if (W.st) hipStreamDestroy(W.st);
    memset(&W, 0, sizeof W);
}

static void ws_reserve(int n, int nnz)
{
    if (!W.st) {
        HCK(hipStreamCreate(&W.st));
        HCK(hipMalloc((void **)&W.part, 4 * NPART * sizeof(double)));
        HCK(hipHostMalloc((void **)&W.hpart, 4 * NPART * sizeof(double), 0));
    }
    if (n > W.cap_n) {
        double **vs[] = {&W.b, &W.x, &W.r, &W.p, &W.Ap, &W.Mp, &W.q, &W.s, &W.As, &W.Ms, &W.x0, &W.scale, &W.tmp};
        for (size_t i = 0; i < sizeof vs / sizeof vs[0]; i++) {
            if (*vs[i]) hipFree(*vs[i]);
            HCK(hipMalloc((void **)vs[i], (size_t)n * sizeof(double)));
        }
        if (W.indptr) hipFree(W.indptr);
        if (W.minv) hipFree(W.minv);
        HCK(hipMalloc((void **)&W.indptr, ((size_t)n + 1) * sizeof(int)));
        HCK(hipMalloc((void **)&W.minv, (size_t)n * 3 * sizeof(double)));
        W.cap_n = n;
    }
    if (nnz > W.cap_nnz) {
        if (W.cols) hipFree(W.cols);
        if (W.vals) hipFree(W.vals);
        HCK(hipMalloc((void **)&W.cols, (size_t)nnz * sizeof(int)));
        HCK(hipMalloc((void **)&W.vals, (size_t)nnz * sizeof(double)));
        W.cap_nnz = nnz;
    }
}

static void upload_csr(const int *indptr, const int *cols, const double *vals, int n)
{
    int nnz = indptr[n];
    ws_reserve(n, nnz > 0 ? nnz : 1);
    W.n = n;
    W.nnz = nnz;
    HCK(hipMemcpyAsync(W.indptr, indptr, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice, W.st));
    if (nnz > 0) {
        HCK(hipMemcpyAsync(W.cols, cols, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, W.st));
        HCK(hipMemcpyAsync(W.vals, vals, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, W.st));
    }
}

static void up(double *d, const double *h, int n)
{
    HCK(hipMemcpyAsync(d, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, W.st));
}
static void down(double *h, const double *d, int n)
{
    HCK(hipMemcpyAsync(h, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, W.st));
    HCK(hipStreamSynchronize(W.st));
}

/* ---- kernels ---- */
__global__ void kc_spmv(int n, const int *ip, const int *ci, const double *v, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = ip[i]; j < ip[i + 1]; j++) s += v[j] * x[ci[j]];
    y[i] = s;
}

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

/* y += a x */
__global__ void kc_axpy(int n, double a, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void kc_scal(int n, double a, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] *= a;
}
__global__ void kc_copy(int n, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}
__global__ void kc_unscale(int n, double *x, const double *x0, const double *s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (x[i] + x0[i]) * s[i];
}
/* per-workgroup partial dot (mode 0) or max-abs (mode 1) */
__global__ void kc_partial(int n, const double *x, const double *y, double *part, int mode)
{
    __shared__ double sh[TPB];
    double a = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        a = mode ? fmax(a, fabs(x[i])) : a + x[i] * y[i];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = mode ? fmax(sh[threadIdx.x], sh[threadIdx.x + o])
                                                    : sh[threadIdx.x] + sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static int nblk(int n) { return (n + TPB - 1) / TPB; }
static int npart(int n)
{
    int g = nblk(n);
    return g > NPART ? NPART : (g < 1 ? 1 : g);
}

static double reduce(const double *x, const double *y, int n, int mode)
{
    int g = npart(n);
    hipLaunchKernelGGL(kc_partial, dim3(g), dim3(TPB), 0, W.st, n, x, y, W.part, mode);
    HCK(hipGetLastError());
    HCK(hipMemcpyAsync(W.hpart, W.part, (size_t)g * sizeof(double), hipMemcpyDeviceToHost, W.st));
    HCK(hipStreamSynchronize(W.st));
    double a = 0.0;
    for (int i = 0; i < g; i++) a = mode ? fmax(a, W.hpart[i]) : a + W.hpart[i];
    return a;
}
static double ddot(const double *x, const double *y, int n) { return reduce(x, y, n, 0); }
static double dinf(const double *x, int n) { return reduce(x, x, n, 1); }

static void spmv(const double *x, double *y, int n)
{
    hipLaunchKernelGGL(kc_spmv, dim3(nblk(n)), dim3(TPB), 0, W.st, n, W.indptr, W.cols, W.vals, x, y);
    HCK(hipGetLastError());
}
static void precond(const double *b, double *x, int n)
{
    int nb = n / W.prep_bs;
    hipLaunchKernelGGL(kc_bj_apply, dim3(nblk(nb)), dim3(TPB), 0, W.st, nb, W.prep_bs, W.minv, b, x);
    HCK(hipGetLastError());
}
static void axpy(double a, const double *x, double *y, int n)
{
    hipLaunchKernelGGL(kc_axpy, dim3(nblk(n)), dim3(TPB), 0, W.st, n, a, x, y);
    HCK(hipGetLastError());
}
static void scal(double a, double *y, int n)
{
    hipLaunchKernelGGL(kc_scal, dim3(nblk(n)), dim3(TPB), 0, W.st, n, a, y);
    HCK(hipGetLastError());
}
static void copy(const double *x, double *y, int n)
{
    hipLaunchKernelGGL(kc_copy, dim3(nblk(n)), dim3(TPB), 0, W.st, n, x, y);
    HCK(hipGetLastError());
}
static void zero(double *y, int n) { scal(0.0, y, n); }

static void check_single_rank(const int *indptr_off_diag, int nghost)
{
    if (indptr_off_diag != NULL || nghost != 0) {
        fprintf(stderr, "adhk la: split-CSR off-diagonal blocks (multi-rank) are not supported on this path\n");
        abort();
    }
}

extern "C" void split_CSR_mat_vec_mult(double *Ax, int *indptr_diag, int *cols_diag, double *vals_diag,
                                       int *indptr_off_diag, int *cols_off_diag, double *vals_off_diag, double *x,
                                       int nrows, int *ghosts, int nghost)
{
    (void)cols_off_diag;
    (void)vals_off_diag;
    (void)ghosts;
    check_single_rank(indptr_off_diag, nghost);
    upload_csr(indptr_diag, cols_diag, vals_diag, nrows);
    up(W.x, x, nrows);
    spmv(W.x, W.b, nrows);
    down(Ax, W.b, nrows);
}

extern "C" void scale_linear_system(int *indptr_diag, int *cols_diag, double *vals_diag, int *indptr_off_diag,
                                    int *cols_off_diag, double *vals_off_diag, double *b, double *x,
                                    double *scale_vect, int local_size, int size, int rank, int *ghosts, int nghost)
{
    (void)cols_off_diag;
    (void)vals_off_diag;
    (void)rank;
    (void)ghosts;
    check_single_rank(indptr_off_diag, nghost);
    if (size != local_size) {
        fprintf(stderr, "adhk la: size != local_size needs ghost dofs (multi-rank)\n");
        abort();
    }
    upload_csr(indptr_diag, cols_diag, vals_diag, local_size);
    up(W.b, b, local_size);
    up(W.x, x, local_size);
    hipLaunchKernelGGL(kc_scale1, dim3(nblk(local_size)), dim3(TPB), 0, W.st, local_size, W.indptr, W.vals, W.scale,
                       W.b, W.x);
    HCK(hipGetLastError());
    hipLaunchKernelGGL(kc_scale2, dim3(nblk(local_size)), dim3(TPB), 0, W.st, local_size, W.indptr, W.cols, W.vals,
                       W.scale);
    HCK(hipGetLastError());
    down(b, W.b, local_size);
    down(x, W.x, local_size);
    down(scale_vect, W.scale, local_size);
    HCK(hipMemcpyAsync(vals_diag, W.vals, (size_t)W.nnz * sizeof(double), hipMemcpyDeviceToHost, W.st));
    HCK(hipStreamSynchronize(W.st));
}

extern "C" void unscale_linear_system(double *x, double *x0, double *scale_vect, int local_size)
{
    ws_reserve(local_size, 1);
    up(W.x, x, local_size);
    up(W.x0, x0, local_size);
    up(W.scale, scale_vect, local_size);
    hipLaunchKernelGGL(kc_unscale, dim3(nblk(local_size)), dim3(TPB), 0, W.st, local_size, W.x, W.x0, W.scale);
    HCK(hipGetLastError());
    down(x, W.x, local_size);
}

extern "C" int prep_umfpack(int *indptr_diag, int *cols_diag, double *vals_diag, int nrow)
{
    upload_csr(indptr_diag, cols_diag, vals_diag, nrow);
    int bs = (g_bs == 3 && nrow % 3 == 0) ? 3 : 1;
    int nb = nrow / bs;
    hipLaunchKernelGGL(kc_bj_prep, dim3(nblk(nb)), dim3(TPB), 0, W.st, nb, bs, W.indptr, W.cols, W.vals, W.minv);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(W.st));
    W.prep_n = nrow;
    W.prep_bs = bs;
    return 0;
}

extern "C" int solve_umfpack(double *x, int *indptr_diag, int *cols_diag, double *vals_diag, double *b, int nrow)
{
    (void)indptr_diag;
    (void)cols_diag;
    (void)vals_diag;
    if (W.prep_n != nrow) {
        fprintf(stderr, "adhk la: solve_umfpack before prep_umfpack for this size\n");
        abort();
    }
    up(W.b, b, nrow);
    precond(W.b, W.x, nrow);
    down(x, W.x, nrow);
    return 0;
}

extern "C" void umfpack_clear(void) { W.prep_n = 0; }
extern "C" void free_bcgstab(void) { ws_free(); }

extern "C" void Screen_print_CSR(int *indptr, int *indices, double *data, int n)
{
    for (int i = 0; i < n; i++)
        for (int j = indptr[i]; j < indptr[i + 1]; j++) printf("(%d,%d) %.17g\n", i, indices[j], data[j]);
}

/* solve_linear_sys_bcgstab, bcgstab_solver.c:65-335 */
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
    check_single_rank(indptr_off_diag, nghost);
    const int n = local_size;
    if (W.prep_n != n) {
        fprintf(stderr, "adhk la: solve_linear_sys_bcgstab before prep_umfpack for this size\n");
        abort();
    }
    const double tol = 1e-11; /* min_conv_tol, conv_tol (:20-21) */
    const int max_it = 100;   /* :22 */
    upload_csr(indptr_diag, cols_diag, vals_diag, n);
    up(W.x0, x, n);
    up(W.tmp, b, n);
    zero(W.x, n);
    zero(W.p, n);
    zero(W.Ap, n);
    zero(W.Mp, n);
    precond(W.tmp, W.b, n); /* b <- M^-1 b */
    copy(W.b, W.x, n);
    spmv(W.x, W.Mp, n);
    precond(W.Mp, W.Ap, n);
    copy(W.b, W.r, n);
    axpy(-1, W.Ap, W.r, n);
    zero(W.Mp, n);
    zero(W.Ap, n);
    copy(W.r, W.q, n);
    double rnorm = dinf(W.r, n), bnorm = dinf(W.b, n);
    double rhop = 1.0, alpha = 1.0, omega = 1.0, beta, gamma, rho = ddot(W.r, W.q, n);
    int it = 0;
    while (rnorm > (tol + bnorm * tol) && it < max_it) {
        it++;
        beta = (rho * alpha) / (rhop * omega);
        axpy(-omega, W.Ap, W.p, n);
        scal(beta, W.p, n);
        axpy(1, W.r, W.p, n);
        spmv(W.p, W.Mp, n);
        precond(W.Mp, W.Ap, n);
        gamma = ddot(W.q, W.Ap, n);
        alpha = rho / gamma;
        copy(W.r, W.s, n);
        axpy(-alpha, W.Ap, W.s, n);
        spmv(W.s, W.Ms, n);
        precond(W.Ms, W.As, n);
        omega = ddot(W.As, W.s, n);
        gamma = ddot(W.As, W.As, n);
        omega = omega / gamma;
        rhop = rho;
        rho = -1.0 * omega * ddot(W.As, W.q, n);
        axpy(alpha, W.p, W.x, n);
        axpy(omega, W.s, W.x, n);
        copy(W.s, W.r, n);
        axpy(-omega, W.As, W.r, n);
        rnorm = dinf(W.r, n);
    }
    up(W.scale, scale_vect, n);
    hipLaunchKernelGGL(kc_unscale, dim3(nblk(n)), dim3(TPB), 0, W.st, n, W.x, W.x0, W.scale);
    HCK(hipGetLastError());
    down(b, W.b, n); /* the reference leaves M^-1 b in b (:135-137) */
    down(x, W.x, n);
    return 0;
}
