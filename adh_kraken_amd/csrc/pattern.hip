/*
 * pattern.hip -- CSR sparsity pattern of the SW2 system built on the GPU
 * (replaces slin_sys_init_sparsity_mono, slin_sys_init_sparsity_mono.c:28-494).
 *
 * Every element emits its node pairs (row node a, col node b) in element order
 * (2-D elements, then 1-D).  A stable radix sort on key = a*N + b gives, in one
 * pass, (1) the sorted, duplicate-free node adjacency of each row -- the
 * reference's per-row sort + unique, at node-block granularity -- and (2) for
 * every 3x3 block the list of elemental contributions in element order, which
 * drives the atomic-free assembly.  The scalar CSR of the reference is the
 * node-block pattern expanded by dof(node, var) = 3*node + var; cols stay
 * ascending because each block row lists node columns in ascending order.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "adhk_internal.h"

#define CHECK(x)                                                                                                  \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s", __FILE__, __LINE__, hipGetErrorString(e_));                                \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)

__global__ void k_emit_pairs(int N, int E, int S, const int *tri, const int *seg, unsigned long long *keys, int *vals)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long ntri = 9L * E;
    if (t < ntri) {
        int e = (int)(t / 9), a = (int)((t % 9) / 3), b = (int)(t % 3);
        unsigned long long na = (unsigned long long)tri[a * E + e], nb = (unsigned long long)tri[b * E + e];
        keys[t] = na * (unsigned long long)N + nb;
        vals[t] = (int)t;
    } else if (t < ntri + 4L * S) {
        long u = t - ntri;
        int s = (int)(u / 4), a = (int)((u % 4) / 2), b = (int)(u % 2);
        unsigned long long na = (unsigned long long)seg[2 * s + a], nb = (unsigned long long)seg[2 * s + b];
        keys[t] = na * (unsigned long long)N + nb;
        vals[t] = (int)t;
    }
}

__global__ void k_head_flags(long n, const unsigned long long *keys, int *flags)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_blocks(long n, int N, const unsigned long long *keys, const int *flags, const int *incl,
                         int *cptr, int *bcol, int *brow_of)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    int b = incl[i] - 1;
    cptr[b] = (int)i;
    bcol[b] = (int)(keys[i] % (unsigned long long)N);
    brow_of[b] = (int)(keys[i] / (unsigned long long)N);
}

/* brow[n] = first block whose row >= n (lower bound); diag_blk[n] = block (n, n) or -1 */
__global__ void k_rowptr(int N, int nnzb, const int *brow_of, const int *bcol, int *brow, int *diag_blk)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n > N) return;
    int lo = 0, hi = nnzb;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (brow_of[mid] < n) lo = mid + 1; else hi = mid;
    }
    brow[n] = lo;
    if (n < N) {
        int d = -1;
        for (int j = lo; j < nnzb && brow_of[j] == n; j++)
            if (bcol[j] == n) { d = j; break; }
        diag_blk[n] = d;
    }
}

extern "C" int adhk_build_pattern(adhk_ctx *c)
{
    long np = 9L * c->E + 4L * c->S;
    unsigned long long *keys = NULL, *keys_s = NULL;
    int *vals = NULL, *flags = NULL, *incl = NULL, *brow_of = NULL, *hlast = NULL;
    void *tmp = NULL;
    size_t tmp_b = 0, tmp_b2 = 0;
    int nnzb = 0;
    int bits = 1;
    unsigned long long maxkey = (unsigned long long)c->N * (unsigned long long)c->N;
    while (bits < 64 && (1ULL << bits) < maxkey) bits++;
    if (np <= 0 || np >= 2147483647L) {
        adhk_set_error("adhk_build_pattern: unsupported pair count %ld", np);
        return -1;
    }
    c->npairs = (int)np;
    CHECK(hipMalloc((void **)&keys, np * sizeof(unsigned long long)));
    CHECK(hipMalloc((void **)&keys_s, np * sizeof(unsigned long long)));
    CHECK(hipMalloc((void **)&vals, np * sizeof(int)));
    CHECK(hipMalloc((void **)&c->contrib, np * sizeof(int)));
    CHECK(hipMalloc((void **)&flags, np * sizeof(int)));
    CHECK(hipMalloc((void **)&incl, np * sizeof(int)));
    hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, c->N, c->E, c->S,
                       c->tri, c->seg, keys, vals);
    CHECK(hipGetLastError());
    CHECK(hipcub::DeviceRadixSort::SortPairs(NULL, tmp_b, keys, keys_s, vals, c->contrib, (int)np, 0, bits,
                                             c->stream));
    CHECK(hipcub::DeviceScan::InclusiveSum(NULL, tmp_b2, flags, incl, (int)np, c->stream));
    if (tmp_b2 > tmp_b) tmp_b = tmp_b2;
    CHECK(hipMalloc(&tmp, tmp_b));
    tmp_b2 = tmp_b;
    CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_b2, keys, keys_s, vals, c->contrib, (int)np, 0, bits,
                                             c->stream));
    hipLaunchKernelGGL(k_head_flags, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, np, keys_s, flags);
    CHECK(hipGetLastError());
    tmp_b2 = tmp_b;
    CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tmp_b2, flags, incl, (int)np, c->stream));
    CHECK(hipMemcpyAsync(&nnzb, incl + (np - 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CHECK(hipStreamSynchronize(c->stream));
    c->nnzb = nnzb;
    CHECK(hipMalloc((void **)&c->cptr, ((size_t)nnzb + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&c->bcol, (size_t)nnzb * sizeof(int)));
    CHECK(hipMalloc((void **)&brow_of, (size_t)nnzb * sizeof(int)));
    CHECK(hipMalloc((void **)&c->brow, ((size_t)c->N + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&c->diag_blk, (size_t)c->N * sizeof(int)));
    hipLaunchKernelGGL(k_blocks, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, np, c->N, keys_s, flags,
                       incl, c->cptr, c->bcol, brow_of);
    CHECK(hipGetLastError());
    CHECK(hipMemcpyAsync(c->cptr + nnzb, &c->npairs, sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_rowptr, dim3((c->N + 1 + 255) / 256), dim3(256), 0, c->stream, c->N, nnzb, brow_of, c->bcol,
                       c->brow, c->diag_blk);
    CHECK(hipGetLastError());
    CHECK(hipStreamSynchronize(c->stream));
    hipFree(keys);
    hipFree(keys_s);
    hipFree(vals);
    hipFree(flags);
    hipFree(incl);
    hipFree(brow_of);
    hipFree(tmp);
    (void)hlast;
    return 0;
fail:
    hipFree(keys);
    hipFree(keys_s);
    hipFree(vals);
    hipFree(flags);
    hipFree(incl);
    hipFree(brow_of);
    hipFree(tmp);
    return -1;
}

/* scalar CSR export: row 3n+r lists, for each block column m of node row n (ascending), 3m, 3m+1, 3m+2 */
__global__ void k_csr_pattern(int N, int nnzb, const int *brow, const int *bcol, int *indptr, int *cols)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int b0 = brow[n], deg = brow[n + 1] - b0;
    for (int r = 0; r < 3; r++) {
        int rs = 9 * b0 + 3 * r * deg;
        indptr[3 * n + r] = rs;
        for (int j = 0; j < deg; j++)
            for (int cc = 0; cc < 3; cc++) cols[rs + 3 * j + cc] = 3 * bcol[b0 + j] + cc;
    }
    if (n == N - 1) indptr[3 * N] = 9 * nnzb;
}

__global__ void k_csr_vals(int N, const int *brow, const double *bvals, double *vals)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int b0 = brow[n], deg = brow[n + 1] - b0;
    for (int r = 0; r < 3; r++) {
        int rs = 9 * b0 + 3 * r * deg;
        for (int j = 0; j < deg; j++)
            for (int cc = 0; cc < 3; cc++) vals[rs + 3 * j + cc] = bvals[9 * (size_t)(b0 + j) + 3 * r + cc];
    }
}

extern "C" int adhk_launch_csr_pattern(adhk_ctx *c, int *d_indptr, int *d_cols)
{
    hipLaunchKernelGGL(k_csr_pattern, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->nnzb, c->brow,
                       c->bcol, d_indptr, d_cols);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        adhk_set_error("csr_pattern: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

extern "C" int adhk_launch_csr_vals(adhk_ctx *c, double *d_vals)
{
    hipLaunchKernelGGL(k_csr_vals, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->brow, c->bvals,
                       d_vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        adhk_set_error("csr_vals: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}
