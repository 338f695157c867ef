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

/* cpos[code] = position of the elemental block `code` in the sorted contribution list */
__global__ void k_cpos(long n, const int *contrib, int *cpos)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cpos[contrib[i]] = (int)i;
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

/* SELL-64: slice width = widest block row among the slice's 64 nodes */
__global__ void k_slice_width(int N, int nslice, const int *brow, int *sw)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslice) return;
    int w = 0;
    for (int l = 0; l < 64; l++) {
        int n = s * 64 + l;
        if (n < N) {
            int d = brow[n + 1] - brow[n];
            if (d > w) w = d;
        }
    }
    sw[s] = w;
}

/* one thread per (slice, lane): the node's diagonal block in slot 0, its other blocks in column order,
 * then zero padding (the Krylov SpMV takes the folded diagonal block as the identity without reading it) */
__global__ void k_sell_fill(int N, int nslice, const int *slice_ptr, const int *brow, const int *bcol, int *scol,
                            int *sell2blk, int *blk2sell, int *diag_slot)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslice * 64) return;
    int s = t >> 6, l = t & 63, n = t;
    int c0 = slice_ptr[s], w = slice_ptr[s + 1] - c0;
    int b0 = 0, d = 0, bd = -1;
    if (n < N) {
        b0 = brow[n];
        d = brow[n + 1] - b0;
        for (int k = 0; k < d; k++)
            if (bcol[b0 + k] == n) bd = k;
        diag_slot[n] = bd >= 0 ? 0 : -1;
    }
    for (int k = 0; k < w; k++) {
        int pos = (c0 + k) * 64 + l;
        if (k < d) {
            /* slot k holds block: the diagonal first, then the others in column order */
            const int kk = bd < 0 ? k : (k == 0 ? bd : (k <= bd ? k - 1 : k));
            int b = b0 + kk;
            scol[pos] = bcol[b];
            sell2blk[pos] = b;
            blk2sell[b] = pos;
        } else {
            scol[pos] = (n < N) ? n : 0;
            sell2blk[pos] = -1;
        }
    }
}

static int build_sell(adhk_ctx *c)
{
    int *sw = NULL;
    void *tmp = NULL;
    size_t tb = 0;
    c->nslice = (c->N + 63) / 64;
    CHECK(hipMalloc((void **)&sw, ((size_t)c->nslice + 1) * sizeof(int)));
    CHECK(hipMemsetAsync(sw, 0, ((size_t)c->nslice + 1) * sizeof(int), c->stream));
    CHECK(hipMalloc((void **)&c->slice_ptr, ((size_t)c->nslice + 1) * sizeof(int)));
    hipLaunchKernelGGL(k_slice_width, dim3((c->nslice + 255) / 256), dim3(256), 0, c->stream, c->N, c->nslice, c->brow,
                       sw);
    CHECK(hipGetLastError());
    CHECK(hipcub::DeviceScan::ExclusiveSum(NULL, tb, sw, c->slice_ptr, c->nslice + 1, c->stream));
    CHECK(hipMalloc(&tmp, tb));
    CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, sw, c->slice_ptr, c->nslice + 1, c->stream));
    CHECK(hipMemcpyAsync(&c->nslot, c->slice_ptr + c->nslice, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CHECK(hipStreamSynchronize(c->stream));
    if ((long)c->nslot * 64 * 9 >= 2147483647L) {
        adhk_set_error("SELL layout too large for 32-bit indices");
        goto fail;
    }
    CHECK(hipMalloc((void **)&c->scol, (size_t)c->nslot * 64 * sizeof(int)));
    CHECK(hipMalloc((void **)&c->sell2blk, (size_t)c->nslot * 64 * sizeof(int)));
    CHECK(hipMalloc((void **)&c->blk2sell, (size_t)c->nnzb * sizeof(int)));
    CHECK(hipMalloc((void **)&c->diag_slot, (size_t)c->N * sizeof(int)));
    CHECK(hipMalloc((void **)&c->diag_id, (size_t)(c->N > 0 ? c->N : 1)));
    CHECK(hipMemsetAsync(c->diag_id, 0, (size_t)(c->N > 0 ? c->N : 1), c->stream));
    hipLaunchKernelGGL(k_sell_fill, dim3((c->nslice * 64 + 255) / 256), dim3(256), 0, c->stream, c->N, c->nslice,
                       c->slice_ptr, c->brow, c->bcol, c->scol, c->sell2blk, c->blk2sell, c->diag_slot);
    CHECK(hipGetLastError());
    CHECK(hipStreamSynchronize(c->stream));
    hipFree(sw);
    hipFree(tmp);
    return 0;
fail:
    hipFree(sw);
    hipFree(tmp);
    return -1;
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
    CHECK(hipMalloc((void **)&c->cpos, np * sizeof(int)));
    hipLaunchKernelGGL(k_cpos, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, np, c->contrib, c->cpos);
    CHECK(hipGetLastError());
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
    return build_sell(c);
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

__global__ void k_csr_vals(int N, const int *brow, const int *blk2sell, const double *bvals, double *vals)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int b0 = brow[n], deg = brow[n + 1] - b0;
    for (int r = 0; r < 3; r++) {
        int rs = 9 * b0 + 3 * r * deg;
        for (int j = 0; j < deg; j++) {
            int pos = blk2sell[b0 + j];
            size_t base = (size_t)(pos >> 6) * 9 * 64 + (pos & 63);
            for (int cc = 0; cc < 3; cc++) vals[rs + 3 * j + cc] = bvals[base + (size_t)(3 * r + cc) * 64];
        }
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
    hipLaunchKernelGGL(k_csr_vals, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->brow, c->blk2sell,
                       c->bvals, d_vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        adhk_set_error("csr_vals: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}
