/*
 * pattern.hip -- CSR sparsity pattern of the SW2 system built on the GPU
 * (replaces slin_sys_init_sparsity_mono, slin_sys_init_sparsity_mono.c:28-494).
 *
 * Every element emits its node pairs (row node a, col node b): contribution t is pair
 * (e*3 + a)*3 + b of triangle e, then 9E + (s*2 + a)*2 + b of segment s, so t runs in
 * element order (2-D elements, then 1-D).  Each row node gathers the pairs it heads into its
 * own run of a node-incidence list (counts, an exclusive scan, a fill), and one thread per
 * node sorts its run by (col node, t): a run holds the pairs of the few elements around one
 * node (about 20 on the meshes here), so no global sort is needed.  The sorted runs give
 * (1) the sorted, duplicate-free node adjacency of each row -- the reference's per-row
 * sort + unique, at node-block granularity -- and (2) for every 3x3 block the list of its
 * elemental contributions in element order, which drives the atomic-free assembly; the
 * diagonal blocks' lists give (3) every node's run of elemental residual slots.  The
 * scalar CSR of the reference is the node-block pattern expanded by
 * dof(node, var) = 3*node + var; cols stay ascending because each block row lists node
 * columns in ascending order.  The integer counts use atomics; the order of the fill they
 * give is undone by the per-node sort, whose keys are unique, so the result is
 * deterministic.
 */
#include <hip/hip_runtime.h>

#include "adhk_internal.h"

#define CHECK(x)                                                                                                  \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s", __FILE__, __LINE__, hipGetErrorString(e_));                                \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)

#define SCAN_TPB 256
#define SCAN_ITEMS 4 /* elements per thread: one workgroup scans 1024 */

/* ---- exclusive prefix sum of int[n] (n < 2^31), three passes: workgroup scans, a scan of the workgroup
 *      totals by one workgroup, the carries added back ---- */
__device__ int wg_exclusive_scan(int v, int *total)
{
    __shared__ int sh[SCAN_TPB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_TPB; o <<= 1) {
        int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const int incl = sh[threadIdx.x];
    *total = sh[SCAN_TPB - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_TPB) void k_scan_local(long n, const int *in, int *out, int *wsum)
{
    const long base = (long)blockIdx.x * SCAN_TPB * SCAN_ITEMS + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int run = wg_exclusive_scan(s, &tot);
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) wsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_TPB) void k_scan_sums(int nw, int *wsum)
{
    int carry = 0;
    for (int c0 = 0; c0 < nw; c0 += SCAN_TPB) {
        const int i = c0 + threadIdx.x;
        const int v = i < nw ? wsum[i] : 0;
        int tot;
        const int ex = wg_exclusive_scan(v, &tot);
        if (i < nw) wsum[i] = ex + carry;
        carry += tot;
    }
}

__global__ __launch_bounds__(SCAN_TPB) void k_scan_add(long n, int *out, const int *wsum)
{
    const long base = (long)blockIdx.x * SCAN_TPB * SCAN_ITEMS + (long)threadIdx.x * SCAN_ITEMS;
    const int c = wsum[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) out[base + k] += c;
}

/* out = exclusive scan of in (n entries); *total (host) = sum of in */
static int exclusive_scan(hipStream_t st, long n, const int *in, int *out, int *total)
{
    const long per = (long)SCAN_TPB * SCAN_ITEMS;
    const int nw = (int)((n + per - 1) / per);
    int *wsum = NULL, last_in = 0, last_out = 0;
    if (n <= 0) {
        *total = 0;
        return 0;
    }
    CHECK(hipMalloc((void **)&wsum, (size_t)nw * sizeof(int)));
    hipLaunchKernelGGL(k_scan_local, dim3(nw), dim3(SCAN_TPB), 0, st, n, in, out, wsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_TPB), 0, st, nw, wsum);
    hipLaunchKernelGGL(k_scan_add, dim3(nw), dim3(SCAN_TPB), 0, st, n, out, wsum);
    CHECK(hipGetLastError());
    CHECK(hipMemcpyAsync(&last_in, in + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    CHECK(hipMemcpyAsync(&last_out, out + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    (void)hipFree(wsum);
    *total = last_in + last_out;
    return 0;
fail:
    (void)hipFree(wsum);
    return -1;
}

/* row and col node of contribution t */
__device__ __forceinline__ void pair_nodes(long t, int E, const int *tri, const int *seg, int *a, int *b)
{
    const long ntri = 9L * E;
    if (t < ntri) {
        const int e = (int)(t / 9), la = (int)((t % 9) / 3), lb = (int)(t % 3);
        *a = tri[la * E + e];
        *b = tri[lb * E + e];
    } else {
        const long u = t - ntri;
        const int s = (int)(u / 4), la = (int)((u % 4) / 2), lb = (int)(u % 2);
        *a = seg[2 * s + la];
        *b = seg[2 * s + lb];
    }
}

__global__ void k_row_count(long np, int E, const int *tri, const int *seg, int *cnt)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    int a, b;
    pair_nodes(t, E, tri, seg, &a, &b);
    atomicAdd(cnt + a, 1);
}

/* key = col node << 32 | t, placed in the row node's run (in arrival order) */
__global__ void k_row_fill(long np, int E, const int *tri, const int *seg, const int *roff, int *fill,
                           unsigned long long *key)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    int a, b;
    pair_nodes(t, E, tri, seg, &a, &b);
    const int pos = roff[a] + atomicAdd(fill + a, 1);
    key[pos] = ((unsigned long long)(unsigned)b << 32) | (unsigned long long)(unsigned)t;
}

/* one thread per node: insertion sort of its run by key (col node, then t), then the run's distinct
 * col nodes (the block row's degree) */
__global__ void k_row_sort(int N, const int *roff, unsigned long long *key, int *deg)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lo = roff[n], hi = roff[n + 1];
    for (int i = lo + 1; i < hi; i++) {
        const unsigned long long k = key[i];
        int j = i - 1;
        while (j >= lo && key[j] > k) {
            key[j + 1] = key[j];
            j--;
        }
        key[j + 1] = k;
    }
    int d = 0;
    for (int i = lo; i < hi; i++)
        if (i == lo || (key[i] >> 32) != (key[i - 1] >> 32)) d++;
    deg[n] = d;
}

/* contribution list, block columns, block starts and the diagonal block of every row node */
__global__ void k_row_blocks(int N, const int *roff, const unsigned long long *key, const int *brow, int *contrib,
                             int *bcol, int *cptr, int *diag_blk)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lo = roff[n], hi = roff[n + 1];
    int blk = brow[n] - 1, d = -1;
    for (int i = lo; i < hi; i++) {
        const int col = (int)(key[i] >> 32);
        contrib[i] = (int)(unsigned)(key[i] & 0xffffffffULL);
        if (i == lo || col != (int)(key[i - 1] >> 32)) {
            blk++;
            bcol[blk] = col;
            cptr[blk] = i;
            if (col == n) d = blk;
        }
    }
    diag_blk[n] = d;
}

/* cpos[code] = position of the elemental block `code` in the sorted contribution list */
__global__ void k_cpos(long n, const int *contrib, int *cpos)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cpos[contrib[i]] = (int)i;
}

/* dd[n] = contributions to node n's diagonal block = the elements (2-D and 1-D) that hold n */
__global__ void k_diag_count(int N, const int *diag_blk, const int *cptr, int *dd)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n > N) return;
    const int b = (n < N) ? diag_blk[n] : -1;
    dd[n] = (b >= 0) ? cptr[b + 1] - cptr[b] : 0;
}

/* rpos of element node (e, a) / (s, a): its rank among node n's diagonal contributions (which run in
 * element order) placed in n's residual run */
__global__ void k_rpos(int E, int S, const int *tri, const int *seg, const int *diag_blk, const int *cptr,
                       const int *cpos, const int *rptr, int *rpos)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3L * E + 2L * S) return;
    int n, code;
    if (t < 3L * E) {
        const int e = (int)(t / 3), a = (int)(t % 3);
        n = tri[a * (long)E + e];
        code = 9 * e + 4 * a;
    } else {
        const int u = (int)(t - 3L * E), s = u / 2, a = u % 2;
        n = seg[2 * s + a];
        code = 9 * E + 4 * s + 3 * a;
    }
    rpos[t] = rptr[n] + (cpos[code] - cptr[diag_blk[n]]);
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
    c->nslice = (c->N + 63) / 64;
    CHECK(hipMalloc((void **)&sw, ((size_t)c->nslice + 1) * sizeof(int)));
    CHECK(hipMemsetAsync(sw, 0, ((size_t)c->nslice + 1) * sizeof(int), c->stream));
    CHECK(hipMalloc((void **)&c->slice_ptr, ((size_t)c->nslice + 1) * sizeof(int)));
    hipLaunchKernelGGL(k_slice_width, dim3((c->nslice + 255) / 256), dim3(256), 0, c->stream, c->N, c->nslice, c->brow,
                       sw);
    CHECK(hipGetLastError());
    if (exclusive_scan(c->stream, (long)c->nslice + 1, sw, c->slice_ptr, &c->nslot)) goto fail;
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
    (void)hipFree(sw);
    return 0;
fail:
    (void)hipFree(sw);
    return -1;
}

extern "C" int adhk_build_pattern(adhk_ctx *c)
{
    const long np = 9L * c->E + 4L * c->S;
    unsigned long long *key = NULL;
    int *cnt = NULL, *roff = NULL, *fill = NULL, *deg = NULL;
    int total = 0, nnzb = 0;
    if (np <= 0 || np >= 2147483647L) {
        adhk_set_error("adhk_build_pattern: unsupported pair count %ld", np);
        return -1;
    }
    c->npairs = (int)np;
    const dim3 gp((unsigned)((np + 255) / 256)), gn((unsigned)((c->N + 255) / 256)), tb(256);
    CHECK(hipMalloc((void **)&cnt, ((size_t)c->N + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&roff, ((size_t)c->N + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&fill, (size_t)c->N * sizeof(int)));
    CHECK(hipMalloc((void **)&deg, ((size_t)c->N + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&key, (size_t)np * sizeof(unsigned long long)));
    CHECK(hipMemsetAsync(cnt, 0, ((size_t)c->N + 1) * sizeof(int), c->stream));
    CHECK(hipMemsetAsync(fill, 0, (size_t)c->N * sizeof(int), c->stream));
    CHECK(hipMemsetAsync(deg, 0, ((size_t)c->N + 1) * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_row_count, gp, tb, 0, c->stream, np, c->E, c->tri, c->seg, cnt);
    CHECK(hipGetLastError());
    if (exclusive_scan(c->stream, (long)c->N + 1, cnt, roff, &total)) goto fail;
    if (total != c->npairs) {
        adhk_set_error("adhk_build_pattern: %d of %d pairs counted (a node id out of range?)", total, c->npairs);
        goto fail;
    }
    hipLaunchKernelGGL(k_row_fill, gp, tb, 0, c->stream, np, c->E, c->tri, c->seg, roff, fill, key);
    hipLaunchKernelGGL(k_row_sort, gn, tb, 0, c->stream, c->N, roff, key, deg);
    CHECK(hipGetLastError());
    CHECK(hipMalloc((void **)&c->brow, ((size_t)c->N + 1) * sizeof(int)));
    if (exclusive_scan(c->stream, (long)c->N + 1, deg, c->brow, &nnzb)) goto fail;
    c->nnzb = nnzb;
    CHECK(hipMalloc((void **)&c->contrib, (size_t)np * sizeof(int)));
    CHECK(hipMalloc((void **)&c->cptr, ((size_t)nnzb + 1) * sizeof(int)));
    CHECK(hipMalloc((void **)&c->bcol, (size_t)nnzb * sizeof(int)));
    CHECK(hipMalloc((void **)&c->diag_blk, (size_t)c->N * sizeof(int)));
    hipLaunchKernelGGL(k_row_blocks, gn, tb, 0, c->stream, c->N, roff, key, c->brow, c->contrib, c->bcol, c->cptr,
                       c->diag_blk);
    CHECK(hipGetLastError());
    CHECK(hipMemcpyAsync(c->cptr + nnzb, &c->npairs, sizeof(int), hipMemcpyHostToDevice, c->stream));
    CHECK(hipMalloc((void **)&c->cpos, (size_t)np * sizeof(int)));
    hipLaunchKernelGGL(k_cpos, gp, tb, 0, c->stream, np, c->contrib, c->cpos);
    CHECK(hipGetLastError());
    {
        /* node runs of the elemental residuals: one slot per element node, a node's slots in the
         * order of its diagonal block's contributions */
        const long nr = 3L * c->E + 2L * c->S;
        int rtot = 0;
        CHECK(hipMalloc((void **)&c->rptr, ((size_t)c->N + 1) * sizeof(int)));
        hipLaunchKernelGGL(k_diag_count, dim3((unsigned)((c->N + 1 + 255) / 256)), tb, 0, c->stream, c->N, c->diag_blk,
                           c->cptr, deg);
        CHECK(hipGetLastError());
        if (exclusive_scan(c->stream, (long)c->N + 1, deg, c->rptr, &rtot)) goto fail;
        if (rtot != nr) {
            adhk_set_error("adhk_build_pattern: %d diagonal contributions for %ld element nodes (an element that "
                           "repeats a node?)", rtot, nr);
            goto fail;
        }
        CHECK(hipMalloc((void **)&c->rpos, (size_t)nr * sizeof(int)));
        hipLaunchKernelGGL(k_rpos, dim3((unsigned)((nr + 255) / 256)), tb, 0, c->stream, c->E, c->S, c->tri, c->seg,
                           c->diag_blk, c->cptr, c->cpos, c->rptr, c->rpos);
        CHECK(hipGetLastError());
    }
    CHECK(hipStreamSynchronize(c->stream));
    (void)hipFree(key);
    (void)hipFree(cnt);
    (void)hipFree(roff);
    (void)hipFree(fill);
    (void)hipFree(deg);
    return build_sell(c);
fail:
    (void)hipFree(key);
    (void)hipFree(cnt);
    (void)hipFree(roff);
    (void)hipFree(fill);
    (void)hipFree(deg);
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
