/*
 * la_kernels.hip -- linear algebra of the Newton step on the node-block CSR
 * matrix (gfx950): check_diag, Dirichlet, row scaling, 3x3 block-Jacobi
 * preconditioner and a device-resident, left-preconditioned BiCGStab.
 *
 * Reference: src/la/matrix_utilities.c (check_diag :120, apply_Dirichlet_BC :42),
 * src/la/scale_linear_system.c (:31, :137), src/la/bcgstab_solver.c (:65-335).
 * The reference preconditions with a full UMFPACK LU of the matrix
 * (bcgstab_solver.c:371-424); this path uses the 3x3 nodal block-Jacobi
 * inverse instead.
 *
 * Every BiCGStab iteration is two HBM-bound kernels over the SELL-64 block matrix:
 *   k_iter_a : finish the previous iteration (x += alpha p + omega s, r = s - omega As, max|r| and the
 *              stop test), form p = r + beta (p - omega Ap) -- recomputed on the fly for the columns,
 *              so no extra pass -- and Ap = M^-1 A p with q.Ap -> alpha
 *   k_iter_s : s = r - alpha Ap, As = M^-1 A s with As.s, As.As, As.q -> omega, rho, beta
 * M^-1 A is formed once per Newton iteration (k_scale_prec folds each row's 3x3 inverse into the
 * row), so each SpMV pass reads one matrix and no preconditioner.
 * The multicolour symmetric Gauss-Seidel preconditioner (further below) runs each product in colour
 * phases over colour-permuted U and L parts of that folded matrix.
 * Scalars live in device memory; every kernel exits at once when the solver
 * has converged, so the host launches iterations in chunks and reads the
 * "done" flag between chunks.  Reductions are deterministic: per-workgroup
 * partials, summed in index order by a one-workgroup finalize kernel (in the
 * SGS solve, by every workgroup of the kernel that uses the result).
 * The block matrix and vectors are templated on the storage type (double for
 * the FP64 path, float for the mixed-precision path); reductions and scalars
 * are always double.
 */
#include <hip/hip_runtime.h>

#include "adhk_internal.h"
#include "xcd.h"

#define TPB 256
#define MAX_GRID 2048
#define D_SMALL 2.220446049250313080847e-16

#define CHECK_LAUNCH()                                                                                            \
    do {                                                                                                          \
        hipError_t e_ = hipGetLastError();                                                                        \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s: %s", __func__, hipGetErrorString(e_));                                            \
            return -1;                                                                                            \
        }                                                                                                         \
    } while (0)

static inline int grid_for(long n)
{
    long g = (n + TPB - 1) / TPB;
    if (g > MAX_GRID) g = MAX_GRID;
    if (g < 1) g = 1;
    return (int)g;
}

/* Reductions and the SELL kernels work on blocks of up to 256 consecutive nodes that never straddle
 * two instances of an ensemble (c->blk_node, c->blk_inst): every partial sum belongs to one block, so
 * every scalar can be per instance, and a mesh solved alone or inside an ensemble sums in the same order.
 *
 * XCD-aware block ranges.  Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
 * XCD w % 8), and each XCD has its own L2.  Logical workgroup (w % 8) * (G / 8) + w / 8 takes a
 * contiguous run of blocks, so every XCD owns a contiguous eighth of the nodes and the neighbour
 * gathers of a slice (nodes +-npy away) hit in that XCD's L2 instead of being fetched by several. */
#define NXCD 8
static inline int grid_blk(int nblk)
{
    long g = ((long)nblk + NXCD - 1) / NXCD * NXCD;
    if (g > MAX_GRID) g = MAX_GRID;
    if (g < NXCD) g = NXCD;
    return (int)g;
}
/* Block metadata and the scalars are passed as separate const __restrict__ kernel arguments: the
 * compiler then knows no store of the kernel can change them and reads them through the scalar
 * cache (one s_load per block instead of dependent vector loads in front of every workgroup). */
#define BLK_ARGS                                                                                                \
    const int *__restrict__ blk_node, const int *__restrict__ blk_inst, int nblk, int n_single, int xcd_map
#define BLK_PASS(c) (c)->blk_node, (c)->blk_inst, (c)->nblk, ((c)->ninst == 1 ? (c)->N : 0), (c)->tune.xcd_map

/* bounds and instance of block j; one mesh alone (n_single = its node count): block j is nodes
 * [256 j, 256 j + 256) of instance 0, with no loads */
__device__ __forceinline__ void blk_get(const int *__restrict__ blk_node, const int *__restrict__ blk_inst,
                                        int n_single, int j, int &nb0, int &nb1, int &inst)
{
    if (n_single) {
        nb0 = j << 8;
        nb1 = min(nb0 + 256, n_single);
        inst = 0;
    } else {
        nb0 = blk_node[j];
        nb1 = blk_node[j + 1];
        inst = blk_inst[j];
        asm volatile("" ::"s"(nb0), "s"(nb1), "s"(inst)); /* keep the three loads ahead of the branch */
    }
}
__device__ __forceinline__ void blk_range(int nblk, int xcd_map, int &j0, int &j1)
{
    const int G = gridDim.x, w = blockIdx.x;
    const int lb = xcd_map ? (w % NXCD) * (G / NXCD) + w / NXCD : w;
    const int pb = (nblk + G - 1) / G;
    j0 = lb * pb;
    j1 = min(nblk, j0 + pb);
}

/* ---- deterministic block reductions ---- */
template <int K, bool MAX> __device__ __forceinline__ void block_reduce(double (&v)[K], double *sh)
{
    /* wave (64 lanes) butterfly, then across the 4 waves of the workgroup through LDS */
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double o = __shfl_xor(v[k], off, 64);
            v[k] = MAX ? fmax(v[k], o) : v[k] + o;
        }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[wv * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double a = sh[k];
            for (int w = 1; w < TPB / 64; w++) a = MAX ? fmax(a, sh[w * K + k]) : a + sh[w * K + k];
            v[k] = a;
        }
    }
    __syncthreads(); /* sh is free again for the next reduction */
}

/* Cross-workgroup reductions are deterministic and fence-free: thread 0 stores the K results of
 * block j at partials[k * stride + j]; a finalize kernel with one workgroup per instance, which runs
 * next on the stream, sums the instance's blocks in index order (the kernel boundary makes them
 * visible). */
template <int K, bool MAX>
__device__ __forceinline__ void block_partial(double (&v)[K], double *partials, int stride, int j)
{
    __shared__ double sh[(TPB / 64) * K];
    block_reduce<K, MAX>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) partials[k * stride + j] = v[k];
}

/* in a finalize workgroup: totals over blocks [b0, b1), valid in thread 0 */
template <int K, bool MAX>
__device__ __forceinline__ void sum_blocks(const double *partials, int stride, int b0, int b1, double (&w)[K])
{
    __shared__ double sh[(TPB / 64) * K];
#pragma unroll
    for (int k = 0; k < K; k++) w[k] = 0.0;
    /* a thread combines blocks b0 + tid, b0 + tid + TPB, ... in that order; the loads of eight of
     * them are issued together, so the finalizer waits for one memory round trip per 8 * TPB blocks
     * instead of one per TPB */
    constexpr int U = 8;
    for (int b = b0 + threadIdx.x; b < b1; b += U * TPB) {
        double t[U][K];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < K; k++) t[u][k] = b + u * TPB < b1 ? partials[k * stride + b + u * TPB] : 0.0;
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < K; k++)
                if (b + u * TPB < b1) w[k] = MAX ? fmax(w[k], t[u][k]) : w[k] + t[u][k];
    }
    block_reduce<K, MAX>(w, sh);
}

/* partial regions, in units of the block count (stride) */
#define PR_A 0 /* 4 rows */
#define PR_B 4 /* 4 rows */

/* ---- SELL-64 block row product y = A_n x (column values from a functor) ---- */
/* kSelect: in a backward phase's gathers, choose between this functor and the stored t per lane by
 * selecting addresses (one load either way) rather than by a branch; kGroup: the slot-group size of the
 * narrow SGS phase kernels (groups of more slots keep more loads in flight, and hold more registers) */
template <typename T> struct ColPlain {
    static constexpr bool kSelect = true;
    static constexpr int kGroup = 6;
    const T *x;
    __device__ __forceinline__ void get(int m, double &a, double &b, double &c) const
    {
        a = (double)x[3 * m];
        b = (double)x[3 * m + 1];
        c = (double)x[3 * m + 2];
    }
    /* own ? x[m] : alt[m], without a branch */
    __device__ __forceinline__ void get_or(int m, bool own, const T *alt, double &a, double &b, double &c) const
    {
        const T *y = own ? x : alt;
        a = (double)y[3 * m];
        b = (double)y[3 * m + 1];
        c = (double)y[3 * m + 2];
    }
};

/* BiCGStab vector updates with explicit fused multiply-adds, so that a value recomputed for a
 * neighbouring column is bit-identical to the value its owner row stores */
template <typename T> __device__ __forceinline__ T upd_s(double r, double Ap, double alpha)
{
    return (T)__fma_rn(-alpha, Ap, r); /* s = r - alpha Ap            (bcgstab_solver.c:247-248) */
}
template <typename T> __device__ __forceinline__ T upd_r(double s, double As, double omega)
{
    return (T)__fma_rn(-omega, As, s); /* r = s - omega As            (:300-301) */
}
template <typename T>
__device__ __forceinline__ T upd_p(double p, double Ap, double r_new, double omega, double beta)
{
    return (T)__fma_rn(__fma_rn(-omega, Ap, p), beta, r_new); /* p = (p - omega Ap) beta + r (:219-221) */
}

/* x += alpha p + omega s (:294-297).  x is touched once per iteration and next read an iteration
 * later, after both products have streamed the matrix: non-temporal access keeps it from displacing
 * matrix lines in the MALL. */
template <typename T> __device__ __forceinline__ void upd_x(T *x, double s, double p, double alpha, double omega)
{
    const double xi = (double)__builtin_nontemporal_load(x);
    __builtin_nontemporal_store((T)__fma_rn(omega, s, __fma_rn(alpha, p, xi)), x);
}

/* s = r - alpha Ap evaluated on the fly for the columns */
template <typename T> struct ColS {
    static constexpr bool kSelect = false;
    static constexpr int kGroup = 4;
    const T *r, *Ap;
    double alpha;
    __device__ __forceinline__ T one(int i) const { return upd_s<T>((double)r[i], (double)Ap[i], alpha); }
    __device__ __forceinline__ void get(int m, double &a, double &b, double &c) const
    {
        a = (double)one(3 * m);
        b = (double)one(3 * m + 1);
        c = (double)one(3 * m + 2);
    }
    __device__ __forceinline__ void get_or(int m, bool own, const T *alt, double &a, double &b, double &c) const
    {
        const T *x = own ? r : alt, *y = own ? Ap : alt;
        double o[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T xv = x[3 * m + k], yv = y[3 * m + k];
            o[k] = own ? (double)upd_s<T>((double)xv, (double)yv, alpha) : (double)xv;
        }
        a = o[0];
        b = o[1];
        c = o[2];
    }
};
/* next direction p' = (p - omega Ap) beta + (s - omega As), evaluated on the fly for the columns */
template <typename T> struct ColP {
    static constexpr bool kSelect = false;
    static constexpr int kGroup = 4;
    const T *s, *As, *p, *Ap;
    double omega, beta;
    __device__ __forceinline__ void one(int i, T &rr, T &pp) const
    {
        rr = upd_r<T>((double)s[i], (double)As[i], omega);
        pp = upd_p<T>((double)p[i], (double)Ap[i], (double)rr, omega, beta);
    }
    __device__ __forceinline__ void get(int m, double &a, double &b, double &c) const
    {
        T rr, pp;
        one(3 * m, rr, pp);
        a = (double)pp;
        one(3 * m + 1, rr, pp);
        b = (double)pp;
        one(3 * m + 2, rr, pp);
        c = (double)pp;
    }
    __device__ __forceinline__ void get_or(int m, bool own, const T *alt, double &a, double &b, double &c) const
    {
        const T *x0 = own ? s : alt, *x1 = own ? As : alt, *x2 = own ? p : alt, *x3 = own ? Ap : alt;
        double o[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const T sv = x0[3 * m + k], asv = x1[3 * m + k], pv = x2[3 * m + k], apv = x3[3 * m + k];
            const T rr = upd_r<T>((double)sv, (double)asv, omega);
            const T pp = upd_p<T>((double)pv, (double)apv, (double)rr, omega, beta);
            o[k] = own ? (double)pp : (double)sv;
        }
        a = o[0];
        b = o[1];
        c = o[2];
    }
};

struct Sell {
    const int *slice_ptr, *scol;
    const unsigned char *diag_id; /* [N]: 1 = the folded diagonal block is the identity */
};

/* a matrix load; NT marks it non-temporal, so that a matrix streamed once per product does not evict
 * the gathered vector lines from L2 (worth it only when the matrix cannot stay in the 256 MB MALL) */
template <bool NT, typename T> __device__ __forceinline__ T ldm(const T *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

/* Slot range [c0, c1) of node n's slice.  UNI: the 64 lanes of the wave hold the 64 nodes of one slice
 * (blocks start at multiples of 64), so the slice index is wave-uniform and the loads are scalar. */
template <bool UNI>
__device__ __forceinline__ void slice_bounds(int n, const Sell &S, int &c0, int &c1)
{
    const int sl = UNI ? __builtin_amdgcn_readfirstlane(n >> 6) : (n >> 6);
    c0 = S.slice_ptr[sl];
    c1 = S.slice_ptr[sl + 1];
}

/* y = (row n of M^-1 A) x over the slots [c0, c1) of n's slice.  Slot c0 of every row holds its
 * diagonal block, which the block-Jacobi fold makes the identity: y starts at x_n and the slot is not
 * read (1/7 of the matrix traffic on a triangle mesh), unless the block was singular and kept the
 * fallback inverse (diag_id 0), when it is multiplied like the others.  The off-diagonal slots go in
 * groups of eight (a triangle-mesh row is rarely wider): the eight column indices first, then the eight
 * neighbours' values, so a group costs two memory latencies instead of eight dependent pairs; the
 * products still accumulate in slot order.  Every multiply-add here and in the dot products is an
 * explicit fused multiply-add, so each kernel that uses them rounds identically whatever the compiler
 * contracts. */
template <typename T, typename Col, bool NT>
__device__ __forceinline__ void block_row_nt(int n, int c0, int c1, const Sell &S, const T *__restrict__ bv,
                                             const Col &col, double &y0, double &y1, double &y2)
{
    const int lane = n & 63;
    col.get(n, y0, y1, y2);
    if (!S.diag_id[n]) {
        const T *a = bv + (size_t)c0 * 9 * 64 + lane;
        const double x0 = y0, x1 = y1, x2 = y2;
        y0 = __fma_rn((double)ldm<NT>(a + 128), x2, __fma_rn((double)ldm<NT>(a + 64), x1, (double)ldm<NT>(a) * x0));
        y1 = __fma_rn((double)ldm<NT>(a + 320), x2,
                      __fma_rn((double)ldm<NT>(a + 256), x1, (double)ldm<NT>(a + 192) * x0));
        y2 = __fma_rn((double)ldm<NT>(a + 512), x2,
                      __fma_rn((double)ldm<NT>(a + 448), x1, (double)ldm<NT>(a + 384) * x0));
    }
    for (int cs = c0 + 1; cs < c1; cs += 8) {
        int m[8];
#pragma unroll
        for (int u = 0; u < 8; u++) m[u] = (cs + u < c1) ? ldm<NT>(S.scol + (cs + u) * 64 + lane) : n;
        double xv[8][3];
#pragma unroll
        for (int u = 0; u < 8; u++) col.get(m[u], xv[u][0], xv[u][1], xv[u][2]);
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (cs + u >= c1) break;
            const T *a = bv + (size_t)(cs + u) * 9 * 64 + lane;
            const double x0 = xv[u][0], x1 = xv[u][1], x2 = xv[u][2];
            y0 = __fma_rn((double)ldm<NT>(a), x0, y0);
            y0 = __fma_rn((double)ldm<NT>(a + 64), x1, y0);
            y0 = __fma_rn((double)ldm<NT>(a + 128), x2, y0);
            y1 = __fma_rn((double)ldm<NT>(a + 192), x0, y1);
            y1 = __fma_rn((double)ldm<NT>(a + 256), x1, y1);
            y1 = __fma_rn((double)ldm<NT>(a + 320), x2, y1);
            y2 = __fma_rn((double)ldm<NT>(a + 384), x0, y2);
            y2 = __fma_rn((double)ldm<NT>(a + 448), x1, y2);
            y2 = __fma_rn((double)ldm<NT>(a + 512), x2, y2);
        }
    }
}

template <typename T, typename Col, bool NT = false>
__device__ __forceinline__ void block_row_cs(int n, int c0, int c1, const Sell &S, const T *__restrict__ bv,
                                             const Col &col, double &y0, double &y1, double &y2)
{
    block_row_nt<T, Col, NT>(n, c0, c1, S, bv, col, y0, y1, y2);
}

template <typename T, typename Col, bool UNI = false, bool NT = false>
__device__ __forceinline__ void block_row(int n, const Sell &S, const T *__restrict__ bv, const Col &col, double &y0,
                                          double &y1, double &y2)
{
    int c0, c1;
    slice_bounds<UNI>(n, S, c0, c1);
    block_row_cs<T, Col, NT>(n, c0, c1, S, bv, col, y0, y1, y2);
}

template <typename T>
__device__ __forceinline__ void apply_minv(const T *__restrict__ mi, int n, double y0, double y1, double y2,
                                           double &z0, double &z1, double &z2)
{
    const T *m = mi + (size_t)(n >> 6) * 9 * 64 + (n & 63);
    z0 = (double)m[0] * y0 + (double)m[64] * y1 + (double)m[128] * y2;
    z1 = (double)m[192] * y0 + (double)m[256] * y1 + (double)m[320] * y2;
    z2 = (double)m[384] * y0 + (double)m[448] * y1 + (double)m[512] * y2;
}

__device__ __forceinline__ size_t sell_val(int cs, int e, int lane) { return ((size_t)cs * 9 + e) * 64 + lane; }

/* ---------------- Newton-level kernels ---------------- */
/* get_residual_norms, newton_tools.c:303-373: max|r|, sum r^2, max|dsol| over non-Dirichlet dofs, per
 * instance (a NaN anywhere in an instance's residual is flagged) */
__global__ __launch_bounds__(TPB) void k_norms(BLK_ARGS, const double *__restrict__ res,
                                               const double *__restrict__ dsol, const int *__restrict__ bc,
                                               double *partials)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        double vmax[3] = {0.0, 0.0, 0.0}; /* max|r|, max|dsol|, nan indicator */
        double vsum[1] = {0.0};
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const int n = nb0 + threadIdx.x;
        if (n < nb1)
            for (int k = 0; k < 3; k++) {
                const int i = 3 * n + k;
                double r = res[i];
                if (isnan(r)) vmax[2] = 1.0;
                if (bc[i] != 1) {
                    vmax[0] = fmax(vmax[0], fabs(r));
                    vmax[1] = fmax(vmax[1], fabs(dsol[i]));
                    vsum[0] = __fma_rn(r, r, vsum[0]);
                }
            }
        block_partial<3, true>(vmax, partials + PR_A * nblk, nblk, j);
        block_partial<1, false>(vsum, partials + PR_B * nblk, nblk, j);
    }
}
/* one workgroup per instance */
__global__ __launch_bounds__(TPB) void k_norms_fin(int nblk, const int *__restrict__ inst_blk, const double *partials,
                                                   adhk_scalars *sc)
{
    const int i = blockIdx.x;
    double vmax[3], vsum[1];
    sum_blocks<3, true>(partials + PR_A * nblk, nblk, inst_blk[i], inst_blk[i + 1], vmax);
    sum_blocks<1, false>(partials + PR_B * nblk, nblk, inst_blk[i], inst_blk[i + 1], vsum);
    if (threadIdx.x == 0) {
        sc[i].rmax = vmax[0];
        sc[i].incmax = vmax[1];
        sc[i].nanflag = (vmax[2] != 0.0) ? 1 : 0;
        sc[i].rl2sum = vsum[0];
    }
}

extern "C" int adhk_launch_norms(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_norms, dim3(grid_blk(c->nblk)), dim3(TPB), 0, c->stream, BLK_PASS(c), c->residual, c->dsol,
                       c->bc_mask, c->partials);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_norms_fin, dim3(c->ninst), dim3(TPB), 0, c->stream, c->nblk, c->inst_blk, c->partials,
                       c->sc);
    CHECK_LAUNCH();
    return 0;
}

/* check_diag, matrix_utilities.c:120-152 */
template <typename TB>
__global__ void k_check_diag(int N, const int *__restrict__ slice_ptr, const int *__restrict__ diag_slot,
                             const TB *__restrict__ bv, int *__restrict__ bc, const int *__restrict__ node_inst,
                             const int *__restrict__ work)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || !work[node_inst[n]]) return;
    int d = diag_slot[n];
    int cs = slice_ptr[n >> 6] + d;
    for (int r = 0; r < 3; r++) {
        double a = (d >= 0) ? (double)bv[sell_val(cs, 4 * r, n & 63)] : 0.0;
        if (fabs(a) < D_SMALL) bc[3 * n + r] = 1;
    }
}

/* f32: the matrix is the FP32 one the mixed-precision Newton iteration assembles */
extern "C" int adhk_launch_check_diag_p(adhk_ctx *c, int f32)
{
    adhk_kt_begin(c, ADHK_KT_PREP);
    if (f32)
        hipLaunchKernelGGL(k_check_diag<float>, dim3((c->N + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->N,
                           c->slice_ptr, c->diag_slot, c->bvals32, c->bc_mask, c->node_inst, c->work);
    else
        hipLaunchKernelGGL(k_check_diag<double>, dim3((c->N + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->N,
                           c->slice_ptr, c->diag_slot, c->bvals, c->bc_mask, c->node_inst, c->work);
    CHECK_LAUNCH();
    adhk_kt_end(c);
    return 0;
}
extern "C" int adhk_launch_check_diag(adhk_ctx *c) { return adhk_launch_check_diag_p(c, 0); }

/* apply_Dirichlet_BC (matrix_utilities.c:42-105): Dirichlet columns -> 0, Dirichlet rows -> identity,
 * residual -> 0; then (do_scale) the row maxima of scale_linear_system (scale_linear_system.c:31-121):
 * s = 1/sqrt(max|a|) (1/sqrt(eps) for a vanishing row), b *= s, x = dsol = 0 */
__global__ __launch_bounds__(TPB) void k_dirichlet_scale(int N, const int *__restrict__ brow,
                                                         const int *__restrict__ slice_ptr,
                                                         const int *__restrict__ scol, double *__restrict__ bv,
                                                         const int *__restrict__ bc, double *__restrict__ res,
                                                         double *__restrict__ dsol, double *__restrict__ scale,
                                                         int do_scale)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lane = n & 63, c0 = slice_ptr[n >> 6], deg = brow[n + 1] - brow[n];
    int brc[3] = {bc[3 * n], bc[3 * n + 1], bc[3 * n + 2]};
    double mx[3] = {0., 0., 0.};
    for (int k = 0; k < deg; k++) {
        const int cs = c0 + k;
        const int m = scol[cs * 64 + lane];
        int bcc[3] = {bc[3 * m], bc[3 * m + 1], bc[3 * m + 2]};
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                double *ap = bv + sell_val(cs, 3 * r + cc, lane);
                double v = *ap;
                if (brc[r] == 1)
                    v = (m == n && cc == r) ? 1.0 : 0.0;
                else if (bcc[cc] == 1)
                    v = 0.0;
                *ap = v;
                double t = fabs(v);
                if (t > mx[r]) mx[r] = t;
            }
    }
    if (!do_scale) {
        for (int r = 0; r < 3; r++)
            if (brc[r] == 1) res[3 * n + r] = 0.0;
        return;
    }
    const double small_factor = 1.0 / sqrt(D_SMALL);
    for (int r = 0; r < 3; r++) {
        int i = 3 * n + r;
        double s = (mx[r] < D_SMALL) ? small_factor : 1.0 / sqrt(fabs(mx[r]));
        scale[i] = s;
        double b = (brc[r] == 1) ? 0.0 : res[i];
        res[i] = b * s;
        dsol[i] = 0.0;
    }
}

extern "C" int adhk_launch_dirichlet(adhk_ctx *c, int do_scale)
{
    adhk_kt_begin(c, ADHK_KT_PREP);
    hipLaunchKernelGGL(k_dirichlet_scale, dim3((c->N + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->N, c->brow,
                       c->slice_ptr, c->scol, c->bvals, c->bc_mask, c->residual, c->dsol, c->scale, do_scale);
    CHECK_LAUNCH();
    adhk_kt_end(c);
    return 0;
}

/* A <- S A S (scale_linear_system.c:104-110), the 3x3 block-Jacobi inverse M^-1 of the scaled
 * diagonal block (adjugate; |det| <= 1e-300 falls back to the inverted diagonal), and the
 * left-preconditioned matrix M^-1 A that the Krylov kernels multiply by.  With T = float, M^-1 A and
 * M^-1 are also written in FP32 for the mixed-precision Krylov solve. */
template <typename T>
__global__ __launch_bounds__(TPB) void k_scale_prec(int N, const int *__restrict__ brow,
                                                    const int *__restrict__ slice_ptr, const int *__restrict__ scol,
                                                    const int *__restrict__ diag_slot, double *__restrict__ bv,
                                                    const double *__restrict__ scale, double *__restrict__ minv,
                                                    T *__restrict__ bv_t, T *__restrict__ minv_t,
                                                    unsigned char *__restrict__ diag_id)
{
    int n = xcd_block() * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lane = n & 63, c0 = slice_ptr[n >> 6], w = slice_ptr[(n >> 6) + 1] - c0, deg = brow[n + 1] - brow[n];
    double si[3] = {scale[3 * n], scale[3 * n + 1], scale[3 * n + 2]};
    for (int k = 0; k < w; k++) {
        const int cs = c0 + k;
        const int m = scol[cs * 64 + lane];
        double sj[3] = {scale[3 * m], scale[3 * m + 1], scale[3 * m + 2]};
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                const size_t ix = sell_val(cs, 3 * r + cc, lane);
                double v = (k < deg) ? bv[ix] * (si[r] * sj[cc]) : 0.0;
                if (k < deg) bv[ix] = v;
            }
    }
    int d = diag_slot[n];
    double a[9];
    for (int k = 0; k < 9; k++) a[k] = (d >= 0) ? bv[sell_val(c0 + d, k, lane)] : 0.0;
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double m[9];
    diag_id[n] = fabs(det) > 1e-300 ? 1 : 0; /* the fold below then makes the diagonal block I */
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
        for (int k = 0; k < 9; k++) m[k] = 0.0;
        for (int k = 0; k < 3; k++) m[4 * k] = (a[4 * k] != 0.0) ? 1.0 / a[4 * k] : 1.0;
    }
    double *mo = minv + (size_t)(n >> 6) * 9 * 64 + lane;
    for (int k = 0; k < 9; k++) {
        mo[k * 64] = m[k];
        if (minv_t) minv_t[(size_t)(n >> 6) * 9 * 64 + lane + k * 64] = (T)m[k];
    }
    /* fold the inverse into the row: the Krylov kernels apply M^-1 A as one SpMV.  FP64: in place;
     * mixed: into the FP32 copy (the FP64 values stay the scaled matrix) */
    for (int k = 0; k < w; k++) {
        const int cs = c0 + k;
        double b[9];
        for (int e = 0; e < 9; e++) b[e] = (k < deg) ? bv[sell_val(cs, e, lane)] : 0.0;
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                const double f = m[3 * r] * b[cc] + m[3 * r + 1] * b[3 + cc] + m[3 * r + 2] * b[6 + cc];
                const size_t ix = sell_val(cs, 3 * r + cc, lane);
                if (bv_t)
                    bv_t[ix] = (T)f;
                else if (k < deg)
                    bv[ix] = f;
            }
    }
}

extern "C" int adhk_launch_scale_prec(adhk_ctx *c, int mixed)
{
    adhk_kt_begin(c, ADHK_KT_PREP);
    if (mixed)
        hipLaunchKernelGGL(k_scale_prec<float>, dim3((c->N + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->N, c->brow,
                           c->slice_ptr, c->scol, c->diag_slot, c->bvals, c->scale, c->minv, c->bvals32, c->minv32,
                           c->diag_id);
    else
        hipLaunchKernelGGL(k_scale_prec<double>, dim3((c->N + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->N,
                           c->brow, c->slice_ptr, c->scol, c->diag_slot, c->bvals, c->scale, c->minv, (double *)NULL,
                           (double *)NULL, c->diag_id);
    CHECK_LAUNCH();
    adhk_kt_end(c);
    return 0;
}

/* ---- the Newton system in two matrix passes ----
 * apply_Dirichlet_BC + scale_linear_system + the block-Jacobi fold, as k_dirichlet_scale followed by
 * k_scale_prec compute them (same entries, same scale factors, same inverse), but with the Dirichlet
 * rows and columns applied on the fly: pass 1 reads the matrix for the row maxima (no write), pass 2
 * reads it once more and writes the final operator in place.  Four matrix passes become two.  In mixed
 * precision both passes work on the FP32 matrix the assembly wrote, with FP64 arithmetic. */
__device__ __forceinline__ double dirichlet_entry(double v, int n, int m, int r, int cc, const int *brc,
                                                  const int *bcc)
{
    if (brc[r] == 1) return (m == n && cc == r) ? 1.0 : 0.0;
    if (bcc[cc] == 1) return 0.0;
    return v;
}

template <typename TB>
__global__ __launch_bounds__(TPB) void k_prep_rowmax(int N, const int *__restrict__ brow,
                                                     const int *__restrict__ slice_ptr,
                                                     const int *__restrict__ scol, const TB *__restrict__ bv,
                                                     const int *__restrict__ bc, double *__restrict__ res,
                                                     double *__restrict__ dsol, double *__restrict__ scale)
{
    int n = xcd_block() * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lane = n & 63, c0 = slice_ptr[n >> 6], deg = brow[n + 1] - brow[n];
    const int brc[3] = {bc[3 * n], bc[3 * n + 1], bc[3 * n + 2]};
    double mx[3] = {0., 0., 0.};
    for (int k = 0; k < deg; k++) {
        const int cs = c0 + k;
        const int m = scol[cs * 64 + lane];
        const int bcc[3] = {bc[3 * m], bc[3 * m + 1], bc[3 * m + 2]};
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                const double t =
                    fabs(dirichlet_entry((double)bv[sell_val(cs, 3 * r + cc, lane)], n, m, r, cc, brc, bcc));
                if (t > mx[r]) mx[r] = t;
            }
    }
    const double small_factor = 1.0 / sqrt(D_SMALL);
    for (int r = 0; r < 3; r++) {
        const int i = 3 * n + r;
        const double sc = (mx[r] < D_SMALL) ? small_factor : 1.0 / sqrt(fabs(mx[r]));
        scale[i] = sc;
        const double b = (brc[r] == 1) ? 0.0 : res[i];
        res[i] = b * sc;
        dsol[i] = 0.0;
    }
}

/* TB: the matrix storage, folded in place (FP64, or FP32 in mixed precision); minv in the same type */
template <typename TB>
__global__ __launch_bounds__(TPB) void k_prep_fold(int N, const int *__restrict__ brow,
                                                   const int *__restrict__ slice_ptr, const int *__restrict__ scol,
                                                   const int *__restrict__ bc, TB *__restrict__ bv,
                                                   const double *__restrict__ scale, TB *__restrict__ minv,
                                                   unsigned char *__restrict__ diag_id)
{
    int n = xcd_block() * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int lane = n & 63, c0 = slice_ptr[n >> 6], w = slice_ptr[(n >> 6) + 1] - c0, deg = brow[n + 1] - brow[n];
    const int brc[3] = {bc[3 * n], bc[3 * n + 1], bc[3 * n + 2]};
    const double si[3] = {scale[3 * n], scale[3 * n + 1], scale[3 * n + 2]};
    /* the scaled diagonal block (slot 0) and its inverse */
    double a[9];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++)
            a[3 * r + cc] =
                (deg > 0) ? dirichlet_entry((double)bv[sell_val(c0, 3 * r + cc, lane)], n, n, r, cc, brc, brc) *
                                (si[r] * si[cc])
                          : 0.0;
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double m[9];
    diag_id[n] = fabs(det) > 1e-300 ? 1 : 0;
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
        for (int k = 0; k < 9; k++) m[k] = 0.0;
        for (int k = 0; k < 3; k++) m[4 * k] = (a[4 * k] != 0.0) ? 1.0 / a[4 * k] : 1.0;
    }
    TB *mo = minv + (size_t)(n >> 6) * 9 * 64 + lane;
    for (int k = 0; k < 9; k++) mo[k * 64] = (TB)m[k];
    /* every slot: Dirichlet, scaling and fold, written in place as M^-1 A */
    for (int k = 0; k < w; k++) {
        const int cs = c0 + k;
        double b[9];
        if (k < deg) {
            const int mm = scol[cs * 64 + lane];
            const int bcc[3] = {bc[3 * mm], bc[3 * mm + 1], bc[3 * mm + 2]};
            const double sj[3] = {scale[3 * mm], scale[3 * mm + 1], scale[3 * mm + 2]};
            for (int r = 0; r < 3; r++)
                for (int cc = 0; cc < 3; cc++)
                    b[3 * r + cc] = dirichlet_entry((double)bv[sell_val(cs, 3 * r + cc, lane)], n, mm, r, cc, brc, bcc) *
                                    (si[r] * sj[cc]);
        } else {
            for (int e = 0; e < 9; e++) b[e] = 0.0;
        }
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                const double f = __fma_rn(m[3 * r + 2], b[6 + cc], __fma_rn(m[3 * r + 1], b[3 + cc], m[3 * r] * b[cc]));
                bv[sell_val(cs, 3 * r + cc, lane)] = (TB)f;
            }
    }
}

extern "C" int adhk_launch_prep(adhk_ctx *c, int mixed)
{
    adhk_kt_begin(c, ADHK_KT_PREP);
    const dim3 g((c->N + TPB - 1) / TPB);
    if (mixed) {
        hipLaunchKernelGGL(k_prep_rowmax<float>, g, dim3(TPB), 0, c->stream, c->N, c->brow, c->slice_ptr, c->scol,
                           c->bvals32, c->bc_mask, c->residual, c->dsol, c->scale);
        CHECK_LAUNCH();
        hipLaunchKernelGGL(k_prep_fold<float>, g, dim3(TPB), 0, c->stream, c->N, c->brow, c->slice_ptr, c->scol,
                           c->bc_mask, c->bvals32, c->scale, c->minv32, c->diag_id);
    } else {
        hipLaunchKernelGGL(k_prep_rowmax<double>, g, dim3(TPB), 0, c->stream, c->N, c->brow, c->slice_ptr, c->scol,
                           c->bvals, c->bc_mask, c->residual, c->dsol, c->scale);
        CHECK_LAUNCH();
        hipLaunchKernelGGL(k_prep_fold<double>, g, dim3(TPB), 0, c->stream, c->N, c->brow, c->slice_ptr, c->scol,
                           c->bc_mask, c->bvals, c->scale, c->minv, c->diag_id);
    }
    CHECK_LAUNCH();
    if (c->pc == ADHK_PC_SGS && adhk_launch_sgs_copy(c, mixed)) return -1;
    adhk_kt_end(c);
    return 0;
}

/* ---------------- BiCGStab (bcgstab_solver.c:65-335) ----------------
 * Kernel sequence: init1, init2, A<first>, S, then {A, S} per iteration.  A of iteration k+1
 * finishes iteration k (x and r updates and the ||r||_inf test) and starts k+1 (new direction p and
 * Ap = M^-1 A p); S computes s and As = M^-1 A s.  p and Ap are double-buffered. */

/* b <- M^-1 b (kept in s until the first S overwrites it), x0 = dsol, x = M^-1 b   (:128-147) */
template <typename T>
__global__ __launch_bounds__(TPB) void k_kinit1(int N, const T *__restrict__ minv, const double *__restrict__ b,
                                                const double *__restrict__ dsol, double *__restrict__ x0,
                                                T *__restrict__ x, T *__restrict__ bp)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double z0, z1, z2;
    apply_minv<T>(minv, n, b[3 * n], b[3 * n + 1], b[3 * n + 2], z0, z1, z2);
    double z[3] = {z0, z1, z2};
    for (int k = 0; k < 3; k++) {
        int i = 3 * n + k;
        x0[i] = dsol[i];
        T zz = (T)z[k];
        x[i] = zz;
        bp[i] = zz;
    }
}

/* r = M^-1 b - (M^-1 A) x, q = r, rnorm, bnorm, rho = r.q  (:155-207) */
template <typename T>
__global__ __launch_bounds__(TPB) void k_kinit2(BLK_ARGS, Sell S, const T *__restrict__ bv,
                                                const T *__restrict__ x, const T *__restrict__ bp, T *__restrict__ r,
                                                T *__restrict__ q, double *partials)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        double vmax[2] = {0.0, 0.0}, vsum[1] = {0.0};
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const int n = nb0 + threadIdx.x;
        if (n < nb1) {
            double y0, y1, y2;
            block_row<T>(n, S, bv, ColPlain<T>{x}, y0, y1, y2); /* bv holds M^-1 A */
            double z[3] = {y0, y1, y2};
            for (int k = 0; k < 3; k++) {
                int i = 3 * n + k;
                double bb = (double)bp[i];
                T rr = (T)(bb + (-1) * z[k]);
                r[i] = rr;
                q[i] = rr;
                vmax[0] = fmax(vmax[0], fabs((double)rr));
                vmax[1] = fmax(vmax[1], fabs(bb));
                vsum[0] = __fma_rn((double)rr, (double)rr, vsum[0]);
            }
        }
        block_partial<2, true>(vmax, partials + PR_A * nblk, nblk, j);
        block_partial<1, false>(vsum, partials + PR_B * nblk, nblk, j);
    }
}

/* one workgroup per instance; an instance outside the work mask starts (and stays) converged */
__global__ __launch_bounds__(TPB) void k_kinit3(int nblk, const int *__restrict__ inst_blk, const double *partials,
                                                adhk_scalars *sc_all, const int *__restrict__ work, double tol,
                                                int max_it)
{
    const int inst = blockIdx.x;
    adhk_scalars *sc = sc_all + inst;
    double vmax[2], vsum[1];
    sum_blocks<2, true>(partials + PR_A * nblk, nblk, inst_blk[inst], inst_blk[inst + 1], vmax);
    sum_blocks<1, false>(partials + PR_B * nblk, nblk, inst_blk[inst], inst_blk[inst + 1], vsum);
    if (threadIdx.x == 0) {
        sc->rnorm = vmax[0];
        sc->bnorm = vmax[1];
        sc->rho = vsum[0];
        sc->tol = tol;
        sc->max_it = max_it;
        sc->rhop = 1.0;
        sc->alpha = 1.0;
        sc->omega = 1.0;
        sc->it = 0;
        sc->beta = (sc->rho * 1.0) / (1.0 * 1.0);
        sc->done = !(sc->rnorm > (tol + sc->bnorm * tol) && 0 < max_it) || !work[inst];
    }
}

/* A: [FIRST] p1 = r0 (p = Ap = 0 at start); otherwise finish iteration k -- x += alpha p + omega s,
 * r = s - omega As, ||r||_inf and the stop test (:294-307, :210) -- and form p' = (p - omega Ap) beta + r.
 * Then Ap' = M^-1 A p' and q.Ap' -> alpha = rho / (q.Ap')  (:215-246).  Blocks of converged
 * instances are skipped. */
template <typename T, bool FIRST, bool NT>
__device__ __forceinline__ void iter_a_body(BLK_ARGS, Sell S, const T *__restrict__ bv, T *__restrict__ x,
                                                T *__restrict__ r, const T *__restrict__ s,
                                                const T *__restrict__ As, const T *__restrict__ p_old,
                                                const T *__restrict__ Ap_old, T *__restrict__ p_new,
                                                T *__restrict__ Ap_new, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const adhk_scalars *si = sc + inst;
        if (si->done) continue;
        const double alpha = si->alpha, omega = si->omega, beta = si->beta;
        double vmax[1] = {0.0}, vsum[1] = {0.0};
        const int n = nb0 + threadIdx.x;
        if (n < nb1) {
            double y0, y1, y2;
            /* slice bounds and q before the updates: their stores must not delay the gathers */
            int c0, c1;
            if (n_single)
                slice_bounds<true>(n, S, c0, c1);
            else
                slice_bounds<false>(n, S, c0, c1);
            const double q0 = (double)q[3 * n], q1 = (double)q[3 * n + 1], q2 = (double)q[3 * n + 2];
            if (FIRST) {
                for (int k = 0; k < 3; k++) p_new[3 * n + k] = r[3 * n + k];
                block_row_cs<T, ColPlain<T>, NT>(n, c0, c1, S, bv, ColPlain<T>{r}, y0, y1, y2);
            } else {
                ColP<T> col{s, As, p_old, Ap_old, omega, beta};
                for (int k = 0; k < 3; k++) {
                    const int i = 3 * n + k;
                    T rr, pp;
                    col.one(i, rr, pp);
                    upd_x(x + i, (double)s[i], (double)p_old[i], alpha, omega);
                    r[i] = rr;
                    p_new[i] = pp;
                    vmax[0] = fmax(vmax[0], fabs((double)rr));
                }
                block_row_cs<T, ColP<T>, NT>(n, c0, c1, S, bv, col, y0, y1, y2);
            }
            T a0 = (T)y0, a1 = (T)y1, a2 = (T)y2; /* bv holds M^-1 A */
            Ap_new[3 * n] = a0;
            Ap_new[3 * n + 1] = a1;
            Ap_new[3 * n + 2] = a2;
            vsum[0] = __fma_rn(q2, (double)a2, __fma_rn(q1, (double)a1, __fma_rn(q0, (double)a0, vsum[0])));
        }
        if (!FIRST) block_partial<1, true>(vmax, partials + PR_A * nblk, nblk, j);
        block_partial<1, false>(vsum, partials + PR_B * nblk, nblk, j);
    }
}
template <typename T, bool FIRST>
__global__ __launch_bounds__(TPB) void k_iter_a(BLK_ARGS, Sell S, const T *__restrict__ bv, T *__restrict__ x,
                                                T *__restrict__ r, const T *__restrict__ s,
                                                const T *__restrict__ As, const T *__restrict__ p_old,
                                                const T *__restrict__ Ap_old, T *__restrict__ p_new,
                                                T *__restrict__ Ap_new, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    iter_a_body<T, FIRST, false>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, x, r, s, As, p_old, Ap_old,
                                 p_new, Ap_new, q, partials, sc);
}

template <typename T, bool FIRST>
__global__ __launch_bounds__(TPB) void k_iter_a_nt(BLK_ARGS, Sell S, const T *__restrict__ bv, T *__restrict__ x,
                                                T *__restrict__ r, const T *__restrict__ s,
                                                const T *__restrict__ As, const T *__restrict__ p_old,
                                                const T *__restrict__ Ap_old, T *__restrict__ p_new,
                                                T *__restrict__ Ap_new, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    iter_a_body<T, FIRST, true>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, x, r, s, As, p_old, Ap_old,
                                p_new, Ap_new, q, partials, sc);
}


/* Split schedule of the A step: k_upd finishes iteration k and stores the new direction, then k_spmv_p
 * multiplies it, gathering one vector per column instead of the four that k_iter_a recomputes p from.
 * Same functions, same bits as k_iter_a. */
template <typename T>
__global__ __launch_bounds__(TPB) void k_upd(BLK_ARGS, T *__restrict__ x, T *__restrict__ r, const T *__restrict__ s,
                                             const T *__restrict__ As, const T *__restrict__ p_old,
                                             const T *__restrict__ Ap_old, T *__restrict__ p_new, double *partials,
                                             const adhk_scalars *__restrict__ sc)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const adhk_scalars *si = sc + inst;
        if (si->done) continue;
        const double alpha = si->alpha, omega = si->omega, beta = si->beta;
        double vmax[1] = {0.0};
        /* element-wise over the block's dofs, consecutive dofs on consecutive lanes (every load and store
         * of a wave is one contiguous 512-byte run) */
        ColP<T> col{s, As, p_old, Ap_old, omega, beta};
        for (int i = 3 * nb0 + threadIdx.x; i < 3 * nb1; i += TPB) {
            T rr, pp;
            col.one(i, rr, pp);
            upd_x(x + i, (double)s[i], (double)p_old[i], alpha, omega);
            r[i] = rr;
            p_new[i] = pp;
            vmax[0] = fmax(vmax[0], fabs((double)rr));
        }
        block_partial<1, true>(vmax, partials + PR_A * nblk, nblk, j);
    }
}

template <typename T, bool NT>
__device__ __forceinline__ void spmv_p_body(BLK_ARGS, Sell S, const T *__restrict__ bv, const T *__restrict__ p,
                                                T *__restrict__ Ap, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const adhk_scalars *si = sc + inst;
        if (si->done) continue;
        double vsum[1] = {0.0};
        const int n = nb0 + threadIdx.x;
        if (n < nb1) {
            int c0, c1;
            if (n_single)
                slice_bounds<true>(n, S, c0, c1);
            else
                slice_bounds<false>(n, S, c0, c1);
            const double q0 = (double)q[3 * n], q1 = (double)q[3 * n + 1], q2 = (double)q[3 * n + 2];
            double y0, y1, y2;
            block_row_cs<T, ColPlain<T>, NT>(n, c0, c1, S, bv, ColPlain<T>{p}, y0, y1, y2); /* bv holds M^-1 A */
            T a0 = (T)y0, a1 = (T)y1, a2 = (T)y2;
            Ap[3 * n] = a0;
            Ap[3 * n + 1] = a1;
            Ap[3 * n + 2] = a2;
            vsum[0] = __fma_rn(q2, (double)a2, __fma_rn(q1, (double)a1, __fma_rn(q0, (double)a0, vsum[0])));
        }
        block_partial<1, false>(vsum, partials + PR_B * nblk, nblk, j);
    }
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_spmv_p(BLK_ARGS, Sell S, const T *__restrict__ bv, const T *__restrict__ p,
                                                T *__restrict__ Ap, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    spmv_p_body<T, false>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, p, Ap, q, partials, sc);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_spmv_p_nt(BLK_ARGS, Sell S, const T *__restrict__ bv, const T *__restrict__ p,
                                                T *__restrict__ Ap, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    spmv_p_body<T, true>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, p, Ap, q, partials, sc);
}


/* finalize A, one workgroup per instance: [not FIRST] ||r||_inf of the completed iteration and the
 * stop test (:210); gamma, alpha */
template <bool FIRST>
__global__ __launch_bounds__(TPB) void k_fin_a(int nblk, const int *__restrict__ inst_blk, const double *partials,
                                               adhk_scalars *sc_all)
{
    adhk_scalars *sc = sc_all + blockIdx.x;
    const int done = sc->done; /* tested after the sums, so that their loads do not wait for it */
    const int b0 = inst_blk[blockIdx.x], b1 = inst_blk[blockIdx.x + 1];
    double vmax[1], vsum[1];
    if (!FIRST) sum_blocks<1, true>(partials + PR_A * nblk, nblk, b0, b1, vmax);
    sum_blocks<1, false>(partials + PR_B * nblk, nblk, b0, b1, vsum);
    if (done) return;
    if (threadIdx.x == 0) {
        if (!FIRST) {
            sc->rnorm = vmax[0];
            sc->done = !(vmax[0] > (sc->tol + sc->bnorm * sc->tol) && sc->it < sc->max_it);
        }
        sc->gamma = vsum[0];
        sc->alpha = sc->rho / vsum[0];
    }
}

/* S: s = r - alpha Ap, As = M^-1 A s, omega = (As.s)/(As.As), rho' = -omega (As.q),
 * beta' = (rho' alpha)/(rho omega), iteration count + 1   (:247-291) */
template <typename T, bool NT>
__device__ __forceinline__ void iter_s_body(BLK_ARGS, Sell S, const T *__restrict__ bv,
                                                const T *__restrict__ r, const T *__restrict__ Ap, T *__restrict__ s,
                                                T *__restrict__ As, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    int j0, j1;
    blk_range(nblk, xcd_map, j0, j1);
    for (int j = j0; j < j1; j++) {
        int nb0, nb1, inst;
        blk_get(blk_node, blk_inst, n_single, j, nb0, nb1, inst);
        const adhk_scalars *si = sc + inst;
        if (si->done) continue;
        ColS<T> col{r, Ap, si->alpha};
        double v[3] = {0.0, 0.0, 0.0};
        const int n = nb0 + threadIdx.x;
        if (n < nb1) {
            /* the row's own r, Ap and q first, so their loads overlap the gathers */
            T sv[3], qv[3];
            for (int k = 0; k < 3; k++) {
                sv[k] = col.one(3 * n + k);
                qv[k] = q[3 * n + k];
            }
            double y0, y1, y2;
            if (n_single)
                block_row<T, ColS<T>, true, NT>(n, S, bv, col, y0, y1, y2); /* bv holds M^-1 A */
            else
                block_row<T, ColS<T>, false, NT>(n, S, bv, col, y0, y1, y2);
            double z[3] = {y0, y1, y2};
            for (int k = 0; k < 3; k++) {
                int i = 3 * n + k;
                T ai = (T)z[k];
                s[i] = sv[k];
                As[i] = ai;
                v[0] = __fma_rn((double)ai, (double)sv[k], v[0]);
                v[1] = __fma_rn((double)ai, (double)ai, v[1]);
                v[2] = __fma_rn((double)ai, (double)qv[k], v[2]);
            }
        }
        block_partial<3, false>(v, partials + PR_A * nblk, nblk, j);
    }
}
template <typename T>
__global__ __launch_bounds__(TPB) void k_iter_s(BLK_ARGS, Sell S, const T *__restrict__ bv,
                                                const T *__restrict__ r, const T *__restrict__ Ap, T *__restrict__ s,
                                                T *__restrict__ As, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    iter_s_body<T, false>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, r, Ap, s, As, q, partials, sc);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_iter_s_nt(BLK_ARGS, Sell S, const T *__restrict__ bv,
                                                const T *__restrict__ r, const T *__restrict__ Ap, T *__restrict__ s,
                                                T *__restrict__ As, const T *__restrict__ q, double *partials,
                                                const adhk_scalars *__restrict__ sc)
{
    iter_s_body<T, true>(blk_node, blk_inst, nblk, n_single, xcd_map, S, bv, r, Ap, s, As, q, partials, sc);
}


/* finalize S, one workgroup per instance: omega = (As.s)/(As.As), rho' = -omega (As.q),
 * beta' = (rho' alpha)/(rho omega), it + 1 */
__global__ __launch_bounds__(TPB) void k_fin_s(int nblk, const int *__restrict__ inst_blk, const double *partials,
                                               adhk_scalars *sc_all)
{
    adhk_scalars *sc = sc_all + blockIdx.x;
    const int done = sc->done;
    double v[3];
    sum_blocks<3, false>(partials + PR_A * nblk, nblk, inst_blk[blockIdx.x], inst_blk[blockIdx.x + 1], v);
    if (done) return;
    if (threadIdx.x == 0) {
        double omega = v[0] / v[1];
        double rhop = sc->rho;
        double rho = -1.0 * omega * v[2];
        sc->omega = omega;
        sc->rhop = rhop;
        sc->rho = rho;
        sc->beta = (rho * sc->alpha) / (rhop * omega);
        sc->it = sc->it + 1;
    }
}

/* unscale_linear_system (scale_linear_system.c:137-146): dsol = (x + x0) * s */
template <typename T>
__global__ void k_kfinish(int ndofs, const T *__restrict__ x, const double *__restrict__ x0,
                          const double *__restrict__ scale, double *__restrict__ dsol)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    double t = (double)x[i] + x0[i];
    dsol[i] = t * scale[i];
}

template <typename T> struct KV {
    T *r, *p[2], *Ap[2], *s, *As, *q, *x;
    const T *bv, *mi;
};
static KV<double> kv64(adhk_ctx *c)
{
    KV<double> k;
    k.r = c->r;
    k.p[0] = c->pv;
    k.p[1] = c->pv2;
    k.Ap[0] = c->Ap;
    k.Ap[1] = c->Ap2;
    k.s = c->s;
    k.As = c->As;
    k.q = c->q;
    k.x = c->xk;
    k.bv = c->bvals;
    k.mi = c->minv;
    return k;
}
static KV<float> kv32(adhk_ctx *c)
{
    KV<float> k;
    k.r = c->r32;
    k.p[0] = c->p32;
    k.p[1] = c->p32b;
    k.Ap[0] = c->Ap32;
    k.Ap[1] = c->Ap32b;
    k.s = c->s32;
    k.As = c->As32;
    k.q = c->q32;
    k.x = c->xk32;
    k.bv = c->bvals32;
    k.mi = c->minv32;
    return k;
}

/* non-temporal matrix loads: tune.stream_matrix 0/1, or (-1) when the matrix in the value type of
 * size vb exceeds the 256 MB MALL that otherwise keeps it between products */
#define ADHK_MALL_BYTES (256L << 20)
static int stream_matrix(const adhk_ctx *c, size_t vb)
{
    if (c->tune.stream_matrix >= 0) return c->tune.stream_matrix;
    return (long)c->nslot * 64 * (9 * (long)vb + 4) > ADHK_MALL_BYTES;
}
extern "C" int adhk_stream_matrix_on(adhk_ctx *c, int mixed)
{
    return stream_matrix(c, mixed ? sizeof(float) : sizeof(double));
}

static Sell sell_of(adhk_ctx *c)
{
    Sell S;
    S.slice_ptr = c->slice_ptr;
    S.scol = c->scol;
    S.diag_id = c->diag_id;
    return S;
}

template <typename T> static int krylov_init(adhk_ctx *c, KV<T> k, double tol, int max_it)
{
    int gn = (c->N + TPB - 1) / TPB;
    hipLaunchKernelGGL(k_kinit1<T>, dim3(gn), dim3(TPB), 0, c->stream, c->N, k.mi, c->residual, c->dsol, c->x0, k.x,
                       k.s);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_kinit2<T>, dim3(grid_blk(c->nblk)), dim3(TPB), 0, c->stream, BLK_PASS(c), sell_of(c), k.bv,
                       k.x, k.s, k.r, k.q, c->partials);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_kinit3, dim3(c->ninst), dim3(TPB), 0, c->stream, c->nblk, c->inst_blk, c->partials, c->sc,
                       c->work, tol, max_it);
    CHECK_LAUNCH();
    return 0;
}

template <typename T, bool FIRST> static int launch_a(adhk_ctx *c, KV<T> k, int cur)
{
    /* reads direction buffer cur, writes 1 - cur (FIRST writes buffer 1) */
    int nw = FIRST ? 1 : 1 - cur;
    const bool nt = stream_matrix(c, sizeof(T));
    if (!FIRST && c->tune.split) {
        adhk_kt_begin(c, ADHK_KT_UPDATE);
        hipLaunchKernelGGL(k_upd<T>, dim3(grid_blk(c->nblk)), dim3(TPB), 0, c->stream, BLK_PASS(c), k.x, k.r, k.s, k.As,
                           k.p[cur], k.Ap[cur], k.p[nw], c->partials, c->sc);
        adhk_kt_end(c);
        CHECK_LAUNCH();
        adhk_kt_begin(c, ADHK_KT_SPMV_P);
        hipLaunchKernelGGL(nt ? k_spmv_p_nt<T> : k_spmv_p<T>, dim3(grid_blk(c->nblk)), dim3(TPB), 0, c->stream,
                           BLK_PASS(c), sell_of(c),
                           k.bv, k.p[nw], k.Ap[nw], k.q, c->partials, c->sc);
        adhk_kt_end(c);
        CHECK_LAUNCH();
    } else {
        /* under the split schedule the class times k_spmv_p alone: its first iteration's k_iter_a is not timed */
        if (!c->tune.split) adhk_kt_begin(c, ADHK_KT_SPMV_P);
        hipLaunchKernelGGL((nt ? k_iter_a_nt<T, FIRST> : k_iter_a<T, FIRST>), dim3(grid_blk(c->nblk)), dim3(TPB), 0,
                           c->stream, BLK_PASS(c),
                           sell_of(c), k.bv, k.x, k.r, k.s, k.As, k.p[cur], k.Ap[cur], k.p[nw], k.Ap[nw], k.q,
                           c->partials, c->sc);
        adhk_kt_end(c);
        CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_fin_a<FIRST>, dim3(c->ninst), dim3(TPB), 0, c->stream, c->nblk, c->inst_blk, c->partials,
                       c->sc);
    CHECK_LAUNCH();
    return 0;
}

template <typename T> static int launch_s(adhk_ctx *c, KV<T> k, int cur)
{
    adhk_kt_begin(c, ADHK_KT_SPMV_S);
    hipLaunchKernelGGL(stream_matrix(c, sizeof(T)) ? k_iter_s_nt<T> : k_iter_s<T>, dim3(grid_blk(c->nblk)),
                       dim3(TPB), 0, c->stream, BLK_PASS(c), sell_of(c), k.bv,
                       k.r, k.Ap[cur], k.s, k.As, k.q, c->partials, c->sc);
    adhk_kt_end(c);
    CHECK_LAUNCH();
    hipLaunchKernelGGL(k_fin_s, dim3(c->ninst), dim3(TPB), 0, c->stream, c->nblk, c->inst_blk, c->partials, c->sc);
    CHECK_LAUNCH();
    return 0;
}

template <typename T> static int krylov_finish(adhk_ctx *c, KV<T> k)
{
    hipLaunchKernelGGL(k_kfinish<T>, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, k.x, c->x0,
                       c->scale, c->dsol);
    CHECK_LAUNCH();
    return 0;
}

/* the multicolour SGS solve (defined below) */
static int sgs_dispatch(adhk_ctx *c, int mixed, int what, double tol, int max_it);
enum { SGD_INIT, SGD_FIRST, SGD_ITER, SGD_LAST, SGD_FINISH };

extern "C" int adhk_launch_krylov_init(adhk_ctx *c, int mixed, double tol, int max_it)
{
    if (c->pc == ADHK_PC_SGS) return sgs_dispatch(c, mixed, SGD_INIT, tol, max_it);
    return mixed ? krylov_init<float>(c, kv32(c), tol, max_it) : krylov_init<double>(c, kv64(c), tol, max_it);
}
/* first iteration: A<first> then S; the current direction buffer is then 1 */
extern "C" int adhk_launch_krylov_first(adhk_ctx *c, int mixed)
{
    if (c->pc == ADHK_PC_SGS) return sgs_dispatch(c, mixed, SGD_FIRST, 0.0, 0);
    if (mixed) {
        if (launch_a<float, true>(c, kv32(c), 0)) return -1;
        return launch_s<float>(c, kv32(c), 1);
    }
    if (launch_a<double, true>(c, kv64(c), 0)) return -1;
    return launch_s<double>(c, kv64(c), 1);
}
/* one more iteration from direction buffer cur: A then S; the current buffer becomes 1 - cur */
extern "C" int adhk_launch_krylov_iter(adhk_ctx *c, int mixed, int cur)
{
    if (c->pc == ADHK_PC_SGS) return sgs_dispatch(c, mixed, SGD_ITER, 0.0, 0);
    if (mixed) {
        if (launch_a<float, false>(c, kv32(c), cur)) return -1;
        return launch_s<float>(c, kv32(c), 1 - cur);
    }
    if (launch_a<double, false>(c, kv64(c), cur)) return -1;
    return launch_s<double>(c, kv64(c), 1 - cur);
}
/* the A kernel that completes the last iteration (x, r updates and the stop test) */
extern "C" int adhk_launch_krylov_last(adhk_ctx *c, int mixed, int cur)
{
    if (c->pc == ADHK_PC_SGS) return sgs_dispatch(c, mixed, SGD_LAST, 0.0, 0);
    return mixed ? launch_a<float, false>(c, kv32(c), cur) : launch_a<double, false>(c, kv64(c), cur);
}
extern "C" int adhk_launch_krylov_finish(adhk_ctx *c, int mixed)
{
    if (c->pc == ADHK_PC_SGS) return sgs_dispatch(c, mixed, SGD_FINISH, 0.0, 0);
    return mixed ? krylov_finish<float>(c, kv32(c)) : krylov_finish<double>(c, kv64(c));
}

/* ---------------- multicolour symmetric Gauss-Seidel, split form ----------------
 * No counterpart in the reference (its preconditioner is an exact LU, bcgstab_solver.c:371-424).  With
 * F = I + L + U the block-Jacobi-folded matrix in the colour order of sgs_setup.c, BiCGStab runs on
 * (I+L)^-1 F (I+U)^-1 with right-hand side (I+L)^-1 D^-1 b, and one product z = (I+L)^-1 F (I+U)^-1 v is,
 * by Eisenstat's identity F = (I+L) + (I+U) - I,
 *     t = (I+U)^-1 v      backward, colours last to first:  t_i = v_i - sum_U F_ij t_j
 *     u = (I+L)^-1 (v - t) forward, colours first to last:  u_i = (v_i - t_i) - sum_L F_ij u_j,   z = t + u,
 * one pass over F in 2 nc - 2 colour phases: B(nc-2) .. B(1); B(0) fused with F(0) (colour 0 has no L
 * blocks); F(1) .. F(nc-2); the top colour (no U blocks, so t = v there and is never stored).  A row's
 * first phase forms the product's input v (the A step: the previous iteration's x and r updates and
 * p = (p - omega Ap) beta + r; the S step: s = r - alpha Ap); the top colour's, whose own phase comes last,
 * its lower neighbours evaluate on the fly, as the block-Jacobi kernels do for every column.  Vectors are
 * updated in place: a row's old p, Ap and s are read only by its own first phase and, for the top colour,
 * by neighbours in phases before the top one.  Scalars, finalizers and the partial-sum layout are those of
 * the block-Jacobi solver, over the permuted rows' 256-row blocks. */
struct SPart {
    const int *__restrict__ ptr;
    const int *__restrict__ col;
};
/* Eight buffers of 3 (Np + 1) values, two of them shared, so that the iteration's working set is as
 * small as the method allows:
 *   s shares r: the S step forms s_i = r_i - alpha Ap_i at row i's first phase, and no other row reads
 *     r_i afterwards (the top colour's s, formed on the fly by its neighbours, is stored by its own
 *     phase, the product's last); the A step forms r_i = s_i - omega As_i at row i alone.
 *   u shares t: backward phases read t only of higher colours, which have not yet had their forward
 *     phase; a forward phase reads its own row's t, then stores u over it, and gathers u only of lower
 *     colours, whose forward phases are done; colour 0's t is used within its one phase only.
 * The finishing backward sweep (k_sgs_bwd) rewrites t for every colour but the top before reading it. */
template <typename T> struct SgVec {
    T *r, *p, *Ap, *s, *As, *q, *xt, *t, *u, *bt;
};
template <typename T> static SgVec<T> sg_vec(T *base, int Np)
{
    const size_t n = (size_t)3 * ((size_t)Np + 1);
    SgVec<T> v;
    v.r = base;
    v.s = base;
    v.p = base + n;
    v.Ap = base + 2 * n;
    v.As = base + 3 * n;
    v.q = base + 4 * n;
    v.xt = base + 5 * n;
    v.t = base + 6 * n;
    v.u = base + 6 * n;
    v.bt = base + 7 * n;
    return v;
}

/* neighbour values of a backward phase: the stored t of a middle colour, the product's input for the top.
 * For a stored input (ColPlain) the lane selects the address; an input formed on the fly from two or four
 * vectors branches, which costs less than loading them for every lane. */
template <typename T, typename Col> struct GetB {
    static constexpr int kGroup = Col::kGroup;
    const T *t;
    Col col;
    int top0;
    __device__ __forceinline__ void get(int m, double &a, double &b, double &c) const
    {
        if constexpr (Col::kSelect) {
            col.get_or(m, m >= top0, t, a, b, c);
        } else if (m >= top0) {
            col.get(m, a, b, c);
        } else {
            a = (double)t[3 * m];
            b = (double)t[3 * m + 1];
            c = (double)t[3 * m + 2];
        }
    }
};

/* a = sum over the row's slots of F_ij y_j, slot order (ascending natural column); the 64 rows of a wave
 * are one slice, so the slot range is wave-uniform */
/* the slot range [c0, c1) of row i's slice in a part */
struct SRange {
    int c0, c1;
};
__device__ __forceinline__ SRange part_range(int i, const SPart P)
{
    const int sl = __builtin_amdgcn_readfirstlane(i >> 6);
    return SRange{P.ptr[sl], P.ptr[sl + 1]};
}

/* N consecutive slots [cs, cs + N) of the row's slice, in straight-line code that issues every column
 * index and matrix value, then every gather, before the first is used (a load under a branch is waited for
 * at the branch's end, and values loaded in the order the sums use them would be waited for one by one) */
template <int N, typename T, bool NT, typename G>
__device__ __forceinline__ void part_slots(int lane, const int *__restrict__ colp, int cs, const T *__restrict__ val,
                                           const G &g, double &a0, double &a1, double &a2)
{
    int m[N];
    T av[N][9];
#pragma unroll
    for (int u = 0; u < N; u++) m[u] = ldm<NT>(colp + (cs + u) * 64 + lane);
#pragma unroll
    for (int u = 0; u < N; u++)
#pragma unroll
        for (int e = 0; e < 9; e++) av[u][e] = ldm<NT>(val + ((size_t)(cs + u) * 9 + e) * 64 + lane);
    double xv[N][3];
#pragma unroll
    for (int u = 0; u < N; u++) g.get(m[u], xv[u][0], xv[u][1], xv[u][2]);
#pragma unroll
    for (int u = 0; u < N; u++) {
        const double x0 = xv[u][0], x1 = xv[u][1], x2 = xv[u][2];
        a0 = __fma_rn((double)av[u][0], x0, a0);
        a0 = __fma_rn((double)av[u][1], x1, a0);
        a0 = __fma_rn((double)av[u][2], x2, a0);
        a1 = __fma_rn((double)av[u][3], x0, a1);
        a1 = __fma_rn((double)av[u][4], x1, a1);
        a1 = __fma_rn((double)av[u][5], x2, a1);
        a2 = __fma_rn((double)av[u][6], x0, a2);
        a2 = __fma_rn((double)av[u][7], x1, a2);
        a2 = __fma_rn((double)av[u][8], x2, a2);
    }
}

/* a group of min(n, MAXG) slots, the code specialised for each size up to MAXG only */
template <int MAXG, typename T, bool NT, typename G>
__device__ __forceinline__ void part_group(int n, int lane, const int *__restrict__ colp, int cs,
                                           const T *__restrict__ val, const G &g, double &a0, double &a1, double &a2)
{
    if constexpr (MAXG == 1) {
        part_slots<1, T, NT>(lane, colp, cs, val, g, a0, a1, a2);
    } else {
        if (n >= MAXG)
            part_slots<MAXG, T, NT>(lane, colp, cs, val, g, a0, a1, a2);
        else
            part_group<MAXG - 1, T, NT>(n, lane, colp, cs, val, g, a0, a1, a2);
    }
}

/* Two register budgets (the tuning field sgs_wide picks one; the sums are the same bits):
 *   narrow: groups of up to G::kGroup slots (6 for stored inputs, 4 for inputs formed on the fly), code
 *     for each size up to that only: 120-160 VGPRs, 3-4 waves per SIMD.  On the 1.2M-dof flume a colour
 *     phase is about 2,100 waves, just over the 2,048 that 2 waves per SIMD can hold, so the occupancy
 *     decides it;
 *   wide: groups of up to eight slots for every input, every size compiled in each kernel: about 230
 *     VGPRs and 2 waves per SIMD, every load of a group in flight at once.  On the 4.1M-dof channel, whose
 *     colour phases take several rounds of waves anyway, this is 4 % faster than narrow. */
template <bool WIDE, typename T, bool NT = false, typename G>
__device__ __forceinline__ void part_sum(int i, const SPart P, SRange rg, const T *__restrict__ val, const G &g,
                                         double &a0, double &a1, double &a2)
{
    const int lane = i & 63;
    a0 = 0.0;
    a1 = 0.0;
    a2 = 0.0;
    /* groups in slot order; the group size is wave-uniform */
    if constexpr (WIDE) {
        for (int cs = rg.c0; cs < rg.c1; cs += 8) {
            switch (rg.c1 - cs) {
            case 1: part_slots<1, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 2: part_slots<2, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 3: part_slots<3, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 4: part_slots<4, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 5: part_slots<5, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 6: part_slots<6, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            case 7: part_slots<7, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            default: part_slots<8, T, NT>(lane, P.col, cs, val, g, a0, a1, a2); break;
            }
        }
    } else {
        for (int cs = rg.c0; cs < rg.c1; cs += G::kGroup)
            part_group<G::kGroup, T, NT>(rg.c1 - cs, lane, P.col, cs, val, g, a0, a1, a2);
    }
}

/* the product: initial residual, A step, S step, A step whose direction k_sgs_upd has stored */
enum { SG_INIT = 0, SG_A = 1, SG_S = 2, SG_AP = 3 };
enum { PH_B = 0, PH_B0 = 1, PH_F = 2, PH_TOP = 3 }; /* the phase kinds above */

/* the product's input v, evaluated at any row from the previous state */
template <typename T, int OP, bool FIRST> struct SgCol;
template <typename T> struct SgCol<T, SG_INIT, false> {
    using type = ColPlain<T>;
    static __device__ __forceinline__ type make(const SgVec<T> &V, double, double, double) { return type{V.xt}; }
};
template <typename T> struct SgCol<T, SG_A, true> { /* p = r at the start */
    using type = ColPlain<T>;
    static __device__ __forceinline__ type make(const SgVec<T> &V, double, double, double) { return type{V.r}; }
};
template <typename T> struct SgCol<T, SG_A, false> {
    using type = ColP<T>;
    static __device__ __forceinline__ type make(const SgVec<T> &V, double, double omega, double beta)
    {
        return type{V.s, V.As, V.p, V.Ap, omega, beta};
    }
};
template <typename T> struct SgCol<T, SG_AP, false> {
    using type = ColPlain<T>;
    static __device__ __forceinline__ type make(const SgVec<T> &V, double, double, double) { return type{V.p}; }
};
template <typename T> struct SgCol<T, SG_S, false> {
    using type = ColS<T>;
    static __device__ __forceinline__ type make(const SgVec<T> &V, double alpha, double, double)
    {
        return type{V.r, V.Ap, alpha};
    }
};

/* The two BiCGStab finalizers of the SGS solve, folded into the kernels that use their results: every
 * workgroup of the consuming kernel sums the same partials in the same order (the sums of k_fin_a and
 * k_fin_s, bit for bit) and workgroup 0 stores the scalars for the kernels that follow.  A kernel never
 * reads a scalar that it stores itself, so no workgroup can see another's store half-way:
 *   - the first phase of the S product (sgs_fin_a): stop test on the A step's max|r| and alpha =
 *     rho / (q.Ap); stores rnorm, done, gamma, alpha and rhop = rho (the rho this alpha used);
 *   - k_sgs_upd, or under the fused schedule the first phase of the A product (sgs_fin_s):
 *     omega = (As.s)/(As.As), rho' = -omega (As.q), beta = (rho' alpha)/(rhop omega), it + 1; stores
 *     omega, rho, beta, it (the fused product's later phases read omega and beta from there).
 * A workgroup that reads a done flag just stored by workgroup 0 of the same launch exits, as its own sum
 * would have told it.  The S partials sit in rows PR_S.. so that the first S phase, which stores them
 * when the mesh has two colours, never overwrites the max|r| and q.Ap rows it reads. */
enum { SGF_NONE = 0, SGF_A_FIRST = 1, SGF_A = 2, SGF_S = 3 };
/* k_sgs_upd's workgroups (each sums every S partial, then loops over 256-row blocks): two per CU; 2048
 * cost the 4.1M-dof channel 12 us more per launch in partial sums, and the 1.2M-dof flume is no faster */
#define SGS_UPD_GRID 512
#define PR_S (PR_A + 1) /* 3 rows: As.s, As.As, As.q */

/* true when the solve has stopped (the workgroup exits); else the iteration's alpha.  Thread 0 reads the
 * done flag once for the whole workgroup: it is 1 when an earlier launch stopped the solve, and may
 * already be 1 when workgroup 0 of this launch has stored this test's result, which this workgroup's own
 * test then repeats. */
template <bool FIRST>
__device__ __forceinline__ bool sgs_fin_a(const double *partials, int nblk, adhk_scalars *sc, double &alpha)
{
    __shared__ double bc[2];
    double vmax[1] = {0.0}, vsum[1];
    if (!FIRST) sum_blocks<1, true>(partials + PR_A * nblk, nblk, 0, nblk, vmax);
    sum_blocks<1, false>(partials + PR_B * nblk, nblk, 0, nblk, vsum);
    if (threadIdx.x == 0) {
        const int stopped = sc->done;
        const double rho = sc->rho;
        const int done = FIRST ? 0 : !(vmax[0] > (sc->tol + sc->bnorm * sc->tol) && sc->it < sc->max_it);
        const double a = rho / vsum[0];
        bc[0] = a;
        bc[1] = (double)(stopped || done);
        if (blockIdx.x == 0 && !stopped) {
            if (!FIRST) {
                sc->rnorm = vmax[0];
                sc->done = done;
            }
            sc->gamma = vsum[0];
            sc->alpha = a;
            sc->rhop = rho;
        }
    }
    __syncthreads();
    alpha = bc[0];
    return bc[1] != 0.0;
}

/* every workgroup of a launch that calls this has passed a done test on a flag no workgroup of the launch
 * stores */
__device__ __forceinline__ void sgs_fin_s(const double *partials, int nblk, adhk_scalars *sc, double &omega,
                                          double &beta)
{
    __shared__ double bc[2];
    double v[3];
    sum_blocks<3, false>(partials + PR_S * nblk, nblk, 0, nblk, v);
    if (threadIdx.x == 0) {
        const double om = v[0] / v[1];
        const double rhop = sc->rhop;
        const double rho = -1.0 * om * v[2];
        const double be = (rho * sc->alpha) / (rhop * om);
        bc[0] = om;
        bc[1] = be;
        if (blockIdx.x == 0) {
            sc->omega = om;
            sc->rho = rho;
            sc->beta = be;
            sc->it = sc->it + 1;
        }
    }
    __syncthreads();
    omega = bc[0];
    beta = bc[1];
}

/* one colour phase over the permuted rows [r0, r0 + 256 gridDim.x) */
/* the 256-row block of a phase's workgroup: with xcd_map, XCD k (workgroups w = k mod 8) takes the k-th
 * contiguous eighth of the colour's blocks, so neighbour gathers stay in that XCD's L2; -1: none */
__device__ __forceinline__ int sgs_block(int nb, int xcd_map)
{
    const int G = gridDim.x, w = blockIdx.x;
    const int lb = xcd_map ? (w % NXCD) * (G / NXCD) + w / NXCD : w;
    return lb < nb ? lb : -1;
}

template <typename T, int OP, int PH, bool FIRST, int FIN, bool WIDE, bool NT>
__global__ __launch_bounds__(TPB) void k_sgs_phase(int r0, int nb, int xcd_map, int top0, SPart U, SPart L,
                                                   const T *__restrict__ uval, const T *__restrict__ lval, SgVec<T> V,
                                                   double *partials, int nblk, const adhk_scalars *__restrict__ sc,
                                                   adhk_scalars *scw)
{
    double alpha = 0.0, omega = 0.0, beta = 0.0;
    const int lb = sgs_block(nb, xcd_map);
    if (lb < 0) return;
    const int i = r0 + lb * TPB + threadIdx.x;
    /* the slice's slot ranges, loaded before the scalars are tested, so that the two loads share one round
     * trip (the ranges enter the test, which keeps the compiler from moving their loads after it) */
    SRange ru{0, 0}, rl{0, 0};
    if constexpr (PH == PH_B || PH == PH_B0) ru = part_range(i, U);
    if constexpr (PH != PH_B && PH != PH_B0) rl = part_range(i, L);
    if constexpr (FIN == SGF_A || FIN == SGF_A_FIRST) {
        /* the first phase of the S product finishes the A step itself, in every workgroup: the stop test
         * and alpha, as k_fin_a would in a launch of its own.  The scalars, which workgroup 0 stores during
         * this launch, are read through scw alone, by one thread per workgroup (sgs_fin_a), so that every
         * wave of a workgroup takes the same exit. */
        if (sgs_fin_a<FIN == SGF_A_FIRST>(partials, nblk, scw, alpha)) return;
    } else if constexpr (FIN == SGF_S) {
        /* the first phase of the fused A product finishes the S step (omega, beta); the done flag, which
         * no workgroup of this launch stores, is read as in every other phase */
        if (scw->done) return;
        alpha = scw->alpha;
        sgs_fin_s(partials, nblk, scw, omega, beta);
    } else if constexpr (OP != SG_INIT) {
        if (sc->done || ru.c1 < ru.c0 || rl.c1 < rl.c0) return; /* the ranges are never reversed */
        alpha = sc->alpha;
        omega = sc->omega;
        beta = sc->beta;
    }
    const int j = i / TPB; /* reduction block: r0 is a multiple of 256 */
    const typename SgCol<T, OP, FIRST>::type col = SgCol<T, OP, FIRST>::make(V, alpha, omega, beta);
    double pa1[1] = {0.0}, pa2[2] = {0.0, 0.0}, pa3[3] = {0.0, 0.0, 0.0}, pb[1] = {0.0};
    double v[3], t[3], w[3], uu[3];
    /* Every load of the row's own values, and the neighbour sums (which read nothing this phase stores),
     * come before the first store: s shares r's buffer and u shares t's, and the compiler, which cannot tell
     * the vectors apart, would otherwise hold each load after a store until the value stored had returned.
     * st0/st1 hold the first phase's stores ([S] s; [A, first iteration] p = r; [A] r and p), xn the new x. */
    T st0[3], st1[3], xn[3], tt[3], ut[3];
    constexpr bool FIRST_PH = (OP != SG_INIT && OP != SG_AP && PH != PH_F); /* the row's first phase */
    if constexpr (OP == SG_INIT) {
        for (int k = 0; k < 3; k++) v[k] = (double)V.xt[3 * i + k];
    } else if constexpr (!FIRST_PH) { /* stored by k_sgs_upd or the row's first phase */
        const T *vs = (OP == SG_S) ? V.s : V.p;
        for (int k = 0; k < 3; k++) v[k] = (double)vs[3 * i + k];
    } else if constexpr (OP == SG_A && FIRST) {
        for (int k = 0; k < 3; k++) {
            st0[k] = V.r[3 * i + k];
            v[k] = (double)st0[k];
        }
    } else if constexpr (OP == SG_A) {
        /* the previous iteration's x += alpha p + omega s, r = s - omega As, and p = (p - omega Ap) beta + r */
        T s_[3], As_[3], p_[3], Ap_[3], x_[3];
        for (int k = 0; k < 3; k++) {
            const int ii = 3 * i + k;
            s_[k] = V.s[ii];
            As_[k] = V.As[ii];
            p_[k] = V.p[ii];
            Ap_[k] = V.Ap[ii];
            x_[k] = __builtin_nontemporal_load(V.xt + ii);
        }
        for (int k = 0; k < 3; k++) {
            const T rr = upd_r<T>((double)s_[k], (double)As_[k], omega);
            const T pp = upd_p<T>((double)p_[k], (double)Ap_[k], (double)rr, omega, beta);
            xn[k] = (T)__fma_rn(omega, (double)s_[k], __fma_rn(alpha, (double)p_[k], (double)x_[k]));
            st0[k] = rr;
            st1[k] = pp;
            v[k] = (double)pp;
            pa1[0] = fmax(pa1[0], fabs((double)rr));
        }
    } else { /* SG_S: s = r - alpha Ap */
        T r_[3], Ap_[3];
        for (int k = 0; k < 3; k++) {
            r_[k] = V.r[3 * i + k];
            Ap_[k] = V.Ap[3 * i + k];
        }
        for (int k = 0; k < 3; k++) {
            st0[k] = upd_s<T>((double)r_[k], (double)Ap_[k], alpha);
            v[k] = (double)st0[k];
        }
    }
    if constexpr (PH == PH_F)
        for (int k = 0; k < 3; k++) t[k] = (double)V.t[3 * i + k];
    if constexpr (PH != PH_B)
        for (int k = 0; k < 3; k++) w[k] = (double)(OP == SG_INIT ? V.bt[3 * i + k] : V.q[3 * i + k]);
    /* backward: t = v - sum_U F t */
    if constexpr (PH == PH_B || PH == PH_B0) {
        double a[3];
        part_sum<WIDE, T, NT>(i, U, ru, uval, GetB<T, typename SgCol<T, OP, FIRST>::type>{V.t, col, top0}, a[0],
                          a[1], a[2]);
        for (int k = 0; k < 3; k++) {
            tt[k] = (T)(v[k] - a[k]);
            t[k] = (double)tt[k];
        }
    } else if constexpr (PH == PH_TOP) {
        for (int k = 0; k < 3; k++) t[k] = v[k];
    }
    /* forward: u = (v - t) - sum_L F u, then z = t + u */
    if constexpr (PH == PH_B0) {
        for (int k = 0; k < 3; k++) uu[k] = v[k] - t[k];
    } else if constexpr (PH != PH_B) {
        double a[3];
        part_sum<WIDE, T, NT>(i, L, rl, lval, ColPlain<T>{V.u}, a[0], a[1], a[2]);
        for (int k = 0; k < 3; k++) uu[k] = (v[k] - t[k]) - a[k];
    }
    if constexpr (PH == PH_B0 || PH == PH_F)
        for (int k = 0; k < 3; k++) {
            ut[k] = (T)uu[k];
            uu[k] = (double)ut[k];
        }
    /* the stores */
    for (int k = 0; k < 3; k++) {
        const int ii = 3 * i + k;
        if constexpr (FIRST_PH && OP == SG_S) V.s[ii] = st0[k];
        if constexpr (FIRST_PH && OP == SG_A && FIRST) V.p[ii] = st0[k];
        if constexpr (FIRST_PH && OP == SG_A && !FIRST) {
            __builtin_nontemporal_store(xn[k], V.xt + ii);
            V.r[ii] = st0[k];
            V.p[ii] = st1[k];
        }
        if constexpr (PH == PH_B) V.t[ii] = tt[k]; /* colour 0's t is used in its one phase only */
        if constexpr (PH == PH_B0 || PH == PH_F) V.u[ii] = ut[k];
        if constexpr (PH != PH_B) {
            if constexpr (OP == SG_INIT) { /* r = b~ - z, q = r */
                const T rr = (T)(w[k] + (-1) * (t[k] + uu[k]));
                V.r[ii] = rr;
                V.q[ii] = rr;
                pa2[0] = fmax(pa2[0], fabs((double)rr));
                pa2[1] = fmax(pa2[1], fabs(w[k]));
                pb[0] = __fma_rn((double)rr, (double)rr, pb[0]);
            } else if constexpr (OP == SG_A || OP == SG_AP) {
                const T zk = (T)(t[k] + uu[k]);
                V.Ap[ii] = zk;
                pb[0] = __fma_rn(w[k], (double)zk, pb[0]);
            } else {
                const T zk = (T)(t[k] + uu[k]);
                V.As[ii] = zk;
                pa3[0] = __fma_rn((double)zk, v[k], pa3[0]);
                pa3[1] = __fma_rn((double)zk, (double)zk, pa3[1]);
                pa3[2] = __fma_rn((double)zk, w[k], pa3[2]);
            }
        }
    }
    /* partial sums of the row's 256-row block, each written by exactly one phase */
    if constexpr (OP == SG_INIT) {
        if constexpr (PH != PH_B) {
            block_partial<2, true>(pa2, partials + PR_A * nblk, nblk, j);
            block_partial<1, false>(pb, partials + PR_B * nblk, nblk, j);
        }
    } else if constexpr (OP == SG_A || OP == SG_AP) {
        if constexpr (OP == SG_A && !FIRST && PH != PH_F)
            block_partial<1, true>(pa1, partials + PR_A * nblk, nblk, j);
        if constexpr (PH != PH_B) block_partial<1, false>(pb, partials + PR_B * nblk, nblk, j);
    } else {
        if constexpr (PH != PH_B) block_partial<3, false>(pa3, partials + PR_S * nblk, nblk, j);
    }
}

/* Split schedule of the A step: the previous iteration's x and r updates and the new direction
 * p = (p - omega Ap) beta + r as one coalesced pass over the permuted rows (max|r| per 256-row block),
 * then the product reads p as stored (SG_AP).  Same functions, same bits as the fused first phases. */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_upd(int nblk, SgVec<T> V, double *partials, adhk_scalars *sc)
{
    if (sc->done) return;
    const double alpha = sc->alpha;
    double omega, beta;
    sgs_fin_s(partials, nblk, sc, omega, beta); /* the S step's finalizer, in every workgroup */
    for (int j = blockIdx.x; j < nblk; j += gridDim.x) {
        /* the block's three entries per thread: every load before the first store (r shares s's buffer,
         * and the compiler would otherwise hold each entry's loads until the previous entry's stores) */
        T s_[3], As_[3], p_[3], Ap_[3], x_[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int ii = 3 * TPB * j + k * TPB + threadIdx.x;
            s_[k] = V.s[ii];
            As_[k] = V.As[ii];
            p_[k] = V.p[ii];
            Ap_[k] = V.Ap[ii];
            x_[k] = __builtin_nontemporal_load(V.xt + ii);
        }
        double vmax[1] = {0.0};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int ii = 3 * TPB * j + k * TPB + threadIdx.x;
            const T rr = upd_r<T>((double)s_[k], (double)As_[k], omega);
            const T pp = upd_p<T>((double)p_[k], (double)Ap_[k], (double)rr, omega, beta);
            __builtin_nontemporal_store((T)__fma_rn(omega, (double)s_[k], __fma_rn(alpha, (double)p_[k], (double)x_[k])),
                                        V.xt + ii);
            V.r[ii] = rr;
            V.p[ii] = pp;
            vmax[0] = fmax(vmax[0], fabs((double)rr));
        }
        block_partial<1, true>(vmax, partials + PR_A * nblk, nblk, j);
    }
}

/* b~ = (I+L)^-1 b^ in place over bt, colour phase of rows [r0, ...) (colour >= 1); the initial guess xt = b~ */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_fwd(int r0, SPart L, const T *__restrict__ lval, T *bt, T *xt)
{
    const int i = r0 + blockIdx.x * TPB + threadIdx.x;
    double a[3];
    part_sum<false, T>(i, L, part_range(i, L), lval, ColPlain<T>{bt}, a[0], a[1], a[2]);
    for (int k = 0; k < 3; k++) {
        const T w = (T)((double)bt[3 * i + k] - a[k]);
        bt[3 * i + k] = w;
        xt[3 * i + k] = w;
    }
}

/* x = (I+U)^-1 xt, colour phase (colour <= nc - 2) into t; the top colour's x is xt itself */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_bwd(int r0, int top0, SPart U, const T *__restrict__ uval,
                                                 const T *__restrict__ xt, T *t)
{
    const int i = r0 + blockIdx.x * TPB + threadIdx.x;
    double a[3];
    part_sum<false, T>(i, U, part_range(i, U), uval, GetB<T, ColPlain<T>>{t, ColPlain<T>{xt}, top0}, a[0], a[1],
                       a[2]);
    for (int k = 0; k < 3; k++) t[3 * i + k] = (T)((double)xt[3 * i + k] - a[k]);
}

/* b^ = D^-1 b permuted (the block-Jacobi step of k_kinit1), x0 = dsol; padding rows zero */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_in(int Np, const int *__restrict__ prow, const T *__restrict__ minv,
                                                const double *__restrict__ b, const double *__restrict__ dsol,
                                                double *__restrict__ x0, T *__restrict__ bt, T *__restrict__ xt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const int n = prow[i];
    double z[3] = {0.0, 0.0, 0.0};
    if (n >= 0) {
        apply_minv<T>(minv, n, b[3 * n], b[3 * n + 1], b[3 * n + 2], z[0], z[1], z[2]);
        for (int k = 0; k < 3; k++) x0[3 * n + k] = dsol[3 * n + k];
    }
    for (int k = 0; k < 3; k++) {
        bt[3 * i + k] = (T)z[k];
        xt[3 * i + k] = (T)z[k];
    }
}

/* unscale_linear_system (scale_linear_system.c:137-146) from the permuted solution: dsol = (x + x0) s */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_out(int Np, int top0, const int *__restrict__ prow,
                                                 const T *__restrict__ xt, const T *__restrict__ t,
                                                 const double *__restrict__ x0, const double *__restrict__ scale,
                                                 double *__restrict__ dsol)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const int n = prow[i];
    if (n < 0) return;
    const T *xs = (i >= top0) ? xt : t;
    for (int k = 0; k < 3; k++) {
        const double v = (double)xs[3 * i + k] + x0[3 * n + k];
        dsol[3 * n + k] = v * scale[3 * n + k];
    }
}

/* the folded matrix's off-diagonal blocks into the parts' slots: one thread per natural SELL slot, so that the
 * reads are coalesced (a wave reads a slot column's nine 512-byte value rows) and each block is written
 * to its place in U or L.  Gathering by destination slot instead read each of a block's nine values from
 * its own cache line wherever the rows of one colour lie apart in the natural order (7x the source bytes
 * on the 4.1M-dof channel).  XCD-contiguous workgroups keep the destination lines a range of source
 * slices fills within one L2. */
template <typename T>
__global__ __launch_bounds__(TPB) void k_sgs_scatter(int nnat, const int *__restrict__ dst, const T *__restrict__ bv,
                                                     T *__restrict__ uval, T *__restrict__ lval)
{
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= nnat) return;
    const int d = dst[p];
    if (d == -1) return;
    T *out = d >= 0 ? uval : lval;
    const int q = d >= 0 ? d : -(d + 2), slot = q >> 6, lane = q & 63;
    T v[9];
#pragma unroll
    for (int e = 0; e < 9; e++) v[e] = bv[((size_t)(p >> 6) * 9 + e) * 64 + (p & 63)];
#pragma unroll
    for (int e = 0; e < 9; e++) out[((size_t)slot * 9 + e) * 64 + lane] = v[e];
}

/* a folded diagonal block that is not the identity (a singular block's fallback inverse) */
__global__ void k_sgs_flag_clear(adhk_scalars *sc)
{
    if (threadIdx.x == 0) sc->sgs_bad = 0;
}
__global__ __launch_bounds__(TPB) void k_sgs_check(int Np, const int *__restrict__ prow,
                                                   const unsigned char *__restrict__ diag_id, adhk_scalars *sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const int n = prow[i];
    if (n >= 0 && !diag_id[n]) sc->sgs_bad = 1;
}

extern "C" int adhk_launch_sgs_copy(adhk_ctx *c, int mixed)
{
    const adhk_sgs *G = &c->sgs;
    const adhk_sgs_part *P[2] = {&G->U, &G->L};
    hipLaunchKernelGGL(k_sgs_flag_clear, dim3(1), dim3(64), 0, c->stream, c->sc);
    hipLaunchKernelGGL(k_sgs_check, dim3((G->Np + TPB - 1) / TPB), dim3(TPB), 0, c->stream, G->Np, G->prow, c->diag_id,
                       c->sc);
    const int nnat = c->nslot * 64;
    if (mixed)
        hipLaunchKernelGGL(k_sgs_scatter<float>, dim3((nnat + TPB - 1) / TPB), dim3(TPB), 0, c->stream, nnat, G->dst,
                           c->bvals32, P[0]->val32, P[1]->val32);
    else
        hipLaunchKernelGGL(k_sgs_scatter<double>, dim3((nnat + TPB - 1) / TPB), dim3(TPB), 0, c->stream, nnat, G->dst,
                           c->bvals, P[0]->val, P[1]->val);
    CHECK_LAUNCH();
    return 0;
}

template <typename T> struct SgM {
    SgVec<T> V;
    const T *uval, *lval, *minv;
};
static SgM<double> sgm64(adhk_ctx *c)
{
    SgM<double> m;
    m.V = sg_vec<double>(c->sgs.v64, c->sgs.Np);
    m.uval = c->sgs.U.val;
    m.lval = c->sgs.L.val;
    m.minv = c->minv;
    return m;
}
static SgM<float> sgm32(adhk_ctx *c)
{
    SgM<float> m;
    m.V = sg_vec<float>(c->sgs.v32, c->sgs.Np);
    m.uval = c->sgs.U.val32;
    m.lval = c->sgs.L.val32;
    m.minv = c->minv32;
    return m;
}

template <typename T, int OP, int PH, bool FIRST, int FIN = SGF_NONE>
static void sgs_phase(adhk_ctx *c, const SgM<T> &M, int colour)
{
    const adhk_sgs *G = &c->sgs;
    const SPart U{G->U.ptr, G->U.col}, L{G->L.ptr, G->L.col};
    const int r0 = G->cs[colour], nb = (G->cs[colour + 1] - r0) / TPB, xm = c->tune.xcd_map;
    const int grid = xm ? (nb + NXCD - 1) / NXCD * NXCD : nb;
    /* matrix loads non-temporal when the matrix cannot stay in the MALL (the rule of the block-Jacobi kernels) */
    const bool nt = stream_matrix(c, sizeof(T));
    if (c->tune.sgs_wide && nt)
        hipLaunchKernelGGL((k_sgs_phase<T, OP, PH, FIRST, FIN, true, true>), dim3(grid), dim3(TPB), 0, c->stream, r0,
                           nb, xm, G->cs[G->nc - 1], U, L, M.uval, M.lval, M.V, G->partials, G->nblk, c->sc, c->sc);
    else if (c->tune.sgs_wide)
        hipLaunchKernelGGL((k_sgs_phase<T, OP, PH, FIRST, FIN, true, false>), dim3(grid), dim3(TPB), 0, c->stream, r0,
                           nb, xm, G->cs[G->nc - 1], U, L, M.uval, M.lval, M.V, G->partials, G->nblk, c->sc, c->sc);
    else if (nt)
        hipLaunchKernelGGL((k_sgs_phase<T, OP, PH, FIRST, FIN, false, true>), dim3(grid), dim3(TPB), 0, c->stream, r0,
                           nb, xm, G->cs[G->nc - 1], U, L, M.uval, M.lval, M.V, G->partials, G->nblk, c->sc, c->sc);
    else
        hipLaunchKernelGGL((k_sgs_phase<T, OP, PH, FIRST, FIN, false, false>), dim3(grid), dim3(TPB), 0, c->stream, r0,
                           nb, xm, G->cs[G->nc - 1], U, L, M.uval, M.lval, M.V, G->partials, G->nblk, c->sc, c->sc);
}

/* one product in 2 nc - 2 phases; FIN: the first phase also finishes the A step (the S product) */
template <typename T, int OP, bool FIRST, int FIN = SGF_NONE> static int sgs_product(adhk_ctx *c, const SgM<T> &M)
{
    const int nc = c->sgs.nc;
    if (nc >= 3) {
        sgs_phase<T, OP, PH_B, FIRST, FIN>(c, M, nc - 2);
        for (int k = nc - 3; k >= 1; k--) sgs_phase<T, OP, PH_B, FIRST>(c, M, k);
        sgs_phase<T, OP, PH_B0, FIRST>(c, M, 0);
    } else {
        sgs_phase<T, OP, PH_B0, FIRST, FIN>(c, M, 0);
    }
    for (int k = 1; k <= nc - 2; k++) sgs_phase<T, OP, PH_F, FIRST>(c, M, k);
    sgs_phase<T, OP, PH_TOP, FIRST>(c, M, nc - 1);
    CHECK_LAUNCH();
    return 0;
}

template <typename T> static int sgs_init(adhk_ctx *c, const SgM<T> &M, double tol, int max_it)
{
    const adhk_sgs *G = &c->sgs;
    const SPart L{G->L.ptr, G->L.col};
    hipLaunchKernelGGL(k_sgs_in<T>, dim3((G->Np + TPB - 1) / TPB), dim3(TPB), 0, c->stream, G->Np, G->prow, M.minv,
                       c->residual, c->dsol, c->x0, M.V.bt, M.V.xt);
    for (int k = 1; k < G->nc; k++)
        hipLaunchKernelGGL(k_sgs_fwd<T>, dim3((G->cs[k + 1] - G->cs[k]) / TPB), dim3(TPB), 0, c->stream, G->cs[k], L,
                           M.lval, M.V.bt, M.V.xt);
    CHECK_LAUNCH();
    if (sgs_product<T, SG_INIT, false>(c, M)) return -1;
    hipLaunchKernelGGL(k_kinit3, dim3(1), dim3(TPB), 0, c->stream, G->nblk, G->inst_blk, G->partials, c->sc, c->work,
                       tol, max_it);
    CHECK_LAUNCH();
    return 0;
}

/* the A step: [split, not FIRST] k_sgs_upd and the product of the stored p, else the product with p formed
 * in each row's first phase; the S step's finalizer runs in k_sgs_upd or in the product's first phase.  The
 * A step's own finalizer runs in the next S product's first phase; `last` (the step that only completes the
 * last iteration) finalizes in a launch of its own. */
template <typename T, bool FIRST> static int sgs_a(adhk_ctx *c, const SgM<T> &M, bool last = false)
{
    const adhk_sgs *G = &c->sgs;
    int rc;
    if (!FIRST && c->tune.split) {
        adhk_kt_begin(c, ADHK_KT_UPDATE);
        hipLaunchKernelGGL(k_sgs_upd<T>, dim3(G->nblk < SGS_UPD_GRID ? G->nblk : SGS_UPD_GRID), dim3(TPB), 0,
                           c->stream,
                           G->nblk, M.V, G->partials, c->sc);
        adhk_kt_end(c);
        CHECK_LAUNCH();
        adhk_kt_begin(c, ADHK_KT_SPMV_P);
        rc = sgs_product<T, SG_AP, false>(c, M);
    } else {
        adhk_kt_begin(c, ADHK_KT_SPMV_P);
        rc = sgs_product<T, SG_A, FIRST, FIRST ? SGF_NONE : SGF_S>(c, M);
    }
    adhk_kt_end(c);
    if (rc) return -1;
    if (last) {
        hipLaunchKernelGGL(k_fin_a<FIRST>, dim3(1), dim3(TPB), 0, c->stream, G->nblk, G->inst_blk, G->partials,
                           c->sc);
        CHECK_LAUNCH();
    }
    return 0;
}

/* the S step after the A step (FIRST: the first iteration's); its finalizer runs in the next A step */
template <typename T, bool FIRST> static int sgs_s(adhk_ctx *c, const SgM<T> &M)
{
    adhk_kt_begin(c, ADHK_KT_SPMV_S);
    const int rc = sgs_product<T, SG_S, false, FIRST ? SGF_A_FIRST : SGF_A>(c, M);
    adhk_kt_end(c);
    return rc ? -1 : 0;
}

template <typename T> static int sgs_finish(adhk_ctx *c, const SgM<T> &M)
{
    const adhk_sgs *G = &c->sgs;
    const SPart U{G->U.ptr, G->U.col};
    const int top0 = G->cs[G->nc - 1];
    for (int k = G->nc - 2; k >= 0; k--)
        hipLaunchKernelGGL(k_sgs_bwd<T>, dim3((G->cs[k + 1] - G->cs[k]) / TPB), dim3(TPB), 0, c->stream, G->cs[k], top0,
                           U, M.uval, M.V.xt, M.V.t);
    hipLaunchKernelGGL(k_sgs_out<T>, dim3((G->Np + TPB - 1) / TPB), dim3(TPB), 0, c->stream, G->Np, top0, G->prow,
                       M.V.xt, M.V.t, c->x0, c->scale, c->dsol);
    CHECK_LAUNCH();
    return 0;
}

template <typename T> static int sgs_run(adhk_ctx *c, const SgM<T> &M, int what, double tol, int max_it)
{
    switch (what) {
    case SGD_INIT:
        return sgs_init<T>(c, M, tol, max_it);
    case SGD_FIRST:
        if (sgs_a<T, true>(c, M)) return -1;
        return sgs_s<T, true>(c, M);
    case SGD_ITER:
        if (sgs_a<T, false>(c, M)) return -1;
        return sgs_s<T, false>(c, M);
    case SGD_LAST:
        return sgs_a<T, false>(c, M, true);
    default:
        return sgs_finish<T>(c, M);
    }
}
static int sgs_dispatch(adhk_ctx *c, int mixed, int what, double tol, int max_it)
{
    if (!c->sgs.ready || (mixed ? c->sgs.v32 == NULL : c->sgs.v64 == NULL)) {
        adhk_set_error("multicolour SGS solve without its layout (adhk_sgs_ensure)");
        return -1;
    }
    return mixed ? sgs_run<float>(c, sgm32(c), what, tol, max_it) : sgs_run<double>(c, sgm64(c), what, tol, max_it);
}

/* ---------------- Newton updates ---------------- */
/* increment_function (newton_tools.c:261) then dsol = -dsol (fe_newton.c:553-555) */
__global__ void k_increment(int n, double *sol, double *dsol, const int *node_inst, const int *work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !work[node_inst[i / 3]]) return;
    double d = dsol[i];
    sol[i] += d;
    dsol[i] = -d;
}
/* line search step (fe_newton.c:681-741): dsol *= 0.5, sol += dsol */
__global__ void k_linesearch(int n, double *sol, double *dsol, const int *node_inst, const int *work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !work[node_inst[i / 3]]) return;
    double d = dsol[i] * 0.5;
    dsol[i] = d;
    sol[i] += d;
}
/* initialize_system's copy (newton_tools.c:33-37) */
__global__ void k_copy(int n, double *dst, const double *src)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
/* initialize_dirichlet_bc (newton_tools.c:130-142) */
__global__ void k_apply_dirichlet_sol(int n, double *sol, const int *bc, const double *dir)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && bc[i] == 1) sol[i] = dir[i];
}
/* smodel_super_prep_sol (smodel_super_prep_sol.c:2-8) */
__global__ void k_rotate(int n, double *sol, double *sol_old, double *sol_older)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sol_older[i] = sol_old[i];
    sol_old[i] = sol[i];
}

extern "C" int adhk_launch_increment(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_increment, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->dsol, c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}
extern "C" int adhk_launch_linesearch(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_linesearch, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->dsol, c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}
/* sol <- sol_old; wet/dry flags; dry nodes -> u,v Dirichlet; sol[bc] <- dirichlet */
__global__ void k_copy_masked(int n, double *dst, const double *src, const int *node_inst, const int *work);
__global__ void k_dirichlet_sol_masked(int n, double *sol, const int *bc, const double *dir, const int *node_inst,
                                       const int *work);

/* initialize_system + initialize_dirichlet_bc for the instances in the work mask: instances outside it
 * (converged ensemble members while others retake their step in sub-steps) keep their state */
extern "C" int adhk_launch_init_system(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_copy_masked, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->sol_old, c->node_inst, c->work);
    CHECK_LAUNCH();
    if (adhk_launch_wdflag(c)) return -1;
    if (adhk_launch_dry_bc(c)) return -1;
    hipLaunchKernelGGL(k_dirichlet_sol_masked, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs,
                       c->sol, c->bc_mask, c->dirichlet, c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}
/* dst <- src for the dofs of instances in the work mask */
__global__ void k_copy_masked(int n, double *dst, const double *src, const int *node_inst, const int *work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && work[node_inst[i / 3]]) dst[i] = src[i];
}
__global__ void k_dirichlet_sol_masked(int n, double *sol, const int *bc, const double *dir, const int *node_inst,
                                       const int *work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && bc[i] == 1 && work[node_inst[i / 3]]) sol[i] = dir[i];
}
extern "C" int adhk_launch_copy_masked(adhk_ctx *c, double *dst, const double *src)
{
    hipLaunchKernelGGL(k_copy_masked, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, dst, src,
                       c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}
/* sol <- sol_old with the Dirichlet values, for the instances in the work mask */
extern "C" int adhk_launch_start_from_old_masked(adhk_ctx *c)
{
    if (adhk_launch_copy_masked(c, c->sol, c->sol_old)) return -1;
    hipLaunchKernelGGL(k_dirichlet_sol_masked, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs,
                       c->sol, c->bc_mask, c->dirichlet, c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}
extern "C" int adhk_launch_copy(adhk_ctx *c, double *dst, const double *src)
{
    hipLaunchKernelGGL(k_copy, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, dst, src);
    CHECK_LAUNCH();
    return 0;
}
/* sol <- sol_old with the Dirichlet values (the start of initialize_system, without the wet/dry flags) */
extern "C" int adhk_launch_start_from_old(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_copy, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->sol_old);
    CHECK_LAUNCH();
    return adhk_launch_dirichlet_sol(c);
}
extern "C" int adhk_launch_dirichlet_sol(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_apply_dirichlet_sol, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs,
                       c->sol, c->bc_mask, c->dirichlet);
    CHECK_LAUNCH();
    return 0;
}
__global__ void k_rotate_masked(int n, double *sol, double *sol_old, double *sol_older, const int *node_inst,
                                const int *work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !work[node_inst[i / 3]]) return;
    sol_older[i] = sol_old[i];
    sol_old[i] = sol[i];
}

/* prep_sol for the instances in the work mask only */
extern "C" int adhk_launch_rotate_masked(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_rotate_masked, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->sol_old, c->sol_older, c->node_inst, c->work);
    CHECK_LAUNCH();
    return 0;
}

extern "C" int adhk_launch_rotate(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_rotate, dim3((c->ndofs + TPB - 1) / TPB), dim3(TPB), 0, c->stream, c->ndofs, c->sol,
                       c->sol_old, c->sol_older);
    CHECK_LAUNCH();
    return 0;
}

/* ---------------- bandwidth probe ----------------
 * A plain device copy of a known size on the context stream: gives the achievable HBM ceiling and
 * the scale of the FETCH_SIZE / WRITE_SIZE counters for a kernel whose traffic is known exactly. */
__global__ __launch_bounds__(TPB) void k_copy_probe(const double4 *__restrict__ a, double4 *__restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

extern "C" int adhk_copy_probe(adhk_ctx *c, long long nbytes, int reps, double *ms_avg)
{
    double4 *a = NULL, *b = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    long n = (long)(nbytes / (long long)sizeof(double4));
    float ms = 0.f;
    int rc = -1;
    if (!c || n <= 0 || reps <= 0) return -1;
    if (hipMalloc((void **)&a, n * sizeof(double4)) != hipSuccess) goto done;
    if (hipMalloc((void **)&b, n * sizeof(double4)) != hipSuccess) goto done;
    if (hipMemsetAsync(a, 0, n * sizeof(double4), c->stream) != hipSuccess) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(TPB), 0, c->stream, a, b, n);
    if (hipEventRecord(e0, c->stream) != hipSuccess) goto done;
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(TPB), 0, c->stream, a, b, n);
    if (hipEventRecord(e1, c->stream) != hipSuccess) goto done;
    if (hipEventSynchronize(e1) != hipSuccess) goto done;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    if (ms_avg) *ms_avg = ms / reps;
    rc = 0;
done:
    if (rc) adhk_set_error("adhk_copy_probe failed");
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}
