/*
 * sw2_kernels.hip -- SW2 elemental residual, central-FD elemental Jacobian,
 * wet/dry flags, and the deterministic gather assembly into the node-block
 * CSR matrix (gfx950).
 *
 * Assembly is atomic-free: every block (and every residual node) sums its
 * contributions in element order (2-D elements first, then 1-D), which is
 * the order of the reference's sequential scatter
 * (assemble_residual.c:69-92, assemble_jacobian.c:20-106).
 */
#include <hip/hip_runtime.h>

#include "adhk_internal.h"
#include "sw2_device.h"
#include "sw2_wetdry.h"
#include "xcd.h"

#define CHECK_LAUNCH()                                                                                            \
    do {                                                                                                          \
        hipError_t e_ = hipGetLastError();                                                                        \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s: %s", __func__, hipGetErrorString(e_));                                            \
            return -1;                                                                                            \
        }                                                                                                         \
    } while (0)

static inline Sw2Par make_par(const adhk_ctx *c)
{
    Sw2Par P;
    P.g = c->p.g;
    P.dt = c->p.dt;
    P.alpha = c->p.alpha;
    P.cf = c->p.cf_const;
    P.manning = (c->p.fric_mode == ADHK_FRIC_MANNING);
    return P;
}

struct MeshView {
    const double *x, *y, *z, *djac, *grad, *ks;
    const int *tri, *wd;
    const double *sol, *sol_old, *sol_older;
    const int *elem_inst, *work; /* elements of instances outside the work mask are skipped */
    int E;
};

__device__ __forceinline__ void gather_tri(const MeshView &m, int e, bool older, TriIn &in)
{
    /* with alpha == 0 the t(n-1) level is unused; point it at the old level (already in cache)
     * instead of branching, which keeps the element in registers */
    const double *soo = older ? m.sol_older : m.sol_old;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int n = m.tri[a * m.E + e];
        in.x[a] = m.x[n];
        in.y[a] = m.y[n];
        in.z[a] = m.z[n];
        in.h[a] = m.sol[3 * n];
        in.u[a] = m.sol[3 * n + 1];
        in.v[a] = m.sol[3 * n + 2];
        in.ho[a] = m.sol_old[3 * n];
        in.uo[a] = m.sol_old[3 * n + 1];
        in.vo[a] = m.sol_old[3 * n + 2];
        in.hoo[a] = soo[3 * n];
        in.uoo[a] = soo[3 * n + 1];
        in.voo[a] = soo[3 * n + 2];
        in.gx[a] = m.grad[a * m.E + e];
        in.gy[a] = m.grad[(3 + a) * m.E + e];
    }
    in.djac = m.djac[e];
    in.ks = m.ks[e];
    in.wd_flag = m.wd[e];
}

static MeshView make_view(const adhk_ctx *c)
{
    MeshView m;
    m.x = c->x;
    m.y = c->y;
    m.z = c->z;
    m.djac = c->djac;
    m.grad = c->grad;
    m.ks = c->ks;
    m.tri = c->tri;
    m.wd = c->wd_flag;
    m.sol = c->sol;
    m.sol_old = c->sol_old;
    m.sol_older = c->sol_older;
    m.elem_inst = c->elem_inst;
    m.work = c->work;
    m.E = c->E;
    return m;
}

/* ---------------- residual ---------------- */
/* An element whose depth field at any level the residual reads is partially dry goes to the
 * wet/dry queue (wdq[0] = count, wdq[1..] = elements) and is evaluated by k_elem_resid_wd;
 * the queue order does not matter, each element owns its output slots. */
__device__ __forceinline__ bool tri_partial(const TriIn &in, bool older)
{
    const int d = n_dry3(in.h[0], in.h[1], in.h[2]);
    const int dold = n_dry3(in.ho[0], in.ho[1], in.ho[2]);
    const int dolder = older ? n_dry3(in.hoo[0], in.hoo[1], in.hoo[2]) : 0;
    return (d == 1 || d == 2) || (dold == 1 || dold == 2) || (dolder == 1 || dolder == 2);
}

/* the same, with every depth perturbation of the FD Jacobian (h[i] +- eps_i) */
__device__ __forceinline__ bool tri_partial_fd(const TriIn &in, bool older)
{
    if (tri_partial(in, older)) return true;
    bool any = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double eps = num_diff_eps(in.h[i], 1e-4);
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
            const double hp = in.h[i] + (sg ? -eps : eps);
            const double h0 = (i == 0) ? hp : in.h[0], h1 = (i == 1) ? hp : in.h[1], h2 = (i == 2) ? hp : in.h[2];
            const int d = n_dry3(h0, h1, h2);
            any = any || (d == 1 || d == 2);
        }
    }
    return any;
}

__device__ __forceinline__ void wdq_push(int *wdq, int e)
{
    const int slot = atomicAdd(wdq, 1);
    wdq[1 + slot] = e;
}

/* the nine residual entries of triangle e into its three node slots (3 rpos[3e + a]) */
__device__ __forceinline__ void store_tri_resid(double *__restrict__ erhs, const int *__restrict__ rpos, int e,
                                                const double r[9])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double *d = erhs + 3 * (size_t)rpos[3 * (size_t)e + a];
        d[0] = r[3 * a];
        d[1] = r[3 * a + 1];
        d[2] = r[3 * a + 2];
    }
}

__global__ __launch_bounds__(256) void k_elem_resid(MeshView m, Sw2Par P, double *__restrict__ erhs,
                                                    const int *__restrict__ rpos, int *flags, int *wdq)
{
    int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= m.E || !m.work[m.elem_inst[e]]) return;
    TriIn in;
    gather_tri(m, e, P.alpha != 0.0, in);
    if (tri_partial(in, P.alpha != 0.0)) {
        wdq_push(wdq, e);
        return;
    }
    double r[9];
    int partial = sw2_tri_resid(in, P, 0, -1, 0.0, r);
    if (partial) flags[0] = 1; /* cannot happen after tri_partial; kept as a guard */
    store_tri_resid(erhs, rpos, e, r);
}

__global__ void k_wdq_reset(int *wdq)
{
    if (threadIdx.x == 0) wdq[0] = 0;
}

/* queued (partially dry) elements: the general wet/dry residual, grid-stride over the queue */
__global__ __launch_bounds__(128) void k_elem_resid_wd(MeshView m, Sw2Par P, const int *__restrict__ wdq,
                                                       double *__restrict__ erhs, const int *__restrict__ rpos)
{
    const int n = wdq[0];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int e = wdq[1 + i];
        TriIn in;
        gather_tri(m, e, true, in);
        double r[9];
        sw2_tri_resid_general(in, P, 0, -1, 0.0, r);
        store_tri_resid(erhs, rpos, e, r);
    }
}

struct SegView {
    const int *seg, *kind;
    const double *val, *sdjac, *nrml, *sol;
    const int *seg_inst, *work;
    int S;
};

/* rpos_seg: the segments' part of rpos (2 slots per segment) */
__global__ void k_seg_resid(SegView s, Sw2Par P, double *__restrict__ erhs, const int *__restrict__ rpos_seg)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s.S || !s.work[s.seg_inst[e]]) return;
    double h[2], u[2], v[2];
    for (int a = 0; a < 2; a++) {
        int n = s.seg[2 * e + a];
        h[a] = s.sol[3 * n];
        u[a] = s.sol[3 * n + 1];
        v[a] = s.sol[3 * n + 2];
    }
    double er[6];
    sw2_seg_resid(s.kind[e], s.val[2 * e], s.val[2 * e + 1], s.nrml[2 * e], s.nrml[2 * e + 1], s.sdjac[e], h, u, v,
                  P, 0, -1, 0.0, er);
    for (int a = 0; a < 2; a++) {
        double *d = erhs + 3 * (size_t)rpos_seg[2 * (size_t)e + a];
        d[0] = er[3 * a];
        d[1] = er[3 * a + 1];
        d[2] = er[3 * a + 2];
    }
}

static SegView make_segview(const adhk_ctx *c)
{
    SegView s;
    s.seg = c->seg;
    s.kind = c->seg_kind;
    s.val = c->seg_val;
    s.sdjac = c->sdjac;
    s.nrml = c->snrml;
    s.sol = c->sol;
    s.seg_inst = c->seg_inst;
    s.work = c->work;
    s.S = c->S;
    return s;
}

/* workgroups of the queue kernels: enough for a wet/dry front, at most one per 128 elements */
static inline unsigned wd_grid(long nwork)
{
    long g = (nwork + 127) / 128;
    return (unsigned)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
}

extern "C" int adhk_launch_elem_resid(adhk_ctx *c)
{
    adhk_kt_begin(c, ADHK_KT_ELEM_RESID);
    Sw2Par P = make_par(c);
    if (c->E > 0) {
        hipLaunchKernelGGL(k_wdq_reset, dim3(1), dim3(64), 0, c->stream, c->wdq);
        const int tpb = c->tune.elem_tpb;
        hipLaunchKernelGGL(k_elem_resid, dim3((c->E + tpb - 1) / tpb), dim3(tpb), 0, c->stream, make_view(c), P, c->erhs,
                           c->rpos, c->flags, c->wdq);
        hipLaunchKernelGGL(k_elem_resid_wd, dim3(wd_grid(c->E)), dim3(128), 0, c->stream, make_view(c), P, c->wdq,
                           c->erhs, c->rpos);
        CHECK_LAUNCH();
    }
    if (c->S > 0) {
        hipLaunchKernelGGL(k_seg_resid, dim3((c->S + 127) / 128), dim3(128), 0, c->stream, make_segview(c), P,
                           c->erhs, c->rpos + 3 * (size_t)c->E);
        CHECK_LAUNCH();
    }
    adhk_kt_end(c);
    return 0;
}

/* One element's residual, optionally perturbed: the reference's per-element residual call
 * (pde[k].resid: sw2_body_resid.c:44, sw2_bc_h.c:57, sw2_bc_vel.c:58), one thread.  which 0: triangle
 * (the general wet/dry form; wd >= 0 replaces the element's wet/dry flag), 1: segment; for checking the
 * kernels' own forms of an all-wet triangle, 2: the residual kernel's, 3: the element Jacobian's (FD form
 * with its shared values).  The result goes to the first 9 (triangle) or 6 (segment) entries of erhs, which
 * the next residual evaluation overwrites. */
__global__ void k_elem_resid_one(MeshView m, SegView s, Sw2Par P, int which, int e, int pnode, int pvar, double pe,
                                 int wd, double *__restrict__ erhs)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (which == 0 || which == 2 || which == 3) {
        TriIn in;
        gather_tri(m, e, true, in);
        if (wd >= 0) in.wd_flag = wd;
        double r[9];
        if (which == 0) {
            sw2_tri_resid_general(in, P, pnode, pvar, pe, r);
        } else if (which == 2) { /* the residual kernel's all-wet form */
            if (sw2_tri_resid<false>(in, P, pnode, pvar, pe, r)) sw2_tri_resid_general(in, P, pnode, pvar, pe, r);
        } else { /* the element Jacobian's FD form, with its shared values */
            FdShare fs;
            fd_share(in, P, fs);
            if (sw2_tri_resid<true>(in, P, pnode, pvar, pe, r, &fs)) sw2_tri_resid_general(in, P, pnode, pvar, pe, r);
        }
        for (int k = 0; k < 9; k++) erhs[k] = r[k];
        return;
    }
    double h[2], u[2], v[2];
    for (int a = 0; a < 2; a++) {
        const int n = s.seg[2 * e + a];
        h[a] = s.sol[3 * n];
        u[a] = s.sol[3 * n + 1];
        v[a] = s.sol[3 * n + 2];
    }
    double er[6];
    sw2_seg_resid(s.kind[e], s.val[2 * e], s.val[2 * e + 1], s.nrml[2 * e], s.nrml[2 * e + 1], s.sdjac[e], h, u, v, P,
                  pnode, pvar, pe, er);
    for (int k = 0; k < 6; k++) erhs[k] = er[k];
}

extern "C" int adhk_elem_residual_one(adhk_sw2 *c, int which, int ie, double pe, int pnode, int pvar, int wd_flag,
                                      double *out)
{
    const int seg = (which == 1);
    if (!c || !out || which < 0 || which > 3 || ie < 0 || ie >= (seg ? c->S : c->E) || pnode < 0 ||
        pnode > (seg ? 1 : 2)) {
        adhk_set_error("adhk_elem_residual_one: bad arguments (which %d, element %d)", which, ie);
        return -1;
    }
    if (pvar < 0 || pvar > 2) pvar = -1;
    hipLaunchKernelGGL(k_elem_resid_one, dim3(1), dim3(64), 0, c->stream, make_view(c), make_segview(c), make_par(c),
                       which, ie, pnode, pvar, pvar >= 0 ? pe : 0.0, wd_flag, c->erhs);
    CHECK_LAUNCH();
    hipError_t e = hipMemcpyAsync(out, c->erhs, (seg ? 6 : 9) * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        adhk_set_error("adhk_elem_residual_one: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

/* residual[3n+k] = -(sum of node n's elemental residual slots, in element order): the slots of a node
 * are one contiguous run, so consecutive nodes read consecutive memory */
__global__ __launch_bounds__(256) void k_resid_gather(int N, const int *__restrict__ rptr,
                                                      const double *__restrict__ erhs, double *__restrict__ residual)
{
    int n = xcd_block() * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    const int j0 = rptr[n], j1 = rptr[n + 1]; /* empty for a node in no element */
    for (int j = j0; j < j1; j++) {
        const double *er = erhs + 3 * (size_t)j;
        r0 += -1.0 * er[0];
        r1 += -1.0 * er[1];
        r2 += -1.0 * er[2];
    }
    residual[3 * n] = r0;
    residual[3 * n + 1] = r1;
    residual[3 * n + 2] = r2;
}

extern "C" int adhk_launch_resid_gather(adhk_ctx *c)
{
    adhk_kt_begin(c, ADHK_KT_RESID_GATHER);
    hipLaunchKernelGGL(k_resid_gather, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->rptr, c->erhs,
                       c->residual);
    CHECK_LAUNCH();
    adhk_kt_end(c);
    return 0;
}

/* ---------------- FD Jacobian ---------------- */
/* Elemental matrices are stored by 3x3 node block in the order of the contribution list: the block
 * (a, b) of triangle e (code 9e + 3a + b; segment s: 9E + 4s + 2a + b) sits at 9 cpos[code], entry
 * (r, c) at + 3r + c.  The assembly gather then reads each block's contributions as one contiguous
 * run (72 bytes per contribution in FP64); the element kernels, which are compute-bound, absorb the
 * scattered block writes. */
__device__ __forceinline__ size_t emat_at(int pos, int r, int c) { return 9 * (size_t)pos + 3 * r + c; }

/* one thread per (element, node i): the thread takes the three columns (i, k), k = h, u, v, in turn, two
 * residual evaluations (+eps, -eps) each, load_elem_mat/perturb_var assemble_jacobian.c:123-363.  The
 * perturbed variable k is the same in every lane at a time, so the values a k-perturbation leaves unchanged
 * (FdShare: the drag coefficient's logarithm when u or v moves, the speeds when h moves) are formed once
 * per thread instead of once per evaluation (k_elem_jac FP64 on the 1.2M-dof flume: 484 -> 425 us; with
 * FP32 storage 460 -> 362 us). */
template <typename TE>
__global__ __launch_bounds__(256) void k_elem_jac(MeshView m, Sw2Par P, TE *__restrict__ emat,
                                                  const int *__restrict__ cpos, int *flags, int *wdq)
{
    /* the element data are gathered once per element into LDS and shared by its three node threads */
    __shared__ TriIn sh[256 / 3];
    const int epb = blockDim.x / 3, wb = xcd_block();
    const int le = threadIdx.x / 3, i = threadIdx.x % 3;
    const int e = wb * epb + le;
    if (threadIdx.x < epb) {
        const int ee = wb * epb + threadIdx.x;
        if (ee < m.E) gather_tri(m, ee, P.alpha != 0.0, sh[threadIdx.x]);
    }
    __syncthreads();
    if (le >= epb || e >= m.E) return;
    if (!m.work[m.elem_inst[e]]) return;
    const TriIn &in = sh[le];
    /* the three threads of an element reach the same decision; node 0 queues it */
    if (tri_partial_fd(in, P.alpha != 0.0)) {
        if (i == 0) wdq_push(wdq, e);
        return;
    }
    FdShare fs;
    fd_share(in, P, fs);
    int pos[3];
#pragma unroll
    for (int a = 0; a < 3; a++) pos[a] = cpos[9 * e + 3 * a + i];
    int bad = 0;
    for (int k = 0; k < 3; k++) {
        const double var = (k == 0) ? in.h[i] : ((k == 1) ? in.u[i] : in.v[i]);
        const double eps = num_diff_eps(var, 1e-4);
        const double eps2 = 2.0 * eps;
        double rp[9], rm[9];
        bad |= sw2_tri_resid<true>(in, P, i, k, (+1) * eps, rp, &fs);
        bad |= sw2_tri_resid<true>(in, P, i, k, (-1) * eps, rm, &fs);
        /* column (i, k) lies in the blocks (a, i) of the three row nodes a */
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int r = 0; r < 3; r++) emat[emat_at(pos[a], r, k)] = (TE)((rp[3 * a + r] - rm[3 * a + r]) / eps2);
    }
    if (bad) flags[0] = 1; /* cannot happen after tri_partial_fd; kept as a guard */
}

/* queued elements: nine columns each, central differences of the general wet/dry residual */
template <typename TE>
__global__ __launch_bounds__(128) void k_elem_jac_wd(MeshView m, Sw2Par P, const int *__restrict__ wdq,
                                                     TE *__restrict__ emat, const int *__restrict__ cpos)
{
    const long n = 9L * wdq[0];
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int e = wdq[1 + t / 9], col = (int)(t % 9);
        const int i = col / 3, k = col % 3;
        TriIn in;
        gather_tri(m, e, true, in);
        const double var = (k == 0) ? in.h[i] : ((k == 1) ? in.u[i] : in.v[i]);
        const double eps = num_diff_eps(var, 1e-4);
        const double eps2 = 2.0 * eps;
        double rp[9], rm[9];
        sw2_tri_resid_general(in, P, i, k, (+1) * eps, rp);
        sw2_tri_resid_general(in, P, i, k, (-1) * eps, rm);
        for (int a = 0; a < 3; a++) {
            const int pos = cpos[9 * e + 3 * a + i];
            for (int r = 0; r < 3; r++) emat[emat_at(pos, r, k)] = (TE)((rp[3 * a + r] - rm[3 * a + r]) / eps2);
        }
    }
}

template <typename TE>
__global__ void k_seg_jac(SegView s, Sw2Par P, int E, TE *__restrict__ emat, const int *__restrict__ cpos)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * s.S) return;
    int e = t / 6, col = t % 6;
    if (!s.work[s.seg_inst[e]]) return;
    int i = col / 3, k = col % 3;
    double h[2], u[2], v[2];
    for (int a = 0; a < 2; a++) {
        int n = s.seg[2 * e + a];
        h[a] = s.sol[3 * n];
        u[a] = s.sol[3 * n + 1];
        v[a] = s.sol[3 * n + 2];
    }
    double var = (k == 0) ? h[i] : ((k == 1) ? u[i] : v[i]);
    double eps = num_diff_eps(var, 1e-4);
    double eps2 = 2.0 * eps;
    double rp[6], rm[6];
    sw2_seg_resid(s.kind[e], s.val[2 * e], s.val[2 * e + 1], s.nrml[2 * e], s.nrml[2 * e + 1], s.sdjac[e], h, u, v,
                  P, i, k, (+1) * eps, rp);
    sw2_seg_resid(s.kind[e], s.val[2 * e], s.val[2 * e + 1], s.nrml[2 * e], s.nrml[2 * e + 1], s.sdjac[e], h, u, v,
                  P, i, k, (-1) * eps, rm);
    for (int a = 0; a < 2; a++) {
        const int pos = cpos[9 * E + 4 * e + 2 * a + i];
        for (int r = 0; r < 3; r++) emat[emat_at(pos, r, k)] = (TE)((rp[3 * a + r] - rm[3 * a + r]) / eps2);
    }
}

template <typename TE> static int launch_elem_jac(adhk_ctx *c, TE *emat)
{
    adhk_kt_begin(c, ADHK_KT_ELEM_JAC);
    Sw2Par P = make_par(c);
    if (c->E > 0) {
        long nt = 9L * c->E;
        hipLaunchKernelGGL(k_wdq_reset, dim3(1), dim3(64), 0, c->stream, c->wdq);
        const int tpb = c->tune.elem_tpb, epb = tpb / 3; /* whole elements per workgroup */
        hipLaunchKernelGGL(k_elem_jac<TE>, dim3((unsigned)((c->E + epb - 1) / epb)), dim3(tpb), 0, c->stream,
                           make_view(c), P, emat, c->cpos, c->flags, c->wdq);
        hipLaunchKernelGGL(k_elem_jac_wd<TE>, dim3(wd_grid(nt)), dim3(128), 0, c->stream, make_view(c), P, c->wdq,
                           emat, c->cpos);
        CHECK_LAUNCH();
    }
    if (c->S > 0) {
        hipLaunchKernelGGL(k_seg_jac<TE>, dim3((6 * c->S + 127) / 128), dim3(128), 0, c->stream, make_segview(c), P,
                           c->E, emat, c->cpos);
        CHECK_LAUNCH();
    }
    adhk_kt_end(c);
    return 0;
}

/* the elemental Jacobians in FP64 (emat), or rounded to FP32 (emat32) for the mixed-precision solve */
extern "C" int adhk_launch_elem_jac_p(adhk_ctx *c, int f32)
{
    return f32 ? launch_elem_jac<float>(c, c->emat32) : launch_elem_jac<double>(c, c->emat);
}
extern "C" int adhk_launch_elem_jac(adhk_ctx *c) { return adhk_launch_elem_jac_p(c, 0); }

/* one thread per SELL slot: block entry (3r+cc) = sum over the block's contributions (element order)
 * of their entries (r, cc); the contributions of block b are the contiguous run of 3x3 blocks
 * [cptr[b], cptr[b+1]) of emat; padding slots get 0 */
template <typename TE, typename TB>
__global__ __launch_bounds__(256) void k_jac_gather(int nslot64, const int *__restrict__ sell2blk,
                                                    const int *__restrict__ cptr, const TE *__restrict__ emat,
                                                    TB *__restrict__ bvals)
{
    int pos = xcd_block() * blockDim.x + threadIdx.x;
    if (pos >= nslot64) return;
    int b = sell2blk[pos];
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) acc[k] = 0.0;
    const int j0 = (b >= 0) ? cptr[b] : 0, j1 = (b >= 0) ? cptr[b + 1] : 0;
    for (int j = j0; j < j1; j++) {
        const TE *em = emat + 9L * j;
#pragma unroll
        for (int k = 0; k < 9; k++) acc[k] += (double)em[k];
    }
    TB *out = bvals + (size_t)(pos >> 6) * 9 * 64 + (pos & 63);
#pragma unroll
    for (int k = 0; k < 9; k++) out[(size_t)k * 64] = (TB)acc[k];
}

/* block assembly: FP64 values from FP64 elemental matrices, or (mixed precision) FP32 values summed in
 * FP64 from the FP32 elemental matrices, written straight into the FP32 matrix */
extern "C" int adhk_launch_jac_gather_p(adhk_ctx *c, int f32)
{
    adhk_kt_begin(c, ADHK_KT_JAC_GATHER);
    if (f32)
        hipLaunchKernelGGL((k_jac_gather<float, float>), dim3((c->nslot * 64 + 255) / 256), dim3(256), 0, c->stream,
                           c->nslot * 64, c->sell2blk, c->cptr, c->emat32, c->bvals32);
    else
        hipLaunchKernelGGL((k_jac_gather<double, double>), dim3((c->nslot * 64 + 255) / 256), dim3(256), 0,
                           c->stream, c->nslot * 64, c->sell2blk, c->cptr, c->emat, c->bvals);
    CHECK_LAUNCH();
    adhk_kt_end(c);
    return 0;
}
extern "C" int adhk_launch_jac_gather(adhk_ctx *c) { return adhk_launch_jac_gather_p(c, 0); }

/* ---------------- wet/dry flags, fe_sw2_wdflag_legacy sw2_wdflag_legacy.c:34-139 ---------------- */
__global__ void k_wd_zero(int N, int *tmp)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) tmp[n] = 0;
}
__global__ void k_wd_mark(int E, const int *tri, const double *sol, int *tmp)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int wet = 0;
    for (int a = 0; a < 3; a++)
        if (sol[3 * tri[a * E + e]] > 0.0) wet++;
    if (wet != 3)
        for (int a = 0; a < 3; a++) tmp[tri[a * E + e]] = 1; /* every writer stores the same value */
}
__global__ void k_wd_flag(int E, const int *tri, const double *sol, const double *sol_old, const int *tmp, int *wd)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int n0 = tri[e], n1 = tri[E + e], n2 = tri[2 * E + e];
    int f = (tmp[n0] != 0 || tmp[n1] != 0 || tmp[n2] != 0) ? 1 : 0;
    if (sol[3 * n0] < 0 && sol[3 * n1] < 0 && sol[3 * n2] < 0 && sol_old[3 * n0] < 0 && sol_old[3 * n1] < 0 &&
        sol_old[3 * n2] < 0)
        f = 2;
    wd[e] = f;
}

extern "C" int adhk_launch_wdflag(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_wd_zero, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->wd_tmp);
    CHECK_LAUNCH();
    if (c->E > 0) {
        hipLaunchKernelGGL(k_wd_mark, dim3((c->E + 255) / 256), dim3(256), 0, c->stream, c->E, c->tri, c->sol,
                           c->wd_tmp);
        CHECK_LAUNCH();
        hipLaunchKernelGGL(k_wd_flag, dim3((c->E + 255) / 256), dim3(256), 0, c->stream, c->E, c->tri, c->sol,
                           c->sol_old, c->wd_tmp, c->wd_flag);
        CHECK_LAUNCH();
    }
    return 0;
}

/* fe_sw2_init, fe_sw2_init.c:28-53: nodes dry at the old level get u, v Dirichlet */
__global__ void k_dry_bc(int N, const double *sol_old, int *bc_mask)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (sol_old[3 * n] <= 0.) {
        bc_mask[3 * n + 1] = 1;
        bc_mask[3 * n + 2] = 1;
    }
}

extern "C" int adhk_launch_dry_bc(adhk_ctx *c)
{
    hipLaunchKernelGGL(k_dry_bc, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->sol_old, c->bc_mask);
    CHECK_LAUNCH();
    return 0;
}

/* ---- grid mass (tl_find_grid_mass_elem2d, tl_grid_mass.c:199-262) ----
 * Water volume of the triangles at the current depth, with the reference's treatment of partially dry
 * elements: fe_3m2d_dry_wet_factornew (tl_grid_mass.c:442-536) gives the wet area fraction f (a node
 * is dry when h < 0); one dry node: the full-element integral of h plus -(1 - f) times the integral
 * of the negative nodal depths; two dry nodes: f times the integral of the clipped depths; otherwise
 * f times the integral of h.  integrate_triangle_f (fe_integrals_2d.c:25) is c djac (f0+f1+f2)/3.
 * One workgroup sums 256 elements; the host adds the workgroup sums in index order. */
__device__ static double gm_wet_factor(const double *xx, const double *yy, const double *h, double djac,
                                       int &num_dry)
{
    int total = 0, pass = 0;
    for (int i = 0; i < 3; i++) {
        const int d = h[i] < 0.0 ? 1 : 0;
        total += d;
        pass += d << i;
    }
    if (total == 3) {
        num_dry = 3;
        return 0.0;
    }
    if (total == 0) {
        num_dry = 0;
        return 1.0;
    }
    if (total != 1) pass = 7 - pass;
    const int only = pass / 2;
    double nx[3], ny[3];
    for (int i = 0; i < 3; i++) {
        if (i != only) {
            const double xsi = h[i] / (h[i] - h[only]);
            nx[i] = xx[i] + xsi * (xx[only] - xx[i]); /* FIXED_POS, macro.h:756 */
            ny[i] = yy[i] + xsi * (yy[only] - yy[i]);
        } else {
            nx[i] = xx[i];
            ny[i] = yy[i];
        }
    }
    const double a = ((nx[1] - nx[0]) * (ny[2] - ny[0]) - (ny[1] - ny[0]) * (nx[2] - nx[0])) / 2.0; /* TRI_AREA */
    if (a / djac < D_SMALL) {
        const int i = total % 2;
        num_dry = i == 1 ? 0 : 3;
        return (double)i;
    }
    double factor = (total >> 1) - (total % 2);
    num_dry = total;
    factor *= a / djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

__global__ __launch_bounds__(256) void k_grid_mass(MeshView m, double *__restrict__ part)
{
    __shared__ double sh[4];
    const int e = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (e < m.E) {
        double xx[3], yy[3], h[3], hneg[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; a++) {
            const int n = m.tri[a * m.E + e];
            xx[a] = m.x[n];
            yy[a] = m.y[n];
            h[a] = m.sol[3 * n];
        }
        const double djac = m.djac[e];
        int num_dry;
        const double f = gm_wet_factor(xx, yy, h, djac, num_dry);
        double f1 = f, f2 = 0.0;
        if (num_dry == 1) {
            for (int a = 0; a < 3; a++)
                if (h[a] < 0.0) hneg[a] = h[a];
            f1 = 1.0;
            f2 = -1.0 * (1.0 - f);
        } else if (num_dry == 2) {
            for (int a = 0; a < 3; a++)
                if (h[a] < 0.0) {
                    hneg[a] = h[a];
                    h[a] = 0.0;
                }
        }
        v = (f1 * djac * (h[0] + h[1] + h[2]) / 3.0) + (f2 * djac * (hneg[0] + hneg[1] + hneg[2]) / 3.0);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

extern "C" int adhk_sw2_grid_mass(adhk_sw2 *c, double density, double *mass)
{
    if (!c || !mass) return -1;
    const int nb = (c->E + 255) / 256;
    double *part = NULL, *h = NULL;
    int rc = -1;
    if (nb == 0) {
        *mass = 0.0;
        return 0;
    }
    if (hipMalloc((void **)&part, (size_t)nb * sizeof(double)) != hipSuccess) {
        adhk_set_error("grid_mass: out of device memory");
        return -1;
    }
    h = (double *)malloc((size_t)nb * sizeof(double));
    if (h) {
        hipLaunchKernelGGL(k_grid_mass, dim3(nb), dim3(256), 0, c->stream, make_view(c), part);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h, part, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) {
            double s = 0.0;
            for (int b = 0; b < nb; b++) s += h[b];
            *mass = s * density;
            rc = 0;
        } else {
            adhk_set_error("grid_mass: %s", hipGetErrorString(e));
        }
    }
    free(h);
    hipFree(part);
    return rc;
}
