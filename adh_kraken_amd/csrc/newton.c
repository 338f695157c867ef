/*
 * newton.c -- C host driver of the SW2 Newton step on one MI355X.
 *
 * Restates the reference's host control flow -- fe_newton
 * (src/newton/fe_newton.c:44-926), initialize_system / initialize_dirichlet_bc
 * (newton_tools.c:21, :130), get_residual_norms (:303), the time loop
 * (src/timeloop/time_loop.c:14-93) -- over device-resident state.  Every
 * numerical operation runs in a gfx950 kernel; the host only reads back the
 * three residual norms per residual evaluation and the BiCGStab "done" flag
 * once per chunk of iterations, and takes the reference's branch decisions
 * (line search, convergence, time-step cut) on them.
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "adhk_internal.h"

static __thread char g_err[1024];

void adhk_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *adhk_last_error(void) { return g_err; }

void adhk_solver_defaults(adhk_solver *s)
{
    s->tol_nonlin = 1e-8; /* smodel_super_alloc_init.c:63 */
    s->inc_nonlin = 1e-6; /* :62 */
    s->max_nonlin_it = 10; /* :65 */
    s->max_linesearch_cuts = 5; /* :64 */
    s->lin_tol = 1e-11; /* bcgstab_solver.c:20-21 */
    s->lin_max_it = 100; /* bcgstab_solver.c:22 */
    s->precision = ADHK_PREC_FP64;
    s->lin_check_every = 8;
    s->mixed_lin_tol = 1e-6;
    s->warm_gate = 1;
    s->substep_depth = 0;
    s->precond = ADHK_PC_BJ;
}

/* the preconditioner of a solve: multicolour SGS only for one mesh alone */
static int pc_of(const adhk_ctx *c, const adhk_solver *s)
{
    return (s->precond == ADHK_PC_SGS && c->ninst == 1) ? ADHK_PC_SGS : ADHK_PC_BJ;
}

#define HC(x)                                                                                                     \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));                        \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)
#define LC(x)                                                                                                     \
    do {                                                                                                          \
        if ((x) != 0) goto fail;                                                                                  \
    } while (0)

/* fr_manningsn_to_rheight, src/friction/fr_manningsn_to_rheight.c:15-32 */
static double manning_to_rheight(double n, double kn, double g)
{
    if (n < 1.E-8) return 0.0;
    return pow((8.25 * sqrt(g) * n / kn), 6.);
}

static void ctx_free(adhk_ctx *c)
{
    if (!c) return;
    void *ptrs[] = {c->x, c->y, c->z, c->tri, c->djac, c->grad, c->ks, c->mann, c->nstat, c->seg, c->seg_kind, c->seg_val, c->sdjac,
                    c->snrml, c->sol, c->sol_old, c->sol_older, c->dirichlet, c->dsol, c->residual, c->scale,
                    c->bc_mask, c->wd_flag, c->wd_tmp, c->wdq, c->brow, c->bcol, c->diag_blk, c->cptr, c->contrib, c->bvals,
                    c->minv, c->bvals32, c->minv32, c->erhs, c->emat, c->r, c->pv, c->Ap, c->s, c->As, c->q, c->x0,
                    c->pv2, c->Ap2, c->xk, c->r32, c->p32, c->Ap32, c->s32, c->As32, c->q32, c->xk32, c->p32b,
                    c->Ap32b, c->emat32, c->partials, c->counter, c->cpos, c->rptr, c->rpos, c->sc, c->flags, c->slice_ptr, c->scol, c->sell2blk,
                    c->blk2sell, c->diag_slot, c->diag_id, c->node_inst, c->elem_inst, c->seg_inst, c->blk_node,
                    c->blk_inst, c->inst_blk, c->work};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
        if (ptrs[i]) hipFree(ptrs[i]);
    adhk_sgs_free(c);
    free(c->h_inst_node);
    free(c->h_inst_elem);
    free(c->h_inst_seg);
    free(c->h_inst_blk);
    free(c->h_work);
    free(c->inst_stats);
    free(c->alive);
    free(c->steps_done);
    adhk_series_free(c);
    for (int i = 0; i < 2; i++) {
        if (c->gchunk[i]) hipGraphExecDestroy(c->gchunk[i]);
        for (int k = 0; k < 3; k++)
            if (c->gstage[i][k].exec) hipGraphExecDestroy(c->gstage[i][k].exec);
    }
    for (int i = 0; i < 8; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    for (int i = 0; i < ADHK_KT_RING; i++) {
        if (c->kt.a[i]) hipEventDestroy(c->kt.a[i]);
        if (c->kt.b[i]) hipEventDestroy(c->kt.b[i]);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    free(c);
}

void adhk_sw2_destroy(adhk_sw2 *c) { ctx_free(c); }

#define DALLOC(ptr, n, T) HC(hipMalloc((void **)&(ptr), ((size_t)(n) > 0 ? (size_t)(n) : 1) * sizeof(T)))
#define DZERO(ptr, n, T) HC(hipMemsetAsync((ptr), 0, ((size_t)(n) > 0 ? (size_t)(n) : 1) * sizeof(T), c->stream))

static adhk_sw2 *create_impl(int ninst, const int *node_off, const int *elem_off, const int *seg_off, int nnodes,
                             const double *x, const double *y, const double *z, int nelems2d, const int *tri,
                             const double *mann, int nelems1d, const int *seg, const int *seg_kind,
                             const double *seg_val, const adhk_params *p)
{
    adhk_ctx *c = NULL;
    double *h_djac = NULL, *h_grad = NULL, *h_ks = NULL, *h_sdjac = NULL, *h_snrml = NULL;
    int *h_tri = NULL, *h_ni = NULL, *h_ei = NULL, *h_si = NULL, *h_bn = NULL, *h_bi = NULL;
    if (nnodes <= 0 || nelems2d <= 0 || nelems1d < 0 || !x || !y || !z || !tri || !p ||
        (nelems1d > 0 && (!seg || !seg_kind || !seg_val))) {
        adhk_set_error("adhk_sw2_create: invalid arguments");
        return NULL;
    }
    if (ninst < 1 || !node_off || !elem_off || !seg_off || node_off[0] != 0 || elem_off[0] != 0 ||
        seg_off[0] != 0 || node_off[ninst] != nnodes || elem_off[ninst] != nelems2d || seg_off[ninst] != nelems1d) {
        adhk_set_error("adhk_sw2_create: instance offsets must start at 0 and end at the totals");
        return NULL;
    }
    for (int i = 0; i < ninst; i++)
        if (node_off[i + 1] <= node_off[i] || elem_off[i + 1] <= elem_off[i] || seg_off[i + 1] < seg_off[i]) {
            adhk_set_error("adhk_sw2_create: instance %d is empty or its offsets decrease", i);
            return NULL;
        }
    if ((long)nnodes * 3 >= 2147483647L || (long)nelems2d * 81 >= 2147483647L) {
        adhk_set_error("adhk_sw2_create: mesh too large for 32-bit indices");
        return NULL;
    }
    for (int k = 0; k < ninst; k++) {
        for (long i = 3L * elem_off[k]; i < 3L * elem_off[k + 1]; i++)
            if (tri[i] < node_off[k] || tri[i] >= node_off[k + 1]) {
                adhk_set_error("adhk_sw2_create: triangle node index out of its instance's range at %ld", i);
                return NULL;
            }
        for (long i = 2L * seg_off[k]; i < 2L * seg_off[k + 1]; i++)
            if (seg[i] < node_off[k] || seg[i] >= node_off[k + 1]) {
                adhk_set_error("adhk_sw2_create: segment node index out of its instance's range at %ld", i);
                return NULL;
            }
    }
    for (int i = 0; i < nelems1d; i++)
        if (seg_kind[i] < ADHK_BC_DEPTH || seg_kind[i] > ADHK_BC_DEPTH_NODAL) {
            adhk_set_error("adhk_sw2_create: unknown segment kind %d", seg_kind[i]);
            return NULL;
        }
    /* node direction check of selem_2d.c:75-80: a triangle whose area (djac) is below 1e-6 is
     * clockwise or degenerate; the reference stops there, this returns the error.  Segments shorter
     * than 1e-6 have no normal. */
    for (int e = 0; e < nelems2d; e++) {
        const int *t = tri + 3 * e;
        const double area =
            0.5 * ((x[t[1]] - x[t[0]]) * (y[t[2]] - y[t[0]]) - (x[t[2]] - x[t[0]]) * (y[t[1]] - y[t[0]]));
        if (!(area >= 1e-6)) {
            adhk_set_error("adhk_sw2_create: improperly numbered triangle %d (nodes %d %d %d): area %g", e, t[0], t[1],
                           t[2], area);
            return NULL;
        }
    }
    for (int e = 0; e < nelems1d; e++) {
        const int *sg = seg + 2 * e;
        const double dx = x[sg[1]] - x[sg[0]], dy = y[sg[1]] - y[sg[0]];
        if (!(sqrt(dx * dx + dy * dy) >= 1e-6)) {
            adhk_set_error("adhk_sw2_create: degenerate segment %d (nodes %d %d)", e, sg[0], sg[1]);
            return NULL;
        }
    }
    c = (adhk_ctx *)calloc(1, sizeof(adhk_ctx));
    if (!c) {
        adhk_set_error("adhk_sw2_create: out of host memory");
        return NULL;
    }
    c->N = nnodes;
    c->E = nelems2d;
    c->S = nelems1d;
    c->ndofs = 3 * nnodes;
    c->p = *p;
    c->ninst = ninst;
    c->tune.chunk = 0;
    c->tune.graphs = 0;
    c->tune.xcd_map = 1;
    c->tune.elem_tpb = 256;
    c->tune.split = 0;
    c->tune.stream_matrix = -1;
    c->tune.sgs_wide = 0;
    HC(hipStreamCreate(&c->stream));
    for (int i = 0; i < 8; i++) HC(hipEventCreate(&c->ev[i]));

    /* geometry: selem2d_get_triangle_linear_djac_nrml_gradPhi (selem_2d.c:270-351), stored SoA */
    {
        const int E = nelems2d, S = nelems1d;
        h_djac = malloc((size_t)E * sizeof(double));
        h_grad = malloc((size_t)E * 6 * sizeof(double));
        h_ks = malloc((size_t)E * sizeof(double));
        h_tri = malloc((size_t)E * 3 * sizeof(int));
        h_sdjac = malloc((size_t)(S > 0 ? S : 1) * sizeof(double));
        h_snrml = malloc((size_t)(S > 0 ? S : 1) * 2 * sizeof(double));
        if (!h_djac || !h_grad || !h_ks || !h_tri || !h_sdjac || !h_snrml) {
            adhk_set_error("adhk_sw2_create: out of host memory");
            goto fail;
        }
        for (int e = 0; e < E; e++) {
            const int *t = tri + 3 * e;
            double s1x = x[t[1]] - x[t[0]], s1y = y[t[1]] - y[t[0]];
            double s2x = x[t[2]] - x[t[0]], s2y = y[t[2]] - y[t[0]];
            double dj = (s1x * s2y) - (s2x * s1y); /* 2 x area >= 2e-6, checked above */
            double g6[6] = {(s1y - s2y), s2y, -s1y, (s2x - s1x), -s2x, s1x}; /* gx0 gx1 gx2 gy0 gy1 gy2 */
            for (int k = 0; k < 6; k++) h_grad[(size_t)k * E + e] = g6[k] * (1. / dj);
            h_djac[e] = dj * 0.5;
            for (int a = 0; a < 3; a++) h_tri[(size_t)a * E + e] = t[a];
            h_ks[e] = (p->fric_mode == ADHK_FRIC_MANNING && mann) ? manning_to_rheight(mann[e], p->manning_units, p->g)
                                                                   : 0.0;
        }
        /* selem1d_load, selem_1d.c:42-56 */
        for (int e = 0; e < S; e++) {
            const int *s = seg + 2 * e;
            double dx = x[s[1]] - x[s[0]], dy = y[s[1]] - y[s[0]];
            double L = sqrt(dx * dx + dy * dy); /* >= 1e-6, checked above */
            h_sdjac[e] = L;
            h_snrml[2 * e] = (y[s[1]] - y[s[0]]) / L;
            h_snrml[2 * e + 1] = -(x[s[1]] - x[s[0]]) / L;
        }
        const int N = nnodes;
        DALLOC(c->x, N, double);
        DALLOC(c->y, N, double);
        DALLOC(c->z, N, double);
        DALLOC(c->tri, 3L * E, int);
        DALLOC(c->djac, E, double);
        DALLOC(c->grad, 6L * E, double);
        DALLOC(c->ks, E, double);
        DALLOC(c->mann, E, double);
        DALLOC(c->seg, 2L * S, int);
        DALLOC(c->seg_kind, S, int);
        DALLOC(c->seg_val, 2L * S, double);
        DALLOC(c->sdjac, S, double);
        DALLOC(c->snrml, 2L * S, double);
        HC(hipMemcpy(c->x, x, (size_t)N * sizeof(double), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->y, y, (size_t)N * sizeof(double), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->z, z, (size_t)N * sizeof(double), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->tri, h_tri, (size_t)E * 3 * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->djac, h_djac, (size_t)E * sizeof(double), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->grad, h_grad, (size_t)E * 6 * sizeof(double), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->ks, h_ks, (size_t)E * sizeof(double), hipMemcpyHostToDevice));
        if (mann)
            HC(hipMemcpy(c->mann, mann, (size_t)E * sizeof(double), hipMemcpyHostToDevice));
        else
            HC(hipMemset(c->mann, 0, (size_t)E * sizeof(double)));
        if (S > 0) {
            HC(hipMemcpy(c->seg, seg, (size_t)S * 2 * sizeof(int), hipMemcpyHostToDevice));
            HC(hipMemcpy(c->seg_kind, seg_kind, (size_t)S * sizeof(int), hipMemcpyHostToDevice));
            HC(hipMemcpy(c->seg_val, seg_val, (size_t)S * 2 * sizeof(double), hipMemcpyHostToDevice));
            HC(hipMemcpy(c->sdjac, h_sdjac, (size_t)S * sizeof(double), hipMemcpyHostToDevice));
            HC(hipMemcpy(c->snrml, h_snrml, (size_t)S * 2 * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    {
        const long nd = c->ndofs, N = c->N, E = c->E, S = c->S;
        DALLOC(c->sol, nd, double);
        DALLOC(c->sol_old, nd, double);
        DALLOC(c->sol_older, nd, double);
        DALLOC(c->dirichlet, nd, double);
        DALLOC(c->dsol, nd, double);
        DALLOC(c->residual, nd, double);
        DALLOC(c->scale, nd, double);
        DALLOC(c->bc_mask, nd, int);
        DALLOC(c->wd_flag, E, int);
        DALLOC(c->wd_tmp, N, int);
        DALLOC(c->wdq, E + 1, int);
        DZERO(c->wdq, E + 1, int);
        DZERO(c->sol, nd, double);
        DZERO(c->sol_old, nd, double);
        DZERO(c->sol_older, nd, double);
        DZERO(c->dirichlet, nd, double);
        DZERO(c->dsol, nd, double);
        DZERO(c->residual, nd, double);
        DZERO(c->scale, nd, double);
        DZERO(c->bc_mask, nd, int);
        DZERO(c->wd_flag, E, int);
        DALLOC(c->erhs, 9 * E + 6 * S, double);
        DALLOC(c->emat, 81 * E + 36 * S, double);
        DALLOC(c->r, nd, double);
        DALLOC(c->pv, nd, double);
        DALLOC(c->Ap, nd, double);
        DALLOC(c->s, nd, double);
        DALLOC(c->As, nd, double);
        DALLOC(c->q, nd, double);
        DALLOC(c->x0, nd, double);
        DALLOC(c->pv2, nd, double);
        DALLOC(c->Ap2, nd, double);
        DALLOC(c->xk, nd, double);
        DALLOC(c->counter, 8, unsigned int);
        DALLOC(c->sc, ninst, adhk_scalars);
        DALLOC(c->flags, 4, int);
        DZERO(c->counter, 8, unsigned int);
        DZERO(c->sc, ninst, adhk_scalars);
        DZERO(c->flags, 4, int);
    }
    /* instances and their 256-node blocks */
    {
        const int N = c->N, E = c->E, S = c->S;
        int nb = 0;
        for (int k = 0; k < ninst; k++) nb += (node_off[k + 1] - node_off[k] + 255) / 256;
        c->nblk = nb;
        c->h_inst_node = malloc((size_t)(ninst + 1) * sizeof(int));
        c->h_inst_elem = malloc((size_t)(ninst + 1) * sizeof(int));
        c->h_inst_seg = malloc((size_t)(ninst + 1) * sizeof(int));
        c->h_inst_blk = malloc((size_t)(ninst + 1) * sizeof(int));
        c->h_work = malloc((size_t)ninst * sizeof(int));
        c->inst_stats = calloc((size_t)ninst, sizeof(adhk_newton_stats));
        c->alive = malloc((size_t)ninst * sizeof(int));
        c->steps_done = calloc((size_t)ninst, sizeof(int));
        h_ni = malloc((size_t)N * sizeof(int));
        h_ei = malloc((size_t)E * sizeof(int));
        h_si = malloc((size_t)(S > 0 ? S : 1) * sizeof(int));
        h_bn = malloc((size_t)(nb + 1) * sizeof(int));
        h_bi = malloc((size_t)nb * sizeof(int));
        if (!c->h_inst_node || !c->h_inst_elem || !c->h_inst_seg || !c->h_inst_blk || !c->h_work ||
            !c->inst_stats || !c->alive || !c->steps_done || !h_ni || !h_ei || !h_si || !h_bn || !h_bi) {
            adhk_set_error("adhk_sw2_create: out of host memory");
            goto fail;
        }
        int b = 0;
        for (int k = 0; k < ninst; k++) {
            c->h_inst_node[k] = node_off[k];
            c->h_inst_elem[k] = elem_off[k];
            c->h_inst_seg[k] = seg_off[k];
            c->h_inst_blk[k] = b;
            c->h_work[k] = 1;
            c->alive[k] = 1;
            for (int n = node_off[k]; n < node_off[k + 1]; n++) h_ni[n] = k;
            for (int e = elem_off[k]; e < elem_off[k + 1]; e++) h_ei[e] = k;
            for (int e = seg_off[k]; e < seg_off[k + 1]; e++) h_si[e] = k;
            for (int n = node_off[k]; n < node_off[k + 1]; n += 256) {
                h_bn[b] = n;
                h_bi[b] = k;
                b++;
            }
        }
        h_bn[nb] = N;
        c->h_inst_node[ninst] = N;
        c->h_inst_elem[ninst] = E;
        c->h_inst_seg[ninst] = S;
        c->h_inst_blk[ninst] = nb;
        DALLOC(c->node_inst, N, int);
        DALLOC(c->elem_inst, E, int);
        DALLOC(c->seg_inst, S, int);
        DALLOC(c->blk_node, nb + 1, int);
        DALLOC(c->blk_inst, nb, int);
        DALLOC(c->inst_blk, ninst + 1, int);
        DALLOC(c->work, ninst, int);
        HC(hipMemcpy(c->node_inst, h_ni, (size_t)N * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->elem_inst, h_ei, (size_t)E * sizeof(int), hipMemcpyHostToDevice));
        if (S > 0) HC(hipMemcpy(c->seg_inst, h_si, (size_t)S * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->blk_node, h_bn, (size_t)(nb + 1) * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->blk_inst, h_bi, (size_t)nb * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->inst_blk, c->h_inst_blk, (size_t)(ninst + 1) * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(c->work, c->h_work, (size_t)ninst * sizeof(int), hipMemcpyHostToDevice));
        DALLOC(c->partials, 8L * nb, double);
    }
    LC(adhk_build_pattern(c));
    DALLOC(c->bvals, 9L * 64 * c->nslot, double);
    DZERO(c->bvals, 9L * 64 * c->nslot, double);
    DALLOC(c->minv, 9L * 64 * c->nslice, double);
    DZERO(c->minv, 9L * 64 * c->nslice, double);
    HC(hipStreamSynchronize(c->stream));
    free(h_djac);
    free(h_grad);
    free(h_ks);
    free(h_tri);
    free(h_sdjac);
    free(h_snrml);
    free(h_ni);
    free(h_ei);
    free(h_si);
    free(h_bn);
    free(h_bi);
    return c;
fail:
    free(h_djac);
    free(h_grad);
    free(h_ks);
    free(h_tri);
    free(h_sdjac);
    free(h_snrml);
    free(h_ni);
    free(h_ei);
    free(h_si);
    free(h_bn);
    free(h_bi);
    ctx_free(c);
    return NULL;
}

adhk_sw2 *adhk_sw2_create(int nnodes, const double *x, const double *y, const double *z, int nelems2d,
                          const int *tri, const double *mann, int nelems1d, const int *seg, const int *seg_kind,
                          const double *seg_val, const adhk_params *p)
{
    const int no[2] = {0, nnodes}, eo[2] = {0, nelems2d}, so[2] = {0, nelems1d};
    return create_impl(1, no, eo, so, nnodes, x, y, z, nelems2d, tri, mann, nelems1d, seg, seg_kind, seg_val, p);
}

adhk_sw2 *adhk_sw2_create_batch(int ninst, const int *node_off, const int *elem_off, const int *seg_off, int nnodes,
                                const double *x, const double *y, const double *z, int nelems2d, const int *tri,
                                const double *mann, int nelems1d, const int *seg, const int *seg_kind,
                                const double *seg_val, const adhk_params *p)
{
    return create_impl(ninst, node_off, elem_off, seg_off, nnodes, x, y, z, nelems2d, tri, mann, nelems1d, seg,
                       seg_kind, seg_val, p);
}

static int ensure_mixed(adhk_ctx *c)
{
    if (c->bvals32) return 0;
    const long nd = c->ndofs;
    DALLOC(c->bvals32, 9L * 64 * c->nslot, float);
    DALLOC(c->minv32, 9L * 64 * c->nslice, float);
    DALLOC(c->p32b, nd, float);
    DALLOC(c->Ap32b, nd, float);
    DALLOC(c->r32, nd, float);
    DALLOC(c->p32, nd, float);
    DALLOC(c->Ap32, nd, float);
    DALLOC(c->s32, nd, float);
    DALLOC(c->As32, nd, float);
    DALLOC(c->q32, nd, float);
    DALLOC(c->xk32, nd, float);
    if (!c->emat32) DALLOC(c->emat32, 81L * c->E + 36L * c->S + 1, float);
    return 0;
fail:
    return -1;
}

int adhk_sw2_sizes(const adhk_sw2 *c, int *ndofs, int *nnz, int *nnzb)
{
    if (!c) return -1;
    if (ndofs) *ndofs = c->ndofs;
    if (nnz) *nnz = 9 * c->nnzb;
    if (nnzb) *nnzb = c->nnzb;
    return 0;
}

int adhk_device_synchronize(void)
{
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        adhk_set_error("hipDeviceSynchronize: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int adhk_sw2_layout(const adhk_sw2 *c, int *nslice, int *nslot)
{
    if (!c) return -1;
    if (nslice) *nslice = c->nslice;
    if (nslot) *nslot = c->nslot;
    return 0;
}

static int copy_in(adhk_ctx *c, double *dst, const double *src)
{
    if (!src) return 0;
    HC(hipMemcpy(dst, src, (size_t)c->ndofs * sizeof(double), hipMemcpyHostToDevice));
    return 0;
fail:
    return -1;
}
static int copy_out(adhk_ctx *c, double *dst, const double *src)
{
    if (!dst) return 0;
    HC(hipStreamSynchronize(c->stream));
    HC(hipMemcpy(dst, src, (size_t)c->ndofs * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
fail:
    return -1;
}

int adhk_sw2_set_state(adhk_sw2 *c, const double *sol, const double *sol_old, const double *sol_older)
{
    if (!c) return -1;
    for (int i = 0; i < c->ninst; i++) {
        c->alive[i] = 1;
        c->steps_done[i] = 0;
    }
    if (copy_in(c, c->sol, sol) || copy_in(c, c->sol_old, sol_old) || copy_in(c, c->sol_older, sol_older)) return -1;
    return 0;
}
int adhk_sw2_get_state(adhk_sw2 *c, double *sol, double *sol_old, double *sol_older)
{
    if (!c) return -1;
    if (copy_out(c, sol, c->sol) || copy_out(c, sol_old, c->sol_old) || copy_out(c, sol_older, c->sol_older))
        return -1;
    return 0;
}
int adhk_sw2_set_bc(adhk_sw2 *c, const int *bc_mask, const double *dirichlet)
{
    if (!c) return -1;
    if (bc_mask) HC(hipMemcpy(c->bc_mask, bc_mask, (size_t)c->ndofs * sizeof(int), hipMemcpyHostToDevice));
    if (copy_in(c, c->dirichlet, dirichlet)) return -1;
    return 0;
fail:
    return -1;
}
int adhk_sw2_set_seg_values(adhk_sw2 *c, const double *seg_val)
{
    if (!c || (c->S > 0 && !seg_val)) {
        adhk_set_error("adhk_sw2_set_seg_values: bad arguments");
        return -1;
    }
    if (c->S == 0) return 0;
    HC(hipMemcpyAsync(c->seg_val, seg_val, (size_t)c->S * 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    if (c->h_seg_val) memcpy(c->h_seg_val, seg_val, (size_t)c->S * 2 * sizeof(double));
    return 0;
fail:
    return -1;
}
int adhk_sw2_get_bc(adhk_sw2 *c, int *bc_mask)
{
    if (!c || !bc_mask) return -1;
    HC(hipStreamSynchronize(c->stream));
    HC(hipMemcpy(bc_mask, c->bc_mask, (size_t)c->ndofs * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
fail:
    return -1;
}
int adhk_sw2_set_dt(adhk_sw2 *c, double dt)
{
    if (!c || !(dt > 0)) return -1;
    c->p.dt = dt;
    return 0;
}
double adhk_sw2_get_dt(const adhk_sw2 *c) { return c ? c->p.dt : -1.0; }
int adhk_sw2_get_timing(const adhk_sw2 *c, adhk_timing *t)
{
    if (!c || !t) return -1;
    *t = c->timing;
    return 0;
}

/* accumulate the oldest pair of the ring (waiting for it if it has not run yet) and drop it */
static int kt_pop(adhk_ctx *c)
{
    adhk_ktimer *k = &c->kt;
    const int i = k->head;
    float ms = 0.f;
    k->head = (k->head + 1) % ADHK_KT_RING;
    k->n--;
    HC(hipEventSynchronize(k->b[i]));
    HC(hipEventElapsedTime(&ms, k->a[i], k->b[i]));
    k->ms[k->kind[i]] += ms;
    k->cnt[k->kind[i]]++;
    return 0;
fail:
    return -1;
}

static int kt_flush(adhk_ctx *c)
{
    int rc = 0;
    while (c->kt.n > 0)
        if (kt_pop(c)) rc = -1;
    return rc;
}

void adhk_kt_begin(adhk_ctx *c, int kind)
{
    adhk_ktimer *k = &c->kt;
    if (!k->on || !((k->mask >> kind) & 1u)) return;
    if (k->seen[kind]++ % k->stride) return;
    if (k->n == ADHK_KT_RING) (void)kt_pop(c); /* the oldest pair finished long ago */
    const int i = (k->head + k->n) % ADHK_KT_RING;
    k->kind[i] = kind;
    (void)hipEventRecord(k->a[i], c->stream);
    k->pending = 1;
}

void adhk_kt_end(adhk_ctx *c)
{
    adhk_ktimer *k = &c->kt;
    if (!k->pending) return;
    (void)hipEventRecord(k->b[(k->head + k->n) % ADHK_KT_RING], c->stream);
    k->n++;
    k->pending = 0;
}

int adhk_sw2_kernel_timing_mask(adhk_sw2 *c, unsigned mask, int stride)
{
    if (!c || stride < 1) return -1;
    if (mask && !c->kt.a[0])
        for (int i = 0; i < ADHK_KT_RING; i++) {
            HC(hipEventCreate(&c->kt.a[i]));
            HC(hipEventCreate(&c->kt.b[i]));
        }
    if (kt_flush(c)) return -1;
    if (mask) {
        memset(c->kt.ms, 0, sizeof c->kt.ms);
        memset(c->kt.cnt, 0, sizeof c->kt.cnt);
        memset(c->kt.seen, 0, sizeof c->kt.seen);
    }
    c->kt.on = mask ? 1 : 0;
    c->kt.mask = mask;
    c->kt.stride = stride;
    c->kt.pending = 0;
    return 0;
fail:
    return -1;
}

int adhk_sw2_kernel_timing(adhk_sw2 *c, int on)
{
    return adhk_sw2_kernel_timing_mask(c, on ? (1u << ADHK_KT_NKINDS) - 1u : 0u, 1);
}

int adhk_sw2_kernel_times(adhk_sw2 *c, double *ms, long long *count)
{
    if (!c) return -1;
    if (kt_flush(c)) return -1;
    for (int i = 0; i < ADHK_KT_NKINDS; i++) {
        if (ms) ms[i] = c->kt.ms[i];
        if (count) count[i] = c->kt.cnt[i];
    }
    return 0;
}

int adhk_sw2_geometry(adhk_sw2 *c, double *djac, double *grad, double *sdjac, double *snrml)
{
    double *tmp = NULL;
    const int E = c->E;
    HC(hipStreamSynchronize(c->stream));
    if (djac) HC(hipMemcpy(djac, c->djac, (size_t)E * sizeof(double), hipMemcpyDeviceToHost));
    if (grad) { /* device SoA (gx0,gx1,gx2,gy0,gy1,gy2) -> per element (g0x,g0y,g1x,g1y,g2x,g2y) */
        tmp = malloc((size_t)E * 6 * sizeof(double));
        HC(hipMemcpy(tmp, c->grad, (size_t)E * 6 * sizeof(double), hipMemcpyDeviceToHost));
        for (int e = 0; e < E; e++)
            for (int a = 0; a < 3; a++) {
                grad[6 * (size_t)e + 2 * a] = tmp[(size_t)a * E + e];
                grad[6 * (size_t)e + 2 * a + 1] = tmp[(size_t)(3 + a) * E + e];
            }
    }
    if (sdjac && c->S) HC(hipMemcpy(sdjac, c->sdjac, (size_t)c->S * sizeof(double), hipMemcpyDeviceToHost));
    if (snrml && c->S) HC(hipMemcpy(snrml, c->snrml, (size_t)c->S * 2 * sizeof(double), hipMemcpyDeviceToHost));
    free(tmp);
    return 0;
fail:
    free(tmp);
    return -1;
}

/* host flags: [0] unqueued partially dry element (internal guard), [1] NaN in the residual */
static int read_flags(adhk_ctx *c, int *flags)
{
    HC(hipMemcpyAsync(flags, c->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return 0;
fail:
    return -1;
}

static int check_partial(adhk_ctx *c)
{
    int f[4];
    if (read_flags(c, f)) return -1;
    if (f[0]) {
        adhk_set_error("a partially dry element reached the all-wet/all-dry kernel without being queued "
                       "for the wet/dry kernel (internal error)");
        return -1;
    }
    return 0;
}

int adhk_sw2_init(adhk_sw2 *c, int *wd_flag_out)
{
    if (!c) return -1;
    LC(adhk_launch_wdflag(c));
    LC(adhk_launch_dry_bc(c));
    if (wd_flag_out) {
        HC(hipStreamSynchronize(c->stream));
        HC(hipMemcpy(wd_flag_out, c->wd_flag, (size_t)c->E * sizeof(int), hipMemcpyDeviceToHost));
    }
    return 0;
fail:
    return -1;
}

static int residual_dev(adhk_ctx *c)
{
    LC(adhk_launch_elem_resid(c));
    LC(adhk_launch_resid_gather(c));
    c->timing.residual_calls++;
    return 0;
fail:
    return -1;
}

int adhk_assemble_residual(adhk_sw2 *c, double *residual)
{
    if (!c) return -1;
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(residual_dev(c));
    if (check_partial(c)) return -1;
    return copy_out(c, residual, c->residual);
fail:
    return -1;
}

int adhk_elem_residuals(adhk_sw2 *c, double *elem_rhs)
{
    size_t nr;
    double *tmp = NULL;
    int *rpos = NULL;
    if (!c) return -1;
    nr = 3 * (size_t)c->E + 2 * (size_t)c->S;
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(adhk_launch_elem_resid(c));
    if (check_partial(c)) return -1;
    if (!elem_rhs) return 0;
    /* the device holds node runs (slot rpos[element node]); the caller gets element order */
    tmp = malloc(3 * nr * sizeof(double));
    rpos = malloc(nr * sizeof(int));
    if (!tmp || !rpos) {
        adhk_set_error("adhk_elem_residuals: out of host memory");
        goto fail;
    }
    HC(hipMemcpy(tmp, c->erhs, 3 * nr * sizeof(double), hipMemcpyDeviceToHost));
    HC(hipMemcpy(rpos, c->rpos, nr * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t t = 0; t < nr; t++)
        for (int k = 0; k < 3; k++) elem_rhs[3 * t + k] = tmp[3 * (size_t)rpos[t] + k];
    free(tmp);
    free(rpos);
    return 0;
fail:
    free(tmp);
    free(rpos);
    return -1;
}

int adhk_sparsity(adhk_sw2 *c, int *indptr, int *cols)
{
    int *di = NULL, *dc = NULL;
    if (!c) return -1;
    DALLOC(di, c->ndofs + 1, int);
    DALLOC(dc, 9L * c->nnzb, int);
    LC(adhk_launch_csr_pattern(c, di, dc));
    HC(hipStreamSynchronize(c->stream));
    if (indptr) HC(hipMemcpy(indptr, di, ((size_t)c->ndofs + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (cols) HC(hipMemcpy(cols, dc, 9 * (size_t)c->nnzb * sizeof(int), hipMemcpyDeviceToHost));
    hipFree(di);
    hipFree(dc);
    return 0;
fail:
    if (di) hipFree(di);
    if (dc) hipFree(dc);
    return -1;
}

static int export_vals(adhk_ctx *c, double *vals)
{
    double *dv = NULL;
    if (!vals) return 0;
    DALLOC(dv, 9L * c->nnzb, double);
    LC(adhk_launch_csr_vals(c, dv));
    HC(hipStreamSynchronize(c->stream));
    HC(hipMemcpy(vals, dv, 9 * (size_t)c->nnzb * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(dv);
    return 0;
fail:
    if (dv) hipFree(dv);
    return -1;
}

/* the FD Jacobian assembled into the FP64 matrix, or (mixed) into the FP32 one through FP32 elemental
 * matrices */
static int jacobian_dev(adhk_ctx *c, int mixed)
{
    if (mixed && !c->emat32) {
        HC(hipMalloc((void **)&c->emat32, (81 * (size_t)c->E + 36 * (size_t)c->S + 1) * sizeof(float)));
    }
    LC(adhk_launch_elem_jac_p(c, mixed));
    LC(adhk_launch_jac_gather_p(c, mixed));
    c->timing.jacobian_calls++;
    return 0;
fail:
    return -1;
}

int adhk_assemble_jacobian(adhk_sw2 *c, double *vals)
{
    if (!c) return -1;
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(jacobian_dev(c, 0));
    if (check_partial(c)) return -1;
    return export_vals(c, vals);
fail:
    return -1;
}

/* device layout: the 3x3 block (a, b) of triangle e (code 9e + 3a + b; segment s: 9E + 4s + 2a + b) at
 * 9 cpos[code], entry (r, c) at + 3r + c  ->  row-major [elem][row][col] */
static void emat_to_rowmajor(const adhk_ctx *c, const int *cpos, const double *tmp, double *out)
{
    for (size_t e = 0; e < (size_t)c->E; e++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const double *blk = tmp + 9 * (size_t)cpos[9 * e + 3 * a + b];
                for (int r = 0; r < 3; r++)
                    for (int cc = 0; cc < 3; cc++) out[81 * e + 9 * (3 * a + r) + 3 * b + cc] = blk[3 * r + cc];
            }
    size_t off = 81 * (size_t)c->E;
    for (size_t s = 0; s < (size_t)c->S; s++)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const double *blk = tmp + 9 * (size_t)cpos[9 * (size_t)c->E + 4 * s + 2 * a + b];
                for (int r = 0; r < 3; r++)
                    for (int cc = 0; cc < 3; cc++) out[off + 36 * s + 6 * (3 * a + r) + 3 * b + cc] = blk[3 * r + cc];
            }
}

int adhk_elem_jacobians(adhk_sw2 *c, double *out)
{
    double *tmp = NULL;
    int *cpos = NULL;
    if (!c) return -1;
    size_t n = (size_t)c->E * 81 + (size_t)c->S * 36;
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(adhk_launch_elem_jac(c));
    if (check_partial(c)) return -1;
    if (!out) return 0;
    tmp = malloc(n * sizeof(double));
    cpos = malloc((size_t)c->npairs * sizeof(int));
    if (!tmp || !cpos) {
        adhk_set_error("adhk_elem_jacobians: out of host memory");
        goto fail;
    }
    HC(hipMemcpy(tmp, c->emat, n * sizeof(double), hipMemcpyDeviceToHost));
    HC(hipMemcpy(cpos, c->cpos, (size_t)c->npairs * sizeof(int), hipMemcpyDeviceToHost));
    emat_to_rowmajor(c, cpos, tmp, out);
    free(tmp);
    free(cpos);
    return 0;
fail:
    free(tmp);
    free(cpos);
    return -1;
}

int adhk_elem_jacobians_mixed(adhk_sw2 *c, double *out)
{
    float *t32 = NULL;
    double *tmp = NULL;
    int *cpos = NULL;
    if (!c) return -1;
    size_t n = (size_t)c->E * 81 + (size_t)c->S * 36;
    if (!c->emat32) HC(hipMalloc((void **)&c->emat32, (n + 1) * sizeof(float)));
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(adhk_launch_elem_jac_p(c, 1));
    if (check_partial(c)) return -1;
    if (!out) return 0;
    t32 = malloc(n * sizeof(float));
    tmp = malloc(n * sizeof(double));
    cpos = malloc((size_t)c->npairs * sizeof(int));
    if (!t32 || !tmp || !cpos) {
        adhk_set_error("adhk_elem_jacobians_mixed: out of host memory");
        goto fail;
    }
    HC(hipMemcpy(t32, c->emat32, n * sizeof(float), hipMemcpyDeviceToHost));
    HC(hipMemcpy(cpos, c->cpos, (size_t)c->npairs * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) tmp[i] = t32[i];
    emat_to_rowmajor(c, cpos, tmp, out);
    free(t32);
    free(tmp);
    free(cpos);
    return 0;
fail:
    free(t32);
    free(tmp);
    free(cpos);
    return -1;
}

int adhk_check_diag(adhk_sw2 *c)
{
    if (!c) return -1;
    return adhk_launch_check_diag(c);
}
int adhk_apply_dirichlet_bc(adhk_sw2 *c)
{
    if (!c) return -1;
    return adhk_launch_dirichlet(c, 0);
}
int adhk_get_system(adhk_sw2 *c, double *vals, double *residual)
{
    if (!c) return -1;
    if (export_vals(c, vals)) return -1;
    return copy_out(c, residual, c->residual);
}

/* residual norms of get_residual_norms (newton_tools.c:303-373), one set per instance */
/* read back the per-instance scalars (norms) after the norm kernels, and the wet/dry guard flag */
static int norms_read(adhk_ctx *c, adhk_scalars *hs)
{
    int f[4];
    HC(hipMemcpyAsync(hs, c->sc, (size_t)c->ninst * sizeof(adhk_scalars), hipMemcpyDeviceToHost, c->stream));
    HC(hipMemcpyAsync(f, c->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    if (f[0]) {
        adhk_set_error("a partially dry element reached the all-wet/all-dry kernel without being queued "
                       "for the wet/dry kernel (internal error)");
        return -1;
    }
    return 0;
fail:
    return -1;
}

/* upload the work mask (the instances that take part in the following launches) */
static int set_work(adhk_ctx *c, const int *mask)
{
    int same = 1;
    for (int i = 0; i < c->ninst; i++)
        if (c->h_work[i] != mask[i]) same = 0;
    if (same) return 0;
    memcpy(c->h_work, mask, (size_t)c->ninst * sizeof(int));
    HC(hipMemcpyAsync(c->work, c->h_work, (size_t)c->ninst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream)); /* h_work is reused by the next call */
    return 0;
fail:
    return -1;
}

static void drop_graphs(adhk_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (c->gchunk[i]) (void)hipGraphExecDestroy(c->gchunk[i]);
        c->gchunk[i] = NULL;
        c->gchunk_len[i] = 0;
        for (int k = 0; k < 3; k++) {
            if (c->gstage[i][k].exec) (void)hipGraphExecDestroy(c->gstage[i][k].exec);
            c->gstage[i][k].exec = NULL;
        }
    }
}

/* the schedule fields a captured launch sequence depends on */
static int same_schedule(const adhk_tuning *a, const adhk_tuning *b)
{
    return a->chunk == b->chunk && a->split == b->split && a->xcd_map == b->xcd_map && a->elem_tpb == b->elem_tpb &&
           a->stream_matrix == b->stream_matrix && a->sgs_wide == b->sgs_wide;
}
static int same_params(const adhk_params *a, const adhk_params *b)
{
    return a->g == b->g && a->dt == b->dt && a->alpha == b->alpha && a->cf_const == b->cf_const &&
           a->manning_units == b->manning_units && a->fric_mode == b->fric_mode;
}

/* capture `chunk` BiCGStab iterations (direction buffer 1 at the start; chunk is even, so also at the
 * end) as one graph; the kernels read every scalar and flag from device memory, so one capture serves
 * every replay */
static int capture_chunk(adhk_ctx *c, int mixed, int chunk)
{
    hipGraph_t g = NULL;
    int cur = 1, rc = 0;
    if (c->gchunk[mixed] && c->gchunk_len[mixed] == chunk && same_schedule(&c->gchunk_tune[mixed], &c->tune) &&
        c->gchunk_pc[mixed] == c->pc)
        return 0;
    if (c->gchunk[mixed]) (void)hipGraphExecDestroy(c->gchunk[mixed]);
    c->gchunk[mixed] = NULL;
    c->gchunk_len[mixed] = 0;
    HC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk && rc == 0; i++) {
        rc = adhk_launch_krylov_iter(c, mixed, cur);
        cur = 1 - cur;
    }
    {
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) goto fail;
        HC(e);
    }
    HC(hipGraphInstantiate(&c->gchunk[mixed], g, NULL, NULL, 0));
    HC(hipGraphDestroy(g));
    c->gchunk_len[mixed] = chunk;
    c->gchunk_tune[mixed] = c->tune;
    c->gchunk_pc[mixed] = c->pc;
    return 0;
fail:
    if (g) (void)hipGraphDestroy(g);
    return -1;
}

/* BiCGStab iterations from the state adhk_launch_krylov_init left, for the instances in the work mask:
 * the host reads the stop flags between chunks.  On return hs holds every instance's final scalars.
 * Every kernel exits at once when its instance has converged, so the launch pattern changes no result,
 * only the launches wasted after convergence and the read-backs, each of which drains the GPU:
 *   - no read-back after the start: a system already converged there runs one chunk of kernels that exit
 *     at once, and the SGS fallback (a folded diagonal block that is not the identity) is seen at the
 *     first read-back, which then restarts the solve with block-Jacobi;
 *   - the Newton iteration nidx (0, 1, ...; -1: none) of a step usually needs about as many BiCGStab
 *     iterations as the same Newton iteration of the previous step, so for one mesh alone with eager
 *     launches the first chunk runs to a little below that count and the rest goes in quarter chunks
 *     (an ensemble's members differ too much for one count to fit them). */
#define KRY_HIST 16
static int krylov_iterate(adhk_ctx *c, const adhk_solver *s, int mixed, adhk_scalars *hs, int nidx)
{
    const int K = c->ninst;
    int chunk = c->tune.chunk > 0 ? c->tune.chunk : (s->lin_check_every > 0 ? s->lin_check_every : 8), all_done = 1;
    const unsigned krylov_kt = (1u << ADHK_KT_SPMV_P) | (1u << ADHK_KT_SPMV_S) | (1u << ADHK_KT_UPDATE);
    const int use_graph = c->tune.graphs && !(c->kt.on && (c->kt.mask & krylov_kt)) && chunk % 2 == 0;
    const int hi = nidx < 0 ? -1 : (nidx < KRY_HIST ? nidx : KRY_HIST - 1);
    int first = chunk, tail = chunk;
    if (!use_graph && hi >= 0 && K == 1 && c->kry_hist[mixed][hi] > 0) {
        tail = chunk / 4 > 2 ? chunk / 4 : 2;
        first = c->kry_hist[mixed][hi] - tail;
        if (first < tail) first = tail;
    }
    for (;;) {
        /* iteration 1, then chunks; the A kernel of iteration k+1 completes iteration k and tests it */
        int cur = 1, launched = 1, restart = 0;
        LC(adhk_launch_krylov_first(c, mixed));
        while (launched < s->lin_max_it) {
            int k = launched == 1 ? first : tail;
            if (launched + k > s->lin_max_it) k = s->lin_max_it - launched;
            if (use_graph && k == chunk && cur == 1) {
                LC(capture_chunk(c, mixed, chunk));
                HC(hipGraphLaunch(c->gchunk[mixed], c->stream));
            } else {
                for (int i = 0; i < k; i++) {
                    LC(adhk_launch_krylov_iter(c, mixed, cur));
                    cur = 1 - cur;
                }
            }
            launched += k;
            HC(hipMemcpyAsync(hs, c->sc, (size_t)K * sizeof(adhk_scalars), hipMemcpyDeviceToHost, c->stream));
            HC(hipStreamSynchronize(c->stream));
            if (c->pc == ADHK_PC_SGS && hs[0].sgs_bad) {
                /* a singular diagonal block kept its fallback inverse: this solve takes block-Jacobi */
                c->pc = ADHK_PC_BJ;
                LC(adhk_launch_krylov_init(c, mixed, mixed ? s->mixed_lin_tol : s->lin_tol, s->lin_max_it));
                restart = 1;
                break;
            }
            all_done = 1;
            for (int i = 0; i < K; i++) all_done &= hs[i].done;
            if (all_done) break;
        }
        if (restart) continue;
        if (!all_done) {
            LC(adhk_launch_krylov_last(c, mixed, cur));
            HC(hipMemcpyAsync(hs, c->sc, (size_t)K * sizeof(adhk_scalars), hipMemcpyDeviceToHost, c->stream));
            HC(hipStreamSynchronize(c->stream));
        }
        break;
    }
    if (hi >= 0) {
        int mx = 0;
        for (int i = 0; i < K; i++)
            if (c->h_work[i] && hs[i].it > mx) mx = hs[i].it;
        c->kry_hist[mixed][hi] = mx;
    }
    return 0;
fail:
    return -1;
}

/* solve_linear_sys_bcgstab on the device for the instances in the work mask; its[i] = iterations */
static int krylov(adhk_ctx *c, const adhk_solver *s, int mixed, adhk_scalars *hs, int *its)
{
    LC(adhk_launch_krylov_init(c, mixed, mixed ? s->mixed_lin_tol : s->lin_tol, s->lin_max_it));
    LC(krylov_iterate(c, s, mixed, hs, -1));
    LC(adhk_launch_krylov_finish(c, mixed));
    HC(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->ninst; i++) its[i] = c->h_work[i] ? hs[i].it : 0;
    c->timing.krylov_calls++;
    return 0;
fail:
    return -1;
}

/* Newton stages: the launches between two host decisions of fe_newton.
 *   STG_RES  residual and its norms (start of a step, learned start, line search)
 *   STG_SYS  FD Jacobian, check_diag, Dirichlet/scaling/block-Jacobi fold, Krylov initial state
 *   STG_UPD  Krylov finish (dsol), Newton increment, residual and its norms
 * With graphs on (and no kernel timer recording events), each is captured once per precision and
 * replayed; every kernel reads state, flags and work masks from device memory, so a capture stays valid
 * until a by-value argument changes (model parameters, schedule, Krylov tolerance), which re-captures. */
enum { STG_RES, STG_SYS, STG_UPD };
static int stage_launch(adhk_ctx *c, const adhk_solver *s, int mixed, int which)
{
    if (which == STG_SYS) {
        LC(adhk_launch_elem_jac_p(c, mixed));
        LC(adhk_launch_jac_gather_p(c, mixed));
        LC(adhk_launch_check_diag_p(c, mixed));
        LC(adhk_launch_prep(c, mixed));
        LC(adhk_launch_krylov_init(c, mixed, mixed ? s->mixed_lin_tol : s->lin_tol, s->lin_max_it));
        return 0;
    }
    if (which == STG_UPD) {
        LC(adhk_launch_krylov_finish(c, mixed));
        LC(adhk_launch_increment(c));
    }
    LC(adhk_launch_elem_resid(c));
    LC(adhk_launch_resid_gather(c));
    LC(adhk_launch_norms(c));
    return 0;
fail:
    return -1;
}

static int run_stage(adhk_ctx *c, const adhk_solver *s, int mixed, int which)
{
    const double tol = mixed ? s->mixed_lin_tol : s->lin_tol;
    hipGraph_t g = NULL;
    if (which == STG_SYS) c->timing.jacobian_calls++;
    if (which == STG_UPD) c->timing.krylov_calls++;
    if (which != STG_SYS) c->timing.residual_calls++;
    if (!c->tune.graphs || c->kt.on) return stage_launch(c, s, mixed, which);
    if (mixed && !c->emat32) LC(ensure_mixed(c)); /* no allocation may happen inside a capture */
    {
        adhk_stage_graph *st = &c->gstage[mixed][which];
        if (!st->exec || !same_params(&st->p, &c->p) || !same_schedule(&st->tune, &c->tune) || st->pc != c->pc ||
            (which == STG_SYS && (st->tol != tol || st->max_it != s->lin_max_it))) {
            int rc;
            if (st->exec) (void)hipGraphExecDestroy(st->exec);
            st->exec = NULL;
            HC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = stage_launch(c, s, mixed, which);
            {
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) goto fail;
                HC(e);
            }
            HC(hipGraphInstantiate(&st->exec, g, NULL, NULL, 0));
            HC(hipGraphDestroy(g));
            g = NULL;
            st->p = c->p;
            st->tune = c->tune;
            st->tol = tol;
            st->max_it = s->lin_max_it;
            st->pc = c->pc;
        }
        HC(hipGraphLaunch(st->exec, c->stream));
    }
    return 0;
fail:
    if (g) (void)hipGraphDestroy(g);
    return -1;
}

int adhk_sw2_get_tuning(const adhk_sw2 *c, adhk_tuning *t)
{
    if (!c || !t) return -1;
    *t = c->tune;
    return 0;
}

int adhk_sw2_set_tuning(adhk_sw2 *c, const adhk_tuning *t)
{
    if (!c || !t) return -1;
    if (t->chunk < 0 || t->chunk % 2 || t->chunk > 100 || (t->graphs != 0 && t->graphs != 1) ||
        (t->xcd_map != 0 && t->xcd_map != 1) || (t->elem_tpb != 64 && t->elem_tpb != 128 && t->elem_tpb != 256) ||
        (t->split != 0 && t->split != 1) || t->stream_matrix < -1 || t->stream_matrix > 1 ||
        (t->sgs_wide != 0 && t->sgs_wide != 1)) {
        adhk_set_error("adhk_sw2_set_tuning: chunk even in [0, 100], graphs, xcd_map, split and sgs_wide 0/1, "
                       "elem_tpb 64/128/256, stream_matrix -1/0/1");
        return -1;
    }
    drop_graphs(c);
    c->tune = *t;
    return 0;
}

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Schedule autotuner.  Builds the Newton system of the current state (residual, FD Jacobian,
 * Dirichlet rows, scaling, preconditioner -- work arrays only: sol, sol_old and bc_mask are not
 * touched) and times, on it: the elemental Jacobian kernel at each workgroup size, then complete
 * BiCGStab solves under every (XCD map, chunk, graph) schedule.  Each schedule must reproduce the
 * default's iteration counts (the schedules do not change arithmetic); the fastest is kept. */
int adhk_sw2_autotune(adhk_sw2 *c, const adhk_solver *s, int ninner, adhk_tuning *chosen, double *ms)
{
    static const int tpbs[3] = {256, 128, 64};
    static const int chunks[3] = {8, 4, 16};
    const int K = c ? c->ninst : 0;
    adhk_tuning base, best;
    adhk_scalars *hs = NULL;
    int *its = NULL, *its_ref = NULL, *all = NULL, saved_kt;
    hipEvent_t e0 = NULL, e1 = NULL;
    double best_e = 1e300, best_k = 1e300, base_k = 0.0;
    if (!c || !s || ninner < 1) return -1;
    const int mixed = (s->precision == ADHK_PREC_MIXED);
    if (mixed && ensure_mixed(c)) return -1;
    saved_kt = c->kt.on;
    c->kt.on = 0;
    base = c->tune;
    best = base;
    hs = calloc((size_t)K, sizeof *hs);
    its = calloc((size_t)K, sizeof(int));
    its_ref = calloc((size_t)K, sizeof(int));
    all = calloc((size_t)K, sizeof(int));
    if (!hs || !its || !its_ref || !all) goto fail;
    for (int i = 0; i < K; i++) all[i] = 1;
    LC(set_work(c, all));
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    LC(residual_dev(c));
    /* 1. elemental Jacobian: workgroup size */
    for (int t = 0; t < 3; t++) {
        float f = 0.f;
        c->tune.elem_tpb = tpbs[t];
        LC(adhk_launch_elem_jac(c));
        HC(hipEventRecord(e0, c->stream));
        for (int r = 0; r < 3; r++) LC(adhk_launch_elem_jac(c));
        HC(hipEventRecord(e1, c->stream));
        HC(hipEventSynchronize(e1));
        HC(hipEventElapsedTime(&f, e0, e1));
        if (f / 3.0 < best_e * 0.98) { /* a change must win by 2% */
            best_e = f / 3.0;
            best.elem_tpb = tpbs[t];
        }
    }
    c->tune.elem_tpb = best.elem_tpb;
    /* 2. the Newton system of the current state */
    c->pc = pc_of(c, s);
    if (c->pc == ADHK_PC_SGS) LC(adhk_sgs_ensure(c, mixed));
    LC(jacobian_dev(c, mixed));
    LC(adhk_launch_prep(c, mixed));
    /* 3. BiCGStab schedules: complete solves, best of ninner */
    for (int sp = 0; sp <= 1; sp++)
    for (int xm = 1; xm >= 0; xm--)
        for (int ci = 0; ci < 3; ci++)
            for (int g = 0; g <= 1; g++) {
                double tbest = 1e300;
                c->tune.split = sp;
                c->tune.xcd_map = xm;
                c->tune.chunk = chunks[ci];
                c->tune.graphs = g;
                for (int r = 0; r <= ninner; r++) { /* the first solve is a warm-up (and the capture) */
                    double t0 = now_ms();
                    LC(krylov(c, s, mixed, hs, its));
                    double t1 = now_ms();
                    if (r > 0 && t1 - t0 < tbest) tbest = t1 - t0;
                }
                if (sp == 0 && xm == 1 && ci == 0 && g == 0) {
                    memcpy(its_ref, its, (size_t)K * sizeof(int));
                    base_k = tbest;
                } else {
                    for (int i = 0; i < K; i++)
                        if (its[i] != its_ref[i]) {
                            adhk_set_error("adhk_sw2_autotune: schedule changed an iteration count (internal error)");
                            goto fail;
                        }
                }
                if (tbest < best_k * 0.98) {
                    best_k = tbest;
                    best.split = sp;
                    best.xcd_map = xm;
                    best.chunk = chunks[ci];
                    best.graphs = g;
                }
            }
    /* 4. the chosen schedule's matrix loads: the other policy than the size rule gives, kept only when
     * it wins by 2% (the MALL may still hold part of a matrix larger than itself) */
    {
        double tbest = 1e300;
        c->tune = best;
        const int eff = adhk_stream_matrix_on(c, mixed);
        c->tune.stream_matrix = !eff;
        for (int r = 0; r <= ninner; r++) {
            double t0 = now_ms();
            LC(krylov(c, s, mixed, hs, its));
            double t1 = now_ms();
            if (r > 0 && t1 - t0 < tbest) tbest = t1 - t0;
        }
        for (int i = 0; i < K; i++)
            if (its[i] != its_ref[i]) {
                adhk_set_error("adhk_sw2_autotune: schedule changed an iteration count (internal error)");
                goto fail;
            }
        if (tbest < best_k * 0.98) {
            best_k = tbest;
            best.stream_matrix = !eff;
        } else {
            best.stream_matrix = eff;
        }
    }
    /* 5. the SGS phase kernels' register budget, the other one than the chosen schedule's, on the same 2%
     * margin */
    if (c->pc == ADHK_PC_SGS) {
        double tbest = 1e300;
        c->tune = best;
        c->tune.sgs_wide = !best.sgs_wide;
        for (int r = 0; r <= ninner; r++) {
            double t0 = now_ms();
            LC(krylov(c, s, mixed, hs, its));
            double t1 = now_ms();
            if (r > 0 && t1 - t0 < tbest) tbest = t1 - t0;
        }
        for (int i = 0; i < K; i++)
            if (its[i] != its_ref[i]) {
                adhk_set_error("adhk_sw2_autotune: schedule changed an iteration count (internal error)");
                goto fail;
            }
        if (tbest < best_k * 0.98) {
            best_k = tbest;
            best.sgs_wide = !best.sgs_wide;
        }
    }
    drop_graphs(c);
    c->tune = best;
    c->kt.on = saved_kt;
    if (chosen) *chosen = best;
    if (ms) {
        ms[0] = best_e;
        ms[1] = best_k;
        ms[2] = base_k;
        ms[3] = (double)its_ref[0];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    free(hs);
    free(its);
    free(its_ref);
    free(all);
    return 0;
fail:
    drop_graphs(c);
    c->tune = base;
    c->kt.on = saved_kt;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free(hs);
    free(its);
    free(its_ref);
    free(all);
    return -1;
}

/* fe_newton, fe_newton.c:44-926 (non-PETSc branch), for every instance of the context at once.  Each
 * instance follows its own Newton sequence -- its own norms, line search, stopping test and iteration
 * count -- exactly as when it is solved alone: launches cover all instances, and the work mask keeps
 * the updates of an instance that has stopped (or is not in a line search) from being applied.
 * Returns TRUE(1) when every instance converged, FALSE(0) otherwise. */
/* One Newton solve of every live instance from its previous level, or from the learned start where that
 * has the smaller residual (use_warm); per-instance results in inst_stats.  Returns TRUE when every live
 * instance converged, FALSE otherwise, -1 on error; failures are handled by the caller. */
static int fe_newton_pass(adhk_ctx *c, const adhk_solver *s, int use_warm)
{
    const int K = c->ninst;
    const int mixed = (s->precision == ADHK_PREC_MIXED);
    adhk_scalars *hs = NULL, *hw = NULL;
    int *act = NULL, *ls = NULL, *solv = NULL, *its = NULL, *lits = NULL, *cuts = NULL;
    double *rmax = NULL, *rl2 = NULL, *incmax = NULL, *old_rl2 = NULL;
    int all_ok = 1, newton_it = 0;
    const int pc = pc_of(c, s);
    if (s->precond != ADHK_PC_BJ && s->precond != ADHK_PC_SGS) {
        adhk_set_error("adhk_fe_newton: precond %d (ADHK_PC_BJ or ADHK_PC_SGS)", s->precond);
        return -1;
    }
    if (mixed && ensure_mixed(c)) return -1;
    if (pc == ADHK_PC_SGS && adhk_sgs_ensure(c, mixed)) return -1; /* no allocation inside a capture */
    hs = calloc((size_t)K, sizeof *hs);
    hw = calloc((size_t)K, sizeof *hw);
    act = calloc((size_t)K, sizeof(int));
    ls = calloc((size_t)K, sizeof(int));
    solv = calloc((size_t)K, sizeof(int));
    its = calloc((size_t)K, sizeof(int));
    lits = calloc((size_t)K, sizeof(int));
    cuts = calloc((size_t)K, sizeof(int));
    rmax = calloc((size_t)K, sizeof(double));
    rl2 = calloc((size_t)K, sizeof(double));
    incmax = calloc((size_t)K, sizeof(double));
    old_rl2 = calloc((size_t)K, sizeof(double));
    if (!hs || !hw || !act || !ls || !solv || !its || !lits || !cuts || !rmax || !rl2 || !incmax || !old_rl2) {
        adhk_set_error("adhk_fe_newton: out of host memory");
        goto fail;
    }
    memset(c->inst_stats, 0, (size_t)K * sizeof(adhk_newton_stats));
    for (int i = 0; i < K; i++) {
        act[i] = c->alive[i];
        solv[i] = 1;
    }
    LC(set_work(c, act));
    HC(hipMemsetAsync(c->flags, 0, 4 * sizeof(int), c->stream));
    /* initialize_system + initialize_dirichlet_bc (newton_tools.c:21-53, 130-142) */
    LC(adhk_launch_init_system(c));
    LC(run_stage(c, s, mixed, STG_RES));
    LC(norms_read(c, hs));
    if (c->warm && use_warm) {
        /* learned start: an instance keeps it only when its residual is smaller than the previous
         * level's; the previous level's residual is kept (in the idle Krylov vector r) for the fallback */
        int any_reject = 0;
        LC(adhk_launch_copy(c, c->r, c->residual));
        LC(adhk_mlp_predict_into(c->warm, c));
        LC(adhk_launch_dirichlet_sol(c));
        LC(run_stage(c, s, mixed, STG_RES));
        LC(norms_read(c, hw));
        for (int i = 0; i < K; i++) {
            const int better = s->warm_gate == 1 ? hw[i].rl2sum < hs[i].rl2sum : hw[i].rmax < hs[i].rmax;
            ls[i] = !(!hw[i].nanflag && !hs[i].nanflag && better); /* reject mask */
            if (!ls[i]) {
                hs[i] = hw[i];
                c->inst_stats[i].warm_accepted = 1;
            }
            any_reject |= ls[i];
        }
        if (any_reject) {
            LC(set_work(c, ls));
            LC(adhk_launch_start_from_old_masked(c));
            LC(adhk_launch_copy_masked(c, c->residual, c->r));
            LC(set_work(c, act));
        }
    }
    for (int i = 0; i < K; i++) {
        const double nd = 3.0 * (c->h_inst_node[i + 1] - c->h_inst_node[i]);
        rmax[i] = hs[i].rmax;
        rl2[i] = sqrt(hs[i].rl2sum) / nd;
        incmax[i] = hs[i].incmax;
        if (hs[i].nanflag) solv[i] = 0;
    }
    for (;;) {
        int any = 0;
        for (int i = 0; i < K; i++) any |= act[i];
        if (!any) break;
        LC(set_work(c, act));
        c->pc = pc;
        LC(run_stage(c, s, mixed, STG_SYS));
        LC(krylov_iterate(c, s, mixed, hw, newton_it++));
        for (int i = 0; i < K; i++) lits[i] = c->h_work[i] ? hw[i].it : 0;
        LC(run_stage(c, s, mixed, STG_UPD));
        for (int i = 0; i < K; i++)
            if (act[i]) {
                c->inst_stats[i].lin_its_total += lits[i];
                if (lits[i] > c->inst_stats[i].lin_its_max) c->inst_stats[i].lin_its_max = lits[i];
                if (lits[i] >= s->lin_max_it) c->inst_stats[i].lin_capped++;
                old_rl2[i] = rl2[i];
                cuts[i] = 0;
            }
        LC(norms_read(c, hs));
        for (int i = 0; i < K; i++)
            if (act[i]) {
                const double nd = 3.0 * (c->h_inst_node[i + 1] - c->h_inst_node[i]);
                rmax[i] = hs[i].rmax;
                rl2[i] = sqrt(hs[i].rl2sum) / nd;
                incmax[i] = hs[i].incmax;
                if (hs[i].nanflag || isnan(rl2[i]) || isinf(rl2[i])) solv[i] = 0;
            }
        /* line search (fe_newton.c:681-741), per instance */
        for (;;) {
            int anyls = 0;
            for (int i = 0; i < K; i++) {
                ls[i] = act[i] && solv[i] && (rmax[i] > s->tol_nonlin) && (rl2[i] > old_rl2[i]) &&
                        (cuts[i] < s->max_linesearch_cuts);
                anyls |= ls[i];
            }
            if (!anyls) break;
            LC(set_work(c, ls));
            LC(adhk_launch_linesearch(c));
            LC(run_stage(c, s, mixed, STG_RES));
            LC(norms_read(c, hs));
            for (int i = 0; i < K; i++)
                if (ls[i]) {
                    const double nd = 3.0 * (c->h_inst_node[i + 1] - c->h_inst_node[i]);
                    rmax[i] = hs[i].rmax;
                    rl2[i] = sqrt(hs[i].rl2sum) / nd;
                    incmax[i] = hs[i].incmax;
                    if (hs[i].nanflag) solv[i] = 0;
                    cuts[i]++;
                    c->inst_stats[i].linesearch_cuts_total++;
                }
        }
        for (int i = 0; i < K; i++)
            if (act[i]) {
                its[i]++;
                act[i] = solv[i] && (its[i] < s->max_nonlin_it) && (rmax[i] < 1.0E+20) &&
                         (rmax[i] > s->tol_nonlin && incmax[i] > s->inc_nonlin);
            }
    }
    for (int i = 0; i < K; i++) {
        adhk_newton_stats *t = &c->inst_stats[i];
        if (!c->alive[i]) continue;
        t->newton_its = its[i];
        t->resid_max_norm = rmax[i];
        t->inc_max_norm = incmax[i];
        t->converged = !(!solv[i] || ((rmax[i] > s->tol_nonlin) && (incmax[i] > s->inc_nonlin)) ||
                         (rmax[i] >= 1.0E+20));
        all_ok &= t->converged;
    }
    free(hs); free(hw); free(act); free(ls); free(solv); free(its); free(lits); free(cuts);
    free(rmax); free(rl2); free(incmax); free(old_rl2);
    return all_ok;
fail:
    free(hs); free(hw); free(act); free(ls); free(solv); free(its); free(lits); free(cuts);
    free(rmax); free(rl2); free(incmax); free(old_rl2);
    return -1;
}

/* a failed solve: one mesh cuts dt and restarts from the old level (tc_scale, tc_scale.c:3-19,
 * DT_REDUCE_FACTOR define.h:127); in an ensemble the failed members restart from their old level and
 * leave the following steps */
static int newton_failure(adhk_ctx *c)
{
    const int K = c->ninst;
    int *mask = calloc((size_t)K, sizeof(int));
    if (!mask) {
        adhk_set_error("adhk_fe_newton: out of host memory");
        return -1;
    }
    if (K == 1) {
        mask[0] = 1;
        LC(set_work(c, mask));
        c->p.dt *= 0.25;
        LC(adhk_launch_init_system(c));
    } else {
        for (int i = 0; i < K; i++) {
            mask[i] = c->alive[i] && !c->inst_stats[i].converged;
            if (mask[i]) c->alive[i] = 0;
        }
        LC(set_work(c, mask));
        LC(adhk_launch_start_from_old_masked(c));
    }
    HC(hipStreamSynchronize(c->stream));
    free(mask);
    return 0;
fail:
    free(mask);
    return -1;
}

static void stats_add(adhk_newton_stats *a, const adhk_newton_stats *b)
{
    a->newton_its += b->newton_its;
    a->linesearch_cuts_total += b->linesearch_cuts_total;
    a->lin_its_total += b->lin_its_total;
    if (b->lin_its_max > a->lin_its_max) a->lin_its_max = b->lin_its_max;
    a->resid_max_norm = b->resid_max_norm;
    a->inc_max_norm = b->inc_max_norm;
    a->substeps += b->substeps;
    a->lin_capped += b->lin_capped;
}

/* The members in mask, each at level t0 (sol = sol_old), advance over [t0, t0 + dt] in four steps of
 * dt / 4 (DT_REDUCE_FACTOR, define.h:127); a member whose quarter step fails goes back to the quarter's
 * start and retakes it in four (depth + 1), down to substep_depth.  On return mask[i] is 1 for the members
 * that reached t0 + dt and 0 for those that failed at the deepest level (left at the last sub-step they
 * completed).  Their statistics are added to acc. */
static int substeps_rec(adhk_ctx *c, const adhk_solver *s, int *mask, double t0, double dt, int depth,
                        adhk_newton_stats *acc)
{
    const int K = c->ninst;
    const double h = 0.25 * dt;
    int *sub = calloc((size_t)K, sizeof(int)), *save = calloc((size_t)K, sizeof(int)),
        *again = calloc((size_t)K, sizeof(int));
    if (!sub || !save || !again) {
        adhk_set_error("sub-steps: out of host memory");
        goto fail;
    }
    for (int k = 0; k < 4; k++) {
        int any = 0, anyf = 0;
        for (int i = 0; i < K; i++) any |= mask[i];
        if (!any) break;
        if (k > 0) { /* prep_sol of the members still stepping */
            LC(set_work(c, mask));
            LC(adhk_launch_rotate_masked(c));
        }
        LC(adhk_series_apply(c, t0 + k * h, t0 + (k + 1) * h));
        memcpy(save, c->alive, (size_t)K * sizeof(int));
        memcpy(c->alive, mask, (size_t)K * sizeof(int));
        c->p.dt = h;
        const int ok = fe_newton_pass(c, s, 0);
        memcpy(c->alive, save, (size_t)K * sizeof(int));
        if (ok < 0) goto fail;
        for (int i = 0; i < K; i++) {
            sub[i] = 0;
            if (!mask[i]) continue;
            stats_add(&acc[i], &c->inst_stats[i]);
            acc[i].substeps++;
            sub[i] = !c->inst_stats[i].converged;
            anyf |= sub[i];
        }
        if (!anyf) continue;
        LC(set_work(c, sub));
        LC(adhk_launch_start_from_old_masked(c));
        memcpy(again, sub, (size_t)K * sizeof(int));
        if (depth < s->substep_depth && substeps_rec(c, s, again, t0 + k * h, h, depth + 1, acc)) goto fail;
        for (int i = 0; i < K; i++)
            if (sub[i] && (depth >= s->substep_depth || !again[i])) mask[i] = 0;
        c->p.dt = h;
    }
    c->p.dt = dt;
    free(sub);
    free(save);
    free(again);
    return 0;
fail:
    c->p.dt = dt;
    free(sub);
    free(save);
    free(again);
    return -1;
}

/* after a pass in which some live members failed: each retakes its step in sub-steps (substep_depth > 0).
 * Returns TRUE when every live member completed the step, FALSE otherwise, -1 on error. */
static int substep_recovery(adhk_ctx *c, const adhk_solver *s)
{
    const int K = c->ninst;
    const double dt0 = c->p.dt;
    int rc = -1, all = 1;
    adhk_newton_stats *first = malloc((size_t)K * sizeof *first), *acc = calloc((size_t)K, sizeof *acc);
    int *mask = calloc((size_t)K, sizeof(int)), *failed = calloc((size_t)K, sizeof(int));
    if (!first || !acc || !mask || !failed) {
        adhk_set_error("sub-steps: out of host memory");
        goto done;
    }
    memcpy(first, c->inst_stats, (size_t)K * sizeof *first);
    for (int i = 0; i < K; i++) failed[i] = mask[i] = c->alive[i] && !first[i].converged;
    if (set_work(c, mask) || adhk_launch_start_from_old_masked(c)) goto done;
    if (substeps_rec(c, s, mask, c->t_prev, dt0, 1, acc)) goto done;
    c->p.dt = dt0;
    for (int i = 0; i < K; i++) {
        c->inst_stats[i] = first[i];
        if (!failed[i]) continue;
        stats_add(&c->inst_stats[i], &acc[i]);
        c->inst_stats[i].converged = mask[i];
        if (!mask[i]) c->alive[i] = 0;
    }
    for (int i = 0; i < K; i++)
        if (failed[i] || c->alive[i]) all &= c->inst_stats[i].converged;
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        adhk_set_error("sub-steps: device error");
        goto done;
    }
    rc = all;
done:
    c->p.dt = dt0;
    free(first);
    free(acc);
    free(mask);
    free(failed);
    return rc;
}

int adhk_fe_newton(adhk_sw2 *c, const adhk_solver *s, adhk_newton_stats *st)
{
    if (!c || !s) return -1;
    int ok = fe_newton_pass(c, s, 1);
    if (ok < 0) return -1;
    if (!ok && c->ninst == 1 && c->warm && c->inst_stats[0].warm_accepted) {
        /* plain Newton stays the fallback: a solve that fails from the learned start is run again from the
         * previous level (the pass starts with initialize_system) before dt is cut */
        const adhk_newton_stats first = c->inst_stats[0];
        ok = fe_newton_pass(c, s, 0);
        if (ok < 0) return -1;
        adhk_newton_stats *t = &c->inst_stats[0];
        t->newton_its += first.newton_its;
        t->linesearch_cuts_total += first.linesearch_cuts_total;
        t->lin_its_total += first.lin_its_total;
        t->lin_capped += first.lin_capped;
        if (first.lin_its_max > t->lin_its_max) t->lin_its_max = first.lin_its_max;
        t->warm_accepted = 0;
        t->warm_retried = 1;
    }
    if (!ok && s->substep_depth > 0) {
        ok = substep_recovery(c, s);
        if (ok < 0) return -1;
    } else if (!ok && newton_failure(c)) {
        return -1;
    }
    if (st) {
        adhk_newton_stats tot;
        memset(&tot, 0, sizeof tot);
        for (int i = 0; i < c->ninst; i++) {
            const adhk_newton_stats *t = &c->inst_stats[i];
            tot.newton_its += t->newton_its;
            tot.linesearch_cuts_total += t->linesearch_cuts_total;
            tot.lin_its_total += t->lin_its_total;
            if (t->lin_its_max > tot.lin_its_max) tot.lin_its_max = t->lin_its_max;
            if (t->resid_max_norm > tot.resid_max_norm) tot.resid_max_norm = t->resid_max_norm;
            if (t->inc_max_norm > tot.inc_max_norm) tot.inc_max_norm = t->inc_max_norm;
            tot.warm_accepted += t->warm_accepted;
            tot.warm_retried += t->warm_retried;
            tot.substeps += t->substeps;
            tot.lin_capped += t->lin_capped;
        }
        tot.converged = ok;
        *st = tot;
    }
    return ok;
}

/* time_loop / advance_time (time_loop.c:14-93): prep_sol then fe_newton per step; per-instance
 * totals accumulate in inst_stats (adhk_sw2_instance_stats) */
int adhk_time_loop(adhk_sw2 *c, const adhk_solver *s, int nsteps, adhk_newton_stats *acc)
{
    adhk_newton_stats tot, st;
    adhk_newton_stats *run = NULL;
    int ok = 1, k, forcing_failed = 0;
    memset(&tot, 0, sizeof tot);
    if (!c || !s) return -1000000000;
    run = calloc((size_t)c->ninst, sizeof *run);
    if (!run) return -1000000000;
    for (k = 0; k < nsteps && (ok == 1 || c->ninst > 1); k++) {
        int any = 0;
        for (int i = 0; i < c->ninst; i++) any |= c->alive[i];
        if (!any) break;
        /* forcing over the step (sseries_setall_ts_valuesAVG, time_loop.c:76); a step the series do
         * not cover is not taken */
        const double dt_step = c->p.dt;
        if (adhk_series_apply(c, c->t_prev, c->t_prev + dt_step)) {
            ok = 0;
            forcing_failed = 1;
            break;
        }
        if (adhk_launch_rotate(c)) {
            free(run);
            return -1000000000;
        }
        ok = adhk_fe_newton(c, s, &st);
        if (ok < 0) {
            free(run);
            return -1000000000;
        }
        if (ok == 1 || c->ninst > 1) c->t_prev += dt_step;
        for (int i = 0; i < c->ninst; i++) {
            const adhk_newton_stats *t = &c->inst_stats[i];
            run[i].newton_its += t->newton_its;
            run[i].linesearch_cuts_total += t->linesearch_cuts_total;
            run[i].lin_its_total += t->lin_its_total;
            if (t->lin_its_max > run[i].lin_its_max) run[i].lin_its_max = t->lin_its_max;
            run[i].resid_max_norm = t->resid_max_norm;
            run[i].inc_max_norm = t->inc_max_norm;
            run[i].warm_accepted += t->warm_accepted;
            run[i].warm_retried += t->warm_retried;
            run[i].substeps += t->substeps;
            run[i].lin_capped += t->lin_capped;
            run[i].converged = c->alive[i] ? t->converged : 0;
            if (c->alive[i] && t->converged) c->steps_done[i]++;
        }
        tot.newton_its += st.newton_its;
        tot.linesearch_cuts_total += st.linesearch_cuts_total;
        tot.lin_its_total += st.lin_its_total;
        if (st.lin_its_max > tot.lin_its_max) tot.lin_its_max = st.lin_its_max;
        tot.resid_max_norm = st.resid_max_norm;
        tot.inc_max_norm = st.inc_max_norm;
        tot.warm_accepted += st.warm_accepted;
        tot.warm_retried += st.warm_retried;
        tot.substeps += st.substeps;
        tot.lin_capped += st.lin_capped;
    }
    memcpy(c->inst_stats, run, (size_t)c->ninst * sizeof *run);
    free(run);
    if (c->ninst > 1 && !forcing_failed) { /* an ensemble completes when every member has */
        ok = 1;
        for (int i = 0; i < c->ninst; i++) ok &= c->alive[i];
        if (!ok) {
            int mn = nsteps;
            for (int i = 0; i < c->ninst; i++)
                if (c->steps_done[i] < mn) mn = c->steps_done[i];
            k = mn;
        }
    }
    tot.converged = ok;
    if (acc) *acc = tot;
    return ok == 1 ? k : -k;
}

int adhk_sw2_instance_status(const adhk_sw2 *c, int *alive, int *steps_done)
{
    if (!c) return -1;
    if (alive) memcpy(alive, c->alive, (size_t)c->ninst * sizeof(int));
    if (steps_done) memcpy(steps_done, c->steps_done, (size_t)c->ninst * sizeof(int));
    return 0;
}

int adhk_sw2_instances(const adhk_sw2 *c) { return c ? c->ninst : -1; }

int adhk_sw2_instance_stats(const adhk_sw2 *c, adhk_newton_stats *out)
{
    if (!c || !out) return -1;
    memcpy(out, c->inst_stats, (size_t)c->ninst * sizeof *out);
    return 0;
}
