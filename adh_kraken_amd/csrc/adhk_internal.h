/*
 * adhk_internal.h -- device context shared by the C host driver (newton.c)
 * and the gfx950 kernels (*.hip).  Plain C so both sides see one layout.
 *
 * HBM layout (N nodes, E triangles, S boundary segments):
 *   node coordinates      x, y, z                   SoA double [N]
 *   triangles             tri                       SoA int [3][E]
 *   geometry              djac [E], grad [6][E]     (gx0,gx1,gx2,gy0,gy1,gy2)
 *   state                 sol, sol_old, sol_older   AoS double [3N] (h,u,v per node)
 *   pattern               block CSR of 3x3 node blocks: brow [N+1], bcol [nnzb], diag_blk [N]
 *   matrix values         sliced ELL of 3x3 blocks (SELL-64): node n sits in slice n/64 at lane
 *                         n%64; slice s owns slot columns slice_ptr[s] .. slice_ptr[s+1]-1 (its
 *                         widest block row); block (slot column c, lane l) has column node
 *                         scol[c*64 + l] and entry e (row-major 3x3) at bvals[(c*9 + e)*64 + l],
 *                         so a wavefront reads 64 consecutive values per load.  Padding slots
 *                         hold zeros and their own node as column.  sell2blk / blk2sell map
 *                         slots to block-CSR indices; diag_slot [N] is the diagonal's slot.
 *   block-Jacobi          minv[(s*9 + e)*64 + l]
 *   assembly lists        cptr [nnzb+1], contrib [npairs]: contributions to each
 *                         block in element order (2-D first, then 1-D); cpos [npairs]: the list
 *                         position of each elemental block (code 9e + 3a + b, 9E + 4s + 2a + b)
 *   residual lists        rptr [N+1], rpos [3E + 2S]: element node (e, a) -> 3e + a, segment node
 *                         (s, a) -> 3E + 2s + a; its residual slot rpos[...] lies in node n's run
 *                         rptr[n] .. rptr[n+1]-1, which lists n's elements in element order
 *   elemental buffers     erhs [3 (3E + 2S)]: the three residual entries of element node slot j at
 *                         3j, so a node's contributions are one contiguous run; emat [npairs*9]:
 *                         the 3x3 block of list position j at 9j (entry 3r + c), likewise
 *   wet/dry queue         wdq [E+1]: count, then the partially dry elements of the last launch
 *   instances             node_inst [N], elem_inst [E], seg_inst [S]; 256-node blocks blk_node/blk_inst;
 *                         per-instance scalars sc [ninst] and work mask [ninst]
 */
#ifndef ADHK_INTERNAL_H
#define ADHK_INTERNAL_H

#include <hip/hip_runtime_api.h>

#include "../../include/adhk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* device-resident scalars of the Krylov solver and the Newton norms */
typedef struct {
    double rho, rhop, alpha, omega, gamma, beta;
    double rnorm, bnorm, tol;
    int it, max_it, done, pad0;
    double rmax, rl2sum, incmax;
    int nanflag, sgs_bad; /* sgs_bad: a folded diagonal block is not the identity (SGS falls back) */
} adhk_scalars;

/* Multicolour SGS layout (one mesh alone): rows permuted by colour (within a colour, sorted by their U and L
 * widths inside windows of 1024 rows), each colour's run padded to a multiple of 256 rows, so a 256-row reduction block and a 64-row slice hold one
 * colour; row Np is a zero row that padding slots point at.  The folded matrix's off-diagonal blocks
 * split into U (neighbour of a higher colour) and L (lower), each its own SELL-64 over the permuted rows
 * (blocks of a row in ascending natural column order); its padding slots hold zero blocks, written once. */
#define ADHK_SGS_MAXC 64
typedef struct {
    int *ptr, *col; /* device [nslice + 1], [nslot * 64] */
    int nslot;
    double *val;
    float *val32;
} adhk_sgs_part;
typedef struct {
    int ready, nc, Np, nblk;
    int cs[ADHK_SGS_MAXC + 1]; /* host: permuted row range of colour c is [cs[c], cs[c+1]) */
    int *colour;               /* host [N] */
    int *prow;                 /* device [Np]: natural node of a permuted row, -1 for padding */
    adhk_sgs_part U, L;
    int *dst; /* device [natural nslot * 64]: where a natural SELL slot's block goes, q >= 0 in U, -(q + 2) in
               * L, -1 nowhere (the diagonal blocks and the natural layout's padding) */
    int *inst_blk;             /* device {0, nblk} */
    double *partials;          /* device [8 * nblk] */
    double *v64;               /* device 8 vectors of 3 (Np + 1) doubles: r (= s) p Ap As q xt t (= u) bt */
    float *v32;                /* the same in FP32 */
} adhk_sgs;

/* contribution encoding: tri: (e*3 + a)*3 + b  (a,b local nodes, row a, col b)
 *                        seg: CONTRIB_SEG_BASE + (s*2 + a)*2 + b                */
#define ADHK_NPARTIAL_MAX 8192

/* per-kernel device timers: hipEvent pairs around launches on the context stream, kept in a ring
 * that is drained from its oldest end, so that reading a pair back never waits on recent work */
#define ADHK_KT_RING 4096
typedef struct {
    int on, head, n, pending, stride;
    unsigned mask;                    /* classes timed: bit k for class k */
    long long seen[ADHK_KT_NKINDS];   /* launches per class since enabled (1 in stride is timed) */
    int kind[ADHK_KT_RING];
    hipEvent_t a[ADHK_KT_RING], b[ADHK_KT_RING];
    double ms[ADHK_KT_NKINDS];
    long long cnt[ADHK_KT_NKINDS];
} adhk_ktimer;

/* a captured Newton stage and the by-value kernel arguments it was captured with */
typedef struct {
    hipGraphExec_t exec;
    adhk_params p;
    adhk_tuning tune;
    double tol;
    int max_it;
    int pc; /* preconditioner (ADHK_PC_*) */
} adhk_stage_graph;

struct adhk_sw2 {
    int N, E, S, ndofs, nnzb, npairs;
    adhk_params p;
    hipStream_t stream;
    /* mesh */
    double *x, *y, *z;
    int *tri;
    double *djac, *grad, *ks, *mann;
    float *nstat; /* per-node bed slope x, y and Manning n [3][N], built on first use by the learned start */
    adhk_mlp *warm; /* learned warm start, or NULL */
    int *seg, *seg_kind;
    double *seg_val, *sdjac, *snrml;
    /* state */
    double *sol, *sol_old, *sol_older, *dirichlet, *dsol, *residual, *scale;
    int *bc_mask, *wd_flag, *wd_tmp;
    int *wdq; /* partially dry elements of the last elemental launch: [0] count, [1..] elements */
    /* block CSR pattern and assembly lists */
    int *brow, *bcol, *diag_blk, *cptr, *contrib;
    /* SELL-64 layout of the matrix values */
    int nslice, nslot;
    int *slice_ptr, *scol, *sell2blk, *blk2sell, *diag_slot; /* a row's diagonal block is its slot 0 */
    unsigned char *diag_id; /* [N]: 1 = the folded diagonal block of M^-1 A is the identity */
    /* matrix values and preconditioner */
    double *bvals, *minv;
    float *bvals32, *minv32;
    /* elemental buffers */
    double *erhs, *emat;
    int *cpos; /* [npairs]: contribution-list position of each elemental block */
    int *rptr, *rpos; /* node runs of the elemental residuals: [N+1], [3E + 2S] */
    float *emat32; /* FP32 elemental Jacobians of the mixed-precision Newton iteration */
    /* Krylov vectors (FP64 and FP32 sets); p and Ap are double-buffered: one BiCGStab kernel
     * reads the previous direction while it writes the next */
    double *r, *pv, *Ap, *s, *As, *q, *x0, *pv2, *Ap2, *xk;
    float *r32, *p32, *Ap32, *s32, *As32, *q32, *xk32, *p32b, *Ap32b;
    /* reductions */
    double *partials;
    unsigned int *counter;
    adhk_scalars *sc; /* [ninst]: BiCGStab scalars and Newton norms of each instance */
    /* preconditioner of the current solve (ADHK_PC_*) and the multicolour SGS layout */
    int pc;
    adhk_sgs sgs;
    /* instances: an ensemble of independent meshes solved as one block-diagonal system (1 = one mesh) */
    int ninst;
    int *h_inst_node, *h_inst_elem, *h_inst_seg, *h_inst_blk; /* host [ninst + 1] offsets */
    int *node_inst, *elem_inst, *seg_inst;                     /* device: instance of each node / element */
    /* blocks of up to 256 consecutive nodes of one instance: the unit of every reduction */
    int nblk;
    int *blk_node, *blk_inst, *inst_blk; /* device [nblk + 1], [nblk], [ninst + 1] */
    int *work, *h_work;                  /* [ninst]: 1 = the instance takes part in the next launches */
    adhk_newton_stats *inst_stats;       /* host [ninst]: per-instance statistics of the last solve */
    int *alive, *steps_done;             /* host [ninst]: ensemble members still stepping, steps completed */
    /* schedule choices (adhk_sw2_autotune); none of them changes a result */
    adhk_tuning tune;
    hipGraphExec_t gchunk[2]; /* captured BiCGStab chunk (FP64, mixed) and its length */
    int gchunk_len[2];
    adhk_tuning gchunk_tune[2]; /* the schedule the chunk was captured with */
    int gchunk_pc[2];           /* and the preconditioner */
    int kry_hist[2][16];        /* BiCGStab iterations of the last step's Newton iteration k (per precision) */
    /* captured Newton stages (per precision): the launches between two host decisions, with the
     * by-value kernel arguments they were captured with */
    adhk_stage_graph gstage[2][3];
    int *flags; /* [0] = unqueued partially dry element (internal guard), [1] = NaN seen */
    /* time-series boundary values (series.c): tables, segment bindings, time of the old level */
    double t_prev;
    int nseries;
    struct adhk_series_tab *series;
    int *h_seg_ser;                  /* host [2S]: series driving each segment value, -1 = none */
    double *h_seg_scale, *h_seg_val; /* host [2S], [2S] */
    /* timing */
    adhk_timing timing;
    hipEvent_t ev[8];
    adhk_ktimer kt;
};
typedef struct adhk_sw2 adhk_ctx;
typedef struct adhk_series_tab adhk_series_tab;

/* ---- time-series boundary values (series.c) ---- */
int adhk_series_apply(adhk_ctx *c, double t1, double t2);
void adhk_series_free(adhk_ctx *c);

/* ---- launchers (sw2_kernels.hip) ---- */
int adhk_launch_wdflag(adhk_ctx *c);
int adhk_launch_dry_bc(adhk_ctx *c);
int adhk_launch_elem_resid(adhk_ctx *c);
int adhk_launch_resid_gather(adhk_ctx *c);
int adhk_launch_elem_jac(adhk_ctx *c);
int adhk_launch_jac_gather(adhk_ctx *c);
int adhk_launch_elem_jac_p(adhk_ctx *c, int f32);   /* f32: FP32 elemental matrices (emat32) */
int adhk_launch_jac_gather_p(adhk_ctx *c, int f32); /* f32: emat32 -> the FP32 matrix bvals32 */
int adhk_launch_check_diag_p(adhk_ctx *c, int f32);

/* ---- launchers (pattern.hip) ---- */
int adhk_build_pattern(adhk_ctx *c);
int adhk_launch_csr_pattern(adhk_ctx *c, int *d_indptr, int *d_cols);
int adhk_launch_csr_vals(adhk_ctx *c, double *d_vals);

/* ---- launchers (la_kernels.hip) ---- */
int adhk_launch_norms(adhk_ctx *c);
int adhk_launch_check_diag(adhk_ctx *c);
int adhk_launch_dirichlet(adhk_ctx *c, int do_scale);
int adhk_launch_scale_prec(adhk_ctx *c, int mixed);
int adhk_launch_prep(adhk_ctx *c, int mixed); /* Dirichlet + scaling + fold in two matrix passes */
int adhk_stream_matrix_on(adhk_ctx *c, int mixed); /* the products' matrix loads are non-temporal */
int adhk_launch_krylov_init(adhk_ctx *c, int mixed, double tol, int max_it);
int adhk_launch_krylov_first(adhk_ctx *c, int mixed);
int adhk_launch_krylov_iter(adhk_ctx *c, int mixed, int cur);
int adhk_launch_krylov_last(adhk_ctx *c, int mixed, int cur);
int adhk_launch_krylov_finish(adhk_ctx *c, int mixed);
int adhk_sgs_setup(adhk_ctx *c);              /* colouring and the permuted L/U layout (sgs_setup.c) */
int adhk_sgs_ensure(adhk_ctx *c, int mixed);  /* the value arrays and vectors of a precision */
void adhk_sgs_free(adhk_ctx *c);
int adhk_launch_sgs_copy(adhk_ctx *c, int mixed); /* folded matrix -> L/U values, diagonal check */
int adhk_launch_increment(adhk_ctx *c);
int adhk_launch_linesearch(adhk_ctx *c);
int adhk_launch_init_system(adhk_ctx *c);
int adhk_launch_rotate(adhk_ctx *c);
int adhk_launch_rotate_masked(adhk_ctx *c);
int adhk_launch_start_from_old(adhk_ctx *c);
int adhk_launch_copy(adhk_ctx *c, double *dst, const double *src);
int adhk_launch_copy_masked(adhk_ctx *c, double *dst, const double *src);
int adhk_launch_start_from_old_masked(adhk_ctx *c);
int adhk_launch_dirichlet_sol(adhk_ctx *c);

/* ---- vector helpers of the struct-level surface (la_csr.hip): y += x on host arrays through the GPU;
 *      get_residual_norms' maxima (first dof on ties), l2 sum and NaN count ---- */
int adhk_la_vec_add(double *y, const double *x, int n);
int adhk_la_residual_norms(int n, const double *residual, const double *dsol, const int *bc_mask,
                           const int *include_dof, double *rmax, double *l2sum, double *dmax, int *irmax, int *idmax,
                           int *nnan);

/* ---- learned warm start (mlp.hip) ---- */
int adhk_mlp_predict_into(adhk_mlp *m, adhk_ctx *c);

/* error reporting and kernel timers (newton.c) */
void adhk_set_error(const char *fmt, ...);
void adhk_kt_begin(adhk_ctx *c, int kind);
void adhk_kt_end(adhk_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
