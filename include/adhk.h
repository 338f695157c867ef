/*
 * adhk.h -- C-ABI of libadhk.so, the MI355X (gfx950) implementation of the
 * AdH-Kraken SW2 Newton step.  Plain pointers and sizes only.
 *
 * Two groups of entry points:
 *
 *  1. adhk_* : the reference's struct-based call surface (SMODEL_SUPER*,
 *     SGRID*) restated over an opaque context that owns the mesh and the
 *     device-resident state.  Each function names the reference function it
 *     replaces.  A shim compiled inside the reference tree maps the reference
 *     structs onto these calls.
 *
 *  2. The reference's plain-pointer linear-algebra surface (src/la/la.h) with
 *     the reference's exact names and signatures, operating on host CSR arrays.
 *
 * Unknown layout everywhere: dof(node, var) = 3*node + var, var 0 = h (depth),
 * 1 = u, 2 = v (src/structs/sivar_position/sivar_position.c:61-80).
 *
 * Error convention: functions returning int return >= 0 on success and -1 on
 * a device or argument error; adhk_last_error() then describes it.  The
 * Newton/time-loop entries return the reference's TRUE(1)/FALSE(0).
 */
#ifndef ADHK_H
#define ADHK_H

#ifdef __cplusplus
extern "C" {
#endif

/* 1-D boundary element kinds */
#define ADHK_BC_DEPTH 0 /* tailwater depth, fe_sw2_bc_h (src/models/sw2/sw2_bc_h.c:57); val0 = depth            */
#define ADHK_BC_FLUX  1 /* explicit flux, fe_sw2_1d_explicit_flow (src/fe/fe_sw2_boundary_integrals.c:56);
                           val0 = normal flux per unit width, negative = into the domain                         */
#define ADHK_BC_VEL   2 /* weak velocity, fe_sw2_bc_vel (src/models/sw2/sw2_bc_vel.c:58); val0, val1 = u, v    */
#define ADHK_BC_DEPTH_NODAL 3 /* fe_sw2_bc_h with a depth per node: val0 at the segment's first node, val1 at
                                 its second (a water-surface elevation minus the bed: exact for a lake at rest
                                 on a bed that varies along the boundary, where one depth per segment is not) */

/* friction law */
#define ADHK_FRIC_CONST   0 /* constant drag coefficient (sw2_body_resid.c:241 hard-codes 0.002) */
#define ADHK_FRIC_MANNING 1 /* Manning n -> roughness height -> log-law drag coefficient
                               (fr_manningsn_to_rheight.c:15, fr_bedshstr_drag_coef.c:18)       */

/* precision of the linear system (Jacobian storage, SpMV, BiCGStab) */
#define ADHK_PREC_FP64  0
#define ADHK_PREC_MIXED 1 /* FP32 matrix/preconditioner/Krylov vectors; FP64 residual, update, test */

typedef struct adhk_sw2 adhk_sw2;

typedef struct {
    double g;             /* gravity (smodel_super_alloc_init.c:92: 9.8) */
    double dt;            /* time step (SDT.dt) */
    double alpha;         /* tau_temporal (sdt.c:24: 0 = backward Euler) */
    double cf_const;      /* drag coefficient for ADHK_FRIC_CONST */
    double manning_units; /* Manning's units constant, 1.0 = SI (sstr_value.c:288) */
    int fric_mode;        /* ADHK_FRIC_* */
} adhk_params;

typedef struct {
    double tol_nonlin;       /* smodel_super_alloc_init.c:63 (1e-8) */
    double inc_nonlin;       /* :62 (1e-6) */
    int max_nonlin_it;       /* :65 (10) */
    int max_linesearch_cuts; /* :64 (5) */
    double lin_tol;          /* bcgstab_solver.c:20-21 (1e-11, absolute and relative) */
    int lin_max_it;          /* bcgstab_solver.c:22 (100) */
    int precision;           /* ADHK_PREC_* */
    int lin_check_every;     /* BiCGStab iterations launched between convergence read-backs (>=1) */
    double mixed_lin_tol;    /* inner (FP32) BiCGStab tolerance of ADHK_PREC_MIXED; the FP64 Newton
                                residual test at tol_nonlin is unchanged (iterative refinement) */
    int warm_gate;           /* learned start kept when its residual is below the previous level's in
                                1: l2 norm (default), 0: max norm (no counterpart in the reference) */
    int substep_depth;       /* 0 (default): a failed step cuts dt by 0.25 and ends (tc_scale.c:3-19, time_loop.c:
                                25-35), an ensemble member drops out.  n > 0 (no counterpart in the reference): the
                                failed model or member retakes the step as four quarter steps from its previous
                                level, a failed quarter again as four, down to dt / 4^n; the step then completes
                                at full dt and the next one runs at dt.  A member that fails at dt / 4^n stops at
                                the last sub-step it completed and drops out. */
    int precond;             /* ADHK_PC_BJ (default) or ADHK_PC_SGS; an ensemble (more than one instance)
                                always takes ADHK_PC_BJ */
} adhk_solver;

/* preconditioner of the BiCGStab solve.  The reference's is an exact LU (bcgstab_solver.c:371-424);
 * neither choice here has a counterpart there. */
#define ADHK_PC_BJ  0 /* 3x3 nodal block-Jacobi, folded into the matrix (M^-1 A) */
#define ADHK_PC_SGS 1 /* multicolour symmetric Gauss-Seidel on the block-Jacobi-folded matrix, applied on both
                         sides through Eisenstat's identity (one pass over the matrix per product); falls back
                         to ADHK_PC_BJ for a solve whose folded diagonal blocks are not all the identity */

typedef struct {
    int newton_its;
    int linesearch_cuts_total;
    int lin_its_total;
    int lin_its_max;
    double resid_max_norm;
    double inc_max_norm;
    int converged;     /* fe_newton return value */
    int warm_accepted; /* Newton solves that started from the learned iterate (adhk_sw2_set_warm_start) */
    int warm_retried;  /* solves that failed from the learned iterate and were solved again from the
                          previous level (no counterpart in the reference) */
    int substeps;      /* sub-steps taken to recover failed steps (adhk_solver.substep_depth) */
    int lin_capped;    /* linear solves that stopped at lin_max_it iterations (bcgstab_solver.c:22) */
} adhk_newton_stats;

/* per-stage device time of the last adhk_time_loop / adhk_fe_newton call, ms */
typedef struct {
    double residual_ms, jacobian_ms, prep_ms, krylov_ms, update_ms, total_ms;
    int residual_calls, jacobian_calls, krylov_calls;
} adhk_timing;

/* Schedule choices of a context.  None changes a computed value: the iterations, the reduction order
 * and every operand are the same under all of them, only their launch and read-back pattern differs.
 * Defaults: chunk 0 (use adhk_solver.lin_check_every), graphs 0, xcd_map 1, elem_tpb 256, split 0,
 * stream_matrix -1, sgs_wide 0. */
typedef struct {
    int chunk;    /* BiCGStab iterations launched between "done" read-backs (even; 0 = the solver's) */
    int graphs;   /* 1 = replay each chunk as a captured hipGraph (when per-kernel timers are off) */
    int xcd_map;  /* 1 = XCD-contiguous block ranges for the matrix kernels, 0 = workgroup order */
    int elem_tpb; /* workgroup size of the elemental residual / Jacobian kernels (64, 128 or 256) */
    int split;    /* 1 = BiCGStab's A step as two kernels: store the new direction, then multiply it
                     (one gathered vector per column instead of four) */
    int stream_matrix; /* matrix loads of the BiCGStab products: 1 non-temporal, 0 cached, -1 (default)
                          non-temporal when the matrix exceeds the 256 MB MALL */
    int sgs_wide; /* multicolour SGS phase kernels: 1 = groups of eight slots with every load in flight
                     (2 waves per SIMD), 0 = smaller groups at 3-4 waves per SIMD */
} adhk_tuning;

/* kernel classes of the per-kernel device timers */
#define ADHK_KT_SPMV_P 0      /* BiCGStab: Ap = M^-1 A p (+ q.Ap)              */
#define ADHK_KT_SPMV_S 1      /* BiCGStab: As = M^-1 A s (+ 3 dots)            */
#define ADHK_KT_UPDATE 2      /* BiCGStab: x, r, p update (+ max|r|), split A   */
#define ADHK_KT_ELEM_JAC 3    /* elemental central-FD Jacobians                */
#define ADHK_KT_JAC_GATHER 4  /* block assembly of the Jacobian                */
#define ADHK_KT_ELEM_RESID 5  /* elemental residuals                           */
#define ADHK_KT_RESID_GATHER 6 /* residual assembly                            */
#define ADHK_KT_PREP 7        /* check_diag, Dirichlet, scaling, block inverse */
#define ADHK_KT_NKINDS 8

const char *adhk_last_error(void);
/* waits for all work of every context on the current device (hipDeviceSynchronize) */
int adhk_device_synchronize(void);
void adhk_solver_defaults(adhk_solver *s);

/* ---- context: mesh upload, geometry (selem_2d.c:270-351, selem_1d.c:42-56),
 *      CSR pattern on the GPU (slin_sys_init_sparsity_mono.c:28-494) ---- */
adhk_sw2 *adhk_sw2_create(int nnodes, const double *x, const double *y, const double *z, int nelems2d,
                          const int *tri, const double *mann, int nelems1d, const int *seg, const int *seg_kind,
                          const double *seg_val, const adhk_params *p);
/* An ensemble of independent meshes (instances) in one context: arrays are the instances' arrays
 * concatenated, with node ids already offset; instance i owns nodes [node_off[i], node_off[i+1]),
 * triangles [elem_off[i], elem_off[i+1]) and segments [seg_off[i], seg_off[i+1]) (offsets [ninst+1],
 * starting at 0).  Every launch covers all instances; each instance keeps its own BiCGStab scalars,
 * Newton norms, line search and stopping test, so its iterates are those of the instance solved alone.
 * No counterpart in the reference (one SMODEL_SUPER per call); for many small meshes at once. */
adhk_sw2 *adhk_sw2_create_batch(int ninst, const int *node_off, const int *elem_off, const int *seg_off, int nnodes,
                                const double *x, const double *y, const double *z, int nelems2d, const int *tri,
                                const double *mann, int nelems1d, const int *seg, const int *seg_kind,
                                const double *seg_val, const adhk_params *p);
int adhk_sw2_instances(const adhk_sw2 *c);
/* Ensemble members are independent: a member whose Newton solve fails is reset to its previous level
 * (as fe_newton's failure path does, fe_newton.c:869-915) and drops out of the following steps while
 * the others continue.  alive [ninst], steps_done [ninst] (either may be NULL).  adhk_sw2_set_state
 * re-admits every member.  (A single mesh keeps the reference's behaviour: dt cut, time loop ends.) */
int adhk_sw2_instance_status(const adhk_sw2 *c, int *alive, int *steps_done);
/* per-instance statistics [ninst] of the last adhk_fe_newton (or totals of the last adhk_time_loop) */
int adhk_sw2_instance_stats(const adhk_sw2 *c, adhk_newton_stats *out);
void adhk_sw2_destroy(adhk_sw2 *c);
int adhk_sw2_sizes(const adhk_sw2 *c, int *ndofs, int *nnz, int *nnz_blocks);
/* SELL-64 layout of the device matrix: 64-node slices and 3x3-block slot columns */
int adhk_sw2_layout(const adhk_sw2 *c, int *nslice, int *nslot);
int adhk_sw2_set_state(adhk_sw2 *c, const double *sol, const double *sol_old, const double *sol_older);
int adhk_sw2_get_state(adhk_sw2 *c, double *sol, double *sol_old, double *sol_older);
int adhk_sw2_set_bc(adhk_sw2 *c, const int *bc_mask, const double *dirichlet_data);
int adhk_sw2_get_bc(adhk_sw2 *c, int *bc_mask);
int adhk_sw2_set_dt(adhk_sw2 *c, double dt);
/* boundary values of every segment, [nelems1d][2] (the meaning of each per ADHK_BC_*); the reference's
 * boundary residuals read theirs from the time series at each call (sw2_bc_h.c:155-157) */
int adhk_sw2_set_seg_values(adhk_sw2 *c, const double *seg_val);
double adhk_sw2_get_dt(const adhk_sw2 *c);
int adhk_sw2_get_timing(const adhk_sw2 *c, adhk_timing *t);
/* per-kernel device timers: on = 1 resets and enables (hipEvent pairs around every launch of the
 * classes above, on the context's stream), on = 0 disables.  kernel_times flushes and returns the
 * accumulated milliseconds and launch counts per class ([ADHK_KT_NKINDS] each). */
int adhk_sw2_kernel_timing(adhk_sw2 *c, int on);
/* as kernel_timing, for the classes whose bits are set in mask (bit k = class k; 0 disables), timing
 * the first launch of each class and then every stride-th one (stride >= 1; 1 times every launch).
 * Each timed launch adds two event records to the stream; sampling keeps that cost off the rest. */
int adhk_sw2_kernel_timing_mask(adhk_sw2 *c, unsigned mask, int stride);
int adhk_sw2_kernel_times(adhk_sw2 *c, double *ms, long long *count);
/* device copy of nbytes (read nbytes + write nbytes) repeated reps times on the context's stream:
 * average milliseconds per copy (the achievable HBM ceiling, and a counter-scale reference) */
int adhk_copy_probe(adhk_sw2 *c, long long nbytes, int reps, double *ms_avg);
int adhk_sw2_geometry(adhk_sw2 *c, double *djac, double *grad_shp, double *seg_djac, double *seg_nrml);

/* schedule choices (see adhk_tuning); no counterpart in the reference */
int adhk_sw2_get_tuning(const adhk_sw2 *c, adhk_tuning *t);
int adhk_sw2_set_tuning(adhk_sw2 *c, const adhk_tuning *t);
/* Autotuner: times each candidate schedule on this context's own system -- the Jacobian and Krylov
 * solve of the current state, ninner BiCGStab iterations per candidate -- and keeps the fastest.
 * The state is restored afterwards.  chosen (or NULL) receives the choice, ms[4] (or NULL) the best
 * times of the element kernels and of one BiCGStab iteration. */
int adhk_sw2_autotune(adhk_sw2 *c, const adhk_solver *s, int ninner, adhk_tuning *chosen, double *ms);

/* node colouring of the multicolour SGS preconditioner (ADHK_PC_SGS; no counterpart in the reference):
 * DSatur (Brelaz 1979) over the block pattern -- the uncoloured node with the most distinct neighbour
 * colours first, ties by the most uncoloured neighbours, then the lowest index; it takes the smallest colour
 * none of its neighbours holds; then Kempe-chain exchanges drop the top colour while every node of it can be
 * moved (components of two colours up to 4096 nodes).  colour [nnodes] (or NULL) and the colour count;
 * layout [3] (or NULL): the permuted
 * row count (each colour padded to a multiple of 256) and the SELL-64 slot columns of the U and L parts.
 * One mesh alone. */
int adhk_sw2_colouring(adhk_sw2 *c, int *colour, int *ncolour, int *layout);

/* replaces fe_sw2_init (src/models/sw2/fe_sw2_init.c:28) incl. fe_sw2_wdflag_legacy (sw2_wdflag_legacy.c:34) */
int adhk_sw2_init(adhk_sw2 *c, int *wd_flag_out /* [nelems2d] or NULL */);

/* replaces assemble_residual(SMODEL_SUPER*, SGRID*) (src/residual/residual.h:5, assemble_residual.c:69):
 * residual = -sum of elemental SW2 residuals at the context's sol. residual may be NULL (device only). */
int adhk_assemble_residual(adhk_sw2 *c, double *residual);

/* elemental residual for one element (the pde[k].resid pointer, sw2_body_resid.c:44 / sw2_bc_h.c:57 /
 * sw2_bc_vel.c:58) evaluated on the GPU for a batch: elem_rhs [nelems2d*9 + nelems1d*6] */
int adhk_elem_residuals(adhk_sw2 *c, double *elem_rhs);

/* one element's residual at sol, optionally perturbed (the reference's pde[k].resid call in its element
 * loops, sw2_body_resid.c:44, sw2_bc_h.c:57, sw2_bc_vel.c:58): which 0 = triangle ie -> out[9], 1 = segment
 * ie -> out[6]; the variable pvar (0 h, 1 u, 2 v; -1 none) of local node pnode is moved by pe.  wd_flag >= 0
 * replaces the triangle's wet/dry flag (else the one of the last adhk_sw2_init).  which 2 and 3 evaluate a
 * triangle with the all-wet forms of the residual kernel and of the element Jacobian kernel (the latter
 * multiplies by 1/3 and reuses values shared by the +-eps pair), for checking that they agree. */
int adhk_elem_residual_one(adhk_sw2 *c, int which, int ie, double pe, int pnode, int pvar, int wd_flag,
                           double *out);

/* replaces slin_sys_init_sparsity_mono (slin_sys_init_sparsity_mono.c:28): CSR pattern built on the GPU */
int adhk_sparsity(adhk_sw2 *c, int *indptr /* ndofs+1 */, int *cols /* nnz */);

/* replaces assemble_jacobian(SMODEL_SUPER*) (src/jacobian/jacobian.h:5, assemble_jacobian.c:20):
 * central-FD Jacobian at sol, stored in the context; vals (CSR order of adhk_sparsity) may be NULL. */
int adhk_assemble_jacobian(adhk_sw2 *c, double *vals);
/* elemental FD Jacobians (load_elem_mat, assemble_jacobian.c:123): [nelems2d*81 + nelems1d*36], row-major */
int adhk_elem_jacobians(adhk_sw2 *c, double *elem_mat);
/* No counterpart in the reference: the elemental matrices of the mixed-precision solve (the same central
 * differences, computed in FP64 and stored in FP32), widened to double in the layout of
 * adhk_elem_jacobians. */
int adhk_elem_jacobians_mixed(adhk_sw2 *c, double *elem_mat);

/* replace check_diag(SMODEL_SUPER*) and apply_Dirichlet_BC(SMODEL_SUPER*) (la.h, matrix_utilities.c:120, :42)
 * on the context's system */
int adhk_check_diag(adhk_sw2 *c);
int adhk_apply_dirichlet_bc(adhk_sw2 *c);
int adhk_get_system(adhk_sw2 *c, double *vals /* nnz or NULL */, double *residual /* ndofs or NULL */);

/* replaces fe_newton(SMODEL_SUPER*) (src/newton/newton.h:19, fe_newton.c:44).  Returns TRUE(1)/FALSE(0), -1 on error.
 * For an ensemble: TRUE when every instance converged; st holds sums over the instances. */
int adhk_fe_newton(adhk_sw2 *c, const adhk_solver *s, adhk_newton_stats *st);

/* replaces time_loop (src/timeloop/time_loop.c:14) for one super model: nsteps x (prep_sol + fe_newton).
 * Returns the number of steps completed (negative if a step failed), -1e9 on error. */
int adhk_time_loop(adhk_sw2 *c, const adhk_solver *s, int nsteps, adhk_newton_stats *acc);

/* ---- reference src/la/la.h surface, host CSR arrays (exact names and signatures) ---- */
void split_CSR_mat_vec_mult(double *Ax, int *indptr_diag, int *cols_diag, double *vals_diag, int *indptr_off_diag,
                            int *cols_off_diag, double *vals_off_diag, double *x, int nrows, int *ghosts,
                            int nghost); /* la.h:13, matrix_utilities.c:177 */
void scale_linear_system(int *indptr_diag, int *cols_diag, double *vals_diag, int *indptr_off_diag,
                         int *cols_off_diag, double *vals_off_diag, double *b, double *x, double *scale_vect,
                         int local_size, int size, int rank, int *ghosts,
                         int nghost); /* la.h:26, scale_linear_system.c:31 */
void unscale_linear_system(double *x, double *x0, double *scale_vect,
                           int local_size); /* la.h:29, scale_linear_system.c:137 */
int prep_umfpack(int *indptr_diag, int *cols_diag, double *vals_diag,
                 int nrow); /* la.h:21, bcgstab_solver.c:371 -- builds the block-Jacobi preconditioner */
int solve_umfpack(double *x, int *indptr_diag, int *cols_diag, double *vals_diag, double *b,
                  int nrow); /* la.h:22, bcgstab_solver.c:415 -- applies it */
void umfpack_clear(void);   /* la.h:23, bcgstab_solver.c:351 */
void free_bcgstab(void);    /* la.h:19, bcgstab_solver.c:438 */
int solve_linear_sys_bcgstab(double *x, int *indptr_diag, int *cols_diag, double *vals_diag, int *indptr_off_diag,
                             int *cols_off_diag, double *vals_off_diag, double *b, double *scale_vect,
                             int local_size, int size, int rank, int *ghosts,
                             int nghost); /* la.h:16, bcgstab_solver.c:65 */
void Screen_print_CSR(int *indptr, int *indices, double *data, int n); /* la.h:11 (declared, never defined there) */
/* block size used by prep_umfpack/solve_umfpack (3 = nodal SW2 blocks; 1 = point Jacobi) */
void adhk_la_set_block_size(int bs);
/* No counterpart in the reference (whose la.h calls end in tl_error -> exit on failure): the status of this
 * thread's last la.h call, 0 or -1 (adhk_last_error() then holds the message; the outputs are untouched),
 * and the BiCGStab iterations of its last solve_linear_sys_bcgstab.  Each host thread owns its own la.h
 * workspace (device buffers, stream, preconditioner); free_bcgstab() releases the calling thread's. */
int adhk_la_last_status(void);
int adhk_la_last_iterations(void);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): prep_umfpack + solve_linear_sys_bcgstab
 * in one call on DEVICE arrays (indptr, cols, vals, b, scale_vect, x of nrow rows), with the same kernels
 * and results, and no hipMalloc/hipFree or allocating library call.  Its working vectors live in a fixed
 * scratch array of the library, one per GPU (about 21 MB), which limits nrow to ADHK_NOALLOC_MAX_N; larger
 * systems return -1.  bs: 3 (nodal 3x3 blocks, when nrow % 3 == 0) or 1.  stream: a hipStream_t, or NULL
 * for the null stream; the call returns after the solve has finished on it.  x: in x0, out (x + x0) * scale;
 * b: out M^-1 b, as solve_linear_sys_bcgstab leaves them; its (or NULL): BiCGStab iterations.  Every call
 * on a GPU uses that GPU's one scratch array: calls that overlap on one GPU (from several threads or
 * streams) race and must be serialised by the caller. */
#define ADHK_NOALLOC_MAX_N 262144
int adhk_bcgstab_device_noalloc(double *x, const int *indptr, const int *cols, const double *vals, double *b,
                                const double *scale_vect, int nrow, int bs, void *stream, int *its);

/* ---- synthetic flume meshes (restates create_rectangular_grid.c:4-445 and adds boundary
 *      segments, seeded Gaussian bed bumps and per-element Manning n) ---- */
typedef struct {
    double xmin, xmax, ymin, ymax;
    int npx, npy;
    double slope;          /* bed z = -slope * (x - xmin) + bumps */
    int nbumps;            /* seeded Gaussian bumps */
    double bump_amp_max, bump_sigma_min, bump_sigma_max;
    unsigned long long seed;
    double manning_n;
    double inflow_q;       /* total discharge at x = xmin (m^3/s) */
    double tailwater;      /* depth at x = xmax */
    int ic_flat_wse;       /* initial state: 1 = flat water surface at the tailwater level, 0 = depth = tailwater */
} adhk_flume_spec;
int adhk_flume_counts(const adhk_flume_spec *f, int *nnodes, int *nelems2d, int *nelems1d);
int adhk_flume_build(const adhk_flume_spec *f, double *x, double *y, double *z, int *tri, double *mann, int *seg,
                     int *seg_kind, double *seg_val);
/* initial state with u = v = 0: ic_flat_wse = 1 -> h = (tailwater - slope*(xmax - xmin)) - z, a flat water
 * surface at the tailwater level of the bump-free bed; ic_flat_wse = 0 -> h = tailwater everywhere */
int adhk_flume_initial_state(const adhk_flume_spec *f, int nnodes, const double *z, double *sol);

/* ---- diagnostics.  tl_find_grid_mass_elem2d (tl_grid_mass.c:199-262): the water volume of the
 *      context's triangles at the current depth, times density, with the reference's weighting of
 *      partially dry elements; summed over all instances of an ensemble */
int adhk_sw2_grid_mass(adhk_sw2 *c, double density, double *mass);

/* ---- time-series boundary values (sseries.c, time_loop.c:74-76).  A series is a piecewise-linear
 *      table of (time, value); each step of adhk_time_loop sets every segment bound to a series to
 *      scale * (the series' average over the step [t, t + dt]), as sseries_set_values_average
 *      (sseries.c:622-668) does for the reference's boundary strings.  The context's time t starts at 0
 *      and advances by dt with every completed step.  A step whose interval leaves a series' time range
 *      fails (returns as a Newton failure would, with the message in adhk_last_error). */
int adhk_sw2_set_series(adhk_sw2 *c, int id, int npts, const double *t, const double *v);
/* bind value `component` (0 or 1) of the segments to series id (-1 unbinds: the value stays as it
 * was last set); it becomes scale * average.  Value 0 is the depth of ADHK_BC_DEPTH, the flux per
 * unit width of ADHK_BC_FLUX and u of ADHK_BC_VEL; value 1 is v of ADHK_BC_VEL */
int adhk_sw2_bind_series(adhk_sw2 *c, int nsegs, const int *segs, int component, int id, double scale);
int adhk_sw2_set_time(adhk_sw2 *c, double t);
double adhk_sw2_get_time(const adhk_sw2 *c);
/* average over [t1, t2] of the piecewise-linear series (t, v) (its value at t1 when t2 == t1): the
 * evaluation the time loop uses, on host arrays */
int adhk_series_average(int npts, const double *t, const double *v, double t1, double t2, double *value);

/* ---- AdH .geo mesh files (sgrid_read.c:84-130, :540-617): ND, TRI and SEG cards, 1-based ids.
 *      counts: nother = QUAD/TET/PRISM cards (3-D models, not read); returns -2 when node ids are not
 *      1..N.  read: nodes by id, elements in file order, 0-based; *_flag = the bflag field or -1.
 *      Returns 0, -1 on bad arguments or counts, or -(line + 2) at a malformed card. */
int adhk_geo_counts(const char *path, int *nnodes, int *ntri, int *nseg, int *nother);
int adhk_geo_read(const char *path, int nnodes, double *x, double *y, double *z, int ntri, int *tri, int *tri_flag,
                  int nseg, int *seg, int *seg_flag);

/* ---- unstructured channel meshes (Poisson-disk points, Delaunay by the caller, RCM order).
 *      The reference reads meshes from .geo files (sgrid_read.c) or builds rectangular ones
 *      (create_rectangular_grid.c); these generate the channel-with-floodplain workloads. */
typedef struct {
    double length, width;         /* domain [0, length] x [0, width] (m) */
    double spacing;               /* Poisson-disk minimum point distance (m) */
    double chan_width, chan_depth; /* main channel: width at mid-bank, depth below the floodplain (m) */
    double bank_width;            /* bed ramps linearly from channel to floodplain across this width */
    double meander_amp, meander_wavelength; /* centreline y = width/2 + amp sin(2 pi (x - phase) / wavelength) */
    double slope;                 /* bed falls by slope per m of x */
    int nbumps;                   /* seeded Gaussian bed bumps, as in the flume */
    double bump_amp_max, bump_sigma_min, bump_sigma_max;
    unsigned long long seed;
    double n_channel, n_floodplain; /* Manning n by element centroid (channel when weight >= 1/2) */
    double inflow_q;              /* discharge through the channel section at x = 0 (m^3/s) */
    double fp_depth_out;          /* initial flat water surface: floodplain depth at x = length (m) */
    double meander_phase;         /* x offset of the centreline's sine (m) */
} adhk_channel_spec;
/* upper bound on the number of points adhk_poisson_disk can return for this rectangle */
int adhk_poisson_disk_bound(double x0, double x1, double y0, double y1, double r);
/* Poisson-disk sample (Bridson's algorithm, 30 tries per active point) of [x0,x1] x [y0,y1] with
 * minimum distance r, starting from the corners and evenly spaced edge points; returns the count
 * (-1 on bad arguments or when more than max_n points would be needed) */
int adhk_poisson_disk(double x0, double x1, double y0, double y1, double r, unsigned long long seed, int max_n,
                      double *x, double *y);
/* reverse Cuthill-McKee order of the triangle mesh's node graph: perm[new] = old */
int adhk_rcm_order(int nnodes, int nelems, const int *tri, int *perm);
/* bed elevation of the channel workload at the given points */
int adhk_channel_bed(const adhk_channel_spec *s, int nnodes, const double *x, const double *y, double *z);
/* Manning n per triangle (channel or floodplain by centroid) */
int adhk_channel_roughness(const adhk_channel_spec *s, int nnodes, const double *x, const double *y, int nelems,
                           const int *tri, double *mann);
/* channel section weight at (x, y): 1 in the channel, 0 on the floodplain, linear across the banks */
double adhk_channel_section_weight(const adhk_channel_spec *s, double x, double y);

/* ---- learned warm start of the Newton iterate.  No counterpart in the reference, whose Newton
 *      solve starts from the previous time level (smodel_super_prep_sol.c:2-8, newton_tools.c:21).
 *      A per-node MLP 12 -> 64 -> 64 -> 3 (GELU, BF16 weights, FP32 arithmetic) predicts the change
 *      of (h, u, v) over one step from 12 features: h, hu, hv at t; their change since t-1; the 1-ring
 *      neighbour means of h, hu, hv at t minus the node's own; the bed slope (x, y) and Manning n, both
 *      area-weighted over the node's triangles.
 *      It is trained on the GPU on (features, change) pairs collected from solver trajectories.
 *      With a model attached, fe_newton evaluates the residual at the previous level and at the
 *      predicted iterate and starts from the smaller (l2 norm by default, adhk_solver.warm_gate);
 *      iteration and stop test are unchanged. */
#define ADHK_MLP_NIN 12
#define ADHK_MLP_HID 64
#define ADHK_MLP_NOUT 3
typedef struct adhk_mlp adhk_mlp;
int adhk_mlp_param_count(void); /* 5187: W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] */
adhk_mlp *adhk_mlp_create(unsigned long long seed);
void adhk_mlp_destroy(adhk_mlp *m);
/* FP32 master parameters [param_count] and normalisation [36]: x mean[12], x 1/std[12], y mean[3],
 * y std[3], y 1/std[3]; set_params resets the optimiser state */
int adhk_mlp_get_params(adhk_mlp *m, float *params, float *norm);
int adhk_mlp_set_params(adhk_mlp *m, const float *params, const float *norm);
/* device dataset: reserve capacity (samples), append one sample per node of the context: features of
 * (sol_old, sol_older) -> target sol - sol_old (call after a converged step) */
int adhk_mlp_data_reserve(adhk_mlp *m, long long capacity);
int adhk_mlp_collect(adhk_mlp *m, adhk_sw2 *c);
long long adhk_mlp_data_size(const adhk_mlp *m);
int adhk_mlp_fit_normalization(adhk_mlp *m);
/* Adam on minibatches drawn from the dataset (seeded); loss_hist [steps] or NULL */
int adhk_mlp_train(adhk_mlp *m, int steps, int batch, double lr, unsigned long long seed, double *loss_hist);
/* checking surface: raw features [n][12] -> raw outputs [n][3]; loss and gradient [param_count] of the
 * mean squared error on normalised outputs over n samples (n <= 65536) */
int adhk_mlp_forward(adhk_mlp *m, int n, const float *x, float *y);
int adhk_mlp_loss_grad(adhk_mlp *m, int n, const float *x, const float *t, float *grad, double *loss);
/* features [nnodes][12] of the context's (sol_old, sol_older) */
int adhk_sw2_features(adhk_sw2 *c, float *feat);
/* attach (or with NULL detach) a model as the Newton warm start of this context */
int adhk_sw2_set_warm_start(adhk_sw2 *c, adhk_mlp *m);

#ifdef __cplusplus
}
#endif
#endif
