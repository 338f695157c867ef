/*
 * sw2_oracle.h -- CPU oracle for the AdH SW2 Newton step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (erdc/adh_kraken, src/) used as the checker for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  The pins (checks of
 * this restatement that do not retype it) live in tests/test_oracle_pins.py:
 *   P1 triangle and line integrals, all-wet body residual  -> 7-point / Gauss-Legendre quadrature
 *   wet/dry wrapper (partially dry elements)                -> quadrature over the clipped wet region
 *   boundary segments (depth, flux, velocity)               -> line quadrature
 *   Manning friction chain                                  -> log-law profile integrated numerically
 *   lake at rest, global mass and x-momentum balances       -> invariants
 *   geometry, dof map, CSR pattern, wet/dry flags           -> closed forms and brute force
 *   FD Jacobian                                             -> directional differences of the residual
 *   block-Jacobi, BiCGStab, Newton                          -> dense inverses, direct sparse solve,
 *                                                              converged-state residual
 *   SUPG combination (sw2_body_resid.c:921-1079)             -> weights from the flux Jacobians of the
 *                                                              shallow-water operator taken by central
 *                                                              differences; rotation / reflection
 *                                                              equivariance; zero on uniform normal flow
 *
 * Layout conventions follow the reference: dof(node, var) = 3*node + var with
 * var 0 = h (depth), 1 = u, 2 = v (depth-averaged velocity)
 * (src/structs/sivar_position/sivar_position.c:61-80); elemental residual
 * arrays are node-major [node*3 + {c, x, y}].
 */
#ifndef SW2_ORACLE_H
#define SW2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary segment kinds (1-D elements) */
#define OR_BC_DEPTH 0 /* fe_sw2_bc_h: tailwater depth, src/models/sw2/sw2_bc_h.c:57  */
#define OR_BC_FLUX  1 /* explicit flow, src/fe/fe_sw2_boundary_integrals.c:56 (val0 = flux per unit width, <0 = inflow) */
#define OR_BC_VEL   2 /* fe_sw2_bc_vel: weak velocity, src/models/sw2/sw2_bc_vel.c:58 (val0,val1 = u,v) */
#define OR_BC_DEPTH_NODAL 3 /* fe_sw2_bc_h with a depth per node (val0 first node, val1 second) */

#define OR_FRIC_CONST   0 /* drag coefficient constant (sw2_body_resid.c:241 hard-codes 0.002) */
#define OR_FRIC_MANNING 1 /* Manning n -> roughness height -> log-law drag (src/friction)    */

typedef struct {
    int nnodes, nelems2d, nelems1d;
    const double *x, *y, *z;   /* [nnodes] */
    const int *tri;            /* [nelems2d*3], counter-clockwise */
    const int *seg;            /* [nelems1d*2], domain on the left of (n0 -> n1) */
    const int *seg_kind;       /* [nelems1d] OR_BC_* */
    const double *seg_val;     /* [nelems1d*2] */
    const double *mann;        /* [nelems2d] Manning n (OR_FRIC_MANNING) */
} or_mesh;

typedef struct {
    double g;          /* gravity, smodel_super_alloc_init.c:92 (9.8) */
    double dt;         /* time step */
    double alpha;      /* tau_temporal, sdt.c:24 (0 = backward Euler) */
    double cf_const;   /* drag coefficient for OR_FRIC_CONST (0.002) */
    int fric_mode;     /* OR_FRIC_* */
    double manning_units; /* Manning's units constant (1.0 = SI), sstr_value.c:288 */
} or_params;

typedef struct {
    double tol_nonlin;            /* 1e-8  smodel_super_alloc_init.c:63 */
    double inc_nonlin;            /* 1e-6  :62 */
    int max_nonlin_it;            /* 10    :65 */
    int max_linesearch_cuts;      /* 5     :64 */
    double lin_tol;               /* 1e-11 bcgstab_solver.c:20-21 (abs and rel) */
    int lin_max_it;               /* 100   bcgstab_solver.c:22 */
    int precond;                  /* OR_PREC_BJ, or OR_PREC_SGS (no counterpart in the reference) */
} or_solver;

#define OR_PREC_BJ  0 /* 3x3 nodal block-Jacobi in place of the reference's exact LU */
#define OR_PREC_SGS 1 /* multicolour symmetric Gauss-Seidel on the block-Jacobi-folded matrix, split form */
#define OR_KEMPE_CAP 4096 /* largest two-colour component the SGS colour reduction exchanges */

typedef struct {
    int newton_its;
    int linesearch_cuts_total;
    int lin_its_total;
    int lin_its_max;
    double resid_max_norm;
    double inc_max_norm;
    int converged;     /* fe_newton return value (1 TRUE / 0 FALSE) */
} or_newton_stats;

/* opaque context holding mesh, geometry, state, CSR system */
typedef struct or_ctx or_ctx;

or_ctx *or_create(const or_mesh *m, const or_params *p);
void or_destroy(or_ctx *c);
int or_ndofs(const or_ctx *c);
int or_nnz(const or_ctx *c);
void or_set_threads(int n);

/* geometry (selem_2d.c:270-351, selem_1d.c:42-56) */
void or_get_geometry(const or_ctx *c, double *djac, double *grad_shp, double *seg_djac, double *seg_nrml);

/* state access: sol, sol_old, sol_older [ndofs]; bc_mask [ndofs] (1=YES); dirichlet [ndofs] */
void or_set_state(or_ctx *c, const double *sol, const double *sol_old, const double *sol_older);
void or_get_state(const or_ctx *c, double *sol, double *sol_old, double *sol_older);
void or_set_bc(or_ctx *c, const int *bc_mask, const double *dirichlet);
void or_get_bc(const or_ctx *c, int *bc_mask);
void or_set_dt(or_ctx *c, double dt);

/* elemental residuals (the PDE "resid" function pointers) */
void or_sw2_body_resid(const or_ctx *c, int ie, double perturbation, int pnode, int pvar, int psign,
                       const double *sol, double *elem_rhs /*9, added to*/);
void or_sw2_seg_resid(const or_ctx *c, int ie, double perturbation, int pnode, int pvar, int psign,
                      const double *sol, double *elem_rhs /*6, added to*/);

/* wet/dry flags, fe_sw2_wdflag_legacy (sw2_wdflag_legacy.c:34-139) */
void or_wdflag(or_ctx *c, int *wd_flag_out /* may be NULL */);
double or_grid_mass(const or_ctx *c, double density); /* tl_find_grid_mass_elem2d on the current depth */
/* fe_sw2_init (fe_sw2_init.c:28-53): wd flags + dry nodes -> u,v Dirichlet */
void or_sw2_init(or_ctx *c);

/* global assembly */
void or_assemble_residual(or_ctx *c, double *residual_out /* may be NULL */);
void or_pattern(const or_ctx *c, int *indptr /*ndofs+1*/, int *cols /*nnz*/);
void or_elem_jacobian_tri(const or_ctx *c, int ie, double *mat /*81, row-major*/);
void or_elem_jacobian_seg(const or_ctx *c, int ie, double *mat /*36*/);
void or_assemble_jacobian(or_ctx *c, double *vals_out /* may be NULL */);
void or_check_diag(or_ctx *c);
void or_apply_dirichlet(or_ctx *c);
void or_get_system(const or_ctx *c, double *vals, double *residual);
void or_set_system(or_ctx *c, const double *vals, const double *residual);

/* linear algebra on plain CSR arrays (src/la) */
void or_spmv(int n, const int *indptr, const int *cols, const double *vals, const double *x, double *y);
void or_scale_linear_system(int n, const int *indptr, const int *cols, double *vals,
                            double *b, double *x, double *scale);
void or_unscale(int n, double *x, const double *x0, const double *scale);
void or_block_jacobi_prep(int nnodes, const int *indptr, const int *cols, const double *vals, double *minv /*9*nnodes*/);
void or_block_jacobi_apply(int nnodes, const double *minv, const double *b, double *x);
int  or_bicgstab(int n, const int *indptr, const int *cols, const double *vals, const double *minv,
                 double *x, double *b, const double *scale, double tol, int max_it, double *final_rnorm);
int  or_colour_dsatur(int nnodes, const int *indptr, const int *cols, int *colour);
int  or_colour_reduce(int nnodes, const int *indptr, const int *cols, int *colour, int nc, int cap);
int  or_bicgstab_sgs(int n, const int *indptr, const int *cols, const double *vals, const double *minv,
                     const int *colour, int ncolour, double *x, const double *b, const double *scale, double tol,
                     int max_it, double *final_rnorm);
void or_residual_norms(int n, const double *residual, const double *dsol, const int *bc_mask,
                       double *rmax, double *rl2, double *incmax);

/* nonlinear solver and time loop */
int or_fe_newton(or_ctx *c, const or_solver *s, or_newton_stats *st);
int or_time_loop(or_ctx *c, const or_solver *s, int nsteps, or_newton_stats *accum);

#ifdef __cplusplus
}
#endif
#endif
