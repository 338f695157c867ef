/*
 * ref_shim.c -- the reference's struct-based entry points of the SW2 Newton step (residual.h, jacobian.h,
 * newton.h, la.h, sw2.h, slin_sys.h, dofmaps.h) with their names and signatures, over the device context
 * of include/adhk.h.  Structure layouts are those of include/adhk_ref.h (checked field by field against
 * the reference's headers).
 *
 * The time loop (time_loop, advance_time, update_dt, tc_init, smodel_super_prep_sol), the series averages,
 * the friction laws, the single-rank messg reductions and the dof-map builder run on the host: they move
 * the caller's arrays and scalars and call fe_newton, which runs on the GPU.
 *
 * Each SMODEL_SUPER gets a context on its first call, built from sm->grid: node coordinates, triangles,
 * boundary segments and, per segment, the boundary kind named by the residual function its physics
 * material carries (fe_sw2_bc_h: depth, fe_sw2_bc_vel: velocity, fe_sw2_bc_flux: explicit flux).  The
 * contexts live in a side table keyed by the SMODEL_SUPER pointer (the reference struct has no spare
 * field).  Before each call the caller's state (sol, sol_old, sol_older), boundary masks and Dirichlet
 * data, time step (ts->dt / nsubsteps) and the current values of the boundary series are copied in; what
 * the reference writes (residual, Jacobian values, sol, bc_mask, wet/dry flags) is copied out.  All
 * arithmetic of the residual, the Jacobian, the linear solve, the norms and the increment runs on the GPU;
 * the host only moves the caller's arrays, walks its materials and series, and runs the reference's
 * index and scatter helpers (dofmaps, load_global_*, add_replace_elem_rhs), which the device path itself
 * never calls.
 *
 * Errors: a message in adhk_last_error() and adhk_ref_status() == -1; fe_newton returns FALSE, the other
 * functions leave their outputs as they were.  Thread-compatible: one thread per model at a time.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/adhk.h"
#include "../../include/adhk_ref.h"

void adhk_set_error(const char *fmt, ...);
int adhk_la_vec_add(double *y, const double *x, int n);
int adhk_la_residual_norms(int n, const double *residual, const double *dsol, const int *bc_mask,
                           const int *include_dof, double *rmax, double *l2sum, double *dmax, int *irmax, int *idmax,
                           int *nnan);

#define REF_YES 1             /* YES, define.h */
#define REF_UNSET_INT (-3)    /* UNSET_INT, define.h:442 */
#define REF_DT_REDUCE 0.25    /* DT_REDUCE_FACTOR, define.h:127 */
#define REF_CF_DEFAULT 0.002  /* the drag coefficient sw2_body_resid.c:241 hard-codes */
#define REF_PERTURBATION 1e-4 /* perturb_var, assemble_jacobian.c:314 */
/* variable codes of the SW2 unknowns (sadh_def.c:28-33): depth, depth-averaged x and y velocity */
#define REF_VAR_H 0
#define REF_VAR_UDA 4
#define REF_VAR_VDA 5

typedef struct {
    SMODEL_SUPER *sm;
    adhk_sw2 *c;
    double *mann; /* Manning n per 2-D element (adhk_ref_set_manning), or NULL: the constant drag */
    int S;
    int *kind, *iseries; /* per segment */
    double *val;         /* [2S] */
} ref_ent;

/* entries are allocated one by one, so a pointer to an entry stays valid while the table grows */
static ref_ent **g_tab;
static int g_ntab, g_cap;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static __thread int tl_status;

int adhk_ref_status(void) { return tl_status; }

static int ref_fail(const char *what)
{
    adhk_set_error("%s", what);
    tl_status = -1;
    return -1;
}

static ref_ent *entry(SMODEL_SUPER *sm, int create)
{
    ref_ent *e = NULL;
    pthread_mutex_lock(&g_mu);
    for (int i = 0; i < g_ntab; i++)
        if (g_tab[i]->sm == sm) e = g_tab[i];
    if (!e && create) {
        if (g_ntab == g_cap) {
            int cap = g_cap ? 2 * g_cap : 16;
            ref_ent **t = (ref_ent **)realloc(g_tab, (size_t)cap * sizeof *t);
            if (t) {
                g_tab = t;
                g_cap = cap;
            }
        }
        if (g_ntab < g_cap && (e = (ref_ent *)calloc(1, sizeof *e)) != NULL) {
            e->sm = sm;
            g_tab[g_ntab++] = e;
        }
    }
    pthread_mutex_unlock(&g_mu);
    return e;
}

void adhk_ref_release(SMODEL_SUPER *sm)
{
    ref_ent *e = NULL;
    pthread_mutex_lock(&g_mu);
    for (int i = 0; i < g_ntab; i++)
        if (g_tab[i]->sm == sm) {
            e = g_tab[i];
            g_tab[i] = g_tab[--g_ntab];
            break;
        }
    pthread_mutex_unlock(&g_mu);
    if (e) {
        if (e->c) adhk_sw2_destroy(e->c);
        free(e->mann);
        free(e->kind);
        free(e->iseries);
        free(e->val);
        free(e);
    }
}

int adhk_ref_set_manning(SMODEL_SUPER *sm, const double *manning_n)
{
    if (!sm || !sm->grid || !manning_n) return ref_fail("adhk_ref_set_manning: NULL argument");
    ref_ent *e = entry(sm, 1);
    if (!e) return ref_fail("adhk_ref_set_manning: out of host memory");
    if (e->c) return ref_fail("adhk_ref_set_manning: the model's context exists already (adhk_ref_release first)");
    const int E = sm->grid->nelems2d;
    free(e->mann);
    e->mann = (double *)malloc((size_t)(E > 0 ? E : 1) * sizeof(double));
    if (!e->mann) return ref_fail("adhk_ref_set_manning: out of host memory");
    memcpy(e->mann, manning_n, (size_t)E * sizeof(double));
    tl_status = 0;
    return 0;
}

/* the value k of series id: ivalue[k] of the list entry with that id (sseries_get_value, sseries.c:575) */
static int series_value(const SSERIES *head, int id, int k, double *v)
{
    for (const SSERIES *s = head; s; s = s->next)
        if (s->id == id) {
            if (!s->ivalue) return -1;
            *v = s->ivalue[k];
            return 0;
        }
    return -1;
}

/* boundary kind of segment s from its material's residual functions */
static int seg_kind(SMODEL_SUPER *sm, int s, int *kind, int *iseries)
{
    const int mat = sm->elem1d_physics_mat ? sm->elem1d_physics_mat[s] : sm->grid->elem1d[s].mat;
    if (!sm->mat_physics_elem || mat < 0 || mat >= sm->nmat_physics) return -1;
    const SMAT_PHYSICS *m = &sm->mat_physics_elem[mat];
    for (int k = 0; k < m->npdes; k++) {
        if (m->pde[k].resid == fe_sw2_bc_h) {
            *kind = ADHK_BC_DEPTH;
            *iseries = 0; /* the reference reads series 0 for every depth segment, sw2_bc_h.c:155-157 */
            return 0;
        }
        if (m->pde[k].resid == fe_sw2_bc_vel) {
            *kind = ADHK_BC_VEL;
            *iseries = m->pde[k].iseries;
            return 0;
        }
        if (m->pde[k].resid == fe_sw2_bc_flux) {
            *kind = ADHK_BC_FLUX;
            *iseries = m->pde[k].iseries;
            return 0;
        }
    }
    return -1;
}

/* the segments' values of this step from the series */
static int seg_values(SMODEL_SUPER *sm, ref_ent *e)
{
    for (int s = 0; s < e->S; s++) {
        double a = 0.0, b = 0.0;
        if (series_value(sm->series_head, e->iseries[s], 0, &a)) {
            adhk_set_error("boundary segment %d: series %d not found (or without values)", s, e->iseries[s]);
            return -1;
        }
        if (e->kind[s] == ADHK_BC_VEL && series_value(sm->series_head, e->iseries[s], 1, &b)) {
            adhk_set_error("velocity segment %d: series %d has no second value", s, e->iseries[s]);
            return -1;
        }
        e->val[2 * s] = a;
        e->val[2 * s + 1] = e->kind[s] == ADHK_BC_VEL ? b : 0.0;
    }
    return 0;
}

static int check_dof_map(SMODEL_SUPER *sm)
{
    /* the device layout is dof(node, var) = 3 node + var (sivar_position.c:61-80 builds it so for SW2) */
    const int N = sm->grid->nnodes;
    if (sm->ndofs != 3 * N) return -1;
    if (sm->ivars && sm->ivar_pos.var) {
        const int vh = sm->ivar_pos.var[REF_VAR_H], vu = sm->ivar_pos.var[REF_VAR_UDA],
                  vv = sm->ivar_pos.var[REF_VAR_VDA];
        for (int n = 0; n < N; n++)
            if (sm->ivars[vh][n] != 3 * n || sm->ivars[vu][n] != 3 * n + 1 || sm->ivars[vv][n] != 3 * n + 2)
                return -1;
    }
    return 0;
}

static double step_dt(const SMODEL_SUPER *sm) { return sm->ts->dt / (sm->nsubsteps > 0 ? sm->nsubsteps : 1); }

/* the model's context, created from sm->grid on the first call */
static ref_ent *ctx_of(SMODEL_SUPER *sm)
{
    if (!sm || !sm->grid || !sm->ts) {
        ref_fail("SMODEL_SUPER without grid or time information");
        return NULL;
    }
    ref_ent *e = entry(sm, 1);
    if (!e) {
        ref_fail("out of host memory");
        return NULL;
    }
    if (e->c) return e;
    SGRID *g = sm->grid;
    const int N = g->nnodes, E = g->nelems2d, S = g->nelems1d;
    if (g->nelems3d > 0 || check_dof_map(sm)) {
        ref_fail("the SW2 device path needs a 2-D mesh with dof(node, var) = 3 node + var");
        return NULL;
    }
    double *x = malloc((size_t)N * sizeof(double)), *y = malloc((size_t)N * sizeof(double)),
           *z = malloc((size_t)N * sizeof(double));
    int *tri = malloc((size_t)(E > 0 ? 3 * E : 1) * sizeof(int)), *seg = malloc((size_t)(S > 0 ? 2 * S : 1) * sizeof(int));
    e->kind = malloc((size_t)(S > 0 ? S : 1) * sizeof(int));
    e->iseries = malloc((size_t)(S > 0 ? S : 1) * sizeof(int));
    e->val = calloc((size_t)(S > 0 ? 2 * S : 2), sizeof(double));
    e->S = S;
    int bad = !x || !y || !z || !tri || !seg || !e->kind || !e->iseries || !e->val;
    for (int n = 0; !bad && n < N; n++) {
        x[n] = g->node[n].x;
        y[n] = g->node[n].y;
        z[n] = g->node[n].z;
    }
    for (int i = 0; !bad && i < E; i++) {
        if (g->elem2d[i].nnodes != 3) {
            bad = 2;
            break;
        }
        for (int a = 0; a < 3; a++) tri[3 * i + a] = g->elem2d[i].nodes[a];
    }
    for (int s = 0; !bad && s < S; s++) {
        seg[2 * s] = g->elem1d[s].nodes[0];
        seg[2 * s + 1] = g->elem1d[s].nodes[1];
        if (seg_kind(sm, s, &e->kind[s], &e->iseries[s])) bad = 3;
    }
    if (!bad && seg_values(sm, e)) bad = 4;
    if (!bad) {
        adhk_params p;
        memset(&p, 0, sizeof p);
        p.g = sm->gravity;
        p.dt = step_dt(sm);
        p.alpha = sm->ts->tau_temporal;
        p.cf_const = REF_CF_DEFAULT;
        p.manning_units = 1.0;
        p.fric_mode = e->mann ? ADHK_FRIC_MANNING : ADHK_FRIC_CONST;
        double *mann = e->mann;
        double *zeros = NULL;
        if (!mann) mann = zeros = calloc((size_t)(E > 0 ? E : 1), sizeof(double));
        e->c = mann ? adhk_sw2_create(N, x, y, z, E, tri, mann, S, seg, e->kind, e->val, &p) : NULL;
        free(zeros);
        if (!e->c) bad = 5;
    }
    free(x);
    free(y);
    free(z);
    free(tri);
    free(seg);
    if (bad) {
        const char *why = bad == 2 ? "only triangles are supported"
                          : bad == 3 ? "a boundary segment's material has none of fe_sw2_bc_h, fe_sw2_bc_vel, fe_sw2_bc_flux"
                          : bad == 4 ? adhk_last_error()
                          : bad == 5 ? adhk_last_error() : "out of host memory";
        char msg[512];
        snprintf(msg, sizeof msg, "context for the model: %s", why);
        adhk_ref_release(sm);
        ref_fail(msg);
        return NULL;
    }
    return e;
}

/* the caller's time step, state, boundary masks and series values into the context */
static int sync_in(SMODEL_SUPER *sm, ref_ent *e, int with_bc)
{
    if (adhk_sw2_set_dt(e->c, step_dt(sm)) || adhk_sw2_set_state(e->c, sm->sol, sm->sol_old, sm->sol_older))
        return ref_fail(adhk_last_error());
    if (with_bc && sm->bc_mask && sm->dirichlet_data && adhk_sw2_set_bc(e->c, sm->bc_mask, sm->dirichlet_data))
        return ref_fail(adhk_last_error());
    if (e->S > 0 && (seg_values(sm, e) || adhk_sw2_set_seg_values(e->c, e->val))) return ref_fail(adhk_last_error());
    return 0;
}

/* ---------------------------------------------------------------------------------- residual.h */

void assemble_residual(SMODEL_SUPER *sm, SGRID *grid)
{
    ref_ent *e = ctx_of(sm);
    if (!e) return;
    if (grid != sm->grid) {
        ref_fail("assemble_residual: grid is not the model's grid");
        return;
    }
    if (!sm->lin_sys || !sm->lin_sys->residual) {
        ref_fail("assemble_residual: no lin_sys->residual");
        return;
    }
    if (sync_in(sm, e, 1) || adhk_assemble_residual(e->c, sm->lin_sys->residual)) {
        ref_fail(adhk_last_error());
        return;
    }
    tl_status = 0;
}

/* add_replace_elem_rhs, assemble_residual.c:115-142, for the SW2 equations: their three variables (codes
 * 0, 4, 5) go to the element's positions elem_var_map[code] */
void add_replace_elem_rhs(double *elem_rhs, double *eq_rhs, int elem_nvars, int *elem_var_map, int imod, int nnodes,
                          double scale)
{
    static const int code[3] = {REF_VAR_H, REF_VAR_UDA, REF_VAR_VDA};
    (void)imod;
    for (int i = 0; i < nnodes; i++)
        for (int k = 0; k < 3; k++) elem_rhs[i * elem_nvars + elem_var_map[code[k]]] += scale * eq_rhs[i * 3 + k];
}

/* load_global_resid, assemble_residual.c:160-174 */
void load_global_resid(double *residual, double *elem_rhs, int nnodes, int elem_nvars, int *local_dofs)
{
    for (int i = 0; i < nnodes * elem_nvars; i++) residual[local_dofs[i]] += elem_rhs[i];
}

/* ---------------------------------------------------------------------------------- jacobian.h */

void assemble_jacobian(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e) return;
    if (!sm->lin_sys || !sm->lin_sys->vals_diag) {
        ref_fail("assemble_jacobian: no lin_sys->vals_diag (slin_sys_init_sparsity_mono first)");
        return;
    }
    int ndofs = 0, nnz = 0, nnzb = 0;
    adhk_sw2_sizes(e->c, &ndofs, &nnz, &nnzb);
    if (sm->lin_sys->nnz_diag != nnz) {
        ref_fail("assemble_jacobian: lin_sys pattern does not match the mesh (slin_sys_init_sparsity_mono)");
        return;
    }
    if (sync_in(sm, e, 1) || adhk_assemble_jacobian(e->c, sm->lin_sys->vals_diag)) {
        ref_fail(adhk_last_error());
        return;
    }
    tl_status = 0;
}

/* binary_search_part.c:1 */
static int bsearch_part(const int *arr, int start, int end, int target)
{
    int lo = start, hi = end - 1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (arr[mid] == target) return mid;
        if (arr[mid] < target) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

/* elem_matrix_deriv, assemble_jacobian.c:387-397: column (node_no, var_no) = (local1 - local2) / diff_ep */
void elem_matrix_deriv(double **mat, int node_no, int var_no, int nnodes, int elem_nvars, double *local1,
                       double *local2, double diff_ep)
{
    const int col = node_no * elem_nvars + var_no;
    for (int i = 0; i < nnodes; i++)
        for (int j = 0; j < elem_nvars; j++) {
            const int r = i * elem_nvars + j;
            mat[r][col] = (local1[r] - local2[r]) / diff_ep;
        }
}

/* perturb_var, assemble_jacobian.c:304-363: central differences of the PDEs' residuals (each evaluated on
 * the GPU) in variable ele_var_no of every node of element ie, with the reference's epsilon rule */
void perturb_var(double **elem_mat, SMODEL_SUPER *sm, SPDE *pde, int ie, int mat_id, int nodes_on_element,
                 int nvar_ele, int *elem_var_map, int npdes, int ele_var_no, int *NodeIDs, int DEBUG)
{
    const int code = sm->mat_physics_elem[mat_id].ivar_pos.var_code[ele_var_no];
    const int vpos = code == REF_VAR_H ? 0 : code == REF_VAR_UDA ? 1 : code == REF_VAR_VDA ? 2 : -1;
    const int nd = nvar_ele * nodes_on_element;
    double rp[3 * 3], rm[3 * 3], tp[3 * 3], tm[3 * 3];
    if (vpos < 0 || nodes_on_element > 3 || nvar_ele > 3) {
        ref_fail("perturb_var: not an SW2 variable or element");
        return;
    }
    for (int i = 0; i < nodes_on_element; i++) {
        const double var = sm->sol[3 * NodeIDs[i] + vpos];
        double eps = fabs(var) * (1.0 + REF_PERTURBATION) - fabs(var); /* NUM_DIFF_EPSILON, macro.h:669-673 */
        if (eps < REF_PERTURBATION) eps = REF_PERTURBATION;
        memset(rp, 0, sizeof rp);
        memset(rm, 0, sizeof rm);
        for (int j = 0; j < npdes; j++) {
            memset(tp, 0, sizeof tp);
            memset(tm, 0, sizeof tm);
            pde[j].resid(sm, tp, ie, eps, i, code, +1, DEBUG);
            add_replace_elem_rhs(rp, tp, nvar_ele, elem_var_map, pde[j].imod, nodes_on_element, 1.0);
            pde[j].resid(sm, tm, ie, eps, i, code, -1, DEBUG);
            add_replace_elem_rhs(rm, tm, nvar_ele, elem_var_map, pde[j].imod, nodes_on_element, 1.0);
        }
        (void)nd;
        elem_matrix_deriv(elem_mat, i, ele_var_no, nodes_on_element, nvar_ele, rp, rm, 2.0 * eps);
    }
}

/* load_elem_mat, assemble_jacobian.c:123-138 */
void load_elem_mat(SMODEL_SUPER *sm, int elem_nnodes, int *nodeIDs, int mat_id, int elem_id, double **elem_mat)
{
    const SMAT_PHYSICS *m = &sm->mat_physics_elem[mat_id];
    for (int k = 0; k < m->ivar_pos.n; k++)
        perturb_var(elem_mat, sm, m->pde, elem_id, mat_id, elem_nnodes, m->ivar_pos.n, m->ivar_pos.var, m->npdes, k,
                    nodeIDs, 0);
}

/* load_global_mat_split_CSR, assemble_jacobian.c:164-215 (rows past the local range go to the off-diagonal
 * block by global column) */
void load_global_mat_split_CSR(double *vals, int *indptr, int *indices, double *off_diag_vals, int *off_diag_indptr,
                               int *off_diag_indices, double **elem_mat, int ndofs_ele, int *dofs, int *global_dofs,
                               int *local_range)
{
    const int nrows = local_range[1] - local_range[0];
    for (int i = 0; i < ndofs_ele; i++) {
        const int row = dofs[i];
        if (row >= nrows) continue;
        for (int j = 0; j < ndofs_ele; j++) {
            const int k = bsearch_part(indices, indptr[row], indptr[row + 1], global_dofs[j]);
            if (k >= 0) vals[k] += elem_mat[i][j];
        }
    }
    (void)off_diag_vals;
    (void)off_diag_indptr;
    (void)off_diag_indices;
}

/* load_global_mat_CSR, assemble_jacobian.c:230-256 */
void load_global_mat_CSR(double *vals, int *indptr, int *indices, double **elem_mat, int ndofs_ele, int *global_dofs,
                         int *local_range)
{
    for (int i = 0; i < ndofs_ele; i++) {
        const int gr = global_dofs[i];
        if (gr < local_range[0] || gr >= local_range[1]) continue;
        const int row = gr - local_range[0];
        for (int j = 0; j < ndofs_ele; j++) {
            const int k = bsearch_part(indices, indptr[row], indptr[row + 1], global_dofs[j]);
            if (k >= 0) vals[k] += elem_mat[i][j];
        }
    }
}

/* ------------------------------------------------------------------------------------ newton.h */

/* initialize_system, newton_tools.c:21-53: sol <- sol_old, then every material PDE's init */
void initialize_system(SMODEL_SUPER *sm)
{
    memcpy(sm->sol, sm->sol_old, (size_t)sm->ndofs * sizeof(double));
    for (int i = 0; i < sm->nmat_physics; i++)
        for (int j = 0; j < sm->mat_physics_elem[i].npdes; j++)
            if (sm->mat_physics_elem[i].pde[j].init) sm->mat_physics_elem[i].pde[j].init(sm);
}

/* initialize_dirichlet_bc, newton_tools.c:130-142 */
void initialize_dirichlet_bc(SMODEL_SUPER *sm)
{
    for (int i = 0; i < sm->ndofs; i++)
        if (sm->bc_mask[i] == REF_YES) sm->sol[i] = sm->dirichlet_data[i];
}

void update_function(SMODEL_SUPER *sm) { (void)sm; } /* newton_tools.c:230: empty */

/* increment_function, newton_tools.c:261-268: sol += lin_sys->dsol (on the GPU) */
void increment_function(SMODEL_SUPER *sm)
{
    if (!sm->lin_sys || adhk_la_vec_add(sm->sol, sm->lin_sys->dsol, sm->ndofs)) {
        ref_fail("increment_function: device error or no lin_sys");
        return;
    }
    tl_status = 0;
}

/* get_residual_norms, newton_tools.c:303-373 (single rank): max |r| and max |dsol| over the dofs not in
 * bc_mask (and in include_dof when given), with their first dof; l2 = sqrt(sum r^2) / ndofs as coded.  A
 * NaN in the residual, where the reference exits, gives NaN norms and status -1. */
void get_residual_norms(double *resid_max_norm, double *resid_l2_norm, double *inc_max_norm, int *imax_dof,
                        int *iinc_dof, int *include_dof, int my_ndofs, int ndofs, int macro_ndofs, double *residual,
                        double *dsol, int *bc_mask)
{
    double rmax = 0.0, l2 = 0.0, dmax = 0.0;
    int ir = 0, id = 0, nnan = 0;
    (void)macro_ndofs;
    if (adhk_la_residual_norms(my_ndofs, residual, dsol, bc_mask, include_dof, &rmax, &l2, &dmax, &ir, &id, &nnan)) {
        ref_fail(adhk_last_error());
        return;
    }
    *resid_max_norm = rmax;
    *resid_l2_norm = sqrt(l2) / ndofs;
    *inc_max_norm = dmax;
    *imax_dof = ir;
    *iinc_dof = id;
    if (nnan) {
        *resid_max_norm = *resid_l2_norm = NAN;
        ref_fail("get_residual_norms: NaN in the residual (the reference exits here, newton_tools.c:348)");
        return;
    }
    tl_status = 0;
}

/* fe_newton, fe_newton.c:44-926 (non-PETSc branch), on the device: the Newton iteration, line search, stop
 * rule and failure path of the reference (a failed solve restores sol from sol_old and cuts ts->dt by
 * 0.25, tc_scale.c:3-19).  Returns TRUE (1) or FALSE (0). */
int fe_newton(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e) return 0;
    if (sync_in(sm, e, 1)) return 0;
    adhk_solver s;
    adhk_solver_defaults(&s);
    s.tol_nonlin = sm->tol_nonlin;
    s.inc_nonlin = sm->inc_nonlin;
    s.max_nonlin_it = sm->max_nonlin_it;
    s.max_linesearch_cuts = sm->max_nonlin_linesearch_cuts;
    adhk_newton_stats st;
    const int rc = adhk_fe_newton(e->c, &s, &st);
    if (rc < 0) {
        ref_fail(adhk_last_error());
        return 0;
    }
    if (adhk_sw2_get_state(e->c, sm->sol, NULL, NULL) || (sm->bc_mask && adhk_sw2_get_bc(e->c, sm->bc_mask))) {
        ref_fail(adhk_last_error());
        return 0;
    }
    sm->it_count_nonlin = st.newton_its;
    sm->nonlinear_it_total += st.newton_its;
    if (!rc) {
        sm->ts->dt *= REF_DT_REDUCE;
        sm->it_count_nonlin_failed += st.newton_its;
    }
    tl_status = 0;
    return rc ? 1 : 0;
}

/* ---------------------------------------------------------------------------------------- la.h */

/* the context's system (the last assembled Jacobian and residual) out to lin_sys, with bc_mask */
static int system_out(SMODEL_SUPER *sm, ref_ent *e)
{
    if (adhk_get_system(e->c, sm->lin_sys->vals_diag, sm->lin_sys->residual) ||
        (sm->bc_mask && adhk_sw2_get_bc(e->c, sm->bc_mask)))
        return ref_fail(adhk_last_error());
    return 0;
}

/* check_diag, matrix_utilities.c:120-152, on the system of the last assemble_jacobian / assemble_residual */
void check_diag(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e) return;
    if (sm->bc_mask && sm->dirichlet_data && adhk_sw2_set_bc(e->c, sm->bc_mask, sm->dirichlet_data)) {
        ref_fail(adhk_last_error());
        return;
    }
    if (adhk_check_diag(e->c) || system_out(sm, e)) {
        ref_fail(adhk_last_error());
        return;
    }
    tl_status = 0;
}

/* apply_Dirichlet_BC, matrix_utilities.c:42-105 ("method 2"), on the same system */
void apply_Dirichlet_BC(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e) return;
    if (sm->bc_mask && sm->dirichlet_data && adhk_sw2_set_bc(e->c, sm->bc_mask, sm->dirichlet_data)) {
        ref_fail(adhk_last_error());
        return;
    }
    if (adhk_apply_dirichlet_bc(e->c) || system_out(sm, e)) {
        ref_fail(adhk_last_error());
        return;
    }
    tl_status = 0;
}

/* ---------------------------------------------------------------------------------------- sw2.h */

static int elem_resid(SMODEL_SUPER *mod, double *elem_rhs, int which, int ie, double perturbation, int perturb_node,
                      int perturb_var, int perturb_sign)
{
    ref_ent *e = ctx_of(mod);
    if (!e || sync_in(mod, e, 0)) return -1;
    const int pvar = perturb_var == REF_VAR_H ? 0 : perturb_var == REF_VAR_UDA ? 1 : perturb_var == REF_VAR_VDA ? 2 : -1;
    int wd = -1;
    if (which == 0 && mod->dvars.elem_flags && mod->dvars.elem_flags[0]) wd = mod->dvars.elem_flags[0][ie];
    double r[9];
    if (adhk_elem_residual_one(e->c, which, ie, perturb_sign * perturbation, pvar >= 0 ? perturb_node : 0, pvar, wd,
                               r))
        return ref_fail(adhk_last_error());
    for (int k = 0; k < (which ? 6 : 9); k++) elem_rhs[k] += r[k];
    tl_status = 0;
    return 0;
}

int fe_sw2_body_resid(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                      int perturb_var, int perturb_sign, int DEBUG)
{
    (void)DEBUG;
    elem_resid(mod, elem_rhs, 0, ie, perturbation, perturb_node, perturb_var, perturb_sign);
    return 0;
}
int fe_sw2_bc_h(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node, int perturb_var,
                int perturb_sign, int DEBUG)
{
    (void)DEBUG;
    elem_resid(mod, elem_rhs, 1, ie, perturbation, perturb_node, perturb_var, perturb_sign);
    return 0;
}
int fe_sw2_bc_vel(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                  int perturb_var, int perturb_sign, int DEBUG)
{
    (void)DEBUG;
    elem_resid(mod, elem_rhs, 1, ie, perturbation, perturb_node, perturb_var, perturb_sign);
    return 0;
}
int fe_sw2_bc_flux(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                   int perturb_var, int perturb_sign, int DEBUG)
{
    (void)DEBUG;
    elem_resid(mod, elem_rhs, 1, ie, perturbation, perturb_node, perturb_var, perturb_sign);
    return 0;
}

/* wet/dry flags of every triangle into dvars.elem_flags[0] (the _WDFLAG row), on the GPU */
static int wdflags(SMODEL_SUPER *sm, ref_ent *e)
{
    int *flags = (sm->dvars.elem_flags && sm->dvars.elem_flags[0]) ? sm->dvars.elem_flags[0] : NULL;
    if (sync_in(sm, e, 1) || adhk_sw2_init(e->c, flags)) return ref_fail(adhk_last_error());
    return 0;
}

/* fe_sw2_wdflag_legacy, sw2_wdflag_legacy.c:34-139 */
void fe_sw2_wdflag_legacy(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e || wdflags(sm, e)) return;
    tl_status = 0;
}

/* fe_sw2_init, fe_sw2_init.c:28-53: wet/dry flags, and the velocities of nodes dry at the old level into
 * bc_mask */
int fe_sw2_init(SMODEL_SUPER *sm)
{
    ref_ent *e = ctx_of(sm);
    if (!e || wdflags(sm, e)) return 0;
    if (sm->bc_mask && adhk_sw2_get_bc(e->c, sm->bc_mask)) {
        ref_fail(adhk_last_error());
        return 0;
    }
    tl_status = 0;
    return 0;
}

/* ------------------------------------------------------------------------------------- slin_sys.h */

/* slin_sys_init_sparsity_mono, slin_sys_init_sparsity_mono.c:28-494: the CSR pattern of the grid's
 * triangles and segments, built on the GPU; cols_diag and vals_diag are reallocated to nnz_diag entries and
 * indptr_diag (the caller's, *local_size + 1) is filled.  Materials and ivars: every element carries the
 * three SW2 variables, dof(node, var) = 3 node + var. */
void slin_sys_init_sparsity_mono(SLIN_SYS *lin_sys, int *elem3d_physics_mat_id, int *elem2d_physics_mat_id,
                                 int *elem1d_physics_mat_id, SMAT_PHYSICS *elem_physics_mat, SGRID *grid, int **ivars)
{
    (void)elem3d_physics_mat_id;
    (void)elem2d_physics_mat_id;
    (void)elem1d_physics_mat_id;
    (void)elem_physics_mat;
    (void)ivars;
    if (!lin_sys || !grid || !lin_sys->local_size || !lin_sys->indptr_diag) {
        ref_fail("slin_sys_init_sparsity_mono: NULL argument");
        return;
    }
    const int N = grid->nnodes, E = grid->nelems2d, S = grid->nelems1d;
    if (*lin_sys->local_size != 3 * N || grid->nelems3d > 0) {
        ref_fail("slin_sys_init_sparsity_mono: local_size must be 3 x nnodes of a 2-D grid");
        return;
    }
    double *x = calloc((size_t)N, sizeof(double)), *y = calloc((size_t)N, sizeof(double)),
           *z = calloc((size_t)N, sizeof(double)), *mann = calloc((size_t)(E > 0 ? E : 1), sizeof(double)),
           *val = calloc((size_t)(S > 0 ? 2 * S : 2), sizeof(double));
    int *tri = malloc((size_t)(E > 0 ? 3 * E : 1) * sizeof(int)), *seg = malloc((size_t)(S > 0 ? 2 * S : 1) * sizeof(int)),
        *kind = calloc((size_t)(S > 0 ? S : 1), sizeof(int));
    adhk_sw2 *c = NULL;
    int *ip = NULL, *cols = NULL;
    int ok = x && y && z && mann && val && tri && seg && kind;
    for (int n = 0; ok && n < N; n++) {
        x[n] = grid->node[n].x;
        y[n] = grid->node[n].y;
    }
    for (int i = 0; ok && i < E; i++) {
        if (grid->elem2d[i].nnodes != 3) ok = 0;
        for (int a = 0; ok && a < 3; a++) tri[3 * i + a] = grid->elem2d[i].nodes[a];
    }
    for (int s = 0; ok && s < S; s++) {
        seg[2 * s] = grid->elem1d[s].nodes[0];
        seg[2 * s + 1] = grid->elem1d[s].nodes[1];
        val[2 * s] = 1.0;
    }
    if (ok) {
        adhk_params p;
        memset(&p, 0, sizeof p);
        p.g = 9.8;
        p.dt = 1.0;
        p.manning_units = 1.0;
        p.fric_mode = ADHK_FRIC_CONST;
        c = adhk_sw2_create(N, x, y, z, E, tri, mann, S, seg, kind, val, &p);
    }
    int ndofs = 0, nnz = 0, nnzb = 0;
    if (c && !adhk_sw2_sizes(c, &ndofs, &nnz, &nnzb)) {
        ip = malloc(((size_t)ndofs + 1) * sizeof(int));
        cols = malloc((size_t)nnz * sizeof(int));
        int *nc = ip && cols ? realloc(lin_sys->cols_diag, (size_t)nnz * sizeof(int)) : NULL;
        if (nc) lin_sys->cols_diag = nc;
        double *nv = nc ? realloc(lin_sys->vals_diag, (size_t)nnz * sizeof(double)) : NULL;
        if (nv) lin_sys->vals_diag = nv;
        if (nc && nv && !adhk_sparsity(c, ip, cols)) {
            memcpy(lin_sys->indptr_diag, ip, ((size_t)ndofs + 1) * sizeof(int));
            memcpy(lin_sys->cols_diag, cols, (size_t)nnz * sizeof(int));
            memset(lin_sys->vals_diag, 0, (size_t)nnz * sizeof(double));
            lin_sys->nnz_diag = nnz;
            tl_status = 0;
        } else {
            ref_fail("slin_sys_init_sparsity_mono: out of memory or device error");
        }
    } else {
        ref_fail(ok ? adhk_last_error() : "slin_sys_init_sparsity_mono: only triangles, or out of host memory");
    }
    if (c) adhk_sw2_destroy(c);
    free(x); free(y); free(z); free(mann); free(val); free(tri); free(seg); free(kind); free(ip); free(cols);
}

/* ------------------------------------------------------------------------------------- dofmaps.h */

/* get_cell_dofs_ivars, cg_maps.c:26-45 */
void get_cell_dofs_ivars(int *local_dofs, int **ivars, int nnodes, int *local_node_ids, int elem_nvars,
                         int *elem_var_pos)
{
    int k = 0;
    for (int i = 0; i < nnodes; i++)
        for (int j = 0; j < elem_nvars; j++) local_dofs[k++] = ivars[elem_var_pos[j]][local_node_ids[i]];
}

/* global_to_local_dbl_ivars, cg_maps.c:67-72 */
void global_to_local_dbl_ivars(double *local_vals, int *nodeIDs, int nnodes, int *map_array, double *global_vals)
{
    for (int i = 0; i < nnodes; i++) local_vals[i] = global_vals[map_array[nodeIDs[i]]];
}

/* global_to_local_SVECT2D_ivars, cg_maps.c:96-104 */
void global_to_local_SVECT2D_ivars(SVECT2D *local_vals, int *nodeIDs, int nnodes, int **ivars, int varx, int vary,
                                   double *global_vals)
{
    for (int i = 0; i < nnodes; i++) {
        local_vals[i].x = global_vals[ivars[varx][nodeIDs[i]]];
        local_vals[i].y = global_vals[ivars[vary][nodeIDs[i]]];
    }
}

/* local_dofs_to_global_dofs, general_maps.c:21-36 */
void local_dofs_to_global_dofs(int *global_dofs, int ndofs_on_ele, int *dofs, int *local_range, int *ghosts)
{
    const int local_size = local_range[1] - local_range[0];
    for (int i = 0; i < ndofs_on_ele; i++)
        global_dofs[i] = dofs[i] < local_size ? dofs[i] + local_range[0] : ghosts[dofs[i] - local_size];
}

/* global_to_local, general_maps.c:55-71 */
int global_to_local(int global_col_no, int local_size, int *ghosts, int nghost)
{
    for (int i = 0; i < nghost; i++)
        if (ghosts[i] == global_col_no) return i + local_size;
    return -1;
}

/* ------------------------------------------------------------------------------- sivar_position.h */

#define REF_N_IVARS 25 /* adh_def.n_ivars = MAX_IVARS + MAX_TRNS, sadh_def.c:26 and sadh_def.h:5-6 */

/* sivar_position_build_dof_map, sivar_position.c:61-80: rows ip->var[j] of the active variables j of ipNode
 * get the node's dofs in node-major order, every other entry UNSET_INT */
int sivar_position_build_dof_map(SIVAR_POSITION *ip, int nnodes, SIVAR_POSITION *ipNode, int ***ivars)
{
    if (!ip || !ipNode || !ivars || nnodes < 0 || ip->n < 0 || !ipNode->var || !ip->var) {
        ref_fail("sivar_position_build_dof_map: NULL or negative argument");
        return -1;
    }
    int **iv = (int **)calloc((size_t)(ip->n > 0 ? ip->n : 1), sizeof(int *));
    int bad = !iv;
    for (int r = 0; !bad && r < ip->n; r++) {
        if (!(iv[r] = (int *)malloc((size_t)(nnodes > 0 ? nnodes : 1) * sizeof(int)))) bad = 1;
        for (int n = 0; !bad && n < nnodes; n++) iv[r][n] = REF_UNSET_INT;
    }
    for (int j = 0; !bad && j < REF_N_IVARS; j++)
        if (ipNode->var[j] != REF_UNSET_INT && (ip->var[j] < 0 || ip->var[j] >= ip->n)) bad = 2;
    if (bad) {
        for (int r = 0; iv && r < ip->n; r++) free(iv[r]);
        free(iv);
        ref_fail(bad == 2 ? "sivar_position_build_dof_map: an active variable of ipNode has no row in ip"
                          : "sivar_position_build_dof_map: out of host memory");
        return -1;
    }
    int ndofs = 0;
    for (int n = 0; n < nnodes; n++)
        for (int j = 0; j < REF_N_IVARS; j++)
            if (ipNode->var[j] != REF_UNSET_INT) iv[ip->var[j]][n] = ndofs++;
    *ivars = iv;
    tl_status = 0;
    return ndofs;
}

/* ----------------------------------------------------------------------------------------- messg.h */

/* messg_dmax.c:10, messg_dsum.c:9, messg_imin.c:6 and their kin without _MESSG: one rank, the identity */
double messg_dmax(double x) { return x; }
double messg_dmin(double x) { return x; }
int messg_imin(int x) { return x; }
int messg_imax(int x) { return x; }
double messg_dsum(double x) { return x; }

/* ------------------------------------------------------------------------------------ friction */

/* fr_bedshstr_drag_coef, fr_bedshstr_drag_coef.c:18-43: c_f = 2 (0.4 b / ((b + 1)(ln(b + 1) - 1) + 1))^2,
 * b = 29.7 depth / k_s, with depth and k_s clamped at 1e-8 and depth / k_s at 1 / 29.7 from below */
double fr_bedshstr_drag_coef(double depth, double roughness_height)
{
    depth = fmax(depth, 1.e-8);
    roughness_height = fmax(roughness_height, 1.e-8);
    const double ratio = roughness_height > 29.7 * depth ? 1.0 / 29.7 : depth / roughness_height;
    const double b = 29.7 * ratio;
    const double q = 0.4 * b / ((b + 1.0) * (log(b + 1.0) - 1.0) + 1.0);
    return 2.0 * q * q;
}

/* fr_manningsn_to_rheight, fr_manningsn_to_rheight.c:15-32: k_s = (8.25 sqrt(g) n / k_n)^6, 0 below n = 1e-8 */
double fr_manningsn_to_rheight(double n, double kn, double g)
{
    return n < 1.e-8 ? 0.0 : pow(8.25 * sqrt(g) * n / kn, 6.0);
}

#define REF_UNSET_FLT (-9999999.9) /* UNSET_FLT, define.h:443 */

/* fe_sw2_get_roughness, sw2_get_roughness.c:25-49: the drag coefficient of string id from its roughness
 * height (the bed-shear law; zero without a height).  The vegetation and ice laws of that function are not
 * on this path. */
double fe_sw2_get_roughness(SMODEL_SUPER *mod, double depth, double velocity, int id, int which)
{
    (void)velocity;
    (void)which;
    if (!mod || !mod->str_values || id < 0) {
        ref_fail("fe_sw2_get_roughness: no str_values or a negative string id");
        return NAN;
    }
    const FRICTION *f = &mod->str_values[id].fterms;
    if (f->sav_flag == REF_YES || f->urv_flag == REF_YES || f->icethick != REF_UNSET_FLT) {
        ref_fail("fe_sw2_get_roughness: vegetation and ice drag (fr_sav_drag_coef, fr_urv_drag_coef, "
                 "fr_stationary_ice_coef) are not on the SW2 path this library serves");
        return NAN;
    }
    tl_status = 0;
    return f->eqrheight > DBL_EPSILON ? fr_bedshstr_drag_coef(depth, f->eqrheight) : 0.0;
}

/* ------------------------------------------------------------------------------------ sseries.h */

#define REF_TIME_SERIES 1 /* TIME_SERIES, define.h:492 */

/* sseries_get_interval, sseries.c:746-775: the left end of the piece holding t, or -1 outside the range
 * widened by the relative DBL_EPSILON */
static int ser_interval(const SSERIES *s, double t)
{
    if (!s->entry || s->size < 1) return -1;
    const int last = s->size - 1;
    if (t < s->entry[0].time * (1.0 - DBL_EPSILON) || t > s->entry[last].time * (1.0 + DBL_EPSILON)) return -1;
    int i = 0;
    while (i < last && t >= s->entry[i].time * (1.0 - DBL_EPSILON)) i++;
    return i - 1;
}

/* tc_eval_series, tc_eval_series.c: theta-weighted value on the piece (the last entry's value past it) */
static double ser_value(const SSERIES *s, int interval, double t, int index)
{
    if (interval == s->size - 1) return s->entry[s->size - 1].value[index];
    const double tl = s->entry[interval].time, tr = s->entry[interval + 1].time;
    const double theta = (tr - t) / (tr - tl);
    return theta * s->entry[interval].value[index] + (1.0 - theta) * s->entry[interval + 1].value[index];
}

/* tc_trap_area, tc_trap_area.c */
static double ser_trap(double lv, double rv, double lt, double rt)
{
    const double d = rt - lt;
    return d < DBL_EPSILON ? 0.0 : (lv + rv) * 0.5 * d;
}

/* sseries_set_values_average, sseries.c:622-668 */
static int ser_set_average(SSERIES *s, double t1, double t2)
{
    const int i1 = ser_interval(s, t1), i2 = ser_interval(s, t2);
    if (i1 < 0 || i2 < 0 || !s->ivalue) return -1;
    for (int k = 0; k < s->nvalues; k++) {
        const double v1 = ser_value(s, i1, t1, k), v2 = ser_value(s, i2, t2, k);
        double area = 0.0;
        if (i1 == i2) {
            area = ser_trap(v1, v2, t1, t2);
        } else {
            for (int i = i1 + 1; i < i2; i++) area += s->entry[i].area[k];
            area += ser_trap(v1, s->entry[i1 + 1].value[k], t1, s->entry[i1 + 1].time);
            area += ser_trap(s->entry[i2].value[k], v2, s->entry[i2].time, t2);
        }
        s->ivalue[k] = area / (t2 - t1);
    }
    return 0;
}

/* sseries_setall_ts_valuesAVG, sseries.c:673-682 */
void sseries_setall_ts_valuesAVG(SSERIES *head, double t1, double t2)
{
    for (SSERIES *s = head; s; s = s->next)
        if (s->type == REF_TIME_SERIES && ser_set_average(s, t1, t2)) {
            char msg[160];
            snprintf(msg, sizeof msg, "series %d: time %.6e or %.6e is not in the range of the series", s->id + 1,
                     t1, t2);
            ref_fail(msg);
            return;
        }
    tl_status = 0;
}

/* ---------------------------------------------------------------------------------- time loop */

#define REF_FE_NEWTON 0 /* FE_NEWTON, model_codes.h:77 */

/* smodel_super_prep_sol, smodel_super_prep_sol.c:2-8: sol_older <- sol_old <- sol */
void smodel_super_prep_sol(SMODEL_SUPER *sm)
{
    const size_t nb = (size_t)sm->ndofs * sizeof(double);
    memcpy(sm->sol_older, sm->sol_old, nb);
    memcpy(sm->sol_old, sm->sol, nb);
}

/* smodel_super_update_dirichlet_data, smodel_super_update_dirichlet_data.c:2-28: the reference sets
 * Dirichlet data only for its HEAT2D test model; an SW2 model's stay as they are */
void smodel_super_update_dirichlet_data(SMODEL_SUPER *sm) { (void)sm; }

/* tc_init, tc_init.c:4-33: dt from the dt series at t_prev, shortened to reach t_final exactly (or halved
 * when two steps would overshoot it) */
void tc_init(SMODEL_DESIGN *dm)
{
    dm->ts.dt_old = dm->ts.dt;
    const int i = dm->series_dt ? ser_interval(dm->series_dt, dm->ts.t_prev) : -1;
    if (i < 0) {
        ref_fail("tc_init: no dt series, or t_prev is not in its range");
        return;
    }
    dm->ts.dt = ser_value(dm->series_dt, i, dm->ts.t_prev, 0);
    if (dm->ts.t_prev + dm->ts.dt >= dm->ts.t_final)
        dm->ts.dt = dm->ts.t_final - dm->ts.t_prev;
    else if (dm->ts.t_prev + 2.0 * dm->ts.dt >= dm->ts.t_final)
        dm->ts.dt = 0.5 * (dm->ts.t_final - dm->ts.t_prev);
    if (dm->ts.dt < DBL_EPSILON) {
        ref_fail("tc_init: time step is too small");
        return;
    }
    tl_status = 0;
}

/* update_dt, time_loop.c:108-120: t_prev advances by dt; FALSE once t_final is reached */
int update_dt(SMODEL_DESIGN *dm)
{
    dm->ts.t_prev += dm->ts.dt;
    return dm->ts.t_prev + FLT_EPSILON >= dm->ts.t_final ? 0 : 1;
}

/* advance_time, time_loop.c:56-93 */
int advance_time(SMODEL_DESIGN *dm, int nsuper)
{
    tl_status = 0;
    for (int k = 0; k < nsuper; k++) {
        SMODEL_SUPER *sm = &dm->superModel[k];
        if (sm->forward_step != REF_FE_NEWTON) {
            ref_fail("advance_time: the super model's forward_step is not FE_NEWTON");
            return 0;
        }
        for (int j = 0; j < sm->nsubsteps; j++) {
            smodel_super_prep_sol(sm);
            sseries_setall_ts_valuesAVG(sm->series_head, sm->ts->t_prev - sm->ts->dt, sm->ts->t_prev);
            if (tl_status) return 0;
            smodel_super_update_dirichlet_data(sm);
            if (!fe_newton(sm)) return 0;
        }
    }
    return 1;
}

/* time_loop, time_loop.c:14-41: YES (1) as the reference returns whether or not every step worked; 0 when
 * this library failed (adhk_ref_status() -1 and adhk_last_error()) */
int time_loop(SMODEL_DESIGN *dm)
{
    if (!dm || dm->nSuperModels < 0 || (dm->nSuperModels > 0 && !dm->superModel)) {
        ref_fail("time_loop: no design model");
        return 0;
    }
    tc_init(dm);
    if (tl_status) return 0;
    int more, worked;
    do {
        more = update_dt(dm);
        worked = advance_time(dm, dm->nSuperModels);
        if (tl_status) return 0;
    } while (more && worked);
    return REF_YES;
}
