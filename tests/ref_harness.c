/* Test harness of the struct-level surface (include/adhk_ref.h): builds the reference's structures for a
 * mesh as a reference caller would -- SGRID with SNODE / SELEM_2D / SELEM_1D, SMODEL_SUPER with its SDT,
 * SLIN_SYS, physics materials (a body material carrying fe_sw2_body_resid and fe_sw2_init, a depth and a
 * flux boundary material), the dof map ivars and the boundary series -- and calls libadhk.so through the
 * reference's function names.  Compiled by tests/test_gpu_ref_surface.py; called there through ctypes. */
#include <stdlib.h>
#include <string.h>

#include "adhk_ref.h"

typedef struct {
    SMODEL_SUPER sm;
    SGRID grid;
    SDT ts;
    SLIN_SYS ls;
    SMAT_PHYSICS mat[3];
    SPDE pde[3];
    SSERIES ser[2];
    double ival[2][2];
    int ndofs, local_size, var[6], var_code[3], *ivars[3], *ivar_store, *e2d_mat, *e1d_mat, *wd, *wd_rows[1];
    int *tri_nodes, *seg_nodes;
} H;

/* kind: 0 depth (series 0), 1 explicit flux (series 1); values: ser0 = the depth of every depth segment,
 * ser1 = the flux per unit width */
H *h_create(int N, const double *x, const double *y, const double *z, int E, const int *tri, int S, const int *seg,
            const int *kind, double ser0, double ser1, double dt, double g)
{
    H *h = calloc(1, sizeof *h);
    h->grid.nnodes = N;
    h->grid.nelems2d = E;
    h->grid.nelems1d = S;
    h->grid.ndim = 2;
    h->grid.node = calloc((size_t)N, sizeof(SNODE));
    h->grid.elem2d = calloc((size_t)E, sizeof(SELEM_2D));
    h->grid.elem1d = calloc((size_t)(S > 0 ? S : 1), sizeof(SELEM_1D));
    h->tri_nodes = malloc((size_t)3 * E * sizeof(int));
    h->seg_nodes = malloc((size_t)2 * (S > 0 ? S : 1) * sizeof(int));
    for (int n = 0; n < N; n++) {
        h->grid.node[n].id = n;
        h->grid.node[n].x = x[n];
        h->grid.node[n].y = y[n];
        h->grid.node[n].z = z[n];
    }
    h->e2d_mat = calloc((size_t)E, sizeof(int));
    h->e1d_mat = calloc((size_t)(S > 0 ? S : 1), sizeof(int));
    for (int e = 0; e < E; e++) {
        memcpy(h->tri_nodes + 3 * e, tri + 3 * e, 3 * sizeof(int));
        h->grid.elem2d[e].id = e;
        h->grid.elem2d[e].nnodes = 3;
        h->grid.elem2d[e].nodes = h->tri_nodes + 3 * e;
    }
    for (int s = 0; s < S; s++) {
        memcpy(h->seg_nodes + 2 * s, seg + 2 * s, 2 * sizeof(int));
        h->grid.elem1d[s].id = s;
        h->grid.elem1d[s].nnodes = 2;
        h->grid.elem1d[s].nodes = h->seg_nodes + 2 * s;
        h->e1d_mat[s] = kind[s] == 0 ? 1 : 2;
    }
    /* materials: 0 body, 1 depth boundary, 2 flux boundary; the SW2 variables h, uda, vda (codes 0, 4, 5) */
    for (int i = 0; i < 6; i++) h->var[i] = -3;
    h->var[0] = 0;
    h->var[4] = 1;
    h->var[5] = 2;
    h->var_code[0] = 0;
    h->var_code[1] = 4;
    h->var_code[2] = 5;
    for (int m = 0; m < 3; m++) {
        h->mat[m].id = m;
        h->mat[m].ivar_pos.n = 3;
        h->mat[m].ivar_pos.var = h->var;
        h->mat[m].ivar_pos.var_code = h->var_code;
        h->mat[m].npdes = 1;
        h->mat[m].pde = &h->pde[m];
    }
    h->pde[0].resid = fe_sw2_body_resid;
    h->pde[0].init = fe_sw2_init;
    h->pde[1].resid = fe_sw2_bc_h;
    h->pde[2].resid = fe_sw2_bc_flux;
    h->pde[2].iseries = 1;
    /* dof map: ivars[var][node] = 3 node + var */
    h->ndofs = 3 * N;
    h->ivar_store = malloc((size_t)3 * N * sizeof(int));
    for (int v = 0; v < 3; v++) {
        h->ivars[v] = h->ivar_store + (size_t)v * N;
        for (int n = 0; n < N; n++) h->ivars[v][n] = 3 * n + v;
    }
    /* series 0 (depth) -> 1 (flux) */
    h->ival[0][0] = ser0;
    h->ival[1][0] = ser1;
    h->ser[0].id = 0;
    h->ser[0].ivalue = h->ival[0];
    h->ser[0].next = &h->ser[1];
    h->ser[1].id = 1;
    h->ser[1].ivalue = h->ival[1];
    h->ser[1].prev = &h->ser[0];
    /* time, linear system, model */
    h->ts.dt = dt;
    h->local_size = 3 * N;
    h->ls.local_size = &h->local_size;
    h->ls.size = &h->local_size;
    h->ls.global_size = &h->local_size;
    h->ls.local_range[1] = 3 * N;
    h->ls.indptr_diag = malloc(((size_t)3 * N + 1) * sizeof(int));
    h->ls.residual = calloc((size_t)3 * N, sizeof(double));
    h->ls.dsol = calloc((size_t)3 * N, sizeof(double));
    h->ls.scale_vect = calloc((size_t)3 * N, sizeof(double));
    h->wd = calloc((size_t)E, sizeof(int));
    h->wd_rows[0] = h->wd;
    SMODEL_SUPER *sm = &h->sm;
    sm->grid = &h->grid;
    sm->ts = &h->ts;
    sm->nsubsteps = 1;
    sm->lin_sys = &h->ls;
    sm->tol_nonlin = 1e-8;
    sm->inc_nonlin = 1e-6;
    sm->max_nonlin_it = 10;
    sm->max_nonlin_linesearch_cuts = 5;
    sm->ndofs = sm->my_ndofs = sm->macro_ndofs = 3 * N;
    sm->sol = calloc((size_t)3 * N, sizeof(double));
    sm->sol_old = calloc((size_t)3 * N, sizeof(double));
    sm->sol_older = calloc((size_t)3 * N, sizeof(double));
    sm->bc_mask = calloc((size_t)3 * N, sizeof(int));
    sm->dirichlet_data = calloc((size_t)3 * N, sizeof(double));
    sm->elem2d_physics_mat = h->e2d_mat;
    sm->elem1d_physics_mat = h->e1d_mat;
    sm->nmat_physics = 3;
    sm->mat_physics_elem = h->mat;
    sm->ivar_pos.n = 3;
    sm->ivar_pos.var = h->var;
    sm->ivar_pos.var_code = h->var_code;
    sm->ivars = h->ivars;
    sm->series_head = &h->ser[0];
    sm->gravity = g;
    sm->dvars.elem_flags = h->wd_rows;
    return h;
}

int h_set_manning(H *h, const double *n) { return adhk_ref_set_manning(&h->sm, n); }

void h_set_state(H *h, const double *s, const double *so, const double *soo)
{
    memcpy(h->sm.sol, s, (size_t)h->ndofs * sizeof(double));
    memcpy(h->sm.sol_old, so, (size_t)h->ndofs * sizeof(double));
    memcpy(h->sm.sol_older, soo, (size_t)h->ndofs * sizeof(double));
}

void h_set_bc(H *h, const int *bc_mask, const double *dirichlet)
{
    memcpy(h->sm.bc_mask, bc_mask, (size_t)h->ndofs * sizeof(int));
    memcpy(h->sm.dirichlet_data, dirichlet, (size_t)h->ndofs * sizeof(double));
}

void h_get(H *h, double *sol, double *residual, int *bc_mask, double *dt, int *wd)
{
    if (sol) memcpy(sol, h->sm.sol, (size_t)h->ndofs * sizeof(double));
    if (residual) memcpy(residual, h->ls.residual, (size_t)h->ndofs * sizeof(double));
    if (bc_mask) memcpy(bc_mask, h->sm.bc_mask, (size_t)h->ndofs * sizeof(int));
    if (dt) *dt = h->ts.dt;
    if (wd) memcpy(wd, h->wd, (size_t)h->grid.nelems2d * sizeof(int));
}

/* the pattern: nnz, or -1 */
int h_sparsity(H *h)
{
    slin_sys_init_sparsity_mono(&h->ls, NULL, h->e2d_mat, h->e1d_mat, h->mat, &h->grid, h->ivars);
    return adhk_ref_status() ? -1 : h->ls.nnz_diag;
}
void h_pattern(H *h, int *indptr, int *cols, double *vals)
{
    memcpy(indptr, h->ls.indptr_diag, ((size_t)h->ndofs + 1) * sizeof(int));
    if (cols) memcpy(cols, h->ls.cols_diag, (size_t)h->ls.nnz_diag * sizeof(int));
    if (vals) memcpy(vals, h->ls.vals_diag, (size_t)h->ls.nnz_diag * sizeof(double));
}

int h_init_system(H *h)
{
    initialize_system(&h->sm);
    initialize_dirichlet_bc(&h->sm);
    return adhk_ref_status();
}
int h_residual(H *h)
{
    assemble_residual(&h->sm, &h->grid);
    return adhk_ref_status();
}
int h_jacobian(H *h)
{
    assemble_jacobian(&h->sm);
    return adhk_ref_status();
}
int h_check_diag_dirichlet(H *h)
{
    check_diag(&h->sm);
    if (adhk_ref_status()) return -1;
    apply_Dirichlet_BC(&h->sm);
    return adhk_ref_status();
}

/* one time step as time_loop/advance_time take it: prep_sol (sol_older <- sol_old <- sol), then fe_newton */
int h_step(H *h, int *newton_its)
{
    const size_t nb = (size_t)h->ndofs * sizeof(double);
    memcpy(h->sm.sol_older, h->sm.sol_old, nb);
    memcpy(h->sm.sol_old, h->sm.sol, nb);
    const int ok = fe_newton(&h->sm);
    if (newton_its) *newton_its = h->sm.it_count_nonlin;
    return ok;
}

/* elemental residual of triangle (which 0) or segment (which 1) ie through the material's resid pointer */
int h_elem_resid(H *h, int which, int ie, double pert, int pnode, int pvar_code, int psign, double *out)
{
    memset(out, 0, (which ? 6 : 9) * sizeof(double));
    const int mat = which ? h->e1d_mat[ie] : h->e2d_mat[ie];
    h->pde[mat].resid(&h->sm, out, ie, pert, pnode, pvar_code, psign, 0);
    return adhk_ref_status();
}

/* the elemental FD Jacobian of triangle ie by the reference's load_elem_mat (row-major 9 x 9) */
int h_elem_mat(H *h, int ie, double *out)
{
    double *rows[9];
    for (int i = 0; i < 9; i++) rows[i] = out + 9 * i;
    memset(out, 0, 81 * sizeof(double));
    load_elem_mat(&h->sm, 3, h->grid.elem2d[ie].nodes, 0, ie, rows);
    return adhk_ref_status();
}

int h_norms(H *h, const double *residual, const double *dsol, double *out3, int *idx2)
{
    int im = 0, ii = 0;
    get_residual_norms(&out3[0], &out3[1], &out3[2], &im, &ii, NULL, h->ndofs, h->ndofs, h->ndofs,
                       (double *)residual, (double *)dsol, h->sm.bc_mask);
    idx2[0] = im;
    idx2[1] = ii;
    return adhk_ref_status();
}

int h_increment(H *h, const double *dsol)
{
    memcpy(h->ls.dsol, dsol, (size_t)h->ndofs * sizeof(double));
    increment_function(&h->sm);
    return adhk_ref_status();
}

void h_destroy(H *h)
{
    adhk_ref_release(&h->sm);
    free(h->grid.node);
    free(h->grid.elem2d);
    free(h->grid.elem1d);
    free(h->tri_nodes);
    free(h->seg_nodes);
    free(h->e2d_mat);
    free(h->e1d_mat);
    free(h->ivar_store);
    free(h->ls.indptr_diag);
    free(h->ls.cols_diag);
    free(h->ls.vals_diag);
    free(h->ls.residual);
    free(h->ls.dsol);
    free(h->ls.scale_vect);
    free(h->wd);
    free(h->sm.sol);
    free(h->sm.sol_old);
    free(h->sm.sol_older);
    free(h->sm.bc_mask);
    free(h->sm.dirichlet_data);
    free(h);
}
