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
    SMODEL_DESIGN dm;       /* h_design: the design model whose one super model is sm */
    SSERIES ser_dt;         /* its dt series */
    SSERIES_ENTRY ent[2][64]; /* entries of ser_dt (0) and of the flux series (1) */
    double ent_vals[2][64][3]; /* value, slope, area of each entry */
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

/* the time loop through SMODEL_DESIGN: series 1 (the flux per unit width) becomes a TIME_SERIES of the n
 * points (tq, qv), with the entry areas the reference's series reader computes; the dt series is constant
 * dt over [t0, tf]; sm->ts points at the design's clock as smodel_design_init_no_read sets it */
static void fill_entries(SSERIES_ENTRY *e, double (*vals)[3], int n, const double *t, const double *v)
{
    for (int i = 0; i < n; i++) {
        e[i].time = t[i];
        e[i].value = &vals[i][0];
        e[i].slope = &vals[i][1];
        e[i].area = &vals[i][2];
        vals[i][0] = v[i];
    }
    for (int i = 0; i + 1 < n; i++) {
        vals[i][1] = (v[i + 1] - v[i]) / (t[i + 1] - t[i]);
        vals[i][2] = (v[i] + v[i + 1]) * 0.5 * (t[i + 1] - t[i]);
    }
}

int h_design(H *h, double t0, double tf, double dt, int n, const double *tq, const double *qv)
{
    if (n < 2 || n > 64) return -1;
    const double tt[2] = {t0, tf}, dd[2] = {dt, dt};
    fill_entries(h->ent[0], h->ent_vals[0], 2, tt, dd);
    fill_entries(h->ent[1], h->ent_vals[1], n, tq, qv);
    h->ser_dt.size = 2;
    h->ser_dt.nvalues = 1;
    h->ser_dt.entry = h->ent[0];
    h->ser[1].type = 1; /* TIME_SERIES */
    h->ser[1].size = n;
    h->ser[1].nvalues = 1;
    h->ser[1].entry = h->ent[1];
    memset(&h->dm, 0, sizeof h->dm);
    h->dm.nSuperModels = 1;
    h->dm.superModel = &h->sm;
    h->dm.ts.t_init = h->dm.ts.t_prev = t0;
    h->dm.ts.t_final = tf;
    h->dm.series_dt = &h->ser_dt;
    h->sm.ts = &h->dm.ts;
    h->sm.forward_step = 0; /* FE_NEWTON */
    return 0;
}

/* time_loop(&dm): its return value; the clock and the Newton iterations it took */
int h_time_loop(H *h, double *t_prev, double *dt, int *newton_total)
{
    h->sm.nonlinear_it_total = 0;
    const int rc = time_loop(&h->dm);
    *t_prev = h->dm.ts.t_prev;
    *dt = h->dm.ts.dt;
    *newton_total = h->sm.nonlinear_it_total;
    return adhk_ref_status() ? -1 : rc;
}

/* ---- host-only helpers (no GPU): series averages, the clock, friction, the dof map, prep_sol */

/* a list of two TIME_SERIES over the same n times (values v0 and v1) averaged over (t1, t2) */
int hs_series_avg(int n, const double *t, const double *v0, const double *v1, double t1, double t2, double *out2)
{
    static SSERIES_ENTRY e[2][64];
    static double vals[2][64][3];
    if (n < 1 || n > 64) return -2;
    SSERIES s[3];
    double iv[3] = {0.0, 0.0, 7.0};
    memset(s, 0, sizeof s);
    fill_entries(e[0], vals[0], n, t, v0);
    fill_entries(e[1], vals[1], n, t, v1);
    for (int k = 0; k < 3; k++) {
        s[k].id = k;
        s[k].type = k < 2 ? 1 : 0; /* the third is not a time series and keeps its value */
        s[k].size = n;
        s[k].nvalues = 1;
        s[k].ivalue = &iv[k];
        s[k].entry = k < 2 ? e[k] : e[0];
        s[k].next = k < 2 ? &s[k + 1] : NULL;
    }
    sseries_setall_ts_valuesAVG(&s[0], t1, t2);
    out2[0] = iv[0];
    out2[1] = iv[1];
    return adhk_ref_status() ? -1 : (iv[2] == 7.0 ? 0 : -3);
}

/* tc_init, then update_dt until it reports t_final reached: the dt, the number of steps and the end time */
int hs_clock(double t0, double tf, double dt, double *dt_out, int *nsteps, double *t_end)
{
    SSERIES_ENTRY e[2];
    double vals[2][3];
    const double tt[2] = {t0, tf}, dd[2] = {dt, dt};
    SSERIES s;
    SMODEL_DESIGN dm;
    memset(&s, 0, sizeof s);
    memset(&dm, 0, sizeof dm);
    fill_entries(e, vals, 2, tt, dd);
    s.size = 2;
    s.nvalues = 1;
    s.entry = e;
    dm.ts.t_init = dm.ts.t_prev = t0;
    dm.ts.t_final = tf;
    dm.series_dt = &s;
    tc_init(&dm);
    if (adhk_ref_status()) return -1;
    *dt_out = dm.ts.dt;
    int k = 0, more;
    do {
        more = update_dt(&dm);
        k++;
    } while (more && k < 1000000);
    *nsteps = k;
    *t_end = dm.ts.t_prev;
    return 0;
}

double hs_roughness(double depth, double eqrheight, int sav)
{
    STR_VALUE sv[2];
    SMODEL_SUPER sm;
    memset(sv, 0, sizeof sv);
    memset(&sm, 0, sizeof sm);
    for (int k = 0; k < 2; k++) sv[k].fterms.icethick = -9999999.9; /* UNSET_FLT */
    sv[1].fterms.eqrheight = eqrheight;
    sv[1].fterms.sav_flag = sav ? 1 : 0;
    sm.str_values = sv;
    return fe_sw2_get_roughness(&sm, depth, 0.0, 1, 0);
}

/* the SW2 dof map of nnodes nodes (variables h, uda, vda active: codes 0, 4, 5) into out[3][nnodes] */
int hs_dof_map(int nnodes, int *out)
{
    int var[25], varn[25];
    for (int j = 0; j < 25; j++) var[j] = varn[j] = -3;
    var[0] = 0;
    var[4] = 1;
    var[5] = 2;
    varn[0] = varn[4] = varn[5] = 1;
    SIVAR_POSITION ip = {3, 0, var, NULL}, ipn = {3, 0, varn, NULL};
    int **iv = NULL;
    const int nd = sivar_position_build_dof_map(&ip, nnodes, &ipn, &iv);
    if (nd < 0) return -1;
    for (int r = 0; r < 3; r++) {
        memcpy(out + (size_t)r * nnodes, iv[r], (size_t)nnodes * sizeof(int));
        free(iv[r]);
    }
    free(iv);
    return nd;
}

/* time_loop's refusals, before any device work: which 0 = a dt series that starts after t_prev (tc_init),
 * 1 = a super model whose forward_step is not FE_NEWTON (advance_time); returns time_loop's value */
int hs_time_loop_refusal(int which)
{
    SSERIES_ENTRY e[2];
    double vals[2][3];
    const double tt[2] = {which == 0 ? 5.0 : 0.0, 10.0}, dd[2] = {1.0, 1.0};
    SSERIES s;
    SMODEL_DESIGN dm;
    SMODEL_SUPER sm;
    memset(&s, 0, sizeof s);
    memset(&dm, 0, sizeof dm);
    memset(&sm, 0, sizeof sm);
    fill_entries(e, vals, 2, tt, dd);
    s.size = 2;
    s.nvalues = 1;
    s.entry = e;
    dm.ts.t_init = dm.ts.t_prev = 0.0;
    dm.ts.t_final = 10.0;
    dm.series_dt = &s;
    dm.nSuperModels = 1;
    dm.superModel = &sm;
    sm.ts = &dm.ts;
    sm.nsubsteps = 1;
    sm.forward_step = 3;
    return time_loop(&dm);
}

/* smodel_super_prep_sol on three host arrays */
void hs_prep_sol(int n, double *sol, double *sol_old, double *sol_older)
{
    SMODEL_SUPER sm;
    memset(&sm, 0, sizeof sm);
    sm.ndofs = n;
    sm.sol = sol;
    sm.sol_old = sol_old;
    sm.sol_older = sol_older;
    smodel_super_prep_sol(&sm);
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
