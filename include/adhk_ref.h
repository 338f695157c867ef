/*
 * adhk_ref.h -- the reference's struct-based call surface of the SW2 Newton step, as libadhk.so exports it.
 *
 * Part 1 declares the structures that AdH-Kraken passes along this path, laid out field for field as the
 * reference declares them, so that a caller compiled against the reference's own headers (adh.h) can pass
 * its SMODEL_SUPER / SGRID / SLIN_SYS to these functions unchanged.  Layout configuration: the default
 * single-rank build (no _MESSG, _PETSC, _HDF5 or _ADH_GROUNDWATER).  tests/test_ref_layout.py checks every
 * offset and size below against a table generated from the reference's headers.  Types that are only
 * carried as pointers on this path stay incomplete.
 *
 * Part 2 declares the functions with the reference's names and signatures (residual.h, jacobian.h,
 * newton.h, la.h, sw2.h, slin_sys.h, dofmaps.h, time_loop.h, messg.h and the series, friction and dof-map
 * helpers the time loop and the residual call).  They run on the GPU through the adhk_* context of
 * include/adhk.h, created for each SMODEL_SUPER on its first call (from sm->grid) and kept in a side table
 * keyed by the SMODEL_SUPER pointer; the caller's host arrays (sol, sol_old, sol_older, bc_mask,
 * dirichlet_data, lin_sys->residual / dsol / vals_diag) stay authoritative: each call copies what it reads
 * in and what it writes out.  Errors print a message (adhk_last_error) and, as the reference's tl_error,
 * end in the caller-visible failure value: FALSE from fe_newton, a non-zero status elsewhere (see each
 * function).
 */
#ifndef ADHK_REF_H
#define ADHK_REF_H

#include <stdbool.h>
#include <stdio.h> /* FILE, carried as a pointer by SMODEL_DESIGN */

#ifdef __cplusplus
extern "C" {
#endif

#define ADHK_REF_MAXLINE 300 /* MAXLINE, define.h:69 */

/* ------------------------------------------------------------------------------------ Part 1: layouts */

/* carried as pointers only on this path */
typedef struct adhk_ref_selem_3d SELEM_3D;
typedef struct adhk_ref_squad SQUAD;
typedef struct adhk_ref_smpi SMPI;
typedef struct adhk_ref_smeteor_file SMETEOR_FILE;
typedef struct adhk_ref_swind SWIND;
typedef struct adhk_ref_swave SWAVE;
typedef struct adhk_ref_scoverage SCOVERAGE;
typedef struct adhk_ref_stestcase STESTCASE;
typedef struct adhk_ref_smeteor_station SMETEOR_STATION;

/* src/structs/sdt/sdt.h:7-16 -- time information */
typedef struct {
    double dt, dt_old, dt_err, dt_prev;
    double t_init, t_prev, t_final;
    double tau_temporal; /* 0: backward Euler */
    double time;
    int nt;
    int t_adpt_flag;
    int ientry_out;
    bool green_ampt;
} SDT;

/* src/structs/svect/svect2d.h, svect.h */
typedef struct {
    double x, y;
} SVECT2D;
typedef struct {
    double x, y, z;
} SVECT;

/* src/structs/snode/snode.h -- a mesh node (coordinates x, y, z are what this path reads) */
typedef struct {
    int id, string, edge_string, node_string, original_id;
    int parent[2], parent_res_pe[2], parent_res_id[2];
    int level, block, els_flag, bflag;
    double x, y, z;
    double area;
    int nelems_connected;
    int *elemID;
    int gid, global_surf_id, global_bed_id, myid, resident_pe, resident_id;
    int *ghosted_arr;
    int n_ghosted;
} SNODE;

/* src/structs/selem/selem_1d.h -- supermodel flux coupling entry and a 1-D (boundary) element */
typedef struct {
    int elem, sub, super;
} FLUX_ELEM;
typedef struct {
    int id, gid, resident_pe, id_orig;
    int nnodes;
    int elem2d;
    int bflag;
    double djac;
    SVECT2D nrml;
    int string, mat, nvars;
    int *vars;
    double *grad_shp;
    int *nodes; /* [nnodes] node ids */
    int *levels;
    FLUX_ELEM *flux_elem;
    int flux_elem_tot;
    double length;
} SELEM_1D;

/* src/structs/selem/selem_2d.h:10-40 -- a 2-D element (triangles on this path: nnodes 3) */
typedef struct {
    int id, gid, id_orig, id_3d;
    int nnodes, nnodes_quad;
    double djac, djac3d, djac3d_fixed;
    SVECT nrml;
    int string, mat, nvars;
    int *vars;
    int bflag;
    SVECT2D *grad_shp;
    int *nodes; /* [nnodes] node ids, counter-clockwise */
    int *levels;
    int nedges;
    double area;
    int resident_pe;
    int interface;
    FLUX_ELEM *flux_elem;
    int flux_elem_tot;
} SELEM_2D;

/* src/structs/sflags/sflags.h -- model flags (single-rank configuration without groundwater or HDF5) */
typedef struct {
    int *model, *params, *dvars;
    bool MG, BAROCLINIC;
    int EOS;
    bool GRID_ADAPTION, ADAPTED_THE_GRID, GRID_REFINED, GRID_UNREFINED;
    bool TIME_ADAPT, TIME_ADAPT_FAIL;
    bool TRANSPORT, NS2_TRANSPORT, NS3_TRANSPORT, SW2_TRANSPORT, SW3_TRANSPORT, VORTICITY;
    bool SEDIMENT, SEDLIB;
    bool STEADY_STATE;
    bool ICM, NSM;
    bool WAVE_RADS, WAVE_STRESS, WIND_STRESS, WAVE_STATION, WIND_STATION;
    bool CORIOLIS, MUC, UNITS, CONVEYANCE, PRN_ADPT, OUTPUT, SOLVE_ATF, UMFPACK_FAIL;
    bool ICE, INS, nice_coords;
    bool TIDE, UNREFINE, CSTORM_WSID;
    bool FLUX_WEIGHTED_NORMALS;
} SFLAGS;

/* src/structs/sstr_value/sstr_value.h -- per-string values; fterms.eqrheight is the roughness height that
 * fe_sw2_get_roughness turns into a drag coefficient */
typedef struct {
    int bc_flag, src_flag, isigma, iu_0, iq, ici, ivx, ivy, ivz;
    double ceq_0;
} INPUT_DATA;
typedef struct {
    double manningsn, skfricoef, dragcoef, eqrheight, rheight, hghtstem, diamstem, densstem;
    double icethick, icedense, bedrhght, icerhght;
    int mng_flag, erh_flag, sav_flag, urv_flag, icemors;
} FRICTION;
typedef struct {
    INPUT_DATA flow, heat, ol_flow, ch_flow, pressure, displacement, bed;
    INPUT_DATA *trans, *sed;
    INPUT_DATA Qs, Td;
    FRICTION fterms;
    double roughness, conveyance, height, ref_height, ref_pressure;
    int ps_flag, string_type, ice_string, phys_flag, flux_flag, sed_div_flag, link, weir_num, flap_num, sluice_num;
    double total_area;
} STR_VALUE;

/* src/structs/sivar_position/sivar_position.h:12-31 */
typedef struct {
    int n, ntrns;
    int *var;
    int *var_code;
} SIVAR_POSITION;

/* src/structs/sdvar/sdvar.h -- dependent variables (the wet/dry flag lives in elem_flags) */
typedef struct {
    int n;
    int *var;
    bool *print_flag;
} SDVAR_POSITION;
typedef struct {
    int n_dvar, nnode_dvar;
    SDVAR_POSITION sdvar_pos_node;
    double **nodal_dvar;
    int n_dvar_elem_dbl, n_dvar_elem_int, nelem_dvar;
    SDVAR_POSITION sdvar_pos_elem_dbl, sdvar_pos_elem_int;
    double **elem_dvar;
    int **elem_flags;
    int *dvar_node_map, *dvar_active, *dvar_elem_map;
} SDVAR;

/* src/structs/smat/smat_physics.h:7-25 -- residual function per PDE of a physics material */
typedef struct SMODEL_SUPER SMODEL_SUPER;
typedef struct {
    int imod;
    int iseries;
    int (*resid)(SMODEL_SUPER *, double *, int, double, int, int, int, int); /* sadh_def.h:17 residual_ptr */
    int (*init)(SMODEL_SUPER *);
} SPDE;
typedef struct {
    int id;
    SIVAR_POSITION ivar_pos;
    int *ivar_loc;
    int npdes;
    SPDE *pde;
} SMAT_PHYSICS;

/* src/structs/sseries/sseries.h -- a time series (linked list; ivalue holds the values of the step) */
typedef struct {
    double time;
    double *area, *slope, *value;
} SSERIES_ENTRY;
typedef struct adhk_ref_sts {
    int id, size, type, nvalues, nnodes, mat_id;
    double infact, outfact, tol;
    double *ivalue;
    SMETEOR_STATION *station;
    SSERIES_ENTRY *entry;
    struct adhk_ref_sts *next, *prev;
} SSERIES;

/* src/structs/slin_sys/slin_sys.h:5-47 -- the split-CSR linear system (single rank: diagonal block only) */
typedef struct {
    int *indptr_diag, *cols_diag;
    double *vals_diag;
    int *indptr_off_diag, *cols_off_diag;
    double *vals_off_diag;
    int nnz_diag, nnz_off_diag, nnz_diag_old, nnz_off_diag_old;
    double *residual, *scale_vect, *dsol;
    int *ghosts;
    int nghost;
    int *local_size, *global_size, *size;
    int local_range[2], local_range_old[2];
    int *local_size_old, *size_old, *global_size_old;
} SLIN_SYS;

/* src/structs/sgrid/sgrid.h:28-155 -- the mesh */
typedef struct {
    SMETEOR_FILE *wind_file;
    SWIND *winds;
    SMETEOR_FILE *wave_file;
    SWAVE *waves;
    SMETEOR_FILE *rain_file;
    double *rain;
} NODAL_ATTRIBUTES;
typedef struct {
    char filename[100];
    char type[15];
    int ndim;
    bool columnar, haveSegs, haveTets, havePrisms, haveTris, haveQuads;
    int nnodes, nnodes_sur, nnodes_bed;
    int nelems3d, nelems2d, nelems1d;
    int nTets, nPrisms, nTris, nQuads;
    SNODE *node;
    SELEM_3D *elem3d;
    SELEM_2D *elem2d;
    SELEM_1D *elem1d;
    int macro_nnodes, macro_nnodes_sur, macro_nnodes_bed;
    int macro_nelems3d, macro_nelems2d, macro_nelems1d, macro_nelems2d_bed;
    int macro_nTets, macro_nPrisms, macro_nQuads, macro_nTris, macro_nSegs;
    int orig_macro_nnodes, orig_macro_nnodes_sur, orig_macro_nnodes_bed;
    int orig_macro_nelems1d, orig_macro_nelems2d, orig_macro_nelems3d;
    int my_nnodes, my_nnodes_sur, my_nnodes_bed;
    int my_nelems3d, my_nelems2d, my_nelems1d;
    int my_nSegs, my_nTets, my_nPrisms, my_nQuads, my_nTris;
    int max_nnodes, max_nnodes_sur, max_nnodes_bed;
    int max_nelems3d, max_nelems2d, max_nelems1d;
    int nnodes_old, nnodes_sur_old, nnodes_bed_old;
    int nelems3d_old, nelems2d_old, nelems1d_old;
    int initial_nnodes, initial_nelems, initial_nnodes_bed, initial_nelems_bed;
    int orig_initial_nnodes, orig_initial_nelems;
    double mesh_length, my_mesh_length;
    double mesh_area, my_mesh_area;
    double mesh_area_surface, my_mesh_area_surface;
    double mesh_area_bed, my_mesh_area_bed;
    double mesh_area_sidewalls, my_mesh_area_sidewalls;
    double mesh_volume, my_mesh_volume;
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double my_x_min, my_x_max, my_y_min, my_y_max, my_z_min, my_z_max;
    int quadrature_order, nqp_1d, nqp_2d, nqp_3d;
    SQUAD *quad_seg, *quad_tri, *quad_rect, *quad_tet, *quad_prism;
    int *nodeID_3d_to_2d_sur, *nodeID_3d_to_2d_bed, *nodeID_2d_to_3d_sur, *nodeID_2d_to_3d_bed;
    SMPI *smpi;
    SMPI *part_smpi;
    int *part_map;
    NODAL_ATTRIBUTES nodal_attribute;
    int *inv_per_node;
    int **nd_on_QuadEdge;
    int *n2n_ind_ptr;
    int *n2n_edge_tab;
} SGRID;

/* src/structs/smodel_super/smodel_super.h:7-113 -- one super model: state, solver knobs, materials, BCs */
struct SMODEL_SUPER {
    int myid, npes;
    SFLAGS flags;
    int id, itrns, isSimple;
    char *type;
    SGRID *grid;
    SDT *ts;
    int nsubsteps, meshcode, forward_step;
    char filebase[ADHK_REF_MAXLINE];
    SLIN_SYS *lin_sys;
    double perturbation, inc_nonlin, tol_nonlin;
    int max_nonlin_linesearch_cuts, max_nonlin_it, it_count_nonlin, force_nonlin_it, nonlinear_it_total;
    int LINEAR_PROBLEM, it_count_nonlin_failed, nalloc_inc;
    int my_ndofs, my_ndofs_old, ndofs, ndofs_old, macro_ndofs, macro_ndofs_old;
    double *sol, *sol_old, *sol_older;
    SDVAR dvars;
    int *elem1d_physics_mat, *elem2d_physics_mat, *elem3d_physics_mat;
    int nmat_physics;
    SMAT_PHYSICS *mat_physics_elem;
    SMAT_PHYSICS **mat_physics_node;
    SIVAR_POSITION ivar_pos;
    int **ivars;
    int nseries;
    SSERIES *series_head, *series_curr;
    SSERIES *series_wind_head, *series_wind_curr;
    SSERIES *series_wave_head, *series_wave_curr;
    SSERIES *series_gw_psk_head, *series_gw_psk_curr;
    int *node_strings_bc;
    int *bc_mask;
    double *dirichlet_data;
    double gravity, density;
    int o_flag;
    double initial_grid_mass, grid_mass_error;
    STR_VALUE *str_values;
    STESTCASE *testcase;
};

/* src/structs/smodel_design/smodel_design.h:7-52 -- the design model: super models and the shared time */
typedef struct {
    int model1, model2, type;
} SMODEL_COUPLE;
typedef struct {
    char filebase[ADHK_REF_MAXLINE], filename_design[ADHK_REF_MAXLINE], filename_grid[ADHK_REF_MAXLINE],
        filename_params[ADHK_REF_MAXLINE];
    FILE *xmf;
    int nSuperModels;
    SMODEL_SUPER *superModel; /* [nSuperModels]; each one's ts points at this design's ts */
    SMODEL_COUPLE *superCouple;
    int nMono, nSimple;
    SGRID *grid;
    SDT ts;
    SSERIES *series_dt, *series_out;
    SCOVERAGE *params;
    int nlin_sys;
    int *lin_sys_id, *unique_id;
    SLIN_SYS *lin_sys;
} SMODEL_DESIGN;

/* -------------------------------------------------------------------- Part 2: reference entry points */

/* residual.h:5-7.  residual -> sm->lin_sys->residual (= minus the sum of the elemental residuals). */
void assemble_residual(SMODEL_SUPER *sm, SGRID *grid);
void add_replace_elem_rhs(double *elem_rhs, double *eq_rhs, int elem_nvars, int *elem_var_map, int imod,
                          int nnodes, double scale);
void load_global_resid(double *residual, double *elem_rhs, int nnodes, int elem_nvars, int *local_dofs);

/* jacobian.h:5-11.  assemble_jacobian writes sm->lin_sys->vals_diag in the pattern of
 * slin_sys_init_sparsity_mono; the others are the reference's plain-array helpers. */
void assemble_jacobian(SMODEL_SUPER *sm);
void load_elem_mat(SMODEL_SUPER *sm, int elem_nnodes, int *nodeIDs, int mat_id, int elem_id, double **elem_mat);
void load_global_mat_split_CSR(double *vals, int *indptr, int *indices, double *off_diag_vals, int *off_diag_indptr,
                               int *off_diag_indices, double **elem_mat, int ndofs_ele, int *dofs, int *global_dofs,
                               int *local_range);
void load_global_mat_CSR(double *vals, int *indptr, int *indices, double **elem_mat, int ndofs_ele, int *global_dofs,
                         int *local_range);
void perturb_var(double **elem_mat, SMODEL_SUPER *sm, SPDE *pde, int ie, int mat_id, int nodes_on_element,
                 int nvar_ele, int *elem_var_map, int npdes, int ele_var_no, int *NodeIDs, int DEBUG);
void elem_matrix_deriv(double **mat, int node_no, int var_no, int nnodes, int elem_nvars, double *local1,
                       double *local2, double diff_ep);

/* newton.h:5-19 */
void initialize_system(SMODEL_SUPER *sm);
void initialize_dirichlet_bc(SMODEL_SUPER *sm);
void update_function(SMODEL_SUPER *sm);
void increment_function(SMODEL_SUPER *sm);
void get_residual_norms(double *resid_max_norm, double *resid_l2_norm, double *inc_max_norm, int *imax_dof,
                        int *iinc_dof, int *include_dof, int my_ndofs, int ndofs, int macro_ndofs, double *residual,
                        double *dsol, int *bc_mask);
int fe_newton(SMODEL_SUPER *sm);

/* la.h:12,18 (the plain-array la.h functions are declared in adhk.h) */
void apply_Dirichlet_BC(SMODEL_SUPER *sm);
void check_diag(SMODEL_SUPER *sm);

/* sw2.h:8-25 and the residual_ptr signature (sadh_def.h:17).  The elemental residual of element ie at
 * sm->sol is added to elem_rhs, as the reference's functions add theirs; perturb_var is the variable code
 * (0 depth, 4 and 5 the depth-averaged velocities, anything else no perturbation). */
int fe_sw2_body_resid(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                      int perturb_var, int perturb_sign, int DEBUG);
int fe_sw2_bc_h(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                int perturb_var, int perturb_sign, int DEBUG);
int fe_sw2_bc_vel(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                  int perturb_var, int perturb_sign, int DEBUG);
int fe_sw2_init(SMODEL_SUPER *sm);
void fe_sw2_wdflag_legacy(SMODEL_SUPER *sm);

/* slin_sys.h:56 -- allocates (realloc) cols_diag and vals_diag and fills indptr_diag [local_size + 1]
 * (the caller's) and nnz_diag; the pattern is built on the GPU */
void slin_sys_init_sparsity_mono(SLIN_SYS *lin_sys, int *elem3d_physics_mat_id, int *elem2d_physics_mat_id,
                                 int *elem1d_physics_mat_id, SMAT_PHYSICS *elem_physics_mat, SGRID *grid,
                                 int **ivars);

/* dofmaps.h:9-15 */
void get_cell_dofs_ivars(int *local_dofs, int **ivars, int nnodes, int *local_node_ids, int elem_nvars,
                         int *elem_var_pos);
void global_to_local_dbl_ivars(double *local_vals, int *nodeIDs, int nnodes, int *map_array, double *global_vals);
void global_to_local_SVECT2D_ivars(SVECT2D *local_vals, int *nodeIDs, int nnodes, int **ivars, int varx, int vary,
                                   double *global_vals);
void local_dofs_to_global_dofs(int *global_dofs, int ndofs_on_ele, int *dofs, int *local_range, int *ghosts);
int global_to_local(int global_col_no, int local_size, int *ghosts, int nghost);

/* time_loop.h:7-9, tc_init.c:4-33 and smodel_super.h:126 -- the caller of fe_newton.  time_loop runs
 * tc_init, then update_dt and advance_time until t_final or a failed step, and returns YES (1) as the
 * reference's does; 0 instead marks an error of this library (adhk_last_error).  advance_time takes each
 * super model through its nsubsteps: smodel_super_prep_sol, the series averaged over (t_prev - dt, t_prev),
 * smodel_super_update_dirichlet_data and the time stepper forward_step names (FE_NEWTON = 0, fe_newton:
 * the only one on this path); TRUE, or FALSE after a failed step, which leaves dt cut by fe_newton. */
int time_loop(SMODEL_DESIGN *dm);
int advance_time(SMODEL_DESIGN *dm, int nsuper);
int update_dt(SMODEL_DESIGN *dm);
void tc_init(SMODEL_DESIGN *dm);
void smodel_super_prep_sol(SMODEL_SUPER *sm);
void smodel_super_update_dirichlet_data(SMODEL_SUPER *sm);

/* sseries.h -- every TIME_SERIES (type 1) of the list set to its average over (t1, t2)
 * (sseries_set_values_average, sseries.c:622-668: the full trapezoids between the two times come from
 * entry[i].area, as the series reader fills it).  A time outside a series' range is an error (status -1,
 * values unchanged), where the reference exits. */
void sseries_setall_ts_valuesAVG(SSERIES *head, double t1, double t2);

/* sw2.h:9 and src/friction/fr_defs.h -- the bed-shear drag coefficient of a string (log law over the
 * roughness height fterms.eqrheight; 0 without one); vegetation and ice drag are not on this path and give
 * status -1 and NaN */
double fe_sw2_get_roughness(SMODEL_SUPER *mod, double depth, double velocity, int id, int which);
double fr_bedshstr_drag_coef(double depth, double roughness_height);
double fr_manningsn_to_rheight(double n, double kn, double g);

/* messg.h:5-33 in the single-rank build (no _MESSG): every reduction is the identity */
double messg_dmax(double x);
double messg_dmin(double x);
int messg_imin(int x);
int messg_imax(int x);
double messg_dsum(double x);

/* sivar_position.h:40 -- ivars[ip->var[j]][node] = the node's dofs in node-major order over the active
 * variables j of ipNode (the 25 variable codes of adh_def); allocates *ivars (ip->n rows of nnodes, free
 * each row, then the array); returns the number of dofs */
int sivar_position_build_dof_map(SIVAR_POSITION *ip, int nnodes, SIVAR_POSITION *ipNode, int ***ivars);

/* No counterparts in the reference.
 * fe_sw2_bc_flux: the residual_ptr of an explicit normal flux per unit width (fe_sw2_1d_explicit_flow,
 *   fe_sw2_boundary_integrals.c:56-94, which the reference compiles without a residual function that calls
 *   it); its value is ivalue[0] of the series the PDE names (SPDE.iseries), negative into the domain.
 * adhk_ref_set_manning: Manning's n per 2-D element for a model not yet seen by any call above (the
 *   reference hard-codes a drag coefficient of 0.002, sw2_body_resid.c:241, which is the default here too).
 * adhk_ref_release: destroys the GPU context of a model (a later call creates it again from sm->grid).
 * adhk_ref_status: 0 after a successful last call on this thread, -1 after a failed one. */
int fe_sw2_bc_flux(SMODEL_SUPER *mod, double *elem_rhs, int ie, double perturbation, int perturb_node,
                   int perturb_var, int perturb_sign, int DEBUG);
int adhk_ref_set_manning(SMODEL_SUPER *sm, const double *manning_n);
void adhk_ref_release(SMODEL_SUPER *sm);
int adhk_ref_status(void);

#ifdef __cplusplus
}
#endif
#endif
