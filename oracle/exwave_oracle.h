/*
 * exwave_oracle.h -- CPU FP64 restatement of ExWave's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in exwave_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * What it restates (reference = kronbichler/exwave, paths relative to the
 * reference root):
 *   - DG operator M^-1 A: source/wave_equation_operations.cc:404-685
 *     (evaluate_cell, evaluate_inner_face, evaluate_boundary_face, inverse mass)
 *   - RKVectorUpdater / low-storage RK: source/time_integrators.cc:107-363
 *   - ExplicitEuler / ClassRK4: source/time_integrators.cc:24-102
 *   - ADER Cauchy-Kowalevski: source/wave_equation_operations.cc:939-1381
 *   - initial projection: source/wave_equation_operations.cc:902-935
 *   - L2 errors: source/explicit_wave.cc:439-496
 *   - mesh + transform: include/input_parameters.h:41-88
 *   - TimeControl: source/utilities.cc:31-95, dt: source/explicit_wave.cc:529-545
 *
 * Discretisation: FE_DGQ(k)^(dim+1) with QGauss(k+1) integration.  The
 * oracle stores the polynomial in the Lagrange basis collocated at the k+1
 * Gauss points (the reference's own FE_DGQArbitraryNodes(QGauss(k+1))
 * variant, source/explicit_wave.cc:163); the QGauss(k+1) mass matrix is then
 * diagonal and the discrete solution is the same polynomial as the
 * reference's up to round-off.
 *
 * Vector layout (shared with the HIP library): cell-major; inside a cell
 * field-major (v_0 .. v_{dim-1}, p); inside a field the (k+1)^dim nodes in
 * lexicographic order with xi_0 fastest.
 */
#ifndef EXWAVE_ORACLE_H
#define EXWAVE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OX_MAX_N 14 /* points per direction: degree <= 12 plus the k+2 error rule */

typedef struct ox_basis {
  int n;                           /* k+1 Gauss points per direction */
  double x[OX_MAX_N], w[OX_MAX_N]; /* QGauss(k+1) on [0,1] */
  double D[OX_MAX_N * OX_MAX_N];   /* D[q*n+i] = l_i'(x_q) (collocation derivative) */
  double face[2][OX_MAX_N];        /* l_i(0), l_i(1) */
} ox_basis;

/* A mesh of 2^dim-vertex multilinear cells with conforming faces in the
 * standard (lexicographic) orientation. Face f = 2*d + s is the face with
 * xi_d = s. neighbor[c*2*dim+f] >= 0 is the neighbouring cell (which sees the
 * face as its face f^1), < 0 is -(boundary id). */
typedef struct ox_mesh {
  int dim;
  int n_cells;
  double *vertices; /* n_cells * 2^dim * dim */
  int *neighbor;    /* n_cells * 2*dim */
  double *rho;      /* n_cells */
  double *speed;    /* n_cells */
  double min_vertex_distance_over_c; /* min_K min_vertex_distance(K)/c_K */
  /* Non-conforming (2:1) faces, as deal.II's hanging faces after one local
   * refinement: record h is hanging[h*(2+2^(dim-1)) + ...] = {coarse cell,
   * coarse face, fine cell of subface 0, 1, ...}.  A subface position b has
   * bit j set when the fine face covers the upper half of the j-th tangential
   * direction (the directions other than the face's, ascending); the fine
   * cells see the face as their face (coarse face ^ 1).  Both sides hold
   * OX_NC_FACE in neighbor[]; nc_rec / nc_pos give the record and the subface
   * position (-1 on the coarse side) per cell face. */
  int n_hanging;
  int *hanging;
  int *nc_rec, *nc_pos; /* n_cells * 2*dim, -1 where conforming */
} ox_mesh;
#define OX_NC_FACE (-1000)

/* Exact solution: superposition of membrane modes in a homogeneous medium
 * (include/input_parameters.h:94-128 is the single mode m_d = m, rho=c=1). */
typedef struct ox_exact {
  int dim;
  int n_modes;
  double amp[8];
  int m[8][3];
  double rho, c;
} ox_exact;

/* transform types for ox_mesh_hypercube */
enum {
  OX_TRANSFORM_NONE = 0,
  OX_TRANSFORM_REFERENCE_SINE = 1, /* x0 += f*prod_d sin(pi x_d), include/input_parameters.h:41-50 */
  OX_TRANSFORM_AXIS_2PI = 2,       /* x_i += f*prod_{e!=i} sin(2 pi x_e) (moves the boundary) */
  OX_TRANSFORM_PRODUCT_2PI = 3     /* x_i += f*prod_e sin(2 pi x_e) (keeps the unit cube) */
};

void ox_basis_init(ox_basis *b, int degree);
void ox_gauss(int n, double *x, double *w);
/* values l_i(x) and derivatives of the Lagrange basis on nodes[0..n-1] */
void ox_lagrange(int n, const double *nodes, double x, double *val, double *der);

ox_mesh *ox_mesh_hypercube(int dim, int n_intervals, int n_refinements, int transform,
                           double factor, int boundary_id);
void ox_mesh_set_material(ox_mesh *m, const double *rho, const double *speed);
void ox_mesh_free(ox_mesh *m);
/* The n_intervals^dim box mesh of ox_mesh_hypercube (transform applied to the
 * base vertices) with the base cells flagged in refine[] (lexicographic, xi_0
 * fastest) split once into 2^dim children whose vertices are the parent's
 * multilinear map at the dyadic points.  Cells are numbered base cell by base
 * cell, children lexicographic inside their parent.  Faces between a refined
 * and an unrefined base cell are non-conforming (hanging[]). */
ox_mesh *ox_mesh_box_refined(int dim, int n_intervals, int transform, double factor,
                             int boundary_id, const int *refine);

/* dst = M^-1 A(src) * (multiplied by boundary_fac: +1 RK, -1 second ADER sweep) */
void ox_apply(const ox_mesh *m, int degree, const double *src, double *dst, double boundary_fac);

/* RKVectorUpdater semantics, source/time_integrators.cc:128-155 */
void ox_rk_update(long n, double f1, double f2, int is_last, const double *v1, double *v2,
                  double *v3);

/* integrators: same register semantics as the reference (vec_n is clobbered,
 * result in vec_np); tmp1/tmp2 are scratch vectors of the same size. */
enum {
  OX_EXPL_EULER = 0,
  OX_CLRK4 = 1,
  OX_LSRK45R2 = 2,
  OX_LSRK33R2 = 3,
  OX_LSRK59R2 = 4,
  OX_ADER = 5,
  OX_ADER_LTS = 6, /* ox_run only: ADER with local time stepping */
  OX_LSRK45R3 = 7,
  OX_SSPRK = 8,        /* SSPRK(8 stages, order 4), the reference's 'SSPRK' */
  OX_ADER_ADCONFULL = 9
};
void ox_time_step(const ox_mesh *m, int degree, int integrator, double dt, double *vec_n,
                  double *vec_np, double *tmp1, double *tmp2, int ader_recon);

/* ADER: tempsrc = Cauchy-Kowalevski time integral of src over [t1,t2] about te,
 * recon = (optional) M^-1 A src used for the first time derivative. */
void ox_ader_ck(const ox_mesh *m, int degree, const double *src, const double *recon, double t2,
                double t1, double te, double *tempsrc);
/* dst = result of WaveEquationOperationADER::apply_ader */
void ox_apply_ader(const ox_mesh *m, int degree, const double *src, double *dst, double dt,
                   int recon, double *scratch);

double ox_exact_value(const ox_exact *e, const double *x, double t, int component);
void ox_project(const ox_mesh *m, int degree, const ox_exact *e, double t, double *dst);
/* out[0] = ||p_h - p||, out[1] = ||v_h - v||, out[2] = ||p|| with QGauss(k+2) */
void ox_l2_errors(const ox_mesh *m, int degree, const ox_exact *e, double t, const double *u,
                  double *out);
/* energy 1/2 int rho |v|^2 + p^2/(rho c^2) with QGauss(k+1) */
double ox_energy(const ox_mesh *m, int degree, const double *u);

/* Whole solve following WaveEquationProblem::run (source/explicit_wave.cc:549-813).
 * Writes one line per output tick into `log`: step, time, error p, error v,
 * solution mag p, error p post (-1 unless post_pressure is set).
 * Returns the number of steps performed. */
typedef struct ox_problem {
  int dim, degree, n_intervals, n_refinements, transform;
  double transform_factor;
  int boundary_id;
  int integrator;
  double cfl_number; /* as in the prm file; divided by k^1.5 like source/parameters.cc:195 */
  int max_time_steps;
  double final_time, output_every_time;
  int ader_recon;
  ox_exact exact;
  int lts_max_clusters, lts_max_diff; /* max_n_clusters, max_diff_clusters */
  int post_pressure;                   /* also log 'error p post' (else -1) */
  int n_adaptive_refinements;          /* levels above n_refinements (forest meshes) */
  int adaptive_refinement_interval;
} ox_problem;
int ox_run(const ox_problem *p, char *log, int log_len, double *final_state_energy);

long ox_n_dofs(const ox_mesh *m, int degree);

/* estimate_error (source/wave_equation_operations.cc:871-898): err[c] =
 * sqrt(int_K |grad p_h - v_t|^2) with v_t the velocity of apply(solution). */
void ox_estimate_error(const ox_mesh *m, int degree, const double *solution, double *err);

/* One level of adaptive refinement (adapt_mesh, source/explicit_wave.cc:265-330):
 * the marking (GridRefinement::refine_and_coarsen_fixed_number of deal.II >= 8.5
 * with 0.1 / 0.6, then adapt_mesh's rules; flags 1 refine, -1 coarsen), the new
 * base-cell refinement of ox_mesh_box_refined, and the solution transfer
 * (children: the parent's polynomial; parent: the reference-cell L2 projection
 * of its children, by quadrature). */
void ox_refinement_flags(int n, const double *err, const int *level, int min_level, int max_level,
                         double max_error_init, int *flags);
void ox_adapt_refine(int dim, int n_intervals, const int *refine_old, const int *flags,
                     int *refine_new);
void ox_adapt_transfer(int dim, int degree, int n_intervals, const int *refine_old,
                       const int *refine_new, const double *old_u, double *new_u);

/* Several levels: forests of refinement trees over the n_intervals^dim box
 * (transform applied to the base vertices, children the multilinear map at
 * dyadic points).  tree[] lists the base cells lexicographically, each depth
 * first: 1 = refined (its 2^dim children follow, lexicographic), 0 = an active
 * cell; active cells are numbered in that order and the level of a cell is its
 * depth.  n_tree < 0 reads as far as the trees go.
 * ox_forest_levels: the cells' levels (level may be NULL); returns n_cells.
 * ox_mesh_forest: NULL unless faces differ by at most one level. */
int ox_forest_levels(int dim, int n_intervals, const int *tree, int n_tree, int *level);
ox_mesh *ox_mesh_forest(int dim, int n_intervals, int transform, double factor, int boundary_id,
                        const int *tree, int n_tree);
/* The new forest after adapt_mesh's flags (1 refine, -1 coarsen, per active
 * cell) as p4est-backed deal.II makes it: families are coarsened only when all
 * children are active and flagged; the mesh is kept 2:1 balanced across faces,
 * edges and corners -- the coarser cell of a violation is refined until none is
 * left, and a coarsening that would create one is not done.  Returns the
 * length of tree_new (NULL: the length only), -1 on error or if cap is short. */
int ox_adapt_forest(int dim, int n_intervals, const int *tree, const int *flags, int *tree_new,
                    int cap);
/* the transfer of ox_adapt_transfer between two forests of ox_adapt_forest */
int ox_adapt_forest_transfer(int dim, int degree, int n_intervals, const int *tree_old,
                             const int *tree_new, const double *old_u, double *new_u);

/* Post-processed pressure (compute_post_pressure,
 * source/wave_equation_operations.cc:786-848): per cell the local Neumann
 * problem (grad phi, grad p*) = (grad phi, -rho dv/dt) on Q_{k+1} (Gauss-Lobatto
 * nodal basis, (k+2)^dim values per cell, QGauss(k+2)), its first unknown
 * fixed, the constant set so the cell average equals that of p_h.
 * post: n_cells * (k+2)^dim values. */
void ox_gauss_lobatto(int n, double *x);
void ox_post_pressure(const ox_mesh *m, int degree, const double *solution, double *post);
/* the per-cell CG's stopping rule (default: the reference's 1e-20, 1e-11, 6 n);
 * iterations, if not NULL, receives each cell's iteration count; mode 0 is the
 * reference's first-unknown constraint, mode 1 CG on the zero-sum subspace of
 * the Neumann problem (same p*, better conditioned) */
void ox_post_set_solver(double abs_tol, double rel_tol, int max_iter_factor, int *iterations,
                        int mode);
/* ||p* - p|| with QGauss(k+3) (source/explicit_wave.cc:477-484) */
double ox_post_error(const ox_mesh *m, int degree, const ox_exact *e, double t,
                     const double *post);

/* Point source of pressure (config-4 workload; no counterpart in the
 * reference): dp/dt gains r(t) delta(x - x0) with the Ricker wavelet
 *   r(t) = amplitude (1 - 2 a tau^2) exp(-a tau^2),  a = (pi f0)^2, tau = t - t0.
 * Discretely, delta is the functional q -> q(x0) on the DG space, shared
 * equally by the n_s cells whose closure holds x0 (a point on a face, edge or
 * vertex), so the pressure right-hand side of cell K is
 * (1/n_s) phi_i(x0), and M^-1 of it is b_K.  A time update over [t1, t2]
 * adds (int_{t1}^{t2} r dt) b_K to the pressure of K: the source's exact time
 * integral added to the time-integrated update, as ADER point-source
 * implementations do. */
typedef struct ox_source {
  double x[3];
  double amplitude, f0, t0;
} ox_source;
double ox_ricker(const ox_source *s, double t);
/* closed form: amplitude [tau exp(-a tau^2)] between t1 - t0 and t2 - t0 */
double ox_ricker_integral(const ox_source *s, double t1, double t2);
/* the cells holding x0 (<= 8, returned count n_s) and per cell the nodal
 * vector b (n_s * (k+1)^dim values, pressure field) */
int ox_source_cells(const ox_mesh *m, int degree, const ox_source *s, int *cells, double *b);
/* u += (int_{t1}^{t2} r dt) b on the pressure of the source cells */
void ox_source_add(const ox_mesh *m, int degree, const ox_source *s, double t1, double t2,
                   double *u);

/* ADER with local time stepping: WaveEquationOperationADERLTS
 * (source/wave_equation_operations.cc:1385-1692) + ClusterManager
 * (include/cluster_manager.templates.h:24-1049), per cell.
 * cfl: cfl_number/k^1.5; fastest_dt: the unrounded CFL step;
 * time0: the time at setup. */
typedef struct ox_lts ox_lts;
ox_lts *ox_lts_create(const ox_mesh *m, int degree, double cfl, double fastest_dt,
                      int max_clusters, int max_diff, double time0, int recon);
/* flags: OX_LTS_SPEED_AWARE clusters by cfl * min vertex distance / c (the
 * reference: cfl * min vertex distance).  Meshes with 2:1 hanging faces are
 * taken: their faces are evaluated per subface like any other face. */
#define OX_LTS_SPEED_AWARE 1
ox_lts *ox_lts_create2(const ox_mesh *m, int degree, double cfl, double fastest_dt,
                       int max_clusters, int max_diff, double time0, int recon, int flags);
/* add the point source to every update of its cells (NULL: none) */
void ox_lts_set_source(ox_lts *L, const ox_source *src);
void ox_lts_free(ox_lts *L);
int ox_lts_n_clusters(const ox_lts *L);
int ox_lts_cluster_diff(const ox_lts *L);
int ox_lts_iterations(const ox_lts *L);
double ox_lts_global_dt(const ox_lts *L);
int ox_lts_n_updates(const ox_lts *L);
/* order[n_updates], times[n_updates][2] (absolute, current step) */
void ox_lts_schedule(const ox_lts *L, int *order, double *times);
/* cluster id per cell; flags: 1 = has slower neighbour, 2 = has faster */
void ox_lts_cells(const ox_lts *L, int *cluster, int *flags);
/* perform_time_step: time_now = time after TimeControl::advance_time_step */
void ox_lts_step(ox_lts *L, double time_now, const double *src, double *dst);

#ifdef __cplusplus
}
#endif
#endif
