/*
 * exwave_amd.h -- C-ABI of the MI355X-native ExWave hot path.
 *
 * Plain pointers and sizes only.  Vector arguments are DEVICE pointers to
 * FP64 arrays of exw_op_n_dofs() entries in the layout
 *   [cell][field][node],  field = v_0..v_{dim-1}, p,  node = lexicographic
 *   index of the (k+1)^dim Gauss points of the cell, xi_0 fastest.
 * (The polynomial space and quadrature are the reference's FE_DGQ(k)^(dim+1)
 * with QGauss(k+1); the basis is the Lagrange basis collocated at the Gauss
 * points, the reference's FE_DGQArbitraryNodes(QGauss(k+1)) variant,
 * source/explicit_wave.cc:163.)
 *
 * Every call is asynchronous on the context's stream and returns 0 on
 * success or a negative EXW_E* code; exw_last_error() gives the message.
 * A context is not thread-safe: use one host thread per context.
 *
 * Reference citations are paths in kronbichler/exwave.
 */
#ifndef EXWAVE_AMD_H
#define EXWAVE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EXW_OK 0
#define EXW_EINVAL (-1)   /* invalid argument (shape, degree, boundary id, aliasing) */
#define EXW_EDEVICE (-2)  /* HIP runtime error */
#define EXW_ENOMEM (-3)   /* device allocation failed */
#define EXW_EUNSUP (-4)   /* combination not compiled (dim/degree/precision) */

/* operator arithmetic; vectors, RK accumulation and norms are always FP64 */
#define EXW_FP64 0
#define EXW_FP32 1

/* time integrators: the reference's IntegratorType names
 * (include/parameters.h, source/parameters.cc:150-193) */
#define EXW_EXPL_EULER 0
#define EXW_CLRK4 1
#define EXW_LSRK45R2 2
#define EXW_LSRK33R2 3
#define EXW_LSRK59R2 4
#define EXW_ADER 5
#define EXW_ADER_LTS 6 /* exw_run / exw_lts_*: ADER with local time stepping */
#define EXW_LSRK45R3 7
#define EXW_SSPRK 8          /* SSPRK with 8 stages and order 4 (the reference's 'SSPRK') */
#define EXW_ADER_ADCONFULL 9 /* ADER with the full operator's Taylor series */

typedef struct exw_ctx_s *exw_ctx;
typedef struct exw_op_s *exw_op;

/* Superposition of membrane modes in a homogeneous medium:
 *   p   = sum_j amp_j cos(w_j t) prod_d sin(m_jd pi x_d),  w_j = c pi |m_j|
 *   v_d = -sum_j amp_j m_jd/(rho c |m_j|) sin(w_j t) cos(m_jd pi x_d) prod_{e!=d} sin(m_je pi x_e)
 * One mode with m_jd = M and rho = c = 1 is the reference's ExactSolution
 * (include/input_parameters.h:94-128). */
typedef struct exw_modes {
  int n_modes; /* <= 8 */
  double amp[8];
  int m[8][3];
  double rho, c;
} exw_modes;

/* Fused epilogue applied to k = M^-1 A(src) at every DoF i of cell K:
 *   o1[i] = (b ? b[i] : 0) + a1*s_K*k[i]
 *   o2[i] = (o2_from_o1 ? o1[i] : c[i]) + a2*s_K*k[i]        (if o2 != NULL)
 * with s_K = cell_scale[K] (1 if cell_scale is NULL); cells with s_K == 0
 * are skipped (no reads, no writes).
 * This covers apply (b=NULL, a1=1), RKVectorUpdater (b=v2, o1=v2, o2=v3,
 * o2_from_o1=1), its last stage, explicit Euler, ClassRK4 stages and the
 * ADER update.  o1/o2 may alias b/c but never src. */
typedef struct exw_epilogue {
  const double *b;
  double *o1;
  double a1;
  const double *c;
  double *o2;
  double a2;
  int o2_from_o1;
  const double *cell_scale; /* optional, device, one double per cell */
} exw_epilogue;

const char *exw_last_error(void);
const char *exw_version(void);

/* ---- context ------------------------------------------------------------
 * ctx_create binds a HIP device and a stream (hipStream_t passed as void*;
 * NULL = the device's null stream). */
int exw_ctx_create(int device, void *stream, exw_ctx *out);
int exw_ctx_set_stream(exw_ctx ctx, void *stream);
int exw_ctx_destroy(exw_ctx ctx);
int exw_sync(exw_ctx ctx);
/* device memory helpers (callers may also bring their own allocations) */
int exw_alloc(exw_ctx ctx, size_t bytes, void **dptr);
int exw_free(exw_ctx ctx, void *dptr);
int exw_memcpy_h2d(exw_ctx ctx, void *dst, const void *src, size_t bytes);
int exw_memcpy_d2h(exw_ctx ctx, void *dst, const void *src, size_t bytes);

/* ---- operator (replaces WaveEquationOperation<dim,k>::setup,
 * source/wave_equation_operations.cc:200-250, and reset_data_vectors, :383-400)
 *
 * Host arrays, copied to the device:
 *   vertices [n_cells][2^dim][dim]  multilinear cell corners (lexicographic)
 *   neighbor [n_cells][2*dim]       face f = 2*d+s is the face xi_d = s; >= 0:
 *                                   neighbouring cell (seeing it as face f^1,
 *                                   same tangential orientation); < 0: -(boundary id)
 *                                   with ids 1 soft wall, 2 p=0, 3 absorbing
 *                                   (source/wave_equation_operations.cc:592-617)
 *   rho, speed [n_cells]            material per cell (Material, include/utilities.h:28-34)
 * Axis-aligned box meshes select the Cartesian kernels, anything else the
 * general multilinear kernels (Jacobians recomputed from the vertices).
 * FP64 RK stages with two outputs on box meshes of 3D degree 3-5, and FP64
 * sweeps on 2D degree 6-8, run a software-pipelined kernel when their vectors
 * are 16-byte aligned (3D) and below 4 GiB and there is no c operand;
 * otherwise the one-cell-per-group Cartesian kernel, with bit-identical
 * results.
 * EXW_CART_PIPE=0 in the environment selects the latter throughout.  On box
 * meshes of 3D degree 6 and up, sweeps sum the neighbours' face traces from
 * memory instead of staging them (bit-identical; EXW_CART_GATHER=0 selects
 * the staging kernel). */
int exw_op_create(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                  const int *neighbor, const double *rho, const double *speed, exw_op *out);
/* The checks exw_op_create / exw_op_create_nc run on the connectivity, on
 * the host, no device needed: neighbour indices and boundary ids in range,
 * every conforming neighbour sees the cell back across face f^1, and every
 * conforming face is in standard orientation -- corner v of face f (bit d of
 * v equal to s, f = 2d+s) lies at corner v^(1<<d) of the neighbour (1e-10 of
 * the cell size).  deal.II meshes may carry rotated or flipped faces
 * (MatrixFree's face orientation); such a mesh is rejected with EXW_EINVAL
 * rather than run with wrong face pairings.  EXW_NC_FACE entries are checked
 * against their records by exw_op_create_nc. */
int exw_mesh_check(int dim, int n_cells, const double *vertices, const int *neighbor,
                   int n_hanging, const int *hanging);
int exw_op_destroy(exw_op op);
int64_t exw_op_n_dofs(exw_op op);
int exw_op_n_cells(exw_op op);
int exw_op_is_cartesian(exw_op op);
/* dt = cfl/k^1.5 * min_K min_vertex_distance(K)/c_K without the cfl factor
 * (source/explicit_wave.cc:529-545, source/parameters.cc:195) */
double exw_op_min_vertex_distance_over_c(exw_op op);

/* ---- hot path ---------------------------------------------------------- */
/* dst = M^-1 A(src): WaveEquationOperationBase<dim>::apply
 * (include/wave_equation_operations.h:75-76; source/wave_equation_operations.cc:665-685) */
int exw_op_apply(exw_op op, const double *src, double *dst, int precision);
/* k = M^-1 A(src) with a fused epilogue (the operator + RKVectorUpdater fusion) */
int exw_op_apply_fused(exw_op op, const double *src, const exw_epilogue *epi, int precision);
/* One low-storage RK stage: RKVectorUpdater semantics fused into the operator
 * (source/time_integrators.cc:107-163):
 *   not last: y_out = b + f1*k ; b_out = y_out + f2*k
 *   last    : y_out = b + f1*k
 * with k = M^-1 A(src).  y_out and b_out must not alias src. */
int exw_rk_stage(exw_op op, const double *src, const double *b, double *y_out, double *b_out,
                 double f1, double f2, int is_last, int precision);
/* The unfused vector update RKVectorUpdater::apply_to_subrange
 * (source/time_integrators.cc:128-155) on n DoFs, for callers that compute
 * vector1 = M^-1 A(src) themselves:
 *   not last: vector2 += factor1*vector1; vector3 = vector2 + factor2*vector1
 *   last    : vector2 = vector3 + factor1*vector1 */
int exw_rk_update(exw_ctx ctx, int64_t n, double factor1, double factor2, int is_last,
                  const double *vector1, double *vector2, double *vector3);
/* ADER-CK time integral (integrate_taylor_cauchykovalewski,
 * source/wave_equation_operations.cc:961-1262): tempsrc = int_{t1}^{t2} u dt
 * by the Taylor series about te; recon (optional, may be NULL) replaces the
 * first time derivative (use_ader_post). */
int exw_ader_ck(exw_op op, const double *src, const double *recon, double t2, double t1,
                double te, double *tempsrc, int precision);
/* ArbitraryHighOrderDG::perform_time_step + WaveEquationOperationADER::apply_ader
 * (source/time_integrators.cc:37-50; source/wave_equation_operations.cc:1346-1381):
 * vec_np = vec_n - apply_ader(vec_n).  work: 2*n_dofs doubles (tempsrc, recon). */
int exw_ader_step(exw_op op, const double *vec_n, double *vec_np, double *work, double dt,
                  int recon, int precision);
/* n_steps global ADER steps of size dt on vec in place (no reconstruction, no
 * source): what the reference's time loop does between two output ticks
 * (ExplicitWaveEquation::run, source/explicit_wave.cc:779-799, calling
 * perform_time_step once per step).  On box meshes without hanging faces (3D
 * degree 3-5) each step's update sweep runs fused with the next step's CK
 * time integral in one kernel; the result is bit-identical to n_steps
 * exw_ader_step calls.  work: 2*n_dofs doubles (the CK fields of two
 * consecutive steps). */
int exw_ader_steps(exw_op op, int n_steps, double dt, double *vec, double *work, int precision);
/* ExplicitIntegrator<..>::perform_time_step (include/time_integrators.h:25-39)
 * for the integrator ids above.  Same contract as the reference: vec_n is
 * clobbered, the result is left in vec_np.  work: 3*n_dofs doubles (the
 * reference's vec_tmp1/vec_tmp2 plus one register for the fused stages). */
int exw_time_step(exw_op op, int integrator, double dt, double *vec_n, double *vec_np,
                  double *work, int ader_recon, int precision);
/* number of work doubles exw_time_step needs per DoF */
int exw_time_step_work_vectors(int integrator);

/* ---- setup and norms --------------------------------------------------- */
/* project_initial_field (source/wave_equation_operations.cc:902-935) */
int exw_project(exw_op op, const exw_modes *modes, double t, double *dst);
/* integrate_difference with QGauss(k+2) (source/explicit_wave.cc:452-496):
 * out_host[0] = ||p_h - p||, [1] = ||v_h - v||, [2] = ||p||; synchronises. */
int exw_l2_errors(exw_op op, const exw_modes *modes, double t, const double *u,
                  double *out_host);
/* E = 1/2 int rho|v|^2 + p^2/(rho c^2) with QGauss(k+1); synchronises. */
int exw_energy(exw_op op, const double *u, double *out_host);
/* exw_l2_errors without allocating and without synchronising: the same three
 * norms (same kernels, same summation order) written to out_dev (device, 3
 * doubles) in stream order.  No counterpart in the reference: for callers that
 * cannot allocate device memory or wait -- code under HIP graph capture, tight
 * serving loops.  The partial sums live in one static device array per GPU
 * (3*1024+4 doubles, ~24.6 KB, held until the process ends); the call rejects
 * an operator whose partial-sum count exceeds it.  Every call on a GPU shares
 * that array: calls that overlap on one GPU (different streams or host
 * threads) race and must be ordered by the caller. */
int exw_l2_errors_noalloc(exw_op op, const exw_modes *modes, double t, const double *u,
                          double *out_dev);

/* ---- point source ---------------------------------------------------------
 * No counterpart in the reference (the two-material refined-box workload):
 * dp/dt gains r(t) delta(x - x0) with the Ricker wavelet
 *   r(t) = amplitude (1 - 2 a tau^2) exp(-a tau^2),  a = (pi f0)^2, tau = t - t0.
 * delta acts on the DG space as q -> q(x0), shared equally by the n_s cells
 * whose closure holds x0; cell K's pressure receives M^-1 of (1/n_s) phi_i(x0)
 * per unit time integral.  An update over [t1, t2] adds the wavelet's exact
 * time integral, amplitude [tau exp(-a tau^2)] from t1 - t0 to t2 - t0, times
 * that vector: the time-integrated point source of ADER schemes (it enters the
 * update, not the Cauchy-Kowalevski predictor). */
typedef struct exw_source {
  double x[3];
  double amplitude, f0, t0;
} exw_source;
/* locate x0 (at most 8 cells) and set the operator's source; NULL removes it.
 * EXW_EINVAL if no cell holds x0.  Synchronises. */
int exw_op_set_source(exw_op op, const exw_source *src);
/* u += (int_{t1}^{t2} r dt) M^-1 delta_x0 on the pressure (after an ADER step
 * from t1 to t2); asynchronous */
int exw_source_add(exw_op op, double t1, double t2, double *u);

/* ---- ADER with local time stepping --------------------------------------
 * WaveEquationOperationADERLTS (source/wave_equation_operations.cc:1385-1692)
 * with ClusterManager (include/cluster_manager.templates.h:24-1049): cells are
 * grouped into clusters whose time steps are multiples 1, 1+d, 1+2d, ... of
 * the fastest step; the clusters advance in the reference's update order and
 * exchange face fluxes through a flux memory. */
typedef struct exw_lts_s *exw_lts;
/* iterate_cluster_categorization (include/cluster_manager.templates.h:24-262)
 * on the host, no device needed: cluster[c] and flags[c] (1: has a slower
 * face neighbour, 2: a faster one) per cell; cfl = cfl_number/k^1.5. */
int exw_lts_categorize(int dim, int n_cells, const double *vertices, const int *neighbor,
                       double cfl, int max_clusters, int max_diff, int *cluster, int *flags,
                       int *n_clusters, int *cluster_diff, int *iterations);
/* WaveEquationOperationADERLTS::setup + ClusterManager::setup (:330-484):
 * fastest_dt = the unrounded CFL step, time0 = the time at set-up, recon =
 * use_ader_post.  The operator must outlive the handle. */
int exw_lts_create(exw_op op, double cfl, double fastest_dt, int max_clusters, int max_diff,
                   double time0, int recon, exw_lts *out);
/* exw_lts_create with flags.  EXW_LTS_SPEED_AWARE clusters cells by
 * cfl * min vertex distance / c: the step each cell's own CFL limit allows
 * (the reference clusters by cfl * min vertex distance only,
 * include/cluster_manager.templates.h:57-66, while its global step divides by
 * c, source/explicit_wave.cc:529-545).  Operators with 2:1 hanging faces
 * (exw_op_create_nc) are taken by both entry points: a cell's neighbours
 * across such a face are the coarse cell or the fine cells of the subfaces
 * (include/cluster_manager.templates.h:118-125), and each subface is a face of
 * its own in the update masks.  A point source set on the operator is added
 * to every update of its cells over that update's interval. */
#define EXW_LTS_SPEED_AWARE 1
int exw_lts_create2(exw_op op, double cfl, double fastest_dt, int max_clusters, int max_diff,
                    double time0, int recon, int flags, exw_lts *out);
int exw_lts_destroy(exw_lts L);
/* global_dt = the slowest cluster's step, by which the time control advances */
int exw_lts_info(exw_lts L, int *n_clusters, int *cluster_diff, int *n_updates,
                 double *global_dt);
int exw_lts_cells(exw_lts L, int *cluster, int *flags);
/* ArbitraryHighOrderDGLTS::perform_time_step (source/time_integrators.cc:54-63):
 * dst = state after one global step; time_now = the time after
 * TimeControl::advance_time_step.  src == dst steps in place (no copy of the
 * state); otherwise src is left unchanged. */
int exw_lts_step(exw_lts L, double time_now, const double *src, double *dst);

/* ---- problem setup and the whole solve ---------------------------------- */
/* transforms of the unit hypercube's vertices */
#define EXW_TRANSFORM_NONE 0
#define EXW_TRANSFORM_REFERENCE_SINE 1 /* x0 += f prod_d sin(pi x_d), include/input_parameters.h:41-50 */
#define EXW_TRANSFORM_AXIS_2PI 2       /* x_i += f prod_{e!=i} sin(2 pi x_e) */
#define EXW_TRANSFORM_PRODUCT_2PI 3    /* x_i += f prod_e sin(2 pi x_e) (keeps the unit cube) */

/* GridGenerator::subdivided_hyper_cube(n, 0, 1) + transform of the coarse
 * vertices + refine_global(r) with flat (multilinear) refinement
 * (input_geometry_description, include/input_parameters.h:59-88).  Cells are
 * numbered lexicographically on the fine grid.  Call with vertices = neighbor
 * = NULL to get n_cells; then pass arrays of n_cells*2^dim*dim doubles and
 * n_cells*2*dim ints. */
int exw_mesh_hypercube(int dim, int n_intervals, int n_refinements, int transform, double factor,
                       int boundary_id, int *n_cells, double *vertices, int *neighbor);

/* WaveEquationOperation::estimate_error (source/wave_equation_operations.cc:
 * 871-898), the refinement indicator of adapt_mesh and the 'Error_estimate'
 * output field: tmp <- apply(solution), then per cell
 *   error_per_cell[c] = sqrt( int_K |grad p_h - v_t|^2 ) (QGauss(k+1)),
 * v_t the velocity part of tmp.  error_per_cell: device, n_cells doubles;
 * asynchronous. */
int exw_estimate_error(exw_op op, const double *solution, double *tmp, double *error_per_cell);

/* ---- one level of adaptive refinement (adapt_mesh, explicit_wave.cc:265-330) --
 * Marking: GridRefinement::refine_and_coarsen_fixed_number (deal.II >= 8.5)
 * with the fractions 0.1 / 0.6 of set_refinement_indicators (:255-260), then
 * adapt_mesh's rules (:281-296): no refinement at max_level, no coarsening at
 * min_level, no refinement below 0.1 max_error_init, coarsening below 0.05
 * max_error_init.  flags[c] = 1 refine, -1 coarsen, 0 keep (host arrays). */
int exw_refinement_flags(int n_cells, const double *error_per_cell, const int *level,
                         int min_level, int max_level, double max_error_init, int *flags);
/* The base-cell refinement of exw_mesh_box_refined after the marking: an
 * unrefined cell flagged 1 is refined, a refined base cell is coarsened when all
 * its children are flagged -1 (host arrays, refine_* of n_intervals^dim). */
int exw_adapt_box_refine(int dim, int n_intervals, const int *refine_old, const int *flags,
                         int *refine_new);
/* SolutionTransfer::interpolate between the two meshes of exw_mesh_box_refined
 * (device vectors): children take the parent's polynomial, a coarsened parent
 * the L2 projection of its children on the reference cell, other cells are
 * copied.  Synchronises. */
int exw_adapt_box_transfer(exw_ctx ctx, int dim, int degree, int n_intervals,
                           const int *refine_old, const int *refine_new,
                           const double *old_solution, double *new_solution);

/* ---- several levels of adaptive refinement (adapt_mesh,
 * source/explicit_wave.cc:265-330, levels n_refinements ..
 * n_refinements + n_adaptive_refinements, :272-273) -------------------------
 * A forest of refinement trees over the n_intervals^dim box of
 * exw_mesh_hypercube: tree[] lists the base cells lexicographically, each
 * depth first -- 1 = refined (its 2^dim children follow, lexicographic), 0 =
 * an active cell.  Active cells are numbered in that order; a cell's level is
 * its depth.  Children are the parent's multilinear map at dyadic points, as
 * deal.II refines straight cells.  The one-level box meshes above are the
 * forests of depth <= 1 with the same cell numbering.
 *
 * exw_mesh_forest: vertices, neighbours, levels (level may be NULL) and the
 * hanging-face records of exw_op_create_nc.  Call with vertices = neighbor =
 * hanging = NULL for n_cells and n_hanging.  EXW_EINVAL for malformed trees
 * or faces whose cells differ by more than one level. */
int exw_mesh_forest(int dim, int n_intervals, int transform, double factor, int boundary_id,
                    int n_tree, const int *tree, int *n_cells, double *vertices, int *neighbor,
                    int *level, int *n_hanging, int *hanging);
/* Triangulation::execute_coarsening_and_refinement of p4est-backed deal.II
 * (parallel::distributed::Triangulation) after adapt_mesh's flags
 * (flags[c] = 1 refine, -1 coarsen, 0 keep, per active cell): a family is
 * coarsened only when all its children are active and flagged; the mesh stays
 * 2:1 balanced across faces, edges and corners -- the coarser cell of a
 * violation is refined until none is left, and a coarsening that would create
 * one is not done.  Writes the new trees into tree_new (capacity entries;
 * NULL: only *n_tree_new). */
int exw_adapt_forest(int dim, int n_intervals, int n_tree, const int *tree, const int *flags,
                     int *n_tree_new, int *tree_new, int capacity);
/* SolutionTransfer::interpolate between two forests of exw_adapt_forest
 * (every cell kept, refined once or coarsened once), the rules of
 * exw_adapt_box_transfer, on device vectors. */
int exw_adapt_forest_transfer(exw_ctx ctx, int dim, int degree, int n_intervals, int n_tree_old,
                              const int *tree_old, int n_tree_new, const int *tree_new,
                              const double *old_solution, double *new_solution);

/* output_results (source/explicit_wave.cc:339-436) as a VTU file, host arrays:
 * each cell subdivided degree times per direction (build_patches(mapping,
 * fe_degree)), the fields evaluated at the equispaced points and written as
 * linear sub-cells.  Point data: solution_velocity, solution_pressure and, if
 * given, error_velocity/error_pressure (error: same layout as solution) and
 * post_pressure ((k+2)^dim Gauss-Lobatto values per cell); cell data, if given:
 * Error_estimate, cluster_id, time_step (n_cells values).  ASCII. */
int exw_write_vtu(int dim, int degree, int n_cells, const double *vertices,
                  const double *solution, const double *error, const double *error_estimate,
                  const double *post_pressure, const double *cluster_id, const double *time_step,
                  const char *filename);

/* ---- post-processed pressure ----------------------------------------------
 * WaveEquationOperation::compute_post_pressure
 * (source/wave_equation_operations.cc:786-848): tmp <- apply(solution) (the
 * reference's tmp_vector, n_dofs values), then per cell the local Neumann
 * problem (grad phi, grad p*) = (grad phi, -rho v_t) on Q_{k+1}, with the
 * constant set so the cell average of p* equals that of p.  post receives
 * exw_post_n_values(op) = n_cells * (k+2)^dim values, per cell the nodal values
 * at the (k+2)^dim Gauss-Lobatto points (the reference's fe_post_disp,
 * source/explicit_wave.cc:165), xi_0 fastest.  iterations (device, n_cells
 * ints, may be NULL) receives each cell's CG iteration count.
 * The local CG runs on the zero-sum subspace of the Neumann problem (same p*
 * as the reference's first-unknown constraint, far better conditioned) and
 * stops per cell at |r| < EXW_POST_ABS_TOL, |r|/|r0| < EXW_POST_REL_TOL or
 * after EXW_POST_MAX_ITER_FACTOR * (k+2)^dim + 1 iterations (the reference:
 * 1e-20, 1e-11, 6 n, :798-799).  Asynchronous on the context's stream. */
#define EXW_POST_ABS_TOL 1e-20
#define EXW_POST_REL_TOL 1e-13
#define EXW_POST_MAX_ITER_FACTOR 6
int64_t exw_post_n_values(exw_op op);
int exw_post_pressure(exw_op op, const double *solution, double *tmp, double *post,
                      int *iterations);
/* ||p* - p_exact||_L2 with QGauss(k+3) (the 'error p post' of
 * source/explicit_wave.cc:477-484); synchronises. */
int exw_post_error(exw_op op, const exw_modes *modes, double t, const double *post,
                   double *err_host);

/* ---- meshes with hanging faces ---------------------------------------------
 * One level of local refinement as deal.II makes it (p4est refinement in
 * adapt_mesh, source/explicit_wave.cc:265-330; MatrixFree integrates the 2:1
 * faces on the fine side with the coarse cell evaluated on the subface).
 * neighbor[] holds EXW_NC_FACE on both sides of such a face; each record of
 * hanging[] (2 + 2^(dim-1) ints) is {coarse cell, coarse face, the fine cells
 * by subface position}.  Subface position b has bit j set when the fine face
 * covers the upper half of the j-th tangential direction (the directions other
 * than the face's, ascending); the fine cells see the face as coarse face ^ 1.
 * Child vertices must be the parent's multilinear map at the dyadic points
 * (deal.II's refinement of straight cells). */
#define EXW_NC_FACE (-1000)
/* exw_op_create for a mesh with n_hanging records in hanging[] (checked
 * against the geometry). */
int exw_op_create_nc(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                     const int *neighbor, int n_hanging, const int *hanging, const double *rho,
                     const double *speed, exw_op *out);
/* The n_intervals^dim box of exw_mesh_hypercube with the base cells flagged in
 * refine[] (lexicographic, xi_0 fastest) refined once: cells numbered base cell
 * by base cell, children lexicographic inside their parent.  Call with
 * vertices = neighbor = hanging = NULL for n_cells and n_hanging. */
int exw_mesh_box_refined(int dim, int n_intervals, int transform, double factor, int boundary_id,
                         const int *refine, int *n_cells, double *vertices, int *neighbor,
                         int *n_hanging, int *hanging);

/* Parameters of one solve (the keys of the reference's prm files that the
 * hot path uses, source/parameters.cc:36-100). */
typedef struct exw_problem {
  int dim, degree, n_intervals, n_refinements, transform;
  double transform_factor;
  int boundary_id;
  int integrator;      /* EXW_EXPL_EULER .. EXW_ADER */
  double cfl_number;   /* as written in the prm file (divided by k^1.5 internally) */
  int max_time_steps;  /* 0: derived from final_time */
  double final_time, output_every_time;
  int ader_recon;      /* use_ader_post */
  int precision;       /* EXW_FP64 / EXW_FP32 operator */
  exw_modes modes;     /* initial condition and exact solution */
  int lts_max_clusters, lts_max_diff; /* EXW_ADER_LTS: max_n_clusters, max_diff_clusters */
  int post_pressure;   /* also report 'error p post' at every tick (else -1) */
  /* adapt_mesh every adaptive_refinement_interval steps and
   * n_adaptive_refinements times before the loop (explicit_wave.cc:704-716,
   * 789-791) on the forest of exw_mesh_forest, levels n_refinements ..
   * n_refinements + n_adaptive_refinements; with EXW_ADER_LTS the clusters are
   * rebuilt after every adaptation, as the operator's setup does in make_dofs */
  int n_adaptive_refinements, adaptive_refinement_interval;
} exw_problem;

/* WaveEquationProblem<dim>::run (source/explicit_wave.cc:549-813): mesh,
 * operator, projection, TimeControl with the reference's dt rules
 * (source/utilities.cc:31-95, source/explicit_wave.cc:222), the time loop and
 * the L2 errors at every output tick.  Writes one line per tick
 *   "<step> <time> <error p> <error v> <solution mag p> <error p post>\n"
 * into log (if not NULL).  Returns the number of steps (>= 0) or an error
 * code.  wall_seconds (optional) receives the time-loop wall time excluding
 * the error evaluations. */
int exw_run(exw_ctx ctx, const exw_problem *p, char *log, int log_len, double *wall_seconds);

/* ---- serving: many independent solves stepped together -------------------
 * No counterpart in the reference (a build-only subsystem).  A request is one
 * membrane solve on the unit hypercube with 2^n_refinements Cartesian cells
 * per direction, boundary id 2 (p = 0), homogeneous rho and c, initial data and
 * exact solution `modes` (its rho and c must equal the request's), LSRK45 at
 * dt = cfl_number/k^1.5 * h/c stepped while t < final_time (the reference's
 * time loop, source/explicit_wave.cc:779-799). */
typedef struct exw_request {
  int dim, degree, n_refinements;
  double rho, c;
  exw_modes modes;
  double final_time, cfl_number;
  int precision; /* EXW_FP64 / EXW_FP32 operator, or EXW_PRECISION_AUTO */
} exw_request;

/* EXW_PRECISION_AUTO: the FP32 operator is used when the solve's own
 * discretisation error is far above what FP32 round-off adds over its steps:
 *   E0 / (2^-24 sqrt(n_steps) ||p0||) >= EXW_FP32_SCORE_MIN,
 * with E0 = ||p0 - I p0|| the error of the initial Gauss-point interpolant
 * (exw_request_initial_error) and ||p0|| the norm of the initial pressure;
 * otherwise FP64.  The result then stays within 1.1x of the FP64 error. */
#define EXW_PRECISION_AUTO 2
#define EXW_FP32_SCORE_MIN 10.0

typedef struct exw_result {
  int64_t id;
  int n_steps;
  double time;                 /* final time reached (>= final_time) */
  double err_p, err_v, mag_p;  /* L2 errors against the exact modes, QGauss(k+2) */
  int precision;               /* operator arithmetic the solve ran with */
} exw_result;

typedef struct exw_server_s *exw_server;

/* The serving loop keeps every admitted solve resident in HBM.  Solves of the
 * same (dim, degree, precision) share a device pool of cells and are stepped
 * together: one operator launch per RK stage covers every cell of every solve
 * in the pool (per-cell time steps in the fused epilogue), the 5 launches of a
 * step are captured once per pool in a hipGraph and replayed, each pool runs on
 * its own stream.  Requests are admitted in arrival order while the pools'
 * footprint fits in hbm_budget_bytes; finished solves leave their cells to the
 * next admissions.  The loop itself runs on the calling host thread. */
int exw_server_create(exw_ctx ctx, size_t hbm_budget_bytes, exw_server *out);
int exw_server_destroy(exw_server s);
/* queue a request; *id receives its id */
int exw_server_submit(exw_server s, const exw_request *r, int64_t *id);
/* admit what fits and enqueue up to max_steps batch steps (one time step of
 * every active solve each); returns the number of active solves */
int exw_server_progress(exw_server s, int max_steps);
/* collect finished solves (up to max_results); returns how many were written */
int exw_server_poll(exw_server s, exw_result *out, int max_results);
/* queued (not admitted), active and finishing solves; bytes of HBM in use */
int exw_server_status(exw_server s, int *n_queued, int *n_active, int *n_finishing,
                      size_t *hbm_bytes);
/* The same solve alone, one launch sequence per step (the one-at-a-time path
 * the serving loop is compared with).  Synchronises. */
int exw_solve(exw_ctx ctx, const exw_request *r, exw_result *out);
/* host only: E0 = ||p0 - I p0|| (QGauss(k+2)) and ||p0|| of a request, in
 * closed form (separable modes, no device work); *score = the FP32 score of
 * EXW_PRECISION_AUTO and *precision the operator it selects. */
int exw_request_initial_error(const exw_request *r, double *e0, double *norm_p0, double *score,
                              int *precision);

#ifdef __cplusplus
}
#endif
#endif
