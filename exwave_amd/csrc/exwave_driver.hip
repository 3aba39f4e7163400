// exwave_driver.hip -- problem setup and the whole solve behind the C-ABI:
// the hypercube mesh (input_geometry_description, include/input_parameters.h:
// 59-88), TimeControl (source/utilities.cc:31-95) and the time loop of
// WaveEquationProblem::run (source/explicit_wave.cc:549-813).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "forest.hpp"

namespace {

void transform_vertex(int dim, int type, double f, double *x) {
  const double in[3] = {x[0], x[1], dim > 2 ? x[2] : 0.0};
  if (type == EXW_TRANSFORM_REFERENCE_SINE) {
    double s = f;
    for (int d = 0; d < dim; ++d) s *= std::sin(M_PI * in[d]);
    x[0] = in[0] + s;
  } else if (type == EXW_TRANSFORM_AXIS_2PI) {
    for (int i = 0; i < dim; ++i) {
      double s = f;
      for (int e = 0; e < dim; ++e)
        if (e != i) s *= std::sin(2.0 * M_PI * in[e]);
      x[i] = in[i] + s;
    }
  } else if (type == EXW_TRANSFORM_PRODUCT_2PI) {
    double s = f;
    for (int e = 0; e < dim; ++e) s *= std::sin(2.0 * M_PI * in[e]);
    for (int i = 0; i < dim; ++i) x[i] = in[i] + s;
  }
}

}  // namespace

extern "C" {

int exw_mesh_hypercube(int dim, int n_intervals, int n_refinements, int transform, double factor,
                       int boundary_id, int *n_cells, double *vertices, int *neighbor) {
  if ((dim != 2 && dim != 3) || n_intervals < 1 || n_refinements < 0 || n_refinements > 12 ||
      transform < 0 || transform > 3 || boundary_id < 1 || boundary_id > 3)
    return EXW_EINVAL;
  const int R = 1 << n_refinements;
  const long Nf = (long)n_intervals * R;
  long nc = 1;
  for (int d = 0; d < dim; ++d) nc *= Nf;
  if (nc > 2000000000L) return EXW_EINVAL;
  if (n_cells) *n_cells = (int)nc;
  if (!vertices || !neighbor) return EXW_OK;
  const int nv = 1 << dim;
  for (long c = 0; c < nc; ++c) {
    long idx[3] = {0, 0, 0}, t = c;
    for (int d = 0; d < dim; ++d) {
      idx[d] = t % Nf;
      t /= Nf;
    }
    // coarse parent corners after the transform
    double PV[8][3];
    for (int v = 0; v < nv; ++v) {
      double x[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) x[d] = double(idx[d] / R + ((v >> d) & 1)) / n_intervals;
      transform_vertex(dim, transform, factor, x);
      for (int d = 0; d < dim; ++d) PV[v][d] = x[d];
    }
    // children are the multilinear map of the parent at dyadic points
    for (int v = 0; v < nv; ++v) {
      double xi[3];
      for (int d = 0; d < dim; ++d) xi[d] = double(idx[d] % R + ((v >> d) & 1)) / R;
      double *out = vertices + (c * nv + v) * dim;
      for (int i = 0; i < dim; ++i) out[i] = 0.0;
      for (int u = 0; u < nv; ++u) {
        double s = 1.0;
        for (int d = 0; d < dim; ++d) s *= ((u >> d) & 1) ? xi[d] : 1.0 - xi[d];
        for (int i = 0; i < dim; ++i) out[i] += s * PV[u][i];
      }
    }
    for (int d = 0; d < dim; ++d) {
      long stride = 1;
      for (int e = 0; e < d; ++e) stride *= Nf;
      neighbor[c * 2 * dim + 2 * d] = idx[d] > 0 ? int(c - stride) : -boundary_id;
      neighbor[c * 2 * dim + 2 * d + 1] = idx[d] + 1 < Nf ? int(c + stride) : -boundary_id;
    }
  }
  return EXW_OK;
}

int exw_mesh_box_refined(int dim, int n_intervals, int transform, double factor, int boundary_id,
                         const int *refine, int *n_cells, double *vertices, int *neighbor,
                         int *n_hanging, int *hanging) {
  if ((dim != 2 && dim != 3) || n_intervals < 1 || transform < 0 || transform > 3 ||
      boundary_id < 1 || boundary_id > 3 || !refine)
    return EXW_EINVAL;
  const int N = n_intervals, nv = 1 << dim, nface = 2 * dim, nsub = 1 << (dim - 1);
  long nbase = 1;
  for (int d = 0; d < dim; ++d) nbase *= N;
  if (nbase * nv > 2000000000L) return EXW_EINVAL;
  // first output cell of every base cell; refined ones become 2^dim children
  std::vector<long> first(nbase + 1, 0);
  long nh = 0;
  for (long B = 0; B < nbase; ++B) first[B + 1] = first[B] + (refine[B] ? nv : 1);
  for (long B = 0; B < nbase; ++B) {
    if (refine[B]) continue;
    long t = B, stride = 1;
    for (int d = 0; d < dim; ++d) {
      const long i = t % N;
      t /= N;
      if (i > 0 && refine[B - stride]) ++nh;
      if (i + 1 < N && refine[B + stride]) ++nh;
      stride *= N;
    }
  }
  if (n_cells) *n_cells = (int)first[nbase];
  if (n_hanging) *n_hanging = (int)nh;
  if (!vertices || !neighbor || !hanging) return EXW_OK;
  long h = 0;
  for (long B = 0; B < nbase; ++B) {
    long idx[3] = {0, 0, 0}, t = B;
    for (int d = 0; d < dim; ++d) {
      idx[d] = t % N;
      t /= N;
    }
    double PV[8][3];
    for (int v = 0; v < nv; ++v) {
      double x[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) x[d] = double(idx[d] + ((v >> d) & 1)) / N;
      transform_vertex(dim, transform, factor, x);
      for (int d = 0; d < dim; ++d) PV[v][d] = x[d];
    }
    const int nch = refine[B] ? nv : 1;
    for (int ch = 0; ch < nch; ++ch) {
      const long c = first[B] + ch;
      // children: the parent's multilinear map at the dyadic points
      for (int v = 0; v < nv; ++v) {
        double xi[3];
        for (int d = 0; d < dim; ++d)
          xi[d] = refine[B] ? 0.5 * (((ch >> d) & 1) + ((v >> d) & 1)) : double((v >> d) & 1);
        double *out = vertices + (c * nv + v) * dim;
        for (int i = 0; i < dim; ++i) out[i] = 0.0;
        for (int u = 0; u < nv; ++u) {
          double w = 1.0;
          for (int d = 0; d < dim; ++d) w *= ((u >> d) & 1) ? xi[d] : 1.0 - xi[d];
          for (int i = 0; i < dim; ++i) out[i] += w * PV[u][i];
        }
      }
      long stride = 1;
      for (int d = 0; d < dim; ++d) {
        for (int s = 0; s < 2; ++s) {
          const int f = 2 * d + s;
          int *nb = neighbor + c * nface + f;
          if (refine[B] && ((ch >> d) & 1) != s) {  // sibling
            *nb = int(first[B] + (ch ^ (1 << d)));
            continue;
          }
          const long ni = idx[d] + (s ? 1 : -1);
          if (ni < 0 || ni >= N) {
            *nb = -boundary_id;
            continue;
          }
          const long B2 = B + (s ? stride : -stride);
          if (refine[B] == refine[B2]) {
            *nb = int(refine[B] ? first[B2] + (ch ^ (1 << d)) : first[B2]);
          } else {
            *nb = EXW_NC_FACE;
            if (!refine[B]) {  // coarse side: one record with the touching children
              int *r = hanging + h * (2 + nsub);
              r[0] = int(c);
              r[1] = f;
              for (int b = 0; b < nsub; ++b) {
                int chn = (s ? 0 : 1) << d, j = 0;
                for (int e = 0; e < dim; ++e)
                  if (e != d) chn |= ((b >> j++) & 1) << e;
                r[2 + b] = int(first[B2] + chn);
              }
              ++h;
            }
          }
        }
        stride *= N;
      }
    }
  }
  return EXW_OK;
}

int exw_mesh_forest(int dim, int n_intervals, int transform, double factor, int boundary_id,
                    int n_tree, const int *tree, int *n_cells, double *vertices, int *neighbor,
                    int *level, int *n_hanging, int *hanging) {
  if ((dim != 2 && dim != 3) || n_intervals < 1 || transform < 0 || transform > 3 ||
      boundary_id < 1 || boundary_id > 3 || n_tree < 1 || !tree)
    return EXW_EINVAL;
  exw::Forest F;
  int64_t used = 0;
  if (!F.parse(dim, n_intervals, tree, n_tree, &used) || used != n_tree) return EXW_EINVAL;
  const int nv = 1 << dim, nface = 2 * dim, nsub = 1 << (dim - 1);
  const int64_t nc = (int64_t)F.cells.size();
  if (nc * nv > 2000000000L) return EXW_EINVAL;
  // pass 1 checks the faces and counts the hanging ones; pass 2 writes
  const bool write = vertices && neighbor && hanging;
  int64_t nh = 0;
  for (int64_t c = 0; c < nc; ++c) {
    const exw::ForestCell &X = F.cells[c];
    const int l = X.level;
    if (write) {
      if (level) level[c] = l;
      int64_t B[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) B[d] = X.I[d] >> l;
      double PV[8][3];
      for (int v = 0; v < nv; ++v) {
        double x[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) x[d] = double(B[d] + ((v >> d) & 1)) / n_intervals;
        transform_vertex(dim, transform, factor, x);
        for (int d = 0; d < dim; ++d) PV[v][d] = x[d];
      }
      const double scale = 1.0 / double((int64_t)1 << l);
      for (int v = 0; v < nv; ++v) {
        double xi[3];
        for (int d = 0; d < dim; ++d)
          xi[d] = double(X.I[d] - (B[d] << l) + ((v >> d) & 1)) * scale;
        double *out = vertices + (c * nv + v) * dim;
        for (int i = 0; i < dim; ++i) out[i] = 0.0;
        for (int u = 0; u < nv; ++u) {
          double w = 1.0;
          for (int d = 0; d < dim; ++d) w *= ((u >> d) & 1) ? xi[d] : 1.0 - xi[d];
          for (int i = 0; i < dim; ++i) out[i] += w * PV[u][i];
        }
      }
    }
    for (int f = 0; f < nface; ++f) {
      const int d = f >> 1, s = f & 1;
      int64_t J[3] = {X.I[0], X.I[1], X.I[2]};
      J[d] += s ? 1 : -1;
      int nb;
      if (J[d] < 0 || J[d] >= F.extent(l)) {
        nb = -boundary_id;
      } else {
        const int r = F.find(l, J);
        if (r >= 0) {
          nb = r;  // conforming
        } else if (r == -1) {
          // coarse side: the neighbour's children touching the face must be active
          int fine[4];
          for (int b = 0; b < nsub; ++b) {
            int64_t K[3] = {2 * J[0], 2 * J[1], 2 * J[2]};
            K[d] += s ? 0 : 1;
            int j = 0;
            for (int e = 0; e < dim; ++e)
              if (e != d) K[e] += (b >> j++) & 1;
            fine[b] = F.find(l + 1, K);
            if (fine[b] < 0) return EXW_EINVAL;  // more than one level across a face
          }
          if (write) {
            int *rec = hanging + nh * (2 + nsub);
            rec[0] = int(c);
            rec[1] = f;
            for (int b = 0; b < nsub; ++b) rec[2 + b] = fine[b];
          }
          ++nh;
          nb = EXW_NC_FACE;
        } else {
          // fine side: the parent of the same-level neighbour must be active
          int64_t P[3] = {J[0] >> 1, J[1] >> 1, J[2] >> 1};
          if (l == 0 || F.find(l - 1, P) < 0) return EXW_EINVAL;
          nb = EXW_NC_FACE;
        }
      }
      if (write) neighbor[c * nface + f] = nb;
    }
  }
  if (n_cells) *n_cells = int(nc);
  if (n_hanging) *n_hanging = int(nh);
  return EXW_OK;
}

namespace {

// The mesh, operator and state vectors of a run; replaced by adapt_mesh.
struct RunState {
  exw_ctx ctx = nullptr;
  exw_op op = nullptr;
  double *buf = nullptr, *post = nullptr, *err = nullptr;
  double *sol = nullptr, *tmp = nullptr, *work = nullptr;
  int n_cells = 0;
  std::vector<int> level;  // adaptive runs: the cells' refinement levels
  void release() {
    if (post) exw_free(ctx, post);
    if (err) exw_free(ctx, err);
    if (buf) exw_free(ctx, buf);
    if (op) exw_op_destroy(op);
    op = nullptr;
    buf = post = err = nullptr;
  }
};

// mesh (hypercube, or the forest tree[] when adaptive), operator and
// vectors; the solution is left uninitialised
int run_setup(exw_ctx ctx, const exw_problem *p, bool adaptive, const std::vector<int> &tree,
              int nwork, RunState &st) {
  st.ctx = ctx;
  int n_cells = 0, n_hanging = 0, rc;
  const int nv = 1 << p->dim;
  std::vector<double> V;
  std::vector<int> nb, H;
  if (adaptive) {
    const int nt = (int)tree.size();
    if ((rc = exw_mesh_forest(p->dim, p->n_intervals, p->transform, p->transform_factor,
                              p->boundary_id, nt, tree.data(), &n_cells, nullptr, nullptr,
                              nullptr, &n_hanging, nullptr)))
      return rc;
    V.resize((size_t)n_cells * nv * p->dim);
    nb.resize((size_t)n_cells * 2 * p->dim);
    H.resize((size_t)std::max(n_hanging, 1) * (2 + nv / 2));
    st.level.resize(n_cells);
    if ((rc = exw_mesh_forest(p->dim, p->n_intervals, p->transform, p->transform_factor,
                              p->boundary_id, nt, tree.data(), &n_cells, V.data(), nb.data(),
                              st.level.data(), &n_hanging, H.data())))
      return rc;
  } else {
    if ((rc = exw_mesh_hypercube(p->dim, p->n_intervals, p->n_refinements, p->transform,
                                 p->transform_factor, p->boundary_id, &n_cells, nullptr,
                                 nullptr)))
      return rc;
    V.resize((size_t)n_cells * nv * p->dim);
    nb.resize((size_t)n_cells * 2 * p->dim);
    if ((rc = exw_mesh_hypercube(p->dim, p->n_intervals, p->n_refinements, p->transform,
                                 p->transform_factor, p->boundary_id, &n_cells, V.data(),
                                 nb.data())))
      return rc;
  }
  std::vector<double> rho(n_cells, p->modes.rho), speed(n_cells, p->modes.c);
  if ((rc = exw_op_create_nc(ctx, p->dim, p->degree, n_cells, V.data(), nb.data(), n_hanging,
                             H.data(), rho.data(), speed.data(), &st.op)))
    return rc;
  st.n_cells = n_cells;
  const int64_t N = exw_op_n_dofs(st.op);
  if ((rc = exw_alloc(ctx, sizeof(double) * N * (2 + nwork), (void **)&st.buf))) return rc;
  st.sol = st.buf;
  st.tmp = st.buf + N;
  st.work = st.buf + 2 * N;
  if (p->post_pressure &&
      (rc = exw_alloc(ctx, sizeof(double) * exw_post_n_values(st.op), (void **)&st.post)))
    return rc;
  if (adaptive && (rc = exw_alloc(ctx, sizeof(double) * n_cells, (void **)&st.err))) return rc;
  return EXW_OK;
}

}  // namespace

int exw_run(exw_ctx ctx, const exw_problem *p, char *log, int log_len, double *wall_seconds) {
  if (!ctx || !p) return EXW_EINVAL;
  const bool adaptive = p->n_adaptive_refinements > 0;
  const bool lts = p->integrator == EXW_ADER_LTS;
  if (adaptive && (p->adaptive_refinement_interval < 1 || p->n_refinements < 0 ||
                   p->n_refinements + p->n_adaptive_refinements > exw::Forest::MAX_LEVEL))
    return EXW_EINVAL;
  const int nwork = lts ? 0 : exw_time_step_work_vectors(p->integrator);
  if (nwork < 0) return EXW_EINVAL;
  int nbase = 1;
  for (int d = 0; d < p->dim; ++d) nbase *= p->n_intervals;
  // the forest of the globally refined box: levels min_level = n_refinements
  // .. max_level = min_level + n_adaptive_refinements (explicit_wave.cc:272-273)
  const int min_level = p->n_refinements, max_level = min_level + p->n_adaptive_refinements;
  std::vector<int> tree;
  if (adaptive) {
    std::function<void(int)> uniform = [&](int depth) {
      tree.push_back(depth > 0);
      if (depth > 0)
        for (int ch = 0; ch < (1 << p->dim); ++ch) uniform(depth - 1);
    };
    for (int B = 0; B < nbase; ++B) uniform(min_level);
  }
  RunState st;
  int rc;
  if ((rc = run_setup(ctx, p, adaptive, tree, nwork, st))) {
    st.release();
    return rc;
  }
  // parameters.cc:195 and explicit_wave.cc:529-545
  const double cfl = p->cfl_number / std::pow((double)p->degree, 1.5);
  const double dt_cfl = cfl * exw_op_min_vertex_distance_over_c(st.op);
  // TimeControl::setup (utilities.cc:31-46), then make_dofs resets the step to
  // the unrounded CFL value (explicit_wave.cc:222)
  const double T = p->final_time, tick = p->output_every_time;
  double time = 0.0;
  double dt = (T - time) / std::max(std::round((T - time) / dt_cfl), 1.0);
  unsigned step = 0;
  const unsigned max_step =
      p->max_time_steps == 0 ? (unsigned)std::round((T - time) / dt) : (unsigned)p->max_time_steps;
  dt = dt_cfl;
  exw_lts L = nullptr;
  if (lts) {
    // WaveEquationOperationADERLTS::setup: clusters from the unrounded CFL step;
    // the time control then advances by the slowest cluster's step
    // (include/cluster_manager.templates.h:463)
    if ((rc = exw_lts_create(st.op, cfl, dt_cfl, p->lts_max_clusters, p->lts_max_diff, 0.0,
                             p->ader_recon, &L))) {
      st.release();
      return rc;
    }
    exw_lts_info(L, nullptr, nullptr, nullptr, &dt);
  }
  // adapt_mesh (explicit_wave.cc:265-330): indicator, marking, new mesh,
  // solution transfer, new CFL step
  double max_err_init = -1.0;
  auto adapt = [&]() -> int {
    int r;
    if ((r = exw_estimate_error(st.op, st.sol, st.tmp, st.err))) return r;
    std::vector<double> err(st.n_cells);
    if ((r = exw_memcpy_d2h(ctx, err.data(), st.err, sizeof(double) * st.n_cells))) return r;
    std::vector<int> flags(st.n_cells), tnew;
    int ntn = 0;
    if ((r = exw_refinement_flags(st.n_cells, err.data(), st.level.data(), min_level, max_level,
                                  max_err_init, flags.data())) ||
        (r = exw_adapt_forest(p->dim, p->n_intervals, (int)tree.size(), tree.data(),
                              flags.data(), &ntn, nullptr, 0)))
      return r;
    tnew.resize(ntn);
    if ((r = exw_adapt_forest(p->dim, p->n_intervals, (int)tree.size(), tree.data(),
                              flags.data(), &ntn, tnew.data(), ntn)))
      return r;
    RunState nst;
    if ((r = run_setup(ctx, p, true, tnew, nwork, nst)) ||
        (r = exw_adapt_forest_transfer(ctx, p->dim, p->degree, p->n_intervals, (int)tree.size(),
                                       tree.data(), ntn, tnew.data(), st.sol, nst.sol))) {
      nst.release();
      return r;
    }
    // make_dofs: compute_time_step_size (explicit_wave.cc:222, 529-545); with
    // ADER-LTS the operator's setup builds new clusters from that step,
    // starting at the current time, and the time control takes the slowest
    // cluster's step (wave_equation_operations.cc:353-379,
    // include/cluster_manager.templates.h:320-484), which adapt_mesh keeps
    // (explicit_wave.cc:324-327)
    exw_lts_destroy(L);  // it refers to the old operator
    L = nullptr;
    st.release();
    st = std::move(nst);
    tree = std::move(tnew);
    dt = cfl * exw_op_min_vertex_distance_over_c(st.op);
    if (lts) {
      if ((r = exw_lts_create(st.op, cfl, dt, p->lts_max_clusters, p->lts_max_diff, time,
                              p->ader_recon, &L)))
        return r;
      exw_lts_info(L, nullptr, nullptr, nullptr, &dt);
    }
    return EXW_OK;
  };
  int pos = 0;
  if (log && log_len > 0) log[0] = 0;
  auto output = [&]() -> int {
    double err[3], epp = -1.0;
    int r = exw_l2_errors(st.op, &p->modes, time, st.sol, err);
    if (r) return r;
    // output_results: compute_post_pressure into tmp_solutions (explicit_wave.cc:476);
    // tmp is free between steps (the next step swaps it with sol first)
    if (st.post && ((r = exw_post_pressure(st.op, st.sol, st.tmp, st.post, nullptr)) ||
                    (r = exw_post_error(st.op, &p->modes, time, st.post, &epp))))
      return r;
    if (log && pos < log_len)
      pos += std::snprintf(log + pos, log_len - pos, "%u %.17g %.10e %.10e %.10e %.10e\n", step,
                           time, err[0], err[1], err[2], epp);
    return 0;
  };
  double wall = 0.0;
  rc = exw_project(st.op, &p->modes, 0.0, st.sol);
  if (!rc && adaptive) {
    // the initial adaptation (explicit_wave.cc:704-716): n_adaptive_refinements
    // times adapt and project again, then the largest indicator of the
    // projected field
    for (int left = p->n_adaptive_refinements; left > 0 && !rc; --left)
      if (!(rc = adapt())) rc = exw_project(st.op, &p->modes, 0.0, st.sol);
    if (!rc && !(rc = exw_estimate_error(st.op, st.sol, st.tmp, st.err))) {
      std::vector<double> err(st.n_cells);
      if (!(rc = exw_memcpy_d2h(ctx, err.data(), st.err, sizeof(double) * st.n_cells))) {
        max_err_init = 0.0;
        for (double e : err) max_err_init = std::max(max_err_init, e);
      }
    }
  }
  // global ADER without reconstruction advances the state in place through
  // exw_ader_steps, several steps per call up to the next step at which the
  // loop below does something besides stepping (an adaptation, an output tick,
  // the end), with the same step count, times and results
  const bool multi = p->integrator == EXW_ADER && !p->ader_recon;
  auto steps_to_event = [&]() {
    unsigned s = step, m = 0;
    double t = time;
    while (!(t >= T || s > max_step)) {
      ++s;
      t += dt;
      ++m;
      if (adaptive && s % p->adaptive_refinement_interval == 0) break;
      if ((int)(t / tick) != (int)((t - dt) / tick)) break;
    }
    return m;
  };
  if (!rc && !(rc = output())) {
    while (!(time >= T || step > max_step)) {
      if (multi) {
        const unsigned m = steps_to_event();
        exw_sync(ctx);
        auto t0 = std::chrono::steady_clock::now();
        if ((rc = exw_ader_steps(st.op, (int)m, dt, st.sol, st.work, p->precision)) ||
            (rc = exw_sync(ctx)))
          break;
        wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (unsigned j = 0; j < m; ++j) {
          ++step;
          time += dt;
        }
        if (adaptive && step % p->adaptive_refinement_interval == 0 && (rc = adapt())) break;
        if ((int)(time / tick) != (int)((time - dt) / tick) && (rc = output())) break;
        continue;
      }
      ++step;
      time += dt;
      std::swap(st.sol, st.tmp);
      exw_sync(ctx);
      auto t0 = std::chrono::steady_clock::now();
      if (lts)
        // ArbitraryHighOrderDGLTS (:54-63), in place on the state
        rc = exw_lts_step(L, time, st.tmp, st.tmp), std::swap(st.sol, st.tmp);
      else
        rc = exw_time_step(st.op, p->integrator, dt, st.tmp, st.sol, st.work, p->ader_recon,
                           p->precision);
      if (rc) break;
      if ((rc = exw_sync(ctx))) break;
      wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (adaptive && step % p->adaptive_refinement_interval == 0 && (rc = adapt())) break;
      // TimeControl::at_tick (utilities.cc:89-95), with the step size as any
      // adaptation left it
      if ((int)(time / tick) != (int)((time - dt) / tick) && (rc = output())) break;
    }
  }
  exw_lts_destroy(L);
  st.release();
  if (wall_seconds) *wall_seconds = wall;
  return rc ? rc : (int)step;
}

}  // extern "C"
