// exwave_adapt.hip -- one level of adaptive refinement on box meshes: the
// marking of adapt_mesh (source/explicit_wave.cc:265-330) and the solution
// transfer on the device.
//
// Marking: GridRefinement::refine_and_coarsen_fixed_number (deal.II >= 8.5,
// the reference's requirement, CMakeLists.txt:31) with the fractions 0.1 and
// 0.6 of set_refinement_indicators (explicit_wave.cc:255-260): the
// int(0.1 n) largest indicators are refined (threshold by nth_element, cells
// >= it), the int(0.6 n) smallest coarsened (<= it, unless flagged for
// refinement); then adapt_mesh's level and threshold rules.
//
// Transfer (SolutionTransfer::interpolate for FE_DGQ): a refined cell's
// children take the parent's polynomial at their nodes (exact); a coarsened
// parent takes the L2 projection on the reference cell of its children,
// parent_i = sum_b sum_j R_b(i, j) child_b(j) per direction with
// R_b(i, j) = l_i((x_j + b)/2) w_j / (2 w_i); cells that stay are copied.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <functional>
#include <unordered_set>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "capi_internal.hpp"
#include "forest.hpp"
#include "gauss_cx.hpp"

namespace {

template <int DIM, int N>
__global__ void __launch_bounds__(256)
    transfer_kernel(int n_new, const int *__restrict__ map, const double *__restrict__ tab,
                    const double *__restrict__ old_u, double *__restrict__ new_u) {
  constexpr int ND = DIM == 2 ? N * N : N * N * N, NF = DIM + 1, NCH = 1 << DIM;
  constexpr int REC = 2 + NCH;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)n_new * ND) return;
  const int c = int(gid / ND), q = int(gid % ND);
  const int *r = map + (size_t)c * REC;
  const int kind = r[0];
  const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
  const double *H = tab, *R = tab + 2 * N * N;  // H_b[i*N + j], R_b[i*N + j]
  for (int f = 0; f < NF; ++f) {
    double val = 0.0;
    if (kind == 0) {
      val = old_u[((size_t)r[2] * NF + f) * ND + q];
    } else if (kind == 1) {  // child r[1] of old cell r[2]
      const double *up = old_u + ((size_t)r[2] * NF + f) * ND;
      for (int j = 0; j < ND; ++j) {
        const int ji[3] = {j % N, (j / N) % N, DIM == 3 ? j / (N * N) : 0};
        double w = 1.0;
        for (int d = 0; d < DIM; ++d) w *= H[((r[1] >> d) & 1) * N * N + qi[d] * N + ji[d]];
        val += w * up[j];
      }
    } else {  // parent of the old children r[2..]
      for (int ch = 0; ch < NCH; ++ch) {
        const double *uc = old_u + ((size_t)r[2 + ch] * NF + f) * ND;
        for (int j = 0; j < ND; ++j) {
          const int ji[3] = {j % N, (j / N) % N, DIM == 3 ? j / (N * N) : 0};
          double w = 1.0;
          for (int d = 0; d < DIM; ++d) w *= R[((ch >> d) & 1) * N * N + qi[d] * N + ji[d]];
          val += w * uc[j];
        }
      }
    }
    new_u[((size_t)c * NF + f) * ND + q] = val;
  }
}

template <int DIM, int N>
hipError_t launch_transfer(int n_new, const int *map, const double *tab, const double *old_u,
                           double *new_u, hipStream_t s) {
  constexpr int ND = DIM == 2 ? N * N : N * N * N;
  const size_t total = (size_t)n_new * ND;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL((transfer_kernel<DIM, N>), dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, s, n_new, map, tab, old_u, new_u);
  return hipGetLastError();
}

using TransferLaunch = hipError_t (*)(int, const int *, const double *, const double *, double *,
                                      hipStream_t);
template <int DIM, int N>
void fill_transfer(TransferLaunch *t) {
  if constexpr (N >= 2) {
    t[N] = launch_transfer<DIM, N>;
    fill_transfer<DIM, N - 1>(t);
  }
}

int afail(int code, const std::string &msg) {
  exw::set_last_error(msg);
  return code;
}

// runs the transfer kernel on the records of map[] (2 + 2^dim ints per new
// cell: kind 0 copy old r[2], 1 child r[1] of old r[2], 2 parent of old r[2..])
int run_transfer(exw_ctx ctx, int dim, int degree, const std::vector<int> &map,
                 const double *old_solution, double *new_solution, const char *who) {
  const int N = degree + 1, nch = 1 << dim, rec = 2 + nch;
  const int n_new = int(map.size() / rec);
  // 1D matrices H_b(i, j) = l_j((x_i + b)/2), R_b(i, j) = l_i((x_j + b)/2) w_j / (2 w_i)
  std::vector<double> tab(4 * N * N);
  double x[16], w[16], l[16];
  exw::cx::gauss01(N, x, w);
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < N; ++i) {
      exw::cx::lagrange(N, x, 0.5 * (x[i] + b), l, nullptr);
      for (int j = 0; j < N; ++j) {
        tab[b * N * N + i * N + j] = l[j];
        tab[2 * N * N + b * N * N + j * N + i] = l[j] * w[i] / (2.0 * w[j]);
      }
    }
  static TransferLaunch t2[exw::MAX_N] = {}, t3[exw::MAX_N] = {};
  if (!t2[2]) {
    fill_transfer<2, exw::MAX_DEGREE + 1>(t2);
    fill_transfer<3, exw::MAX_DEGREE_3D + 1>(t3);
  }
  TransferLaunch fn = dim == 2 ? t2[N] : t3[N];
  if (!fn) return afail(EXW_EUNSUP, std::string(who) + ": degree not compiled");
  int *d_map = nullptr;
  double *d_tab = nullptr;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc((void **)&d_map, std::max<size_t>(map.size(), 1) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&d_tab, tab.size() * sizeof(double));
  if (e == hipSuccess && !map.empty())
    e = hipMemcpyAsync(d_map, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = fn(n_new, d_map, d_tab, old_solution, new_solution, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d_map);
  hipFree(d_tab);
  if (e != hipSuccess) return afail(EXW_EDEVICE, std::string(who) + ": " + hipGetErrorString(e));
  return EXW_OK;
}


}  // namespace

extern "C" {

int exw_refinement_flags(int n_cells, const double *error_per_cell, const int *level,
                         int min_level, int max_level, double max_error_init, int *flags) {
  if (n_cells < 1 || !error_per_cell || !level || !flags)
    return afail(EXW_EINVAL, "exw_refinement_flags: invalid argument");
  const int n = n_cells;
  std::fill(flags, flags + n, 0);
  const int refine_cells = int(0.1 * n), coarsen_cells = int(0.6 * n);
  std::vector<double> tmp(error_per_cell, error_per_cell + n);
  if (refine_cells) {
    std::nth_element(tmp.begin(), tmp.begin() + refine_cells - 1, tmp.end(),
                     std::greater<double>());
    const double thr = tmp[refine_cells - 1];
    for (int c = 0; c < n; ++c)
      if (std::fabs(error_per_cell[c]) >= thr) flags[c] = 1;
  }
  if (coarsen_cells) {
    std::nth_element(tmp.begin(), tmp.begin() + (n - coarsen_cells), tmp.end(),
                     std::greater<double>());
    const double thr = tmp[n - coarsen_cells];
    for (int c = 0; c < n; ++c)
      if (std::fabs(error_per_cell[c]) <= thr && flags[c] != 1) flags[c] = -1;
  }
  // adapt_mesh (source/explicit_wave.cc:281-296)
  for (int c = 0; c < n; ++c) {
    if (flags[c] == 1 && level[c] == max_level)
      flags[c] = 0;
    else if (flags[c] == -1 && level[c] == min_level)
      flags[c] = 0;
    if (flags[c] == 1 && error_per_cell[c] < 0.1 * max_error_init) flags[c] = 0;
    if (error_per_cell[c] < 0.05 * max_error_init) flags[c] = -1;
  }
  return EXW_OK;
}

int exw_adapt_box_refine(int dim, int n_intervals, const int *refine_old, const int *flags,
                         int *refine_new) {
  if ((dim != 2 && dim != 3) || n_intervals < 1 || !refine_old || !flags || !refine_new)
    return afail(EXW_EINVAL, "exw_adapt_box_refine: invalid argument");
  const int nch = 1 << dim;
  long nbase = 1;
  for (int d = 0; d < dim; ++d) nbase *= n_intervals;
  long c = 0;
  for (long B = 0; B < nbase; ++B) {
    if (!refine_old[B]) {
      // an unrefined cell: refine if flagged (coarsening a base cell is not possible)
      refine_new[B] = flags[c] == 1 ? 1 : 0;
      ++c;
    } else {
      // children: coarsened only if every child is flagged for coarsening
      bool all = true;
      for (int ch = 0; ch < nch; ++ch) all &= flags[c + ch] == -1;
      refine_new[B] = all ? 0 : 1;
      c += nch;
    }
  }
  return EXW_OK;
}

int exw_adapt_box_transfer(exw_ctx ctx, int dim, int degree, int n_intervals,
                           const int *refine_old, const int *refine_new,
                           const double *old_solution, double *new_solution) {
  if (!ctx || (dim != 2 && dim != 3) || degree < 1 || degree > exw::max_degree(dim) ||
      n_intervals < 1 || !refine_old || !refine_new || !old_solution || !new_solution)
    return afail(EXW_EINVAL, "exw_adapt_box_transfer: invalid argument");
  const int nch = 1 << dim;
  long nbase = 1;
  for (int d = 0; d < dim; ++d) nbase *= n_intervals;
  std::vector<int> map;
  long co = 0;
  for (long B = 0; B < nbase; ++B) {
    const int ro = refine_old[B] ? 1 : 0, rn = refine_new[B] ? 1 : 0;
    if (ro == rn) {
      for (int ch = 0; ch < (ro ? nch : 1); ++ch) {
        map.push_back(0);
        map.push_back(0);
        map.push_back(int(co + ch));
        for (int i = 1; i < nch; ++i) map.push_back(-1);
      }
    } else if (rn) {  // refined now: children of old cell co
      for (int ch = 0; ch < nch; ++ch) {
        map.push_back(1);
        map.push_back(ch);
        map.push_back(int(co));
        for (int i = 1; i < nch; ++i) map.push_back(-1);
      }
    } else {  // coarsened: parent of old cells co .. co + nch - 1
      map.push_back(2);
      map.push_back(0);
      for (int ch = 0; ch < nch; ++ch) map.push_back(int(co + ch));
    }
    co += ro ? nch : 1;
  }
  return run_transfer(ctx, dim, degree, map, old_solution, new_solution,
                      "exw_adapt_box_transfer");
}

int exw_adapt_forest(int dim, int n_intervals, int n_tree, const int *tree, const int *flags,
                     int *n_tree_new, int *tree_new, int capacity) {
  exw::Forest F;
  int64_t used = 0;
  if (!tree || !flags || !n_tree_new || n_tree < 1 ||
      !F.parse(dim, n_intervals, tree, n_tree, &used) || used != n_tree)
    return afail(EXW_EINVAL, "exw_adapt_forest: invalid argument or malformed trees");
  const int nch = 1 << dim;
  using exw::Forest;
  // the new trees: refined nodes -1, active cells >= 0 (the value is unused)
  std::unordered_map<uint64_t, int> T = F.nodes();
  auto exists = [&](int l, const int64_t *I) {
    for (int d = 0; d < dim; ++d)
      if (I[d] < 0 || I[d] >= F.extent(l)) return false;
    return T.count(Forest::key(l, I)) > 0;
  };
  auto refine = [&](int l, const int64_t *I) {
    if (F.extent(l + 1) >= ((int64_t)1 << Forest::COORD_BITS)) return false;
    T[Forest::key(l, I)] = -1;
    for (int ch = 0; ch < nch; ++ch) {
      int64_t J[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) J[d] = 2 * I[d] + ((ch >> d) & 1);
      T[Forest::key(l + 1, J)] = 0;
    }
    return true;
  };
  // a node of level l + 2 in the ring of cells around (l, I) that share a
  // face, an edge or a corner with it, skipping those in 'gone'
  auto finer_around = [&](int l, const int64_t *I,
                          const std::unordered_set<uint64_t> *gone) -> bool {
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      lo[d] = 4 * I[d] - 1;
      hi[d] = 4 * I[d] + 4;
    }
    int64_t K[3] = {0, 0, 0};
    for (K[2] = lo[2]; K[2] <= hi[2]; ++K[2])
      for (K[1] = lo[1]; K[1] <= hi[1]; ++K[1])
        for (K[0] = lo[0]; K[0] <= hi[0]; ++K[0]) {
          bool ring = false;
          for (int d = 0; d < dim; ++d) ring |= K[d] == lo[d] || K[d] == hi[d];
          if (!ring || !exists(l + 2, K)) continue;
          if (gone) {
            int64_t P[3] = {K[0] >> 1, K[1] >> 1, K[2] >> 1};
            if (gone->count(Forest::key(l + 1, P))) continue;
          }
          return true;
        }
    return false;
  };
  const int64_t n = (int64_t)F.cells.size();
  for (int64_t c = 0; c < n; ++c)
    if (flags[c] == 1 && !refine(F.cells[c].level, F.cells[c].I))
      return afail(EXW_EINVAL, "exw_adapt_forest: too many levels");
  // 2:1 balance over faces, edges and corners: refine the coarser cell
  for (bool changed = true; changed;) {
    changed = false;
    std::vector<std::pair<int, std::array<int64_t, 3>>> act;
    for (const auto &kv : T)
      if (kv.second >= 0) {
        uint64_t k = kv.first;
        std::array<int64_t, 3> I;
        for (int d = 2; d >= 0; --d) {
          I[d] = int64_t(k & ((1u << Forest::COORD_BITS) - 1));
          k >>= Forest::COORD_BITS;
        }
        act.push_back({int(k), I});
      }
    for (const auto &a : act)
      if (finer_around(a.first, a.second.data(), nullptr)) {
        if (!refine(a.first, a.second.data()))
          return afail(EXW_EINVAL, "exw_adapt_forest: too many levels");
        changed = true;
      }
  }
  // coarsening: whole families of active cells flagged -1 that are still active
  std::vector<std::pair<int, std::array<int64_t, 3>>> cand;  // parents
  for (int64_t c = 0; c < n; ++c) {
    const exw::ForestCell &X = F.cells[c];
    bool first = X.level > 0 && flags[c] == -1;
    for (int d = 0; d < dim && first; ++d) first = (X.I[d] & 1) == 0;
    if (!first) continue;
    bool ok = true;
    for (int ch = 0; ch < nch && ok; ++ch) {
      int64_t J[3] = {X.I[0], X.I[1], X.I[2]};
      for (int d = 0; d < dim; ++d) J[d] += (ch >> d) & 1;
      const int y = F.find(X.level, J);
      ok = y >= 0 && flags[y] == -1 && T.at(Forest::key(X.level, J)) >= 0;
    }
    if (ok) cand.push_back({X.level - 1, {X.I[0] >> 1, X.I[1] >> 1, X.I[2] >> 1}});
  }
  // drop the coarsenings that would break the balance, until none does
  for (bool dropped = true; dropped;) {
    dropped = false;
    std::unordered_set<uint64_t> gone;
    for (const auto &p : cand) gone.insert(Forest::key(p.first, p.second.data()));
    for (size_t i = 0; i < cand.size();) {
      if (finer_around(cand[i].first, cand[i].second.data(), &gone)) {
        cand[i] = cand.back();
        cand.pop_back();
        dropped = true;
      } else {
        ++i;
      }
    }
  }
  for (const auto &p : cand) {
    T[Forest::key(p.first, p.second.data())] = 0;
    for (int ch = 0; ch < nch; ++ch) {
      int64_t J[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) J[d] = 2 * p.second[d] + ((ch >> d) & 1);
      T.erase(Forest::key(p.first + 1, J));
    }
  }
  const int64_t len = Forest::emit(dim, n_intervals, T, nullptr, 0);
  if (len < 0 || len > 2000000000L) return afail(EXW_EINVAL, "exw_adapt_forest: internal error");
  *n_tree_new = int(len);
  if (!tree_new) return EXW_OK;
  if (capacity < len) return afail(EXW_EINVAL, "exw_adapt_forest: capacity too small");
  Forest::emit(dim, n_intervals, T, tree_new, capacity);
  return EXW_OK;
}

int exw_adapt_forest_transfer(exw_ctx ctx, int dim, int degree, int n_intervals, int n_tree_old,
                              const int *tree_old, int n_tree_new, const int *tree_new,
                              const double *old_solution, double *new_solution) {
  if (!ctx || (dim != 2 && dim != 3) || degree < 1 || degree > exw::max_degree(dim) ||
      n_intervals < 1 || !tree_old || !tree_new || n_tree_old < 1 || n_tree_new < 1 ||
      !old_solution || !new_solution)
    return afail(EXW_EINVAL, "exw_adapt_forest_transfer: invalid argument");
  const int nch = 1 << dim;
  std::vector<int> map;
  int64_t po = 0, pn = 0;
  int co = 0;
  bool ok = true;
  // both trees walked together, depth first
  std::function<void()> walk = [&]() {
    if (!ok || po >= n_tree_old || pn >= n_tree_new) {
      ok = false;
      return;
    }
    const int a = tree_old[po++], b = tree_new[pn++];
    if (a == 0 && b == 0) {
      map.insert(map.end(), {0, 0, co++});
      map.insert(map.end(), nch - 1, -1);
    } else if (a == 0 && b == 1) {
      for (int ch = 0; ch < nch; ++ch) {
        if (pn >= n_tree_new || tree_new[pn++] != 0) {
          ok = false;
          return;
        }
        map.insert(map.end(), {1, ch, co});
        map.insert(map.end(), nch - 1, -1);
      }
      ++co;
    } else if (a == 1 && b == 0) {
      map.insert(map.end(), {2, 0});
      for (int ch = 0; ch < nch; ++ch) {
        if (po >= n_tree_old || tree_old[po++] != 0) {
          ok = false;
          return;
        }
        map.push_back(co++);
      }
    } else if (a == 1 && b == 1) {
      for (int ch = 0; ch < nch && ok; ++ch) walk();
    } else {
      ok = false;
    }
  };
  int64_t nbase = 1;
  for (int d = 0; d < dim; ++d) nbase *= n_intervals;
  for (int64_t B = 0; B < nbase && ok; ++B) walk();
  if (!ok || po != n_tree_old || pn != n_tree_new)
    return afail(EXW_EINVAL,
                 "exw_adapt_forest_transfer: the trees are malformed or differ by more than "
                 "one level");
  return run_transfer(ctx, dim, degree, map, old_solution, new_solution,
                      "exw_adapt_forest_transfer");
}

}  // extern "C"
