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
#include <functional>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "capi_internal.hpp"
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
  if (!ctx || (dim != 2 && dim != 3) || degree < 1 || degree > 8 || n_intervals < 1 ||
      !refine_old || !refine_new || !old_solution || !new_solution)
    return afail(EXW_EINVAL, "exw_adapt_box_transfer: invalid argument");
  const int N = degree + 1, nch = 1 << dim, rec = 2 + nch;
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
  static TransferLaunch t2[10] = {}, t3[10] = {};
  if (!t2[2]) {
    fill_transfer<2, 9>(t2);
    fill_transfer<3, 9>(t3);
  }
  TransferLaunch fn = dim == 2 ? t2[N] : t3[N];
  int *d_map = nullptr;
  double *d_tab = nullptr;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc((void **)&d_map, map.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&d_tab, tab.size() * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_map, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = fn(n_new, d_map, d_tab, old_solution, new_solution, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d_map);
  hipFree(d_tab);
  if (e != hipSuccess)
    return afail(EXW_EDEVICE, std::string("exw_adapt_box_transfer: ") + hipGetErrorString(e));
  return EXW_OK;
}

}  // extern "C"
