// launchers.hpp -- typed launchers and per-(dim,degree,precision,geometry)
// dispatch tables of the MI355X kernels.
#pragma once
#include <stdlib.h>

#include <type_traits>

#include "dg_ader.hpp"
#include "dg_ader_fused.hpp"
#include "dg_lts.hpp"
#include "dg_misc.hpp"
#include "dg_operator.hpp"
#include "dg_operator_cart.hpp"
#include "dg_operator_cartg.hpp"
#include "dg_operator_big.hpp"
#include "dg_operator_gen.hpp"
#include "dg_operator_pipe.hpp"
#include "dg_nc.hpp"
#include "dg_post.hpp"

namespace exw {

// Degrees compiled: 1..12 in 2D (the reference's range,
// source/wave_equation_operations.cc:1695-1802), 1..12 in 3D: up to k = 9 one lane per
// node fits a 1024-lane workgroup; k >= 10 (1331+ nodes) takes the big-cell
// operator kernel, and only the operator, projection and norms exist there.
constexpr int MAX_DEGREE = 12;
constexpr int MAX_DEGREE_3D = 12;
inline int max_degree(int dim) { return dim == 2 ? MAX_DEGREE : MAX_DEGREE_3D; }

// (mesh, src, device basis table, host basis table, epilogue, stream)
using OpLaunch = hipError_t (*)(const DevMesh &, const double *, const double *, const double *,
                                const Epilogue &, hipStream_t);
// (mesh, src, recon, CK table, CK offsets, t2, t1, te, cell selection (NULL:
// all cells, uniform te), out, stream)
using CkLaunch = hipError_t (*)(const DevMesh &, const double *, const double *, const double *,
                                const CkTab &, double, double, double, const CellSel *, double *,
                                hipStream_t);
// (mesh, basis table, sweep arguments, stream)
using LtsLaunch = hipError_t (*)(const DevMesh &, const double *, const LtsArgs &, hipStream_t);
using ProjLaunch = hipError_t (*)(const DevMesh &, const double *, const DevModes &, double,
                                  double *, hipStream_t);
using ErrLaunch = hipError_t (*)(const DevMesh &, const double *, const DevModes &, double,
                                 const double *, double *, int, hipStream_t);
using EnergyLaunch = hipError_t (*)(const DevMesh &, const double *, const double *, double *,
                                    int, hipStream_t);
// (mesh, basis table, solution, apply(solution), per-cell error, stream)
using EstLaunch = hipError_t (*)(const DevMesh &, const double *, const double *, const double *,
                                 double *, hipStream_t);
// (mesh, table, cells with hanging faces, src, epilogue, stream)
using NcLaunch = hipError_t (*)(const DevMesh &, const double *, const NcArgs &, const double *,
                                const Epilogue &, hipStream_t);
// (mesh, post table, solution, apply(solution), post, iterations, solver, stream)
using PostLaunch = hipError_t (*)(const DevMesh &, const double *, const double *, const double *,
                                  double *, int *, const PostSolver &, hipStream_t);
// (mesh, post table, modes, t, post, partials, n blocks, stream)
using PostErrLaunch = hipError_t (*)(const DevMesh &, const double *, const DevModes &, double,
                                     const double *, double *, int, hipStream_t);

// (mesh, first cell, cell count, per-cell flux coefficients out, stream)
using CoefLaunch = hipError_t (*)(const DevMesh &, int, int, double *, hipStream_t);
// (mesh, w_in, u (in place), CK table, CK offsets, dt, w_out, stream)
using AderFusedLaunch = hipError_t (*)(const DevMesh &, const double *, double *, const double *,
                                       const CkTab &, double, double *, hipStream_t);

struct DimKernels {
  CoefLaunch coef;
  OpLaunch op[MAX_DEGREE + 1][2][2];  // [k][fp32][cart]
  CkLaunch ck[MAX_DEGREE + 1][2][2];
  LtsLaunch lts[MAX_DEGREE + 1][2];  // [k][cart]
  ProjLaunch project[MAX_DEGREE + 1];
  ErrLaunch l2err[MAX_DEGREE + 1];
  EnergyLaunch energy[MAX_DEGREE + 1];
  PostLaunch post[MAX_DEGREE + 1];
  NcLaunch nc[MAX_DEGREE + 1];
  EstLaunch estimate[MAX_DEGREE + 1];
  PostErrLaunch post_error[MAX_DEGREE + 1];
  AderFusedLaunch ader_fused[MAX_DEGREE + 1][2];  // [k][fp32], Cartesian; NULL: none
  int block[MAX_DEGREE + 1];
  int cells_per_block[MAX_DEGREE + 1];
  // the FP64 Cartesian sweeps of degree k may run dg_operator_pipe_kernel,
  // which reads the per-cell flux coefficients (DevMesh::coef)
  bool pipe[MAX_DEGREE + 1];
};

// The software-pipelined Cartesian kernel is the default where it applies;
// EXW_CART_PIPE=0 in the environment selects dg_operator_cart_kernel instead.
inline bool cart_pipe_enabled() {
  static const bool on = [] {
    const char *v = getenv("EXW_CART_PIPE");
    return !(v && v[0] == '0');
  }();
  return on;
}

// The general-geometry kernel of one cell per group (dg_operator_gen_kernel)
// where it applies; EXW_GEN_KERNEL=0 selects dg_operator_kernel throughout.
inline bool gen_kernel_enabled() {
  static const bool on = [] {
    const char *v = getenv("EXW_GEN_KERNEL");
    return !(v && v[0] == '0');
  }();
  return on;
}

// The gathered-trace Cartesian kernel (dg_operator_cartg_kernel) where it
// applies; EXW_CART_GATHER=0 selects dg_operator_cart_kernel.  Read at every
// launch (a scan of the environment, far below a launch's cost), so that one
// process can run both kernels and compare them.
inline bool cart_gather_enabled() {
  const char *v = getenv("EXW_CART_GATHER");
  return !(v && v[0] == '0');
}

#ifndef EXW_PIPE_TPG
#define EXW_PIPE_TPG 2
#endif

template <int DIM>
hipError_t launch_cart_coef(const DevMesh &m, int first, int count, double *coef,
                            hipStream_t s) {
  if (first < 0 || count < 0 || first + count > m.n_cells) return hipErrorInvalidValue;
  const int64_t n = (int64_t)count * 2 * DIM;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((cart_coef_kernel<DIM>), dim3(int((n + 255) / 256)), dim3(256), 0, s, m,
                     first, count, coef);
  return hipGetLastError();
}

template <int DIM, int K, typename Number, bool CART>
hipError_t launch_op(const DevMesh &m, const double *src, const double *tab,
                     const double *host_tab, const Epilogue &e, hipStream_t s) {
  // The pipelined kernel where it was measured faster: 2D (k = 6..8) always;
  // 3D k = 3..5 for the low-storage RK stages only (two outputs; their last,
  // single-output stage stays on the same kernel: both measured equal there)
  // -- on the other single-output sweeps (ADER, Euler, apply) the one-cell
  // kernel's higher occupancy wins by 3-4 %, and at 3D k = 6 (one 111 KB group
  // per CU) by 16 %.
  if constexpr (CART && std::is_same_v<Number, double> && PipeShape<DIM, K>::ENABLED &&
                (DIM == 2 || K <= 5)) {
    constexpr uint64_t cell_bytes = uint64_t(PipeShape<DIM, K>::NF) * PipeShape<DIM, K>::ND * 8;
    if (cart_pipe_enabled() && m.coef &&
        (DIM == 2 ? (!e.o2 || e.o2_from_o1) : ((e.o2 && e.o2_from_o1) || (e.rk_stage && !e.o2))) &&
        uint64_t(m.n_cells) * cell_bytes < (uint64_t(1) << 32) &&
        (!PipeShape<DIM, K>::X4 ||
         ((uintptr_t)src % 16 == 0 && (uintptr_t)e.b % 16 == 0))) {
      constexpr int TPG = EXW_PIPE_TPG;
      const int n = e.cell_list ? e.n_list : m.n_cells;
      const int grid = (n + TPG - 1) / TPG;
      if (grid == 0) return hipSuccess;
      if (e.cell_list || e.cell_scale)
        hipLaunchKernelGGL((dg_operator_pipe_kernel<DIM, K, TPG, true>), dim3(grid),
                           dim3(PipeShape<DIM, K>::BLOCK), 0, s, m, src, e);
      else
        hipLaunchKernelGGL((dg_operator_pipe_kernel<DIM, K, TPG, false>), dim3(grid),
                           dim3(PipeShape<DIM, K>::BLOCK), 0, s, m, src, e);
      return hipGetLastError();
    }
  }
  if constexpr (CART && CartGShape<DIM, K>::ENABLED) {
    // large box cells (3D k >= 6): neighbour traces gathered from memory, 7
    // fields of LDS instead of 16 (3D k=6 32^3 LSRK45 stage: 107 -> 142 GDoF/s)
    if (cart_gather_enabled()) {
      const int n = e.cell_list ? e.n_list : m.n_cells;
      if (n == 0) return hipSuccess;
      hipLaunchKernelGGL((dg_operator_cartg_kernel<DIM, K, Number>), dim3(n),
                         dim3(CartGShape<DIM, K>::BLOCK), 0, s, m, src, e);
      return hipGetLastError();
    }
  }
  if constexpr (CART) {
    // box cells: node/line/node kernel with compile-time 1D matrices
    using CS = CartShape<DIM, K>;
    const int n = e.cell_list ? e.n_list : m.n_cells;
    const int grid = (n + CS::CPB * CS::ITER - 1) / (CS::CPB * CS::ITER);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((dg_operator_cart_kernel<DIM, K, Number>), dim3(grid), dim3(CS::BLOCK), 0,
                       s, m, src, e);
  } else {
    using S = Shape<DIM, K>;
    if (e.cell_list) return hipErrorInvalidValue;  // cell lists: Cartesian kernel only
    if constexpr (GenShape<DIM, K>::ENABLED) {
      if (gen_kernel_enabled()) {
        // one cell per group (node/face/node, geometry computed once per cell)
        if (m.n_cells == 0) return hipSuccess;
        hipLaunchKernelGGL((dg_operator_gen_kernel<DIM, K, Number>), dim3(m.n_cells),
                           dim3(GenShape<DIM, K>::BLOCK), 0, s, m, src, tab, e);
        return hipGetLastError();
      }
    }
    const int grid = (m.n_cells + S::CPB - 1) / S::CPB;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((dg_operator_kernel<DIM, K, Number, CART>), dim3(grid), dim3(S::BLOCK), 0,
                       s, m, src, tab, e);
  }
  return hipGetLastError();
}

template <int DIM, int K, typename Number>
hipError_t launch_ader_fused(const DevMesh &m, const double *w_in, double *u, const double *cktab,
                             const CkTab &ck, double dt, double *w_out, hipStream_t s) {
  if constexpr (AderFusedGShape<DIM, K>::ENABLED && std::is_same_v<Number, double>) {
    // gathered neighbour traces (FP64, 3D k = 3..5)
    using G = AderFusedGShape<DIM, K>;
    const int grid = (m.n_cells + G::CPB - 1) / G::CPB;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((ader_sweepg_ck_kernel<DIM, K, Number>), dim3(grid),
                       dim3(AderFusedGShape<DIM, K>::BLOCK), 0, s, m, w_in, u, cktab, ck, dt, w_out);
    return hipGetLastError();
  }
  using F = AderFusedShape<DIM, K>;
  const int grid = (m.n_cells + F::CPB - 1) / F::CPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((ader_sweep_ck_kernel<DIM, K, Number>), dim3(grid), dim3(F::BLOCK), 0, s, m,
                     w_in, u, cktab, ck, dt, w_out);
  return hipGetLastError();
}

// cells of more nodes than a workgroup has lanes (3D k >= 10), any geometry;
// no cell lists (the serving pools stop at 3D k = 9)
template <int DIM, int K, typename Number>
hipError_t launch_op_big(const DevMesh &m, const double *src, const double *tab,
                         const double *host_tab, const Epilogue &e, hipStream_t s) {
  if (e.cell_list) return hipErrorInvalidValue;
  if (m.n_cells == 0) return hipSuccess;
  hipLaunchKernelGGL((dg_operator_big_kernel<DIM, K, Number>), dim3(m.n_cells),
                     dim3(BigShape<DIM, K>::BLOCK), 0, s, m, src, tab, e);
  return hipGetLastError();
}

template <int DIM, int K, typename Number, bool CART>
hipError_t launch_ck(const DevMesh &m, const double *src, const double *recon,
                     const double *cktab, const CkTab &ck, double t2, double t1, double te,
                     const CellSel *sel, double *out, hipStream_t s) {
  using G = CkGroup<DIM, K>;
  CellSel cs{};
  if (sel) cs = *sel;
  const int n = cs.list ? cs.n : m.n_cells;
  const int grid = (n + G::CPB - 1) / G::CPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((ader_ck_kernel<DIM, K, Number, CART>), dim3(grid), dim3(G::BLOCK), 0, s,
                     m, src, recon, cktab, ck, t2, t1, te, cs, out);
  return hipGetLastError();
}

template <int DIM, int K, bool CART>
hipError_t launch_lts(const DevMesh &m, const double *tab, const LtsArgs &a, hipStream_t s) {
  using S = Shape<DIM, K>;
  const int grid = (a.n + S::CPB - 1) / S::CPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((dg_lts_kernel<DIM, K, CART>), dim3(grid), dim3(S::BLOCK), 0, s, m, tab, a);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_project(const DevMesh &m, const double *tab, const DevModes &md, double t,
                          double *dst, hipStream_t s) {
  using S = Shape<DIM, K>;
  const size_t total = (size_t)m.n_cells * S::ND;
  const int grid = int((total + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((project_kernel<DIM, K>), dim3(grid), dim3(256), 0, s, m, tab, md, t, dst);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_l2err(const DevMesh &m, const double *errtab, const DevModes &md, double t,
                        const double *u, double *partial, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((l2_error_kernel<DIM, K>), dim3(nblocks), dim3(256), 0, s, m, errtab, md, t,
                     u, partial);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_energy(const DevMesh &m, const double *tab, const double *u, double *partial,
                         int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((energy_kernel<DIM, K>), dim3(nblocks), dim3(256), 0, s, m, tab, u, partial);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_estimate(const DevMesh &m, const double *tab, const double *sol,
                           const double *dudt, double *err, hipStream_t s) {
  using S = Shape<DIM, K>;
  const int grid = (m.n_cells + S::CPB - 1) / S::CPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((estimate_error_kernel<DIM, K>), dim3(grid), dim3(S::BLOCK), 0, s, m, tab,
                     sol, dudt, err);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_nc(const DevMesh &m, const double *tab, const NcArgs &nc, const double *src,
                     const Epilogue &e, hipStream_t s) {
  if (nc.n == 0) return hipSuccess;
  hipLaunchKernelGGL((nc_face_kernel<DIM, K>), dim3(nc.n), dim3(NcShape<DIM, K>::BLOCK), 0, s, m,
                     tab, nc, src, e);
  return hipGetLastError();
}

// the post-processing kernel puts one lane on each of the (k+2)^dim
// Gauss-Lobatto nodes of a cell (3D k = 9: 1331 > 1024 lanes, not compiled)
template <int DIM, int K>
constexpr bool post_fits() {
  constexpr int P = K + 2;
  return (DIM == 2 ? P * P : P * P * P) <= 1024;
}

template <int DIM, int K>
hipError_t launch_post(const DevMesh &m, const double *tab, const double *sol, const double *dudt,
                       double *post, int *iterations, const PostSolver &cfg, hipStream_t s) {
  using S = PostShape<DIM, K>;
  const int grid = (m.n_cells + S::CPB - 1) / S::CPB;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((post_pressure_kernel<DIM, K>), dim3(grid), dim3(S::BLOCK), 0, s, m, tab,
                     sol, dudt, post, iterations, cfg);
  return hipGetLastError();
}

template <int DIM, int K>
hipError_t launch_post_error(const DevMesh &m, const double *tab, const DevModes &md, double t,
                             const double *post, double *partial, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((post_error_kernel<DIM, K>), dim3(nblocks), dim3(256), 0, s, m, tab, md, t,
                     post, partial);
  return hipGetLastError();
}

template <int DIM, int K>
void fill_dim_kernels(DimKernels &t) {
  t.coef = launch_cart_coef<DIM>;
  if constexpr (K >= 1 && Shape<DIM, K>::ND > 1024) {
    // more nodes than lanes: the operator (any geometry and precision), the
    // projection and the norms; CK, LTS, hanging faces, the error estimator
    // and the post-processing stay unsupported
    static_assert(BigShape<DIM, K>::ENABLED, "big-cell kernel for cells of more than 1024 nodes");
    for (int p = 0; p < 2; ++p)
      for (int c = 0; c < 2; ++c)
        t.op[K][p][c] = p == 0 ? launch_op_big<DIM, K, double> : launch_op_big<DIM, K, float>;
    t.project[K] = launch_project<DIM, K>;
    t.l2err[K] = launch_l2err<DIM, K>;
    t.energy[K] = launch_energy<DIM, K>;
    t.block[K] = BigShape<DIM, K>::BLOCK;
    t.cells_per_block[K] = 1;
    t.pipe[K] = false;
    fill_dim_kernels<DIM, K - 1>(t);
  } else if constexpr (K >= 1) {
    t.op[K][0][0] = launch_op<DIM, K, double, false>;
    t.op[K][0][1] = launch_op<DIM, K, double, true>;
    t.op[K][1][0] = launch_op<DIM, K, float, false>;
    t.op[K][1][1] = launch_op<DIM, K, float, true>;
    t.ck[K][0][0] = launch_ck<DIM, K, double, false>;
    t.ck[K][0][1] = launch_ck<DIM, K, double, true>;
    t.ck[K][1][0] = launch_ck<DIM, K, float, false>;
    t.ck[K][1][1] = launch_ck<DIM, K, float, true>;
    t.lts[K][0] = launch_lts<DIM, K, false>;
    t.lts[K][1] = launch_lts<DIM, K, true>;
    t.project[K] = launch_project<DIM, K>;
    t.l2err[K] = launch_l2err<DIM, K>;
    t.energy[K] = launch_energy<DIM, K>;
    if constexpr (post_fits<DIM, K>()) t.post[K] = launch_post<DIM, K>;
    t.nc[K] = launch_nc<DIM, K>;
    t.estimate[K] = launch_estimate<DIM, K>;
    if constexpr (post_fits<DIM, K>()) t.post_error[K] = launch_post_error<DIM, K>;
    if constexpr (AderFusedShape<DIM, K>::ENABLED) {
      t.ader_fused[K][0] = launch_ader_fused<DIM, K, double>;
      t.ader_fused[K][1] = launch_ader_fused<DIM, K, float>;
    }
    t.block[K] = Shape<DIM, K>::BLOCK;
    t.cells_per_block[K] = Shape<DIM, K>::CPB;
    t.pipe[K] = PipeShape<DIM, K>::ENABLED && (DIM == 2 || K <= 5);
    fill_dim_kernels<DIM, K - 1>(t);
  }
}

const DimKernels &kernels_2d();
const DimKernels &kernels_3d();

}  // namespace exw
