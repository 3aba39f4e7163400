// dg_ader_fused.hpp -- the ADER update sweep of one step fused with the
// Cauchy-Kowalevski time integral of the next.
//
// One ADER step (exw_ader_step; ArbitraryHighOrderDG::perform_time_step,
// source/time_integrators.cc:37-50, with apply_ader's two phases,
// source/wave_equation_operations.cc:1266-1342) is
//   w     = CK(u_n)                    cell-local (ader_ck_kernel)
//   u_n+1 = u_n + M^-1 A(w)            reads w of the face neighbours
// The CK of step n+1 needs only the cell's own u_n+1, which the update sweep
// has in registers: this kernel runs the sweep (dg_operator_cart_kernel's
// arithmetic, b = o1 = u, a1 = 1), stores u_n+1, and continues on the same
// cell with the CK of step n+1 (ader_ck_kernel's arithmetic), writing w'.
// The neighbours still read w while w' is written, so w and w' are two
// vectors.  Both halves are the shared device functions of the two kernels,
// so n steps through this kernel give the bits of n exw_ader_step calls.
//
// What it saves: the CK kernel's read of u_n+1 and one launch per step, and
// -- the point -- the CK's LDS-bound arithmetic overlaps the sweep's memory
// traffic of other groups on the same CU instead of running in a kernel of
// its own.  Cartesian cells whose sweep and CK groups coincide (3D k = 3..5).
#pragma once
#include "dg_ader.hpp"
#include "dg_operator_cart.hpp"

namespace exw {

template <int DIM, int K>
struct AderFusedShape {
  using CS = CartShape<DIM, K>;
  using G = CkGroup<DIM, K>;
  static constexpr int CPB = CS::CPB, BLOCK = CS::BLOCK;
  static constexpr int LOW = CkShape<K>::low_off(DIM, CkShape<K>::N_LEVELS) > 0
                                 ? CkShape<K>::low_off(DIM, CkShape<K>::N_LEVELS)
                                 : 1;
  // the two kernels' groups hold the same cells on the same lanes
  static constexpr bool ENABLED =
      DIM == 3 && K >= 3 && K <= 5 && CS::ITER == 1 && CS::CPB == G::CPB && CS::BLOCK == G::BLOCK;
};

// u (n_cells cells) <- u + M^-1 A(w_in); w_out <- CK(u) over [0, dt] about 0.
// w_in and w_out must differ; u is read and written in place.
template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((AderFusedShape<DIM, K>::BLOCK))
    ader_sweep_ck_kernel(const DevMesh mesh, const double *__restrict__ w_in, double *u,
                         const double *__restrict__ cktab, const CkTab ckt, double dt,
                         double *__restrict__ w_out) {
  using F = AderFusedShape<DIM, K>;
  using S = Shape<DIM, K>;
  using CS = CkShape<K>;
  constexpr int ND = S::ND, NF = S::NF, CPB = F::CPB, LOW = F::LOW;
  // the sweep's LDS, then (after a barrier) the CK's in the same bytes
  // (3D k=4: 16 KB, 10 groups per CU)
  __shared__ union {
    CartLds<DIM, K, Number> sweep;
    struct {
      Number a[CPB][NF][ND];
      Number b[CPB][NF][ND];
      Number c[CPB][NF][LOW];
      Number tab[CS::TAB_SIZE];
    } ck;
  } lds;

  const int tid = threadIdx.x;

  // ---- the update sweep of step n (dg_operator_cart_kernel, b = o1 = u, a1 = 1) ----
  Epilogue epi{};
  epi.b = u;
  epi.o1 = u;
  epi.a1 = 1.0;
  int lc, cell;
  bool swept;
  double cscale, bv[NF];
  Number k[NF];
  cart_cell<DIM, K, Number>(mesh, w_in, epi, xcd_remap(blockIdx.x, gridDim.x), lds.sweep, lc,
                            cell, swept, cscale, bv, k);
  double un[NF];
  if (swept) {
    const int lane = tid - lc * CartShape<DIM, K>::LANES;
    const double ea1 = epi.a1 * cscale;
    const size_t cbase = (size_t)cell * NF * ND;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      un[f] = bv[f] + ea1 * double(k[f]);
      __builtin_nontemporal_store(un[f], u + cbase + f * ND + lane);
    }
  }
  __syncthreads();  // every lane is done with the sweep's LDS
  // the CK tables (visible after ck_cell's first barrier)
  for (int i = tid; i < CS::TAB_SIZE && i < ckt.size; i += blockDim.x)
    lds.ck.tab[i] = Number(cktab[i]);

  // ---- the CK of step n+1 on the same cells (ader_ck_kernel's set-up) ----
  const int lck = tid / ND;
  const int q = tid - lck * ND;
  const bool lane_ok = lck < CPB;
  const int lcc = lane_ok ? lck : 0;
  // the same cells on the same lanes: a node lane of the CK is a swept lane
  const bool active = lane_ok && swept;
  Number rho_inv, rc2, ih[DIM];
  ck_factors<DIM, true>(mesh, cell, active, rho_inv, rc2, ih);
  Number uc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) uc[f] = active ? Number(un[f]) : Number(0);
  ck_cell<DIM, K, Number, true>(mesh, nullptr, lds.ck.tab, ckt, dt, 0.0, 0.0, lds.ck.a, lds.ck.b,
                                lds.ck.c, nullptr, lck, lcc, q, lane_ok, active,
                                (size_t)(active ? cell : 0) * NF * ND, rho_inv, rc2, ih, uc,
                                w_out);
}

}  // namespace exw
