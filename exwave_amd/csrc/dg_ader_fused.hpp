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
// its own.  Cartesian cells whose sweep and CK groups coincide (3D k = 3..5);
// FP64 runs ader_sweepg_ck_kernel (below) instead, whose sweep half
// gathers the neighbour traces instead of staging the neighbours in LDS.
#pragma once
#include "dg_ader.hpp"
#include "dg_operator_cart.hpp"
#include "dg_operator_pipe.hpp"

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

// The same fusion with the sweep half in the gathered-trace form of
// dg_operator_cartg_kernel (FP64, 3D k = 3..5: one cell per group at k = 4, 5,
// two one-wave cells at k = 3): the line lane of line t of direction d (lanes
// 0 .. 3 (k+1)^2 - 1 of the cell's, the directions mixed within a wave) sums
// the neighbour traces of faces 2d, 2d+1 at its face
// points straight from the neighbours' nodes in memory, instead of the
// neighbours' fields being staged in LDS.  The LDS of the sweep shrinks to the
// cell's fields and the lines' dp/dt shares (3D k=4: 7.4 KB; the union is then
// the CK's 9.4 KB instead of 16 KB) and the registers to 68 (k=4), so a CU
// holds 14 groups instead of 10 for the CK's latency-bound phases: 3D k=4 64^3
// 90.1 -> 97.9, k=5 40^3 71.8 -> 84.0, k=3 64^3 115 -> 127 GDoF/s per ADER
// step (FP64).  The FP32
// operator keeps the staged form (117 against 107 GDoF/s at k=4).
// The arithmetic is dg_operator_cart_kernel's expression for expression (the
// traces' chains over the same values, the line phase in the three passes of
// dg_operator_cartg_kernel, the pressure shares summed in the same order): the
// bits of ader_sweep_ck_kernel and of single steps (tested).
template <int DIM, int K>
struct AderFusedGShape {
  using G = CkGroup<DIM, K>;
  static constexpr int N = K + 1, ND = G::ND, NT = N * N, NLINE = DIM * NT, NF = DIM + 1;
  static constexpr int CPB = G::CPB, BLOCK = G::BLOCK, LANES = BLOCK / CPB;  // lanes per cell
  // each cell on whole waves, with a lane for each of its nodes and lines
  static constexpr bool ENABLED = DIM == 3 && K >= 3 && K <= 5 && CPB * LANES == BLOCK &&
                                  LANES % 64 == 0 && ND <= LANES && NLINE <= LANES;
};

// u (n_cells cells, CPB per group) <- u + M^-1 A(w_in); w_out <- CK(u) over
// [0, dt] about 0.  w_in and w_out must differ; u is read and written in place.
template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((AderFusedGShape<DIM, K>::BLOCK))
    ader_sweepg_ck_kernel(const DevMesh mesh, const double *__restrict__ w_in, double *u,
                          const double *__restrict__ cktab, const CkTab ckt, double dt,
                          double *__restrict__ w_out) {
  using F = AderFusedGShape<DIM, K>;
  using CS = CkShape<K>;
  constexpr int N = F::N, ND = F::ND, NT = F::NT, NF = F::NF, NFACE = 2 * DIM;
  constexpr int NLINE = F::NLINE, CPB = F::CPB, LANES = F::LANES;
  constexpr int LOW = AderFusedShape<DIM, K>::LOW;
  constexpr const CartConst<N> &T = CartTable<N>::value;
  __shared__ __align__(16) union {
    struct {
      Number u[CPB][NF][ND];
      Number kp[CPB][DIM][ND];
      Number fc[CPB][NFACE][8];
      Number vc[CPB][DIM][2];
    } sw;
    struct {
      Number a[CPB][NF][ND];
      Number b[CPB][NF][ND];
      Number c[CPB][NF][LOW];
      Number tab[CS::TAB_SIZE];
    } ck;
  } lds;

  // the lane's cell slot (uniform over a wave) and its lane within the cell;
  // a slot past the last cell (the last group) computes on the last cell and
  // stores nothing
  const int tid = threadIdx.x;
  const int lc = __builtin_amdgcn_readfirstlane(tid / LANES);
  const int lt = tid - lc * LANES;
  const int e = xcd_remap(blockIdx.x, gridDim.x) * CPB + lc;
  const bool swept = e < mesh.n_cells;
  const int cell = swept ? e : mesh.n_cells - 1;
  auto &s_u = lds.sw.u[lc];
  auto &s_kp = lds.sw.kp[lc];
  auto &s_fc = lds.sw.fc[lc];
  auto &s_vc = lds.sw.vc[lc];
  int nbr[NFACE];
#pragma unroll
  for (int f = 0; f < NFACE; ++f) nbr[f] = pipe::sld(mesh.neighbor, (size_t)cell * NFACE + f);
  const size_t cbase = (size_t)cell * NF * ND;
  // ---- the update sweep of step n: the cell's fields (LDS-DMA for a 16-byte
  // aligned FP64 vector, else through registers), the lines' neighbour traces ----
  const int nl = lt < ND ? lt : ND - 1;
  constexpr int OWN_BYTES = NF * ND * 8, NDMA = (OWN_BYTES + 1023) / 1024;
  const bool dma = std::is_same_v<Number, double> && (uintptr_t)w_in % 16 == 0;
  double uo[NF];
  if (dma) {
    // the cell slot's waves share the instructions
    const int wave = __builtin_amdgcn_readfirstlane(lt >> 6);
    const uint32_t lbase = pipe::lds_addr(&s_u[0][0]);
    const uint32_t lane16 = uint32_t(tid & 63) * 16u;
    const double *own = w_in + cbase;
    static_for<NDMA>([&](auto I_) __attribute__((always_inline)) {
      constexpr int I = decltype(I_)::value;
      if (wave == I % (LANES / 64) && int(I * 1024 + lane16) < OWN_BYTES)
        pipe::dma_dwordx4(lbase + I * 1024, own, lane16 + I * 1024);
    });
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) uo[f] = w_in[cbase + f * ND + nl];
  }
  const bool line_ok = lt < NLINE;
  const int tl = line_ok ? lt : 0;
  const int d = tl / NT, t = tl - d * NT;
  int base, sd;
  {
    const int t0 = t % N, t1 = t / N;
    base = d == 0 ? t0 * N + t1 * N * N : (d == 1 ? t0 + t1 * N * N : t0 + t1 * N);
    sd = d == 0 ? 1 : (d == 1 ? N : N * N);
  }
  Number tv[2], tp[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int nb = d == 0 ? nbr[s] : (d == 1 ? nbr[2 + s] : nbr[4 + s]);
    const double *un = w_in + (size_t)(nb >= 0 ? nb : cell) * NF * ND;
    const double *uv = un + d * ND + base, *up = un + DIM * ND + base;
    double xv[N], xp[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      xv[i] = uv[i * sd];
      xp[i] = up[i * sd];
    }
    Number vn = Number(0), pn = Number(0);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const Number lo = Number(T.L[(1 - s) * N + i]);
      vn += lo * Number(xv[i]);
      pn += lo * Number(xp[i]);
    }
    tv[s] = vn;
    tp[s] = pn;
  }
  // (the sums taken by an empty asm, so that the loads' registers are free
  // again before the barrier)
  asm volatile("" : "+v"(tv[0]), "+v"(tv[1]), "+v"(tp[0]), "+v"(tp[1]));
  if (!dma && lt < ND) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[f][lt] = Number(uo[f]);
  }
  if (lt < NFACE) {
    // face lane: the linear flux map and the cell term's factors
    // (dg_operator_cart_kernel's face-lane expressions)
    const int f = lt, fd = f >> 1;
    int nb = nbr[0];
#pragma unroll
    for (int g = 1; g < NFACE; ++g)
      if (lt == g) nb = nbr[g];
    const double rho = mesh.rho[cell], cc = mesh.speed[cell];
    const double rho_m = nb >= 0 ? mesh.rho[nb] : rho, c_m = nb >= 0 ? mesh.speed[nb] : cc;
    const double ih = recip(mesh.h[(size_t)cell * DIM + fd]);
    const double sg = (f & 1) ? 1.0 : -1.0;
    const double ihf = nb == -NC_BID ? 0.0 : ih;
    double a[4], b[4];
    flux_matrix<double>(rho, cc, rho_m, c_m, nb >= 0 ? 0 : -nb, a, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double in = (j == 0 || j == 2) ? sg : 1.0;
      s_fc[f][j] = Number(a[j] * in * sg * ihf);
      s_fc[f][4 + j] = Number(b[j] * in * ihf);
    }
    if ((f & 1) == 0) {
      s_vc[fd][0] = Number(-ih * recip(rho));
      s_vc[fd][1] = Number(rho * cc * cc * ih);
    }
  }
  if (dma) pipe::wait_vm<0>();
  __syncthreads();
  double bv[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) bv[f] = __builtin_nontemporal_load(u + cbase + f * ND + nl);

  // the line phase (dg_operator_cartg_kernel's three passes; the memory
  // clobbers make the second and third pass read the line from LDS again)
  if (line_ok) {
    const Number cvv = s_vc[d][0], cvp = s_vc[d][1];
    Number fv[2], fp[2];
    {
      Number vo[2], po[2];
      {
        Number vl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) vl[i] = s_u[d][base + i * sd];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          Number a = Number(0);
#pragma unroll
          for (int i = 0; i < N; ++i) a += Number(T.L[s * N + i]) * vl[i];
          vo[s] = a;
        }
      }
      {
        Number pl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) pl[i] = s_u[DIM][base + i * sd];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          Number a = Number(0);
#pragma unroll
          for (int i = 0; i < N; ++i) a += Number(T.L[s * N + i]) * pl[i];
          po[s] = a;
        }
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const Number *c = s_fc[2 * d + s];
        fv[s] = c[0] * vo[s] + c[1] * po[s] + c[2] * tv[s] + c[3] * tp[s];
        fp[s] = c[4] * vo[s] + c[5] * po[s] + c[6] * tv[s] + c[7] * tp[s];
      }
    }
    asm volatile("" ::: "memory");
    {
      Number vl[N];
#pragma unroll
      for (int i = 0; i < N; ++i) vl[i] = s_u[d][base + i * sd];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Number dv = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) dv += Number(T.DT[i * N + m]) * vl[m];
        Number k = cvp * dv;
        k += Number(T.LW[i]) * fp[0];
        k += Number(T.LW[N + i]) * fp[1];
        s_kp[d][base + i * sd] = k;
      }
    }
    asm volatile("" ::: "memory");
    {
      Number pl[N];
#pragma unroll
      for (int i = 0; i < N; ++i) pl[i] = s_u[DIM][base + i * sd];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Number gp = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) gp += Number(T.D[i * N + m]) * pl[m];
        Number k = cvv * gp;
        k += Number(T.LW[i]) * fv[0];
        k += Number(T.LW[N + i]) * fv[1];
        s_u[d][base + i * sd] = k;
      }
    }
  }
  __syncthreads();
  // u_n+1 at the lane's node (b = o1 = u, a1 = 1)
  const bool node_ok = swept && lt < ND;
  double un[NF];
  if (node_ok) {
    Number k[NF];
    Number kp = s_kp[0][lt];
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd) {
      k[dd] = s_u[dd][lt];
      if (dd > 0) kp += s_kp[dd][lt];
    }
    k[DIM] = kp;
    const double ea1 = 1.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      un[f] = bv[f] + ea1 * double(k[f]);
      __builtin_nontemporal_store(un[f], u + cbase + f * ND + lt);
    }
  }
  __syncthreads();  // every lane is done with the sweep's LDS
  // ---- the CK of step n+1 on the same cell (ader_sweep_ck_kernel's) ----
  for (int i = tid; i < CS::TAB_SIZE && i < ckt.size; i += blockDim.x)
    lds.ck.tab[i] = Number(cktab[i]);
  // the CK's lanes: node q of cell slot lck, the sweep's node lane of the same
  // cell (ND = LANES at 3D k=3; one cell per group at k=4, 5)
  const int lck = tid / ND;
  const int q = tid - lck * ND;
  const bool lane_ok = lck < CPB;
  const int lcc = lane_ok ? lck : 0;
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
