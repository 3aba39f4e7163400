// dg_operator_cart.hpp -- fused operator + RK epilogue for Cartesian (box)
// cells: the hot path of the 3D k=4 benchmark.
//
// On a box cell J = diag(h) and the normal of face 2d+s is +-e_d, so the
// operator splits by direction (evaluate_cell and the face fluxes,
// source/wave_equation_operations.cc:424-551, 569-629):
//   dv_d/dt = -rho^-1 d_d p / h_d                     + lifting of faces +-d
//   dp/dt   = sum_d rho c^2 (weak d_d v_d) / h_d      + lifting of faces +-d
// and every term of direction d only needs v_d and p on lines parallel to e_d.
//
// One workgroup per cell batch, three phases:
//   A (lane = node): coalesced loads of the cell's nodes, of the epilogue
//     operand and of the six neighbours' (v_n, p) nodes, staged in LDS;
//   B (lane = line): each of the dim*N^(dim-1) lines of the cell is read
//     from LDS into registers once; the collocation derivative, the weak
//     derivative, both face traces, both fluxes and the face lifting are
//     register FMAs with wave-uniform 1D matrices taken from the kernel
//     arguments; the line's dv_d/dt and its share of dp/dt go back to LDS over
//     the neighbour staging it has just consumed (only this lane reads it);
//   C (lane = node): sum the pressure shares, fused RK epilogue (FP64,
//     RKVectorUpdater semantics, source/time_integrators.cc:128-155) with
//     streaming stores.
// Global traffic is coalesced exactly like a streaming kernel; LDS traffic is
// one read + one write per value and line instead of N reads per derivative.
#pragma once
#include "dg_common.hpp"
#include "gauss_cx.hpp"

namespace exw {

#ifndef EXW_CART_GROUP
#define EXW_CART_GROUP 128
#endif


template <int DIM, int K>
struct CartShape {
  static constexpr int N = K + 1;
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;
  static constexpr int NT = DIM == 2 ? N : N * N;  // lines per direction
  static constexpr int NF = DIM + 1;
  static constexpr int NLINE = DIM * NT;           // lines per cell
  // node lanes per cell: cells of >= 48 nodes occupy whole waves (their
  // neighbour rows stay wave-uniform scalar loads); smaller cells are packed
  static constexpr bool WAVE_ALIGNED = ND >= 48;
  static constexpr int LANES = WAVE_ALIGNED ? ((ND + 63) / 64) * 64 : (ND > NLINE ? ND : NLINE);
  // cells per workgroup: about EXW_CART_GROUP lanes per group
  static constexpr int CPB = LANES >= EXW_CART_GROUP ? 1 : EXW_CART_GROUP / LANES;
  static constexpr int BLOCK = ((CPB * LANES + 63) / 64) * 64;
  // cells per group one after the other (their loads issued together; one
  // was measured fastest: fewer cells in flight per CU with more)
  static constexpr int ITER = 1;
  static_assert(CPB * NLINE <= BLOCK, "the packed line phase must fit the group");
};


// The LDS of one group of the Cartesian kernel.
template <int DIM, int K, typename Number>
struct CartLds {
  using S = CartShape<DIM, K>;
  Number u[S::CPB][S::NF][S::ND];
  // neighbour (v_n, p) nodes of face f; phase B overwrites [2d][0] with the
  // lines' dv_d/dt and [2d][1] with their dp/dt share
  Number nb[S::CPB][2 * DIM][2][S::ND];
  // per face: the numerical flux as a linear map of the traces (own v_n, own
  // p, neighbour v_n, neighbour p) -> (v lifting, p lifting), with 1/h and the
  // normal's sign folded in
  Number fc[S::CPB][2 * DIM][8];
  // whether each cell slot is active (several cells per group only: with one,
  // every lane knows)
  int act[S::CPB > 1 ? S::CPB : 0];
};

// The operator part of dg_operator_cart_kernel for the group's cells: phases
// A and B, then k = M^-1 A(src) at the lane's node (phase C's first half).
// Returns the lane's cell slot lc, cell, whether it is active (a node lane of
// an active cell), the cell scale, the epilogue operand b at the node (0
// without b) and k.  Three barriers' worth of LDS use; the caller may reuse
// the LDS after a barrier.
template <int DIM, int K, typename Number>
__device__ __forceinline__ void cart_cell(const DevMesh &mesh, const double *__restrict__ src,
                                          const Epilogue &epi, int group,
                                          CartLds<DIM, K, Number> &L, int &lc_out, int &cell_out,
                                          bool &active_out, double &cscale_out,
                                          double (&bv)[CartShape<DIM, K>::NF],
                                          Number (&k)[CartShape<DIM, K>::NF]) {
  using S = CartShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NF = S::NF, LANES = S::LANES, CPB = S::CPB;
  constexpr int NFACE = 2 * DIM;
  // the 1D matrices are compile-time constants (folded into the unrolled loops)
  constexpr const CartConst<N> &T = CartTable<N>::value;
  auto &s_u = L.u;
  auto &s_nb = L.nb;
  auto &s_fc = L.fc;
  auto &s_act = L.act;

  const int tid = threadIdx.x;
  int lc = tid / LANES;
  // a wave-aligned cell slot is uniform over the wave: say so, so that the
  // cell's index, neighbour row and material are scalar loads
  if constexpr (S::WAVE_ALIGNED) lc = __builtin_amdgcn_readfirstlane(lc);
  const int lane = tid - lc * LANES;
  const bool lane_ok = lc < CPB;
  const int lane_c = lane < ND ? lane : ND - 1;

  // ---------------- A: the loads of the group's cells ----------------
  // Every load is unconditional and in bounds (no branch between the loads,
  // so all of them are in flight together): lanes past the last node read the
  // last node, inactive slots read cell 0, and boundary faces read the cell
  // itself -- their neighbour inputs have zero flux coefficients.
  int nbr[NFACE];
  double uo[NF], nv[NFACE][2];
  const int e = group * CPB + lc;
  const int cell = epi.cell_list ? (e < epi.n_list ? epi.cell_list[e] : -1) : e;
  bool active = lane_ok && cell >= 0 && cell < mesh.n_cells;
  double cscale = 1.0;
  if (active && epi.cell_scale) {
    cscale = epi.cell_scale[cell];
    active = cscale != 0.0;
  }
#pragma unroll
  for (int f = 0; f < NFACE; ++f) nbr[f] = active ? mesh.neighbor[(size_t)cell * NFACE + f] : -1;
  const int cell_c = active ? cell : 0;
  const size_t cbase_c = (size_t)cell_c * NF * ND;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    uo[f] = src[cbase_c + f * ND + lane_c];
    bv[f] = 0.0;
  }
  if (epi.b) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      bv[f] = __builtin_nontemporal_load(epi.b + cbase_c + f * ND + lane_c);
  }
#pragma unroll
  for (int f = 0; f < NFACE; ++f) {
    const double *un = src + (size_t)(nbr[f] >= 0 ? nbr[f] : cell_c) * NF * ND;
    nv[f][0] = un[(f >> 1) * ND + lane_c];
    nv[f][1] = un[DIM * ND + lane_c];
  }

  if (CPB > 1 && lane_ok && lane == 0) s_act[lc] = active ? 1 : 0;
  if (active && lane < NFACE) {
    // face lane: columns of the linear flux map (evaluate_inner_face /
    // evaluate_boundary_face, source/wave_equation_operations.cc:480-629)
    const int f = lane, d = f >> 1;
    int nb = nbr[0];
#pragma unroll
    for (int g = 1; g < NFACE; ++g)
      if (lane == g) nb = nbr[g];
    const double rho = mesh.rho[cell], cc = mesh.speed[cell];
    const double rho_m = nb >= 0 ? mesh.rho[nb] : rho, c_m = nb >= 0 ? mesh.speed[nb] : cc;
    const double ih = recip(mesh.h[(size_t)cell * DIM + d]);
    const double sg = (f & 1) ? 1.0 : -1.0;
    // a hanging face contributes nothing here (nc_face_kernel adds it)
    const double ihf = nb == -NC_BID ? 0.0 : ih;
    double a[4], b[4];
    flux_matrix<double>(rho, cc, rho_m, c_m, nb >= 0 ? 0 : -nb, a, b);
    // inputs (v_d, p, neighbour v_d, neighbour p): v.n = sg v_d on both sides
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double in = (j == 0 || j == 2) ? sg : 1.0;
      s_fc[lc][f][j] = Number(a[j] * in * sg * ihf);
      s_fc[lc][f][4 + j] = Number(b[j] * in * ihf);
    }
  }
  if (lane_ok) {  // lanes past the last node store the last node's (same) values
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[lc][f][lane_c] = Number(uo[f]);
#pragma unroll
    for (int f = 0; f < NFACE; ++f) {
      s_nb[lc][f][0][lane_c] = Number(nv[f][0]);
      s_nb[lc][f][1][lane_c] = Number(nv[f][1]);
    }
  }
  __syncthreads();

  // ---------------- B: one lane per line, the group's lines packed ----------------
  const int lcB = tid / S::NLINE, line = tid - lcB * S::NLINE;
  if (lcB < CPB && (CPB > 1 ? s_act[lcB] != 0 : active)) {
    const int d = line / NT, t = line - d * NT;
    int base, sd;
    if constexpr (DIM == 2) {
      base = d == 0 ? t * N : t;
      sd = d == 0 ? 1 : N;
    } else {
      const int t0 = t % N, t1 = t / N;
      base = d == 0 ? t0 * N + t1 * N * N : (d == 1 ? t0 + t1 * N * N : t0 + t1 * N);
      sd = d == 0 ? 1 : (d == 1 ? N : N * N);
    }
    // the cell term's factors of direction d, -1/(rho h_d) and rho c^2/h_d
    // (the line's own; as few bytes of LDS as possible keep 10 groups per CU)
    Number cvv, cvp;
    {
      const int eB = group * CPB + lcB;
      const int cB = CPB > 1 ? (epi.cell_list ? epi.cell_list[eB] : eB) : cell;
      const double rho = mesh.rho[cB], cc = mesh.speed[cB];
      const double ih = recip(mesh.h[(size_t)cB * DIM + d]);
      cvv = Number(-ih * recip(rho));
      cvp = Number(rho * cc * cc * ih);
    }
    Number vl[N], pl[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int n = base + i * sd;
      vl[i] = s_u[lcB][d][n];
      pl[i] = s_u[lcB][DIM][n];
    }
    Number kv[N], kp[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      Number gp = Number(0), dv = Number(0);
#pragma unroll
      for (int m = 0; m < N; ++m) {
        gp += Number(T.D[i * N + m]) * pl[m];
        dv += Number(T.DT[i * N + m]) * vl[m];
      }
      kv[i] = cvv * gp;
      kp[i] = cvp * dv;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int f = 2 * d + s;
      Number vo = Number(0), po = Number(0), vn = Number(0), pn = Number(0);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int n = base + i * sd;
        const Number ls = Number(T.L[s * N + i]), lo = Number(T.L[(1 - s) * N + i]);
        vo += ls * vl[i];
        po += ls * pl[i];
        vn += lo * s_nb[lcB][f][0][n];
        pn += lo * s_nb[lcB][f][1][n];
      }
      const Number *c = s_fc[lcB][f];
      const Number fv = c[0] * vo + c[1] * po + c[2] * vn + c[3] * pn;
      const Number fp = c[4] * vo + c[5] * po + c[6] * vn + c[7] * pn;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const Number lw = Number(T.LW[s * N + i]);
        kv[i] += lw * fv;
        kp[i] += lw * fp;
      }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int n = base + i * sd;
      s_nb[lcB][2 * d][0][n] = kv[i];
      s_nb[lcB][2 * d][1][n] = kp[i];
    }
  }
  __syncthreads();

  // ---------------- C (first half): k at the lane's node ----------------
  active = active && lane < ND;
  if (active) {
    Number kp = s_nb[lc][0][1][lane];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      k[d] = s_nb[lc][2 * d][0][lane];
      if (d > 0) kp += s_nb[lc][2 * d][1][lane];
    }
    k[DIM] = kp;
  }
  lc_out = lc;
  cell_out = cell;
  active_out = active;
  cscale_out = cscale;
}

template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((CartShape<DIM, K>::BLOCK))
    dg_operator_cart_kernel(const DevMesh mesh, const double *__restrict__ src,
                            const Epilogue epi) {
  using S = CartShape<DIM, K>;
  constexpr int ND = S::ND, NF = S::NF, LANES = S::LANES;
  static_assert(S::ITER == 1, "one cell per lane slot and group");
  __shared__ CartLds<DIM, K, Number> lds;
#ifdef EXW_LDS_PAD
  __shared__ double s_pad[EXW_LDS_PAD];
  if (threadIdx.x == 1023) s_pad[0] = 0.0;
#endif
  int lc, cell;
  bool active;
  double cscale, bv[NF];
  Number k[NF];
  cart_cell<DIM, K, Number>(mesh, src, epi, xcd_remap(blockIdx.x, gridDim.x), lds, lc, cell,
                            active, cscale, bv, k);

  // ---------------- C: fused epilogue ----------------
  if (active) {
    const int lane = threadIdx.x - lc * LANES;
    const double ea1 = epi.a1 * cscale, ea2 = epi.a2 * cscale, ea3 = epi.a3 * cscale;
    const size_t cbase = (size_t)cell * NF * ND;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const size_t idx = cbase + f * ND + lane;
      const double kk = double(k[f]);
      const double o1 = bv[f] + ea1 * kk;
      __builtin_nontemporal_store(o1, epi.o1 + idx);
      if (epi.o2) {
        const double cv = epi.o2_from_o1 ? o1 : epi.c[idx];
        __builtin_nontemporal_store(cv + ea2 * kk, epi.o2 + idx);
        if (epi.o3) __builtin_nontemporal_store(cv + ea3 * kk, epi.o3 + idx);
      }
    }
  }
}

}  // namespace exw
