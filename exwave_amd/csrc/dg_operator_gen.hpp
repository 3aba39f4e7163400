// dg_operator_gen.hpp -- fused operator + RK epilogue for general multilinear
// cells of >= 48 nodes (3D k >= 3, 2D k >= 6): one cell per workgroup, one
// lane per node, every global load issued before the first barrier.
//
// The arithmetic is that of evaluate_cell / evaluate_inner_face /
// evaluate_boundary_face (source/wave_equation_operations.cc:424-629) on a
// cell whose Jacobian varies from point to point:
//   dv/dt = -rho^-1 J^-T grad_xi p            + lifting of the face terms
//   dp/dt = weak div of rho c^2 w det(J) J^-1 v + lifting of the face terms
// followed by the inverse mass 1/(w det J) (collocated Gauss basis) and the
// RK epilogue.  How the data moves:
//   A (lane = node): the cell's nodes and the node q of each of the 2 dim
//     neighbours (all fields) are loaded at once; the
//     node's Jacobian is computed once from the cell's corners (scalar loads)
//     and kept in registers; each lane also computes the unit normal and the
//     area element at one or two face points (LDS);
//   A' (lane = node): the neighbours' (v.n, p) at node q, with n the normal
//     at the face point of q's line, are staged in LDS -- the trace of v.n is
//     then a line sum of staged values, as on Cartesian cells; the gradient
//     of p and the test-gradient values rho c^2 w det J (J^-1 v) are formed;
//   F (lane = face point, and lane = node): the face fluxes from LDS traces,
//     and the weak divergence at the node;
//   C (lane = node): lifting, inverse mass, epilogue with streaming stores.
// Four barriers, no atomics; after the loads only the epilogue
// operands are read from global memory (issued after the face phase).
#pragma once
#include <type_traits>

#include "dg_common.hpp"

namespace exw {

template <int DIM, int K>
struct GenShape {
  static constexpr int N = K + 1;
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;
  static constexpr int NT = DIM == 2 ? N : N * N;
  static constexpr int NF = DIM + 1;
  static constexpr int NFACE = 2 * DIM;
  static constexpr int NFP = NFACE * NT;          // face points of a cell
  static constexpr int BLOCK = (ND + 63) / 64 * 64;
  static constexpr int FP_ITER = (NFP + BLOCK - 1) / BLOCK;  // face points per lane
  // cells that fill whole waves and whose LDS fits gfx950's 160 KB
  static constexpr int LDS_DOUBLES = (NF + 2 * NFACE + DIM) * ND + NFACE * NT * (DIM + 1) +
                                     6 * N + 2 * N * N + (1 << DIM) * DIM + 3 * NFACE;
  static constexpr bool ENABLED = ND >= 48 && LDS_DOUBLES * 8 <= 160 * 1024;
};

template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((GenShape<DIM, K>::BLOCK))
    dg_operator_gen_kernel(const DevMesh mesh, const double *__restrict__ src,
                           const double *__restrict__ tab, const Epilogue epi) {
  using S = GenShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NF = S::NF, NFACE = S::NFACE;
  constexpr int NFP = S::NFP, BLOCK = S::BLOCK, NV = 1 << DIM;
  using TO = TabOff<N>;

  __shared__ Number s_u[NF][ND];
  // neighbour f's (v.n, p) at node-aligned positions (n: this cell's outward
  // normal at the face point of the node's line)
  __shared__ Number s_nb[NFACE][2][ND];
  __shared__ Number s_g[DIM][ND];
  // per face point: unit outward normal and area element (A), then the face
  // flux contributions (F writes them over its own entries)
  __shared__ Number s_fp[NFACE][NT][DIM + 1];
  __shared__ Number s_tab[TO::SIZE];
  // the cell's corners and the neighbours' materials (wave-uniform values,
  // kept in LDS rather than in registers of every lane)
  __shared__ double s_V[NV * DIM];
  __shared__ Number s_mat[NFACE][2];
  __shared__ int s_nbr[NFACE];

  const int tid = threadIdx.x;
  const int q = tid < ND ? tid : ND - 1;  // lanes past the last node mirror it
  const int cell = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x));
  double cscale = 1.0;
  if (epi.cell_scale) cscale = epi.cell_scale[cell];
  const bool active = cscale != 0.0;  // uniform over the group
  if (!active) return;
  const size_t cbase = (size_t)cell * NF * ND;

  // ---------------- A: every global load ----------------
  int nbr[NFACE];
#pragma unroll
  for (int f = 0; f < NFACE; ++f) nbr[f] = mesh.neighbor[(size_t)cell * NFACE + f];
  double uo[NF], nu[NFACE][NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) uo[f] = src[cbase + f * ND + q];
#pragma unroll
  for (int f = 0; f < NFACE; ++f) {
    // boundary and hanging faces read the cell itself (unused values)
    const double *un = src + (size_t)(nbr[f] >= 0 ? nbr[f] : cell) * NF * ND;
#pragma unroll
    for (int k = 0; k < NF; ++k) nu[f][k] = un[k * ND + q];
  }
  const double rho_d = mesh.rho[cell], c_d = mesh.speed[cell];
  if (tid < NV * DIM) s_V[tid] = mesh.vertices[(size_t)cell * NV * DIM + tid];
  if (tid < NFACE) {
    int nb = nbr[0];
#pragma unroll
    for (int f = 1; f < NFACE; ++f)
      if (tid == f) nb = nbr[f];
    s_nbr[tid] = nb;
    s_mat[tid][0] = Number(nb >= 0 ? mesh.rho[nb] : rho_d);
    s_mat[tid][1] = Number(nb >= 0 ? mesh.speed[nb] : c_d);
  }
  for (int i = tid; i < TO::SIZE; i += BLOCK) s_tab[i] = Number(tab[i]);
  __syncthreads();  // the table and the corners
  const Number rho = Number(rho_d), cc = Number(c_d);
  const Number rho_inv = Number(1) / rho, rc2 = rho * cc * cc;
  const Number *X = s_tab + TO::X, *W = s_tab + TO::W;

  const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
  // the node's geometry, once (registers)
  Number Ji[DIM * DIM], jxw;
  {
    Number xi[DIM], J[DIM * DIM], wq = Number(1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = X[qi[d]];
      wq *= W[qi[d]];
    }
    jxw = wq * multilinear_jacobian<DIM>(s_V, xi, J, Ji);
  }
  if (tid < ND) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[f][q] = Number(uo[f]);
    // test-gradient values rho c^2 w det J (J^-1 v)_e
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      Number s = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) s += Ji[e * DIM + i] * Number(uo[i]);
      s_g[e][q] = rc2 * jxw * s;
    }
  }
  // face geometry: unit outward normal J^-T e_d / |.| and area element
  // det J |J^-T e_d| (the tangential weights are applied at the lifting)
#pragma unroll 1
  for (int it = 0; it < S::FP_ITER; ++it) {
    const int fp = tid + it * BLOCK;
    if (fp < NFP) {
      const int f = fp / NT, t = fp - f * NT, d = f >> 1, s = f & 1;
      const int td[2] = {t % N, DIM == 3 ? t / N : 0};
      Number xi[DIM], J[DIM * DIM], Jf[DIM * DIM];
      int j = 0;
#pragma unroll
      for (int e = 0; e < DIM; ++e) xi[e] = e == d ? Number(s) : X[td[j++]];
      const Number det = multilinear_jacobian<DIM>(s_V, xi, J, Jf);
      Number na = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) na += Jf[d * DIM + i] * Jf[d * DIM + i];
      na = sqrt(na);
      const Number sg = s ? Number(1) : Number(-1);
#pragma unroll
      for (int i = 0; i < DIM; ++i) s_fp[f][t][i] = sg * Jf[d * DIM + i] / na;
      // a hanging face contributes nothing here (nc_face_kernel adds it)
      s_fp[f][t][DIM] = s_nbr[f] == -NC_BID ? Number(0) : fabs(det) * na;
    }
  }
  __syncthreads();

  // ---------------- A': neighbour staging, cell term ----------------
  const Number *D = s_tab + TO::D, *L = s_tab + TO::L;
  if (tid < ND) {
#pragma unroll
    for (int f = 0; f < NFACE; ++f) {
      const int d = f >> 1;
      int t;
      if constexpr (DIM == 2)
        t = qi[1 - d];
      else
        t = d == 0 ? qi[1] + N * qi[2] : (d == 1 ? qi[0] + N * qi[2] : qi[0] + N * qi[1]);
      Number vn = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) vn += s_fp[f][t][i] * Number(nu[f][i]);
      s_nb[f][0][q] = vn;
      s_nb[f][1][q] = Number(nu[f][DIM]);
    }
  }
  Number kv[DIM];
  {
    Number gref[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      const int se = pow_n<N>(e), q0 = q - qi[e] * se;
      Number g = Number(0);
#pragma unroll
      for (int m = 0; m < N; ++m) g += D[qi[e] * N + m] * s_u[DIM][q0 + m * se];
      gref[e] = g;
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      Number gp = Number(0);
#pragma unroll
      for (int e = 0; e < DIM; ++e) gp += Ji[e * DIM + i] * gref[e];
      kv[i] = -rho_inv * gp;
    }
  }
  __syncthreads();

  // ---------------- F: face fluxes (lane = face point), weak divergence ----
#pragma unroll 1
  for (int it = 0; it < S::FP_ITER; ++it) {
    const int fp = tid + it * BLOCK;
    if (fp < NFP) {
      const int f = fp / NT, t = fp - f * NT, d = f >> 1, s = f & 1;
      int base;
      if constexpr (DIM == 2) {
        base = d == 0 ? t * N : t;
      } else {
        const int t0 = t % N, t1 = t / N;
        base = d == 0 ? t0 * N + t1 * N * N : (d == 1 ? t0 + t1 * N * N : t0 + t1 * N);
      }
      const int sd = d == 0 ? 1 : (d == 1 ? N : N * N);
      const Number *Ls = L + s * N, *Lo = L + (1 - s) * N;
      Number up[NF], nvm = Number(0), pm = Number(0);
      // one pass along the line for all six sums (the order of each sum is
      // unchanged), rolled: an unrolled pass keeps all its LDS loads in flight
      Number acc[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) acc[k] = Number(0);
#pragma unroll 1
      for (int i = 0; i < N; ++i) {
        const Number ls = Ls[i], lo = Lo[i];
        const int o = base + i * sd;
#pragma unroll
        for (int k = 0; k < NF; ++k) acc[k] += ls * s_u[k][o];
        nvm += lo * s_nb[f][0][o];
        pm += lo * s_nb[f][1][o];
      }
#pragma unroll
      for (int k = 0; k < NF; ++k) up[k] = acc[k];
      Number nrm[DIM], nvp = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        nrm[i] = s_fp[f][t][i];
        nvp += up[i] * nrm[i];
      }
      const Number dS = s_fp[f][t][DIM];
      const int nb = s_nbr[f];
      Number val_pn, val_p;
      flux<Number>(nvp, up[DIM], nvm, pm, rho, cc, s_mat[f][0], s_mat[f][1], nb >= 0 ? 0 : -nb,
                   val_pn, val_p);
      // over this lane's own geometry entries (read above)
#pragma unroll
      for (int i = 0; i < DIM; ++i) s_fp[f][t][i] = val_pn * nrm[i] * dS;
      s_fp[f][t][DIM] = val_p * dS;
    }
  }
  // the epilogue operands, in flight during the weak divergence and the last
  // barrier (issued after the face phase, whose traces set the register peak:
  // 166 instead of 222 VGPRs at 3D k=4, 3 waves per SIMD instead of 2)
  double bv[NF], cv[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    bv[f] = epi.b ? __builtin_nontemporal_load(epi.b + cbase + f * ND + q) : 0.0;
    cv[f] = (epi.o2 && !epi.o2_from_o1) ? __builtin_nontemporal_load(epi.c + cbase + f * ND + q)
                                        : 0.0;
  }
  Number kp = Number(0);
#pragma unroll
  for (int e = 0; e < DIM; ++e) {
    const int se = pow_n<N>(e), q0 = q - qi[e] * se;
    Number a = Number(0);
#pragma unroll
    for (int m = 0; m < N; ++m) a += D[m * N + qi[e]] * s_g[e][q0 + m * se];
    kp += a;
  }
  __syncthreads();

  // ---------------- C: lifting, inverse mass, epilogue ----------------
  if (tid < ND) {
    Number fv[DIM], fpv = Number(0);
#pragma unroll
    for (int i = 0; i < DIM; ++i) fv[i] = Number(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      int t;
      Number wt = Number(1);
      if constexpr (DIM == 2) {
        t = qi[1 - d];
        wt = W[qi[1 - d]];
      } else {
        t = d == 0 ? qi[1] + N * qi[2] : (d == 1 ? qi[0] + N * qi[2] : qi[0] + N * qi[1]);
#pragma unroll
        for (int e = 0; e < DIM; ++e)
          if (e != d) wt *= W[qi[e]];
      }
      const Number l0 = L[qi[d]] * wt, l1 = L[N + qi[d]] * wt;
#pragma unroll
      for (int i = 0; i < DIM; ++i) fv[i] += l0 * s_fp[2 * d][t][i] + l1 * s_fp[2 * d + 1][t][i];
      fpv += l0 * s_fp[2 * d][t][DIM] + l1 * s_fp[2 * d + 1][t][DIM];
    }
    const Number ij = Number(1) / jxw;
#pragma unroll
    for (int i = 0; i < DIM; ++i) kv[i] += fv[i] * ij;
    kp = (kp + fpv) * ij;
    const double ea1 = epi.a1 * cscale, ea2 = epi.a2 * cscale, ea3 = epi.a3 * cscale;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const size_t idx = cbase + f * ND + q;
      const double kk = double(f < DIM ? kv[f < DIM ? f : 0] : kp);
      const double o1 = bv[f] + ea1 * kk;
      __builtin_nontemporal_store(o1, epi.o1 + idx);
      if (epi.o2) {
        const double c = epi.o2_from_o1 ? o1 : cv[f];
        __builtin_nontemporal_store(c + ea2 * kk, epi.o2 + idx);
        if (epi.o3) __builtin_nontemporal_store(c + ea3 * kk, epi.o3 + idx);
      }
    }
  }
}

}  // namespace exw
