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
//   A (lane = node): the cell's nodes; a few lanes turn the cell's corners
//     into the coefficients of its Jacobian columns, each a bilinear (3D) or
//     linear (2D) polynomial in the other reference coordinates;
//   G (lane = node, then lane = face point): the node's Jacobian from those
//     coefficients (27 FMAs in 3D instead of a sum over the 8 corners) and its
//     inverse, the test-gradient values rho c^2 w det J (J^-1 v) (LDS), the
//     pressure gradient and dv/dt; per face point the neighbour's nodes on
//     the face point's line are loaded (L2 hits within the tile), the unit
//     normal and area element are formed from the cofactors of the Jacobian
//     there while they arrive (Nanson: det J J^-T e_d = j_{d+1} x j_{d+2}, no
//     inverse), then the neighbour's traces and its v.n with this side's
//     normal (LDS: 2 values per face point instead of 2 dim node-aligned
//     fields);
//   F (lane = face point, and lane = node): the face fluxes, and the weak
//     divergence at the node;
//   C (lane = node): lifting, inverse mass, epilogue with streaming stores.
// Three barriers, no atomics.
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
  static constexpr int LDS_DOUBLES = (NF + DIM) * ND + NFACE * NT * (DIM + 3) + 6 * N +
                                     2 * N * N + DIM * DIM * (1 << (DIM - 1)) + 3 * NFACE;
  static constexpr bool ENABLED = ND >= 48 && LDS_DOUBLES * 8 <= 160 * 1024;
};

// Column e of the Jacobian of a multilinear cell as a polynomial in the other
// reference coordinates: with D_j = V(corner j, xi_e = 1) - V(corner j,
// xi_e = 0) over the 2^(dim-1) corners j of the other coordinates (bit 0 of j:
// the lower other coordinate a, bit 1: the upper one b),
//   dx/dxi_e = D_0 + xi_a (D_1 - D_0) [+ xi_b (D_2 - D_0) + xi_a xi_b (D_3 - D_2 - D_1 + D_0)]
// jc[e][m][i]: coefficient m of component i.  Lane (e, i) of the first DIM^2.
template <int DIM, typename Number>
__device__ __forceinline__ void jacobian_coefficients(const double *V, int e, int i,
                                                      Number (*jc)[1 << (DIM - 1)][DIM]) {
  constexpr int NJ = 1 << (DIM - 1);
  const int a = e == 0 ? 1 : 0, b = e == 2 ? 1 : 2;  // the other coordinates (3D: a < b)
  double D[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int v0 = ((j & 1) << a) | (DIM == 3 ? ((j >> 1) << b) : 0);
    D[j] = V[(v0 | (1 << e)) * DIM + i] - V[v0 * DIM + i];
  }
  jc[e][0][i] = Number(D[0]);
  jc[e][1][i] = Number(D[1] - D[0]);
  if constexpr (DIM == 3) {
    jc[e][2][i] = Number(D[2] - D[0]);
    jc[e][3][i] = Number((D[3] - D[2]) - (D[1] - D[0]));
  }
}

// J[i*DIM+e] = dx_i/dxi_e at xi from the column coefficients
template <int DIM, typename Number>
__device__ __forceinline__ void jacobian_eval(const Number (*jc)[1 << (DIM - 1)][DIM],
                                              const Number *xi, Number *J) {
#pragma unroll
  for (int e = 0; e < DIM; ++e) {
    // one column's coefficients at a time: an empty asm that clobbers memory
    // keeps the compiler from loading all 2 DIM^2 NJC registers of them at
    // once and holding them (3D k=4: 166 instead of 190 VGPRs, 3 waves per
    // SIMD instead of 2)
    asm volatile("" ::: "memory");
    const int a = e == 0 ? 1 : 0, b = e == 2 ? 1 : 2;
    const Number xa = xi[a];
    if constexpr (DIM == 3) {
      const Number xb = xi[b], xab = xa * xb;
#pragma unroll
      for (int i = 0; i < DIM; ++i)
        J[i * DIM + e] = jc[e][0][i] + xa * jc[e][1][i] + xb * jc[e][2][i] + xab * jc[e][3][i];
    } else {
#pragma unroll
      for (int i = 0; i < DIM; ++i) J[i * DIM + e] = jc[e][0][i] + xa * jc[e][1][i];
    }
  }
}

// Jinv[e*DIM+i] = dxi_e/dx_i of J; returns det J (as multilinear_jacobian)
template <int DIM, typename T>
__device__ __forceinline__ T jacobian_invert(const T *J, T *Jinv) {
  T det;
  if constexpr (DIM == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    const T id = T(1) / det;
    Jinv[0] = J[3] * id;
    Jinv[1] = -J[1] * id;
    Jinv[2] = -J[2] * id;
    Jinv[3] = J[0] * id;
  } else {
    const T a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7],
            k = J[8];
    const T A = e * k - f * h, B = f * g - d * k, C = d * h - e * g;
    det = a * A + b * B + c * C;
    const T id = T(1) / det;
    Jinv[0] = A * id;
    Jinv[3] = B * id;
    Jinv[6] = C * id;
    Jinv[1] = (c * h - b * k) * id;
    Jinv[4] = (a * k - c * g) * id;
    Jinv[7] = (b * g - a * h) * id;
    Jinv[2] = (b * f - c * e) * id;
    Jinv[5] = (c * d - a * f) * id;
    Jinv[8] = (a * e - b * d) * id;
  }
  return det;
}

template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((GenShape<DIM, K>::BLOCK))
    dg_operator_gen_kernel(const DevMesh mesh, const double *__restrict__ src,
                           const double *__restrict__ tab, const Epilogue epi) {
  using S = GenShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NF = S::NF, NFACE = S::NFACE;
  constexpr int NFP = S::NFP, BLOCK = S::BLOCK, NV = 1 << DIM;
  using TO = TabOff<N>;

  __shared__ Number s_u[NF][ND];
  // per face point the neighbour's traces (v.n with this side's normal, p)
  __shared__ Number s_tr[NFACE][NT][2];
  __shared__ Number s_g[DIM][ND];
  // per face point: unit outward normal and area element (A), then the face
  // flux contributions (F writes them over its own entries)
  __shared__ Number s_fp[NFACE][NT][DIM + 1];
  __shared__ Number s_tab[TO::SIZE];
  // the neighbours' materials and the Jacobian coefficients (wave-uniform
  // values, kept in LDS rather than in registers of every lane)
  __shared__ Number s_mat[NFACE][2];
  __shared__ int s_nbr[NFACE];
  __shared__ Number s_jc[DIM][1 << (DIM - 1)][DIM];

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
  double uo[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) uo[f] = src[cbase + f * ND + q];
  const double rho_d = mesh.rho[cell], c_d = mesh.speed[cell];
  if (tid < DIM * DIM)
    jacobian_coefficients<DIM, Number>(mesh.vertices + (size_t)cell * NV * DIM, tid / DIM,
                                       tid % DIM, s_jc);
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
  if (tid < ND) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[f][q] = Number(uo[f]);
  }
  __syncthreads();  // the table, the Jacobian coefficients, the cell's fields
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
    jacobian_eval<DIM>(s_jc, xi, J);
    jxw = wq * jacobian_invert<DIM>(J, Ji);
  }
  if (tid < ND) {
    // test-gradient values rho c^2 w det J (J^-1 v)_e
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      Number s = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) s += Ji[e * DIM + i] * Number(uo[i]);
      s_g[e][q] = rc2 * jxw * s;
    }
  }
  // the cell term dv/dt = -rho^-1 J^-T grad_xi p at the node (s_u is complete)
  const Number *D = s_tab + TO::D, *L = s_tab + TO::L;
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
  // (empty asm that takes kv: the compiler would otherwise sink its sums
  // into the last phase and hold the node's inverse Jacobian and gradient
  // through the face phases, 24 registers instead of 6)
#pragma unroll
  for (int i = 0; i < DIM; ++i) asm volatile("" : "+v"(kv[i]));

  // lane = face point: the neighbour's nodes along the face point's line are
  // loaded first (32-bit offsets: vectors below 4 GiB), the face geometry is
  // computed while they arrive, then the neighbour's traces of all fields and
  // its v.n with this side's normal.
  // Face geometry: n dA = sg det J J^-T e_d, the cofactor row d of J: in 3D
  // the cross product j_{d+1} x j_{d+2} of the Jacobian columns, in 2D the
  // rotated column j_{1-d}; the unit normal n = sg sign(det J) c / |c| and
  // the area element |c| (the tangential weights are applied at the
  // lifting).  All cofactor rows are formed and row d is selected: the face
  // direction varies across the lanes, and selecting the operands instead
  // would cost more selects than the two extra rows cost products.
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
      const int nb = s_nbr[f];
      // boundary and hanging faces read the cell itself (unused values)
      const uint32_t un = uint32_t(nb >= 0 ? nb : cell) * uint32_t(NF * ND) + uint32_t(base);
      double gn[NF][N];
#pragma unroll
      for (int k = 0; k < NF; ++k)
#pragma unroll
        for (int i = 0; i < N; ++i) gn[k][i] = src[un + uint32_t(k * ND + i * sd)];

      const Number xa = X[t % N], xb = DIM == 3 ? X[t / N] : Number(0), xs = Number(s);
      Number xi[DIM], J[DIM * DIM];
      if constexpr (DIM == 3) {
        xi[0] = d == 0 ? xs : xa;
        xi[1] = d == 0 ? xa : (d == 1 ? xs : xb);
        xi[2] = d == 2 ? xs : xb;
      } else {
        xi[0] = d == 0 ? xs : xa;
        xi[1] = d == 0 ? xa : xs;
      }
      jacobian_eval<DIM>(s_jc, xi, J);
      Number c[DIM], det;
      if constexpr (DIM == 3) {
        Number C[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const int r1 = (r + 1) % 3, r2 = (r + 2) % 3;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            C[r][i] = J[i1 * 3 + r1] * J[i2 * 3 + r2] - J[i2 * 3 + r1] * J[i1 * 3 + r2];
          }
        }
        det = J[0] * C[0][0] + J[3] * C[0][1] + J[6] * C[0][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) c[i] = d == 0 ? C[0][i] : (d == 1 ? C[1][i] : C[2][i]);
      } else {
        // cofactor row d of the 2x2 Jacobian: d = 0: (J11, -J01); d = 1: (-J10, J00)
        c[0] = d == 0 ? J[3] : -J[2];
        c[1] = d == 0 ? -J[1] : J[0];
        det = J[0] * J[3] - J[1] * J[2];
      }
      Number na = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) na += c[i] * c[i];
      na = sqrt(na);
      const Number sg = (s != 0) == (det > Number(0)) ? Number(1) : Number(-1);
      const Number inv = recip(na);
      Number nrm[DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        nrm[i] = sg * c[i] * inv;
        s_fp[f][t][i] = nrm[i];
      }
      // a hanging face contributes nothing here (nc_face_kernel adds it)
      s_fp[f][t][DIM] = nb == -NC_BID ? Number(0) : na;

      const Number *Lo = L + (1 - s) * N;
      Number an[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        Number a = Number(0);
#pragma unroll
        for (int i = 0; i < N; ++i) a += Lo[i] * Number(gn[k][i]);
        an[k] = a;
      }
      Number nvm = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) nvm += nrm[i] * an[i];
      s_tr[f][t][0] = nvm;
      s_tr[f][t][1] = an[DIM];
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
      const Number *Ls = L + s * N;
      // the own traces: one rolled pass along the line for all the sums (an
      // unrolled pass keeps all its LDS loads in flight); the neighbour's
      // from the gather phase
      Number up[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) up[k] = Number(0);
#pragma unroll 1
      for (int i = 0; i < N; ++i) {
        const Number ls = Ls[i];
        const int o = base + i * sd;
#pragma unroll
        for (int k = 0; k < NF; ++k) up[k] += ls * s_u[k][o];
      }
      const Number nvm = s_tr[f][t][0], pm = s_tr[f][t][1];
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
  // barrier (issued after the face phase)
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
  asm volatile("" : "+v"(kp));  // (the same for the divergence's loaded terms)
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
