// dg_operator.hpp -- fused DG operator + RK epilogue kernel for gfx950.
//
// One workgroup processes CPB cells; one lane owns one Gauss node of one cell
// (all dim+1 fields).  Per cell:
//   1. every independent global load is issued up front: the lane's own node,
//      the epilogue operand(s) (RK accumulator), the cell's neighbour row;
//   2. the neighbours' nodes are loaded node-aligned (coalesced exactly like
//      the own node) into LDS while the cell integral runs from LDS
//      (evaluate_cell, source/wave_equation_operations.cc:424-456):
//        dv/dt += -rho^-1 grad p        (strong, collocation derivative),
//        dp/dt += rho c^2 * weak divergence of v  (transposed derivative);
//   3. face integrals of all 2*dim faces of the cell, evaluated from this
//      cell's side (evaluate_inner_face / evaluate_boundary_face,
//      :480-551, :569-629; the flux is symmetric in +/-, so each cell computes
//      its own half and nothing is scattered: no atomics);
//   4. inverse mass = 1/(w_q det J_q) (diagonal in this basis, :633-661);
//   5. the epilogue writes the RK update (RKVectorUpdater,
//      source/time_integrators.cc:128-155) from registers with streaming stores.
// Only two dependent global-memory round trips per cell (own + neighbour row,
// then neighbour nodes).  Workgroups are remapped so each XCD walks a
// contiguous range of cells (neighbours share an L2).
// Number is the operator arithmetic (double, or float for the mixed path);
// loads, the epilogue and all vectors are FP64.
#pragma once
#include <type_traits>

#include "dg_common.hpp"

namespace exw {

// Waves per SIMD the general kernel is held to (1: no constraint).  Capping
// the registers spills to scratch and was measured slower for every variant
// but 3D k=6 with the FP32 operator: 28.9 against 24.6 GDoF/s per stage at 4
// waves (48^3 and 32^3 non-affine cells).
template <int DIM, int K, typename Number>
struct GeneralWaves {
  static constexpr int value = DIM == 3 && K == 6 && std::is_same_v<Number, float> ? 4 : 1;
};

template <int DIM, int K, typename Number, bool CART>
__global__ void __launch_bounds__((Shape<DIM, K>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(GeneralWaves<DIM, K, Number>::value)))
    dg_operator_kernel(const DevMesh mesh, const double *__restrict__ src,
                       const double *__restrict__ tab, const Epilogue epi) {
  using S = Shape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF, NT = S::NT, CPB = S::CPB;
  constexpr int NFACE = 2 * DIM;
  // traces needed per face: Cartesian faces only see v_n and p
  constexpr int NFF = CART ? 2 : NF;
  using TO = TabOff<N>;

  __shared__ Number s_u[CPB][NF][ND];
  // Cartesian: the neighbours' v_n and p at every node (node-aligned loads)
  __shared__ Number s_nb[CART ? CPB : 1][CART ? NFACE : 1][CART ? 2 : 1][CART ? ND : 1];
  __shared__ Number s_g[CART ? 1 : CPB][CART ? 1 : DIM][CART ? 1 : ND];
  __shared__ Number s_F[CPB][NFACE][NFF][NT];
  __shared__ Number s_tab[TO::SIZE];
  __shared__ double s_V[CART ? 1 : CPB][CART ? 1 : (1 << DIM) * DIM];
  __shared__ int s_nbr[CPB][NFACE];
  __shared__ Number s_nmat[CPB][NFACE][2];

  const int tid = threadIdx.x;
  const int lc = tid / ND;
  const int q = tid - lc * ND;
  const int lblock = xcd_remap(blockIdx.x, gridDim.x);
  const int cell = lblock * CPB + lc;
  const bool lane_ok = lc < CPB;
  bool active = lane_ok && cell < mesh.n_cells;
  double cscale = 1.0;
  if (active && epi.cell_scale) {
    cscale = epi.cell_scale[cell];
    active = cscale != 0.0;
  }
  const double ea1 = epi.a1 * cscale, ea2 = epi.a2 * cscale, ea3 = epi.a3 * cscale;
  const size_t cbase = (size_t)cell * NF * ND;

  // ---------------- phase 1: independent loads ----------------
  double uo[NF], bv[NF], cv[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) uo[f] = bv[f] = cv[f] = 0.0;
  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) uo[f] = src[cbase + f * ND + q];
    if (epi.b) {
#pragma unroll
      for (int f = 0; f < NF; ++f) bv[f] = __builtin_nontemporal_load(epi.b + cbase + f * ND + q);
    }
    if (epi.o2 && !epi.o2_from_o1) {
#pragma unroll
      for (int f = 0; f < NF; ++f) cv[f] = __builtin_nontemporal_load(epi.c + cbase + f * ND + q);
    }
    if (q < NFACE) s_nbr[lc][q] = mesh.neighbor[(size_t)cell * NFACE + q];
  }
  for (int i = tid; i < TO::SIZE; i += blockDim.x) s_tab[i] = Number(tab[i]);
  if constexpr (!CART) {
    constexpr int NVD = (1 << DIM) * DIM;
    for (int i = tid; i < CPB * NVD; i += blockDim.x) {
      const int c2 = lblock * CPB + i / NVD;
      s_V[i / NVD][i % NVD] = c2 < mesh.n_cells ? mesh.vertices[(size_t)c2 * NVD + i % NVD] : 0.0;
    }
  }
  Number rho = Number(1), cc = Number(1), hh[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) hh[d] = Number(1);
  if (active) {
    rho = Number(mesh.rho[cell]);
    cc = Number(mesh.speed[cell]);
    if constexpr (CART) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) hh[d] = Number(mesh.h[(size_t)cell * DIM + d]);
    }
  }
  const Number rho_inv = Number(1) / rho;
  const Number rc2 = rho * cc * cc;
  if (lane_ok) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[lc][f][q] = Number(uo[f]);
  }

  // node multi-index (xi_0 fastest)
  int qi[3];
  qi[0] = q % N;
  qi[1] = (q / N) % N;
  qi[2] = DIM == 3 ? q / (N * N) : 0;
  __syncthreads();

  // ---------------- phase 2: neighbour loads, overlapped with the cell term ----
  double nbv[CART ? NFACE : 1][2];
  if constexpr (CART) {
#pragma unroll
    for (int f = 0; f < NFACE; ++f) {
      const int nb = active ? s_nbr[lc][f] : -1;
      nbv[f][0] = nbv[f][1] = 0.0;
      if (nb >= 0) {
        const double *un = src + (size_t)nb * NF * ND;
        nbv[f][0] = un[(f >> 1) * ND + q];
        nbv[f][1] = un[DIM * ND + q];
      }
    }
  }
  if (active && q < NFACE) {
    const int nb = s_nbr[lc][q];
    s_nmat[lc][q][0] = nb >= 0 ? Number(mesh.rho[nb]) : rho;
    s_nmat[lc][q][1] = nb >= 0 ? Number(mesh.speed[nb]) : cc;
  }

  const Number *D = s_tab + TO::D;
  const Number *DT = s_tab + TO::DT;
  const Number *L = s_tab + TO::L;
  const Number *LW = s_tab + TO::LW;
  const Number *X = s_tab + TO::X;
  const Number *W = s_tab + TO::W;

  Number kv[DIM], kp = Number(0);
  Number jxw_node = Number(1);  // w_q det J_q (general geometry)

  if (lane_ok) {
    if constexpr (CART) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int sd = pow_n<N>(d);
        const int q0 = q - qi[d] * sd;
        Number gp = Number(0), dv = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) {
          gp += D[qi[d] * N + m] * s_u[lc][DIM][q0 + m * sd];
          dv += DT[qi[d] * N + m] * s_u[lc][d][q0 + m * sd];
        }
        const Number ih = Number(1) / hh[d];
        kv[d] = -rho_inv * gp * ih;
        kp += dv * ih;
      }
      kp *= rc2;
#pragma unroll
      for (int f = 0; f < NFACE; ++f) {
        s_nb[lc][f][0][q] = Number(nbv[f][0]);
        s_nb[lc][f][1][q] = Number(nbv[f][1]);
      }
    } else {
      Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM];
      Number wq = Number(1);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        xi[d] = X[qi[d]];
        wq *= W[qi[d]];
      }
      const Number det = multilinear_jacobian<DIM>(s_V[lc], xi, J, Ji);
      jxw_node = wq * det;
      Number gref[DIM];
#pragma unroll
      for (int e = 0; e < DIM; ++e) {
        const int se = pow_n<N>(e);
        const int q0 = q - qi[e] * se;
        Number g = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) g += D[qi[e] * N + m] * s_u[lc][DIM][q0 + m * se];
        gref[e] = g;
      }
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        Number gp = Number(0);
#pragma unroll
        for (int e = 0; e < DIM; ++e) gp += Ji[e * DIM + i] * gref[e];
        kv[i] = -rho_inv * gp;
      }
      // test-gradient values g_e = rho c^2 w det J (J^-1 v)_e
#pragma unroll
      for (int e = 0; e < DIM; ++e) {
        Number s = Number(0);
#pragma unroll
        for (int i = 0; i < DIM; ++i) s += Ji[e * DIM + i] * s_u[lc][i][q];
        s_g[lc][e][q] = rc2 * jxw_node * s;
      }
    }
  }
  __syncthreads();

  // ---------------- faces ----------------
  if (lane_ok) {
    for (int fp = q; fp < NFACE * NT; fp += ND) {
      const int f = fp / NT, t = fp - f * NT, d = f >> 1, s = f & 1;
      // tangential digits (other directions ascending) -> base node
      int base, tdig[2] = {0, 0};
      if constexpr (DIM == 2) {
        tdig[0] = t;
        base = (d == 0) ? t * N : t;
      } else {
        tdig[0] = t % N;
        tdig[1] = t / N;
        base = (d == 0) ? tdig[0] * N + tdig[1] * N * N
                        : (d == 1 ? tdig[0] + tdig[1] * N * N : tdig[0] + tdig[1] * N);
      }
      const int sd = d == 0 ? 1 : (d == 1 ? N : N * N);
      const int nb = active ? s_nbr[lc][f] : -2;
      const Number *Ls = L + s * N, *Lo = L + (1 - s) * N;
      const Number rho_m = active ? s_nmat[lc][f][0] : rho;
      const Number c_m = active ? s_nmat[lc][f][1] : cc;
      const int bid = nb >= 0 ? 0 : -nb;
      if constexpr (CART) {
        Number vd = Number(0), pp = Number(0), vdm = Number(0), pm = Number(0);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          vd += Ls[i] * s_u[lc][d][base + i * sd];
          pp += Ls[i] * s_u[lc][DIM][base + i * sd];
          vdm += Lo[i] * s_nb[lc][f][0][base + i * sd];
          pm += Lo[i] * s_nb[lc][f][1][base + i * sd];
        }
        const Number sg = s ? Number(1) : Number(-1);
        const Number fsc = bid == NC_BID ? Number(0) : Number(1);
        Number val_pn, val_p;
        flux<Number>(sg * vd, pp, sg * vdm, pm, rho, cc, rho_m, c_m, bid, val_pn, val_p);
        s_F[lc][f][0][t] = fsc * sg * val_pn;
        s_F[lc][f][1][t] = fsc * val_p;
      } else {
        Number up[NF], um[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          Number a = Number(0);
#pragma unroll
          for (int i = 0; i < N; ++i) a += Ls[i] * s_u[lc][k][base + i * sd];
          up[k] = a;
          um[k] = Number(0);
        }
        if (nb >= 0) {
          const double *un = src + (size_t)nb * NF * ND;
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            Number a = Number(0);
#pragma unroll
            for (int i = 0; i < N; ++i) a += Lo[i] * Number(un[k * ND + base + i * sd]);
            um[k] = a;
          }
        }
        // geometry at the face point: normal = J^-T e_d / |.|, dS = det J |J^-T e_d|
        Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM];
        {
          int j = 0;
#pragma unroll
          for (int e = 0; e < DIM; ++e) {
            if (e == d)
              xi[e] = Number(s);
            else
              xi[e] = X[tdig[j++]];
          }
        }
        const Number det = multilinear_jacobian<DIM>(s_V[lc], xi, J, Ji);
        Number nrm[DIM], na = Number(0);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          nrm[i] = Ji[d * DIM + i];
          na += nrm[i] * nrm[i];
        }
        na = sqrt(na);
        const Number sg = s ? Number(1) : Number(-1);
#pragma unroll
        for (int i = 0; i < DIM; ++i) nrm[i] = sg * nrm[i] / na;
        // the tangential weights cancel against the node weights later; a
        // hanging face contributes nothing here (nc_face_kernel adds it)
        const Number dS = bid == NC_BID ? Number(0) : fabs(det) * na;
        Number nvp = Number(0), nvm = Number(0);
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          nvp += up[i] * nrm[i];
          nvm += um[i] * nrm[i];
        }
        Number val_pn, val_p;
        flux<Number>(nvp, up[DIM], nvm, um[DIM], rho, cc, rho_m, c_m, bid, val_pn, val_p);
#pragma unroll
        for (int i = 0; i < DIM; ++i) s_F[lc][f][i][t] = val_pn * nrm[i] * dS;
        s_F[lc][f][DIM][t] = val_p * dS;
      }
    }
  }
  __syncthreads();

  // ---------------- gather faces, weak divergence, inverse mass ----------------
  if (lane_ok) {
    if constexpr (CART) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        int t;
        if constexpr (DIM == 2)
          t = qi[1 - d];
        else
          t = d == 0 ? qi[1] + N * qi[2] : (d == 1 ? qi[0] + N * qi[2] : qi[0] + N * qi[1]);
        const Number ih = Number(1) / hh[d];
        const Number l0 = LW[qi[d]] * ih, l1 = LW[N + qi[d]] * ih;
        kv[d] += l0 * s_F[lc][2 * d][0][t] + l1 * s_F[lc][2 * d + 1][0][t];
        kp += l0 * s_F[lc][2 * d][1][t] + l1 * s_F[lc][2 * d + 1][1][t];
      }
    } else {
      // transposed derivative of the test-gradient values
#pragma unroll
      for (int e = 0; e < DIM; ++e) {
        const int se = pow_n<N>(e);
        const int q0 = q - qi[e] * se;
        Number a = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) a += D[m * N + qi[e]] * s_g[lc][e][q0 + m * se];
        kp += a;
      }
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
        for (int i = 0; i < DIM; ++i)
          fv[i] += l0 * s_F[lc][2 * d][i][t] + l1 * s_F[lc][2 * d + 1][i][t];
        fpv += l0 * s_F[lc][2 * d][DIM][t] + l1 * s_F[lc][2 * d + 1][DIM][t];
      }
      const Number ij = Number(1) / jxw_node;
#pragma unroll
      for (int i = 0; i < DIM; ++i) kv[i] += fv[i] * ij;
      kp = (kp + fpv) * ij;
    }
  }

  // ---------------- fused epilogue (FP64, streaming stores) ----------------
  if (active) {
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
