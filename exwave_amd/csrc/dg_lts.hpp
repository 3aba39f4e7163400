// dg_lts.hpp -- the cell sweeps of ADER with local time stepping
// (WaveEquationOperationADERLTS, source/wave_equation_operations.cc:1385-1692,
// driven by ClusterManager, include/cluster_manager.templates.h:488-1049).
//
// One kernel, three roles, each over a list of cells (one lane per Gauss node,
// general multilinear geometry, FP64):
//   LTS_ACC     the face integrals of the listed faces of each listed cell,
//               evaluated from the CK time integrals in `src` on both sides
//               (local_apply_ader_face / _boundary_face, :1514-1557, with
//               boundary_fac = fac), times the inverse mass, added into that
//               cell's row of acc0 (the update's face sum) or acc1 (the flux
//               memory, for a cell of a neighbouring cluster);
//   LTS_UPDATE  the second ADER sweep of the update cells (local_apply_
//               secondader_domain, :1413-1497): the volume term of the CK
//               integral in `src` with the ADER sign, plus the face sum (acc0)
//               and the flux memory (acc1), subtracted from the state; acc0 and
//               the flux memory of the cell are cleared;
//   LTS_APPLY   M^-1 A(src) of the listed cells (reconstruct_div_grad,
//               :1593-1680).
// The reference accumulates face terms before the inverse mass; here every
// partial sum is kept multiplied by the (diagonal) inverse mass, the same
// linear combination.  Each cell is owned by one workgroup and only its own
// rows are written: no atomics.
#pragma once
#include "dg_common.hpp"

namespace exw {

enum { LTS_ACC = 0, LTS_UPDATE = 1, LTS_APPLY = 2 };

struct LtsArgs {
  const int *cells;             // entry -> cell
  const unsigned *mask;         // entry -> faces to evaluate (bit 4 f); NULL: all
  const unsigned char *target;  // LTS_ACC: entry -> 0 (acc0) or 1 (acc1)
  int n;                        // entries
  int mode;
  const double *src;            // ACC: CK integrals; UPDATE: CK over [0,dt]; APPLY: input
  double fac;                   // boundary_fac of the face terms
  double *acc0, *acc1;          // face sum, flux memory
  double *out;                  // UPDATE: state (-=); APPLY: result
  int use_fm;                   // UPDATE: add and clear the flux memory
};

// On box cells (CART) the Jacobian is diag(h) (mesh.h) instead of the
// multilinear map of the corners.
template <int DIM, bool CART>
__device__ __forceinline__ double lts_jacobian(const double *V, const double *h, const double *xi,
                                               double *J, double *Ji) {
  if constexpr (CART) {
    double det = 1.0;
#pragma unroll
    for (int k = 0; k < DIM * DIM; ++k) J[k] = Ji[k] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      J[d * DIM + d] = h[d];
      Ji[d * DIM + d] = 1.0 / h[d];
      det *= h[d];
    }
    return det;
  } else {
    return multilinear_jacobian<DIM>(V, xi, J, Ji);
  }
}

template <int DIM, int K, bool CART>
__global__ void __launch_bounds__((Shape<DIM, K>::BLOCK))
    dg_lts_kernel(const DevMesh mesh, const double *__restrict__ tab, const LtsArgs a) {
  using S = Shape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF, NT = S::NT, CPB = S::CPB;
  constexpr int NFACE = 2 * DIM;
  constexpr int NVD = (1 << DIM) * DIM;
  using TO = TabOff<N>;

  __shared__ double s_u[CPB][NF][ND];
  __shared__ double s_g[CPB][DIM][ND];
  __shared__ double s_F[CPB][NFACE][NF][NT];
  __shared__ double s_tab[TO::SIZE];
  __shared__ double s_V[CPB][NVD];
  __shared__ int s_nbr[CPB][NFACE];

  const int tid = threadIdx.x;
  const int lc = tid / ND;
  const int q = tid - lc * ND;
  const bool lane_ok = lc < CPB;
  const int e = blockIdx.x * CPB + lc;
  const bool active = lane_ok && e < a.n;
  const int cell = active ? a.cells[e] : 0;
  // bit 4f selects face f (the bits 4f+1..4f+3 address the other subfaces of
  // a 2:1 face, which nc_face_kernel evaluates: such faces add nothing here)
  const unsigned mask = active ? (a.mask ? a.mask[e] : 0xFFFFFFFFu) : 0u;
  const bool do_vol = a.mode != LTS_ACC;
  const bool do_face = a.mode != LTS_UPDATE;
  const size_t cbase = (size_t)cell * NF * ND;

  for (int i = tid; i < TO::SIZE; i += blockDim.x) s_tab[i] = tab[i];
  if constexpr (!CART) {
    for (int i = tid; i < CPB * NVD; i += blockDim.x) {
      const int e2 = blockIdx.x * CPB + i / NVD;
      s_V[i / NVD][i % NVD] = e2 < a.n ? mesh.vertices[(size_t)a.cells[e2] * NVD + i % NVD] : 0.0;
    }
  }
  double rho = 1.0, cc = 1.0, hc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) hc[d] = 1.0;
  if (active) {
    rho = mesh.rho[cell];
    cc = mesh.speed[cell];
    if constexpr (CART) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) hc[d] = mesh.h[(size_t)cell * DIM + d];
    }
    if (q < NFACE) s_nbr[lc][q] = mesh.neighbor[(size_t)cell * NFACE + q];
  }
  const double rho_inv = 1.0 / rho, rc2 = rho * cc * cc;
  if (lane_ok) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[lc][f][q] = active ? a.src[cbase + f * ND + q] : 0.0;
  }
  int qi[3];
  qi[0] = q % N;
  qi[1] = (q / N) % N;
  qi[2] = DIM == 3 ? q / (N * N) : 0;
  __syncthreads();

  const double *D = s_tab + TO::D;
  const double *L = s_tab + TO::L;
  const double *X = s_tab + TO::X;
  const double *W = s_tab + TO::W;

  double kv[DIM], kp = 0.0, jxw_node = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) kv[d] = 0.0;
  if (lane_ok) {
    double xi[DIM], J[DIM * DIM], Ji[DIM * DIM], wq = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = X[qi[d]];
      wq *= W[qi[d]];
    }
    const double det = lts_jacobian<DIM, CART>(s_V[lc], hc, xi, J, Ji);
    jxw_node = wq * det;
    if (do_vol) {
      // evaluate_cell (:424-456): strong -rho^-1 grad p, test gradient of p
      double gref[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int sd = pow_n<N>(d);
        const int q0 = q - qi[d] * sd;
        double g = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) g += D[qi[d] * N + m] * s_u[lc][DIM][q0 + m * sd];
        gref[d] = g;
      }
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        double gp = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) gp += Ji[d * DIM + i] * gref[d];
        kv[i] = -rho_inv * gp;
      }
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) s += Ji[d * DIM + i] * s_u[lc][i][q];
        s_g[lc][d][q] = rc2 * jxw_node * s;
      }
    }
  }
  __syncthreads();

  // face points of the listed faces, evaluated from this cell's side
  if (lane_ok && do_face) {
    for (int fp = q; fp < NFACE * NT; fp += ND) {
      const int f = fp / NT, t = fp - f * NT, d = f >> 1, s = f & 1;
      if (!((mask >> (4 * f)) & 1u) || s_nbr[lc][f] == -NC_BID) {
#pragma unroll
        for (int k = 0; k < NF; ++k) s_F[lc][f][k][t] = 0.0;
        continue;
      }
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
      const int nb = s_nbr[lc][f];
      const double *Ls = L + s * N, *Lo = L + (1 - s) * N;
      double up[NF], um[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) v += Ls[i] * s_u[lc][k][base + i * sd];
        up[k] = v;
        um[k] = 0.0;
      }
      double rho_m = rho, c_m = cc;
      if (nb >= 0) {
        const double *un = a.src + (size_t)nb * NF * ND;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double v = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) v += Lo[i] * un[k * ND + base + i * sd];
          um[k] = v;
        }
        rho_m = mesh.rho[nb];
        c_m = mesh.speed[nb];
      }
      double xi[DIM], J[DIM * DIM], Ji[DIM * DIM];
      {
        int j = 0;
#pragma unroll
        for (int e2 = 0; e2 < DIM; ++e2) xi[e2] = e2 == d ? double(s) : X[tdig[j++]];
      }
      const double det = lts_jacobian<DIM, CART>(s_V[lc], hc, xi, J, Ji);
      double nrm[DIM], na = 0.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        nrm[i] = Ji[d * DIM + i];
        na += nrm[i] * nrm[i];
      }
      na = sqrt(na);
      const double sg = s ? 1.0 : -1.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) nrm[i] = sg * nrm[i] / na;
      const double dS = fabs(det) * na * a.fac;
      double nvp = 0.0, nvm = 0.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        nvp += up[i] * nrm[i];
        nvm += um[i] * nrm[i];
      }
      double val_pn, val_p;
      flux<double>(nvp, up[DIM], nvm, um[DIM], rho, cc, rho_m, c_m, nb >= 0 ? 0 : -nb, val_pn,
                   val_p);
#pragma unroll
      for (int i = 0; i < DIM; ++i) s_F[lc][f][i][t] = val_pn * nrm[i] * dS;
      s_F[lc][f][DIM][t] = val_p * dS;
    }
  }
  __syncthreads();

  if (lane_ok) {
    if (do_vol) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int sd = pow_n<N>(d);
        const int q0 = q - qi[d] * sd;
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) v += D[m * N + qi[d]] * s_g[lc][d][q0 + m * sd];
        kp += v;
      }
    }
    double fv[DIM], fpv = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) fv[i] = 0.0;
    if (do_face) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        int t;
        double wt = 1.0;
        if constexpr (DIM == 2) {
          t = qi[1 - d];
          wt = W[qi[1 - d]];
        } else {
          t = d == 0 ? qi[1] + N * qi[2] : (d == 1 ? qi[0] + N * qi[2] : qi[0] + N * qi[1]);
#pragma unroll
          for (int e2 = 0; e2 < DIM; ++e2)
            if (e2 != d) wt *= W[qi[e2]];
        }
        const double l0 = L[qi[d]] * wt, l1 = L[N + qi[d]] * wt;
#pragma unroll
        for (int i = 0; i < DIM; ++i)
          fv[i] += l0 * s_F[lc][2 * d][i][t] + l1 * s_F[lc][2 * d + 1][i][t];
        fpv += l0 * s_F[lc][2 * d][DIM][t] + l1 * s_F[lc][2 * d + 1][DIM][t];
      }
    }
    const double ij = 1.0 / jxw_node;
#pragma unroll
    for (int i = 0; i < DIM; ++i) kv[i] += fv[i] * ij;
    kp = (kp + fpv) * ij;
  }

  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const size_t idx = cbase + f * ND + q;
      const double k = f < DIM ? kv[f < DIM ? f : 0] : kp;
      if (a.mode == LTS_ACC) {
        double *t = (a.target && a.target[e]) ? a.acc1 : a.acc0;
        t[idx] += k;
      } else if (a.mode == LTS_UPDATE) {
        // second ADER sweep: -volume(CK) + face sum (+ flux memory), then
        // local_state -= that (include/cluster_manager.templates.h:1046)
        double tot = -k + a.acc0[idx];
        a.acc0[idx] = 0.0;
        if (a.use_fm) {
          tot += a.acc1[idx];
          a.acc1[idx] = 0.0;
        }
        a.out[idx] -= tot;
      } else {
        a.out[idx] = k;
      }
    }
  }
}

}  // namespace exw
