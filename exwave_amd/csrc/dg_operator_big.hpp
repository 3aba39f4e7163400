// dg_operator_big.hpp -- fused operator + RK epilogue for cells with more Gauss
// nodes than a workgroup has lanes (3D k >= 10: 1331 to 2197 nodes): one cell
// per workgroup of 512 lanes, each lane looping over up to five nodes and two
// face points (the LDS holds one such cell per CU or two, so a larger group
// would add no waves, only halve each lane's registers).
//
// The arithmetic is that of dg_operator_gen_kernel (evaluate_cell /
// evaluate_inner_face / evaluate_boundary_face, source/wave_equation_operations.cc:
// 424-629, multilinear geometry, inverse mass 1/(w det J), RK epilogue); only
// the data movement differs, to fit the LDS:
//   1 (lane = node): the cell's fields into LDS;
//   2 (lane = face point, up to 1014): own face traces from LDS, the
//     neighbour's trace gathered from memory along its normal line, geometry,
//     flux into LDS;
//   3 (lane = node): the test gradients rho c^2 w det J (J^-1 v), written
//     over v (each lane over its own nodes, after every face trace is read);
//   4 (lane = node): the pressure gradient, the weak divergence, lifting,
//     inverse mass, epilogue; the node's geometry is recomputed in each node
//     phase, so that a lane holds nothing across a barrier.
// LDS: (dim+1) N^3 + 2 dim (dim+1) N^2 values (102 KB at k = 12 in FP64).
#pragma once
#include "dg_common.hpp"

namespace exw {

template <int DIM, int K>
struct BigShape {
  static constexpr int N = K + 1;
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;
  static constexpr int NT = DIM == 2 ? N : N * N;
  static constexpr int NF = DIM + 1;
  static constexpr int NFACE = 2 * DIM;
  static constexpr int BLOCK = 512;
  static constexpr int R = (ND + BLOCK - 1) / BLOCK;  // nodes per lane
  static constexpr int NFP = NFACE * NT;
  static constexpr int RF = (NFP + BLOCK - 1) / BLOCK;  // face points per lane
  static constexpr bool ENABLED = ND > 1024;
};

template <int DIM, int K, typename Number>
__global__ void __launch_bounds__(512)
    dg_operator_big_kernel(const DevMesh mesh, const double *__restrict__ src,
                           const double *__restrict__ tab, const Epilogue epi) {
  using S = BigShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NF = S::NF, NFACE = S::NFACE;
  constexpr int BLOCK = S::BLOCK, R = S::R, RF = S::RF, NFP = S::NFP, NV = 1 << DIM;
  using TO = TabOff<N>;

  __shared__ Number s_u[NF][ND];          // fields; v_e replaced by the test gradients
  __shared__ Number s_F[NFACE][NF][NT];   // face flux contributions
  __shared__ Number s_tab[TO::SIZE];
  __shared__ double s_V[NV * DIM];
  __shared__ Number s_mat[NFACE][2];
  __shared__ int s_nbr[NFACE];

  const int tid = threadIdx.x;
  const int cell = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x));
  double cscale = 1.0;
  if (epi.cell_scale) cscale = epi.cell_scale[cell];
  if (cscale == 0.0) return;  // uniform over the group
  const size_t cbase = (size_t)cell * NF * ND;
  const double rho_d = mesh.rho[cell], c_d = mesh.speed[cell];

  // ---------------- 1: the cell's fields, corners, materials ----------------
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int q = tid + r * BLOCK;
    if (q < ND)
#pragma unroll
      for (int f = 0; f < NF; ++f) s_u[f][q] = Number(src[cbase + f * ND + q]);
  }
  if (tid < NV * DIM) s_V[tid] = mesh.vertices[(size_t)cell * NV * DIM + tid];
  if (tid < NFACE) {
    const int nb = mesh.neighbor[(size_t)cell * NFACE + tid];
    s_nbr[tid] = nb;
    s_mat[tid][0] = Number(nb >= 0 ? mesh.rho[nb] : rho_d);
    s_mat[tid][1] = Number(nb >= 0 ? mesh.speed[nb] : c_d);
  }
  for (int i = tid; i < TO::SIZE; i += BLOCK) s_tab[i] = Number(tab[i]);
  __syncthreads();
  const Number rho = Number(rho_d), cc = Number(c_d);
  const Number rho_inv = Number(1) / rho, rc2 = rho * cc * cc;
  const Number *X = s_tab + TO::X, *W = s_tab + TO::W, *D = s_tab + TO::D, *L = s_tab + TO::L;

  // ---------------- 2: face points ----------------
#pragma unroll 1
  for (int rf = 0; rf < RF; ++rf) {
    const int fp = tid + rf * BLOCK;
    if (fp >= NFP) continue;
    const int f = fp / NT, t = fp - f * NT, d = f >> 1, s = f & 1;
    int base, td[2] = {0, 0};
    if constexpr (DIM == 2) {
      td[0] = t;
      base = d == 0 ? t * N : t;
    } else {
      td[0] = t % N;
      td[1] = t / N;
      base = d == 0 ? td[0] * N + td[1] * N * N
                    : (d == 1 ? td[0] + td[1] * N * N : td[0] + td[1] * N);
    }
    const int sd = d == 0 ? 1 : (d == 1 ? N : N * N);
    const Number *Ls = L + s * N, *Lo = L + (1 - s) * N;
    const int nb = s_nbr[f];
    // boundary faces read the cell itself (values unused by the boundary flux)
    const double *un = src + (size_t)(nb >= 0 ? nb : cell) * NF * ND;
    Number up[NF], um[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) up[k] = um[k] = Number(0);
    // the neighbour's nodes in chunks of GC line points: GC * NF loads in
    // flight at a time instead of one point's NF (a rolled loop waited for
    // every point in turn), the sums in the same order
    constexpr int GC = 4;
#pragma unroll 1
    for (int i0 = 0; i0 < N; i0 += GC) {
      double g[GC][NF];
#pragma unroll
      for (int j = 0; j < GC; ++j) {
        const int o = base + (i0 + j < N ? i0 + j : N - 1) * sd;
#pragma unroll
        for (int k = 0; k < NF; ++k) g[j][k] = un[k * ND + o];
      }
#pragma unroll
      for (int j = 0; j < GC; ++j) {
        if (i0 + j < N) {
          const int o = base + (i0 + j) * sd;
          const Number ls = Ls[i0 + j], lo = Lo[i0 + j];
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            up[k] += ls * s_u[k][o];
            um[k] += lo * Number(g[j][k]);
          }
        }
      }
    }
    Number xi[DIM], J[DIM * DIM], Jf[DIM * DIM];
    {
      int j = 0;
#pragma unroll
      for (int e = 0; e < DIM; ++e) xi[e] = e == d ? Number(s) : X[td[j++]];
    }
    const Number det = multilinear_jacobian<DIM>(s_V, xi, J, Jf);
    Number nrm[DIM], na = Number(0);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      nrm[i] = Jf[d * DIM + i];
      na += nrm[i] * nrm[i];
    }
    na = sqrt(na);
    const Number sg = s ? Number(1) : Number(-1);
#pragma unroll
    for (int i = 0; i < DIM; ++i) nrm[i] = sg * nrm[i] / na;
    // a hanging face contributes nothing here
    const Number dS = nb == -NC_BID ? Number(0) : fabs(det) * na;
    Number nvp = Number(0), nvm = Number(0);
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      nvp += up[i] * nrm[i];
      nvm += um[i] * nrm[i];
    }
    Number val_pn, val_p;
    flux<Number>(nvp, up[DIM], nvm, um[DIM], rho, cc, s_mat[f][0], s_mat[f][1],
                 nb >= 0 ? 0 : -nb, val_pn, val_p);
#pragma unroll
    for (int i = 0; i < DIM; ++i) s_F[f][i][t] = val_pn * nrm[i] * dS;
    s_F[f][DIM][t] = val_p * dS;
  }
  __syncthreads();  // every trace of v read

  // ---------------- 3: test gradients over v (each lane its own nodes) ----------------
#pragma unroll 1
  for (int r = 0; r < R; ++r) {
    const int q = tid + r * BLOCK;
    if (q >= ND) continue;
    const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
    Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM], wq = Number(1);
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      xi[e] = X[qi[e]];
      wq *= W[qi[e]];
    }
    const Number jxw = wq * multilinear_jacobian<DIM>(s_V, xi, J, Ji);
    Number gt[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      Number sum = Number(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i) sum += Ji[e * DIM + i] * s_u[i][q];
      gt[e] = rc2 * jxw * sum;
    }
#pragma unroll
    for (int e = 0; e < DIM; ++e) s_u[e][q] = gt[e];
  }
  __syncthreads();

  // ---------------- 4: pressure gradient, weak divergence, lifting, inverse mass,
  // epilogue (the node's geometry recomputed: nothing is held across barriers) ----
  const double ea1 = epi.a1 * cscale, ea2 = epi.a2 * cscale, ea3 = epi.a3 * cscale;
#pragma unroll 1
  for (int r = 0; r < R; ++r) {
    const int q = tid + r * BLOCK;
    if (q >= ND) continue;
    const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
    Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM], wq = Number(1);
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      xi[e] = X[qi[e]];
      wq *= W[qi[e]];
    }
    const Number jxw = wq * multilinear_jacobian<DIM>(s_V, xi, J, Ji);
    Number gref[DIM], kp = Number(0);
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      const int se = pow_n<N>(e), q0 = q - qi[e] * se;
      Number g = Number(0), a = Number(0);
#pragma unroll 1
      for (int m = 0; m < N; ++m) {
        g += D[qi[e] * N + m] * s_u[DIM][q0 + m * se];
        a += D[m * N + qi[e]] * s_u[e][q0 + m * se];
      }
      gref[e] = g;
      kp += a;
    }
    Number kk[NF];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      Number gp = Number(0);
#pragma unroll
      for (int e = 0; e < DIM; ++e) gp += Ji[e * DIM + i] * gref[e];
      kk[i] = -rho_inv * gp;
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
      for (int i = 0; i < DIM; ++i) fv[i] += l0 * s_F[2 * d][i][t] + l1 * s_F[2 * d + 1][i][t];
      fpv += l0 * s_F[2 * d][DIM][t] + l1 * s_F[2 * d + 1][DIM][t];
    }
    const Number ij = Number(1) / jxw;
#pragma unroll
    for (int i = 0; i < DIM; ++i) kk[i] += fv[i] * ij;
    kk[DIM] = (kp + fpv) * ij;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const size_t idx = cbase + f * ND + q;
      const double k = double(kk[f]);
      const double o1 = (epi.b ? epi.b[idx] : 0.0) + ea1 * k;
      const double cv = (epi.o2 && !epi.o2_from_o1) ? epi.c[idx] : 0.0;
      __builtin_nontemporal_store(o1, epi.o1 + idx);
      if (epi.o2) {
        const double c = epi.o2_from_o1 ? o1 : cv;
        __builtin_nontemporal_store(c + ea2 * k, epi.o2 + idx);
        if (epi.o3) __builtin_nontemporal_store(c + ea3 * k, epi.o3 + idx);
      }
    }
  }
}

}  // namespace exw
