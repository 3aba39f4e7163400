// dg_ader.hpp -- ADER Cauchy-Kowalevski time integral as one fused kernel.
//
// Restates integrate_taylor_cauchykovalewski with spectral_evaluation
// (source/wave_equation_operations.cc:961-1262): per cell
//   contrib = (t2-t1) u + 1/2[(t2-te)^2-(t1-te)^2] u'
//             + sum_{j>=2} fac_j [(t2-te)^{j+1}-(t1-te)^{j+1}] u^(j),
// where u^(j) are obtained by applying the local collocation operator
// repeatedly at the Gauss points; every second derivative the field is
// L2-projected (on the unit cell) to two polynomial degrees less
// (source/wave_equation_operations.cc:287-313, 1238-1255), and the
// lower-degree contributions are interpolated back at the end.
//
// GPU layout: one lane per top-level Gauss node; the working field is double
// buffered in LDS; the level-0 contribution stays in registers, lower levels'
// contributions live in LDS.
#pragma once
#include "dg_common.hpp"

namespace exw {

// Offsets of the CK tables (per level l: n_l = K+1-2l points).
struct CkTab {
  int n_levels;
  int off_D[8];  // D_l[q*n_l+i] = l_i'(x_q)
  int off_X[8];  // Gauss points of level l
  int off_P[8];  // P_l (level l-1 -> l), [i*n_{l-1}+q] = l^lo_i(x^hi_q) w^hi_q / w^lo_i
  int off_V[8];  // V_l (level l -> l-1), [q*n_l+i] = l^lo_i(x^hi_q)
  int size;
};

template <int K>
struct CkShape {
  static constexpr int N_STEPS = K >= 2 ? K - 1 : 1;  // steps 0 .. max(0, K-2)
  static constexpr int N_LEVELS = (N_STEPS - 1) / 2 + 1;
  static constexpr int n_of(int l) { return K + 1 - 2 * l; }
  static constexpr int TAB_SIZE = [] {
    int s = 0;
    for (int l = 0; l < N_LEVELS; ++l) {
      const int n = K + 1 - 2 * l;
      s += n * n + n;
      if (l > 0) s += 2 * n * (n + 2);
    }
    return s;
  }();
};

template <int DIM, int N>
__device__ __forceinline__ constexpr int ipow_c() {
  return DIM == 2 ? N * N : N * N * N;
}

template <int DIM, int K, typename Number, bool CART>
__global__ void __launch_bounds__((Shape<DIM, K>::BLOCK))
    ader_ck_kernel(const DevMesh mesh, const double *__restrict__ src,
                   const double *__restrict__ recon, const double *__restrict__ cktab,
                   const CkTab ck, double t2, double t1, double te_all, const CellSel sel,
                   double *__restrict__ tempsrc) {
  using S = Shape<DIM, K>;
  using CS = CkShape<K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF, CPB = S::CPB;
  // sizes of the lower levels' contributions
  constexpr int LOW = [] {
    int s = 0;
    for (int l = 1; l < CS::N_LEVELS; ++l) {
      int n = CS::n_of(l);
      s += DIM == 2 ? n * n : n * n * n;
    }
    return s > 0 ? s : 1;
  }();

  __shared__ Number s_a[CPB][NF][ND];
  __shared__ Number s_b[CPB][NF][ND];
  __shared__ Number s_c[CPB][NF][LOW];
  __shared__ double s_V[CART ? 1 : CPB][CART ? 1 : (1 << DIM) * DIM];
  __shared__ Number s_ck[CS::TAB_SIZE];

  const int tid = threadIdx.x;
  for (int i = tid; i < CS::TAB_SIZE && i < ck.size; i += blockDim.x) s_ck[i] = Number(cktab[i]);
  const int lc = tid / ND;
  const int q = tid - lc * ND;
  const bool lane_ok = lc < CPB;
  // entry -> cell (identity, or through the selection list)
  auto cell_of = [&](int e) { return sel.list ? (e < sel.n ? sel.list[e] : -1) : e; };
  const int cell = cell_of(blockIdx.x * CPB + lc);
  const bool active = lane_ok && cell >= 0 && cell < mesh.n_cells;
  const double te = active && sel.cluster ? sel.te[sel.cluster[cell]] : te_all;

  if constexpr (!CART) {
    constexpr int NVD = (1 << DIM) * DIM;
    for (int i = tid; i < CPB * NVD; i += blockDim.x) {
      const int c2 = cell_of(blockIdx.x * CPB + i / NVD);
      s_V[i / NVD][i % NVD] =
          c2 >= 0 && c2 < mesh.n_cells ? mesh.vertices[(size_t)c2 * NVD + i % NVD] : 0.0;
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
  const Number rho_inv = Number(1) / rho, rc2 = rho * cc * cc;

  const size_t cbase = (size_t)cell * NF * ND;
  Number u[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) u[f] = Number(0);
  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      u[f] = Number(src[cbase + f * ND + q]);
      s_a[lc][f][q] = u[f];
    }
  }
  __syncthreads();

  // Physical derivative of the level field `in` (n points per direction) at
  // the lane's node: out_v[d] = rho^-1 d_d in_p, out_p = rho c^2 div in_v
  auto derive = [&](Number(*in)[ND], int n, const Number *Dl, const Number *Xl, int qq,
                    Number *outv, Number &outp) {
    int qd[3] = {qq % n, (qq / n) % n, DIM == 3 ? qq / (n * n) : 0};
    const int st[3] = {1, n, n * n};
    if constexpr (CART) {
      Number div = Number(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int q0 = qq - qd[d] * st[d];
        Number gp = Number(0), gv = Number(0);
        for (int m = 0; m < n; ++m) {
          gp += Dl[qd[d] * n + m] * in[DIM][q0 + m * st[d]];
          gv += Dl[qd[d] * n + m] * in[d][q0 + m * st[d]];
        }
        const Number ih = Number(1) / hh[d];
        outv[d] = rho_inv * gp * ih;
        div += gv * ih;
      }
      outp = rc2 * div;
    } else {
      Number gr[NF][DIM];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
          const int q0 = qq - qd[e] * st[e];
          Number g = Number(0);
          for (int m = 0; m < n; ++m) g += Dl[qd[e] * n + m] * in[f][q0 + m * st[e]];
          gr[f][e] = g;
        }
      Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) xi[d] = Xl[qd[d]];
      multilinear_jacobian<DIM>(s_V[lc], xi, J, Ji);
      Number div = Number(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        Number pg = Number(0);
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
          pg += gr[DIM][e] * Ji[e * DIM + d];
          div += gr[d][e] * Ji[e * DIM + d];
        }
        outv[d] = rho_inv * pg;
      }
      outp = rc2 * div;
    }
  };

  auto tabv = [&](int off, int i) { return s_ck[off + i]; };

  // contrib (level 0) and the first derivative
  const Number c1 = Number(0.5 * ((t2 - te) * (t2 - te) - (t1 - te) * (t1 - te)));
  Number contrib[NF];
  Number up[NF];
  if (lane_ok) {
    if (recon) {
#pragma unroll
      for (int f = 0; f < NF; ++f) up[f] = active ? Number(recon[cbase + f * ND + q]) : Number(0);
    } else {
      Number ov[DIM], op;
      derive(s_a[lc], N, s_ck + ck.off_D[0], s_ck + ck.off_X[0], q, ov, op);
#pragma unroll
      for (int d = 0; d < DIM; ++d) up[d] = -ov[d];
      up[DIM] = -op;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) contrib[f] = Number(t2 - t1) * u[f] + c1 * up[f];
  }
  __syncthreads();
  if (lane_ok) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_a[lc][f][q] = -up[f];
  }
  __syncthreads();

  // steps
  Number(*cur)[NF][ND] = s_a;
  Number(*nxt)[NF][ND] = s_b;
  int low_off = 0;  // offset of the current level's contribution in s_c
  double fac = -0.5;
  for (int step = 0; step < CS::N_STEPS; ++step) {
    fac /= -double(step + 3);
    const int level = step / 2;
    const int n = K + 1 - 2 * level;
    const int nd = DIM == 2 ? n * n : n * n * n;
    const Number fac_t =
        Number(fac * (pow(t2 - te, double(step + 3)) - pow(t1 - te, double(step + 3))));
    if (lane_ok && q < nd) {
      Number ov[DIM], op;
      derive(cur[lc], n, s_ck + ck.off_D[level], s_ck + ck.off_X[level], q, ov, op);
#pragma unroll
      for (int d = 0; d < DIM; ++d) nxt[lc][d][q] = ov[d];
      nxt[lc][DIM][q] = op;
      if (level == 0) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) contrib[d] += fac_t * ov[d];
        contrib[DIM] += fac_t * op;
      } else {
        const bool first = (step & 1) == 0;  // first step of a reduced level sets
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          s_c[lc][d][low_off + q] = (first ? Number(0) : s_c[lc][d][low_off + q]) + fac_t * ov[d];
        s_c[lc][DIM][low_off + q] =
            (first ? Number(0) : s_c[lc][DIM][low_off + q]) + fac_t * op;
      }
    }
    __syncthreads();
    auto tmp = cur;
    cur = nxt;
    nxt = tmp;
    // project to the next level after an odd step that is not the last one
    if ((step & 1) == 1 && step < K - 2) {
      const int nl = n - 2;
      int sz[3] = {n, n, DIM == 3 ? n : 1};
      const int P = ck.off_P[level + 1];
      // directions 0.. DIM-1, ping-ponging cur/nxt
      for (int dir = 0; dir < DIM; ++dir) {
        if (lane_ok) {
          for (int f = 0; f < NF; ++f) {
            // A[r*a_rs + c*a_cs] = P[r*n + c]
            int sz2[3] = {sz[0], sz[1], sz[2]};
            const int total_out = (sz2[0] * sz2[1] * sz2[2] / sz2[dir]) * nl;
            int szo[3] = {sz2[0], sz2[1], sz2[2]};
            szo[dir] = nl;
            const int sin[3] = {1, sz2[0], sz2[0] * sz2[1]};
            for (int o = q; o < total_out; o += ND) {
              int idx[3] = {o % szo[0], (o / szo[0]) % szo[1], o / (szo[0] * szo[1])};
              const int r = idx[dir];
              idx[dir] = 0;
              const int b = idx[0] * sin[0] + idx[1] * sin[1] + idx[2] * sin[2];
              Number acc = Number(0);
              for (int c = 0; c < n; ++c) acc += tabv(P, r * n + c) * cur[lc][f][b + c * sin[dir]];
              nxt[lc][f][o] = acc;
            }
          }
        }
        sz[dir] = nl;
        __syncthreads();
        auto t3 = cur;
        cur = nxt;
        nxt = t3;
      }
      if (level >= 1) low_off += nd;
    }
  }

  // interpolate the lower levels' contributions back up and add
  {
    // level offsets in s_c
    int offs[8];
    int o = 0;
    for (int l = 1; l < CS::N_LEVELS; ++l) {
      offs[l] = o;
      const int n = K + 1 - 2 * l;
      o += DIM == 2 ? n * n : n * n * n;
    }
    for (int l = CS::N_LEVELS - 1; l >= 1; --l) {
      const int nlo = K + 1 - 2 * l, nhi = nlo + 2;
      const int V = ck.off_V[l];
      // interpolate s_c[level l] (nlo^DIM) to nhi^DIM through s_a/s_b
      Number(*bufA)[NF][ND] = s_a;
      Number(*bufB)[NF][ND] = s_b;
      int sz[3] = {nlo, nlo, DIM == 3 ? nlo : 1};
      for (int dir = 0; dir < DIM; ++dir) {
        if (lane_ok) {
          for (int f = 0; f < NF; ++f) {
            int szo[3] = {sz[0], sz[1], sz[2]};
            szo[dir] = nhi;
            const int total_out = szo[0] * szo[1] * szo[2];
            const int sin[3] = {1, sz[0], sz[0] * sz[1]};
            for (int oo = q; oo < total_out; oo += ND) {
              int idx[3] = {oo % szo[0], (oo / szo[0]) % szo[1], oo / (szo[0] * szo[1])};
              const int r = idx[dir];
              idx[dir] = 0;
              const int b = idx[0] * sin[0] + idx[1] * sin[1] + idx[2] * sin[2];
              Number acc = Number(0);
              for (int c = 0; c < nlo; ++c) {
                const Number v = dir == 0 ? s_c[lc][f][offs[l] + b + c * sin[dir]]
                                          : bufA[lc][f][b + c * sin[dir]];
                acc += tabv(V, r * nlo + c) * v;
              }
              bufB[lc][f][oo] = acc;
            }
          }
        }
        sz[dir] = nhi;
        __syncthreads();
        auto t3 = bufA;
        bufA = bufB;
        bufB = t3;
      }
      // bufA holds the interpolated contribution at level l-1
      const int ndhi = DIM == 2 ? nhi * nhi : nhi * nhi * nhi;
      if (lane_ok && q < ndhi) {
        if (l - 1 == 0) {
#pragma unroll
          for (int f = 0; f < NF; ++f) contrib[f] += bufA[lc][f][q];
        } else {
          for (int f = 0; f < NF; ++f) s_c[lc][f][offs[l - 1] + q] += bufA[lc][f][q];
        }
      }
      __syncthreads();
    }
  }

  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) tempsrc[cbase + f * ND + q] = double(contrib[f]);
  }
}

}  // namespace exw
