// dg_nc.hpp -- the subface integrals of non-conforming (2:1 hanging) faces.
//
// deal.II integrates a hanging face on the fine cell's face with its QGauss(k+1)
// points and evaluates the coarse cell there (FEFaceEvaluation on subfaces of
// the face loop, source/wave_equation_operations.cc:480-551):
//   fine side: own trace and geometry as on any face, the coarse neighbour's
//     trace at the matching points of its face quarter;
//   coarse side: the sum over the 2^(dim-1) subfaces of the same integrals,
//     with this cell's trace, normal, area element and test functions at the
//     mapped points (x_t + b)/2 and the weights w_t / 2 per tangential
//     direction.
// The operator sweeps give these faces no contribution (NC_BID); this kernel
// runs after the sweep over the cells that touch such faces only, one
// workgroup per cell, and adds the lifted subface integrals through the same
// epilogue: every output is linear in the operator value k, so o1 += a1 s dk,
// o2 += (a2 [+ a1]) s dk, o3 += (a3 [+ a1]) s dk.  FP64 throughout.
#pragma once
#include "dg_common.hpp"

namespace exw {

struct NcArgs {
  int n;              // cells with at least one non-conforming face
  const int *cells;   // [n]
  const int *kind;    // [n][2 dim]: 0 conforming, 1 fine side, 2 coarse side
  const int *other;   // [n][2 dim][4]: fine side {coarse cell, subface position},
                      // coarse side: the fine cells by subface position
  // optional [n]: bit 4f+b selects subface b of face f (the fine side's one
  // subface is bit 4f); NULL: every subface (local time stepping masks)
  const unsigned *mask = nullptr;
};

// table: X[N], W[N], L[2][N] (l_i(0), l_i(1)), H[2][N][N] (H_b[t][j] = l_j((x_t + b)/2))
template <int N>
struct NcTab {
  static constexpr int X = 0, W = N, L = 2 * N, H = 4 * N, SIZE = 4 * N + 2 * N * N;
};

template <int DIM, int K>
struct NcShape {
  static constexpr int N = K + 1;
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;
  static constexpr int NT = DIM == 2 ? N : N * N;
  static constexpr int NSUB = 1 << (DIM - 1);
  static constexpr int NF = DIM + 1;
  static constexpr int LANES0 = ND > NSUB * NT ? ND : NSUB * NT;
  static constexpr int BLOCK = (LANES0 + 63) / 64 * 64;
  static_assert(BLOCK <= 1024, "cell does not fit a workgroup");
};

template <int DIM, int K>
__global__ void __launch_bounds__((NcShape<DIM, K>::BLOCK))
    nc_face_kernel(const DevMesh mesh, const double *__restrict__ tab, const NcArgs nc,
                   const double *__restrict__ src, const Epilogue epi) {
  using S = NcShape<DIM, K>;
  using T = NcTab<K + 1>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NSUB = S::NSUB, NF = S::NF;
  constexpr int NFACE = 2 * DIM, NV = 1 << DIM;
  __shared__ double s_tab[T::SIZE];
  __shared__ double s_u[NF][ND];
  __shared__ double s_tr[NF][NT];
  __shared__ double s_Fb[NSUB][NF][NT];
  __shared__ double s_F[NFACE][NF][NT];
  __shared__ double s_V[NV * DIM];

  const int tid = threadIdx.x;
  const int cell = nc.cells[blockIdx.x];
  const double *uc = src + (size_t)cell * NF * ND;
  for (int i = tid; i < T::SIZE; i += S::BLOCK) s_tab[i] = tab[i];
  if (tid < NV * DIM) s_V[tid] = mesh.vertices[(size_t)cell * NV * DIM + tid];
  if (tid < ND)
#pragma unroll
    for (int f = 0; f < NF; ++f) s_u[f][tid] = uc[f * ND + tid];
  const double rho = mesh.rho[cell], cc = mesh.speed[cell];
  __syncthreads();
  const double *X = s_tab + T::X, *W = s_tab + T::W, *L = s_tab + T::L, *H = s_tab + T::H;

  // base node and stride of face point t (tangential digits ascending) on a
  // face of direction d
  auto face_base = [&](int d, int t, int *td) {
    if constexpr (DIM == 2) {
      td[0] = t;
      return d == 0 ? t * N : t;
    } else {
      td[0] = t % N;
      td[1] = t / N;
      return d == 0 ? td[0] * N + td[1] * N * N
                    : (d == 1 ? td[0] + td[1] * N * N : td[0] + td[1] * N);
    }
  };
  // outward unit normal and area element at reference point xi of face (d, s)
  auto face_geom = [&](int d, int s, const double *xi, double *nrm) {
    double J[DIM * DIM], Ji[DIM * DIM];
    const double det = multilinear_jacobian<DIM>(s_V, xi, J, Ji);
    double na = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      nrm[i] = Ji[d * DIM + i];
      na += nrm[i] * nrm[i];
    }
    na = sqrt(na);
    const double sg = s ? 1.0 : -1.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) nrm[i] = sg * nrm[i] / na;
    return fabs(det) * na;
  };
  // value at face point t of a field given by its face-node trace tr (NT
  // values), interpolated to the subface b of the face
  auto sub_interp = [&](const double *tr, int b, const int *td) {
    double s = 0.0;
    if constexpr (DIM == 2) {
      const double *h = H + (b & 1) * N * N + td[0] * N;
#pragma unroll
      for (int j = 0; j < N; ++j) s += h[j] * tr[j];
    } else {
      const double *h0 = H + (b & 1) * N * N + td[0] * N;
      const double *h1 = H + ((b >> 1) & 1) * N * N + td[1] * N;
      for (int j1 = 0; j1 < N; ++j1) {
        double a = 0.0;
#pragma unroll
        for (int j0 = 0; j0 < N; ++j0) a += h0[j0] * tr[j0 + N * j1];
        s += h1[j1] * a;
      }
    }
    return s;
  };

  const int *other = nc.other + (size_t)blockIdx.x * NFACE * 4;
  const unsigned emask = nc.mask ? nc.mask[blockIdx.x] : 0xFFFFFFFFu;
  int kind[NFACE];  // face kinds, 0 where the mask selects no subface
#pragma unroll
  for (int f = 0; f < NFACE; ++f) {
    const int kd = nc.kind[(size_t)blockIdx.x * NFACE + f];
    const unsigned sub = (emask >> (4 * f)) & 0xFu;
    kind[f] = (kd == 2 && sub) || (kd == 1 && (sub & 1u)) ? kd : 0;
  }
  for (int f = 0; f < NFACE; ++f) {
    const int kd = kind[f];
    if (kd == 0) continue;  // uniform over the group
    const int d = f >> 1, s = f & 1, sd = d == 0 ? 1 : (d == 1 ? N : N * N);
    if (kd == 1) {
      // fine side: the coarse neighbour's trace on its face nodes, then at
      // this face's points through the subface interpolation
      const int g = other[f * 4 + 0], pos = other[f * 4 + 1];
      const double *ug = src + (size_t)g * NF * ND;
      if (tid < NT) {
        int td[2];
        const int base = face_base(d, tid, td);
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double a = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) a += L[(1 - s) * N + i] * ug[k * ND + base + i * sd];
          s_tr[k][tid] = a;
        }
      }
      __syncthreads();
      if (tid < NT) {
        int td[2];
        const int base = face_base(d, tid, td);
        double up[NF], um[NF], xi[DIM], nrm[DIM];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double a = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) a += L[s * N + i] * s_u[k][base + i * sd];
          up[k] = a;
          um[k] = sub_interp(s_tr[k], pos, td);
        }
        int j = 0;
#pragma unroll
        for (int e = 0; e < DIM; ++e) xi[e] = e == d ? double(s) : X[td[j++]];
        const double dS = face_geom(d, s, xi, nrm);
        double nvp = 0.0, nvm = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          nvp += up[i] * nrm[i];
          nvm += um[i] * nrm[i];
        }
        double val_pn, val_p;
        flux<double>(nvp, up[DIM], nvm, um[DIM], rho, cc, mesh.rho[g], mesh.speed[g], 0, val_pn,
                     val_p);
#pragma unroll
        for (int i = 0; i < DIM; ++i) s_F[f][i][tid] = val_pn * nrm[i] * dS;
        s_F[f][DIM][tid] = val_p * dS;
      }
    } else {
      // coarse side: own trace on the face nodes, then every subface point
      if (tid < NT) {
        int td[2];
        const int base = face_base(d, tid, td);
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double a = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) a += L[s * N + i] * s_u[k][base + i * sd];
          s_tr[k][tid] = a;
        }
      }
      __syncthreads();
      if (tid < NSUB * NT && !((emask >> (4 * f + tid / NT)) & 1u)) {
        const int b = tid / NT, t = tid - b * NT;  // subface not selected
#pragma unroll
        for (int k = 0; k < NF; ++k) s_Fb[b][k][t] = 0.0;
      } else if (tid < NSUB * NT) {
        const int b = tid / NT, t = tid - b * NT;
        const int nb = other[f * 4 + b];
        const double *un = src + (size_t)nb * NF * ND;
        int td[2];
        const int base = face_base(d, t, td);
        double up[NF], um[NF], xi[DIM], nrm[DIM];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double a = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) a += L[(1 - s) * N + i] * un[k * ND + base + i * sd];
          um[k] = a;
          up[k] = sub_interp(s_tr[k], b, td);
        }
        int j = 0;
        double wt = 1.0 / NSUB;
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
          if (e == d) {
            xi[e] = double(s);
          } else {
            xi[e] = 0.5 * (X[td[j]] + ((b >> j) & 1));
            wt *= W[td[j]];
            ++j;
          }
        }
        const double dS = face_geom(d, s, xi, nrm) * wt;
        double nvp = 0.0, nvm = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          nvp += up[i] * nrm[i];
          nvm += um[i] * nrm[i];
        }
        double val_pn, val_p;
        flux<double>(nvp, up[DIM], nvm, um[DIM], rho, cc, mesh.rho[nb], mesh.speed[nb], 0,
                     val_pn, val_p);
#pragma unroll
        for (int i = 0; i < DIM; ++i) s_Fb[b][i][t] = val_pn * nrm[i] * dS;
        s_Fb[b][DIM][t] = val_p * dS;
      }
      __syncthreads();
      // back to the face nodes: G(t') w_t' = sum_b sum_t H_b(t, t') Fb(t)
      if (tid < NT) {
        int td[2];
        face_base(d, tid, td);
        double wt = W[td[0]];
        if constexpr (DIM == 3) wt *= W[td[1]];
        for (int k = 0; k < NF; ++k) {
          double a = 0.0;
          for (int b = 0; b < NSUB; ++b) {
            if constexpr (DIM == 2) {
              const double *h = H + (b & 1) * N * N;
#pragma unroll
              for (int t = 0; t < N; ++t) a += h[t * N + td[0]] * s_Fb[b][k][t];
            } else {
              const double *h0 = H + (b & 1) * N * N, *h1 = H + ((b >> 1) & 1) * N * N;
              for (int t1 = 0; t1 < N; ++t1) {
                double c = 0.0;
#pragma unroll
                for (int t0 = 0; t0 < N; ++t0) c += h0[t0 * N + td[0]] * s_Fb[b][k][t0 + N * t1];
                a += h1[t1 * N + td[1]] * c;
              }
            }
          }
          s_F[f][k][tid] = a / wt;
        }
      }
    }
    __syncthreads();  // s_tr and s_Fb are reused by the next face
  }

  // lifting to the nodes, inverse mass, epilogue increments
  if (tid < ND) {
    const int q = tid;
    const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
    double xi[DIM], J[DIM * DIM], Ji[DIM * DIM], wq = 1.0;
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      xi[e] = X[qi[e]];
      wq *= W[qi[e]];
    }
    const double jxw = wq * multilinear_jacobian<DIM>(s_V, xi, J, Ji);
    double dr[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) dr[k] = 0.0;
    for (int f = 0; f < NFACE; ++f) {
      if (kind[f] == 0) continue;
      const int d = f >> 1, s = f & 1;
      int t;
      double wt = 1.0;
      if constexpr (DIM == 2) {
        t = qi[1 - d];
        wt = W[qi[1 - d]];
      } else {
        t = d == 0 ? qi[1] + N * qi[2] : (d == 1 ? qi[0] + N * qi[2] : qi[0] + N * qi[1]);
#pragma unroll
        for (int e = 0; e < DIM; ++e)
          if (e != d) wt *= W[qi[e]];
      }
      const double l = L[s * N + qi[d]] * wt;
#pragma unroll
      for (int k = 0; k < NF; ++k) dr[k] += l * s_F[f][k][t];
    }
    double sc = 1.0;
    if (epi.cell_scale) sc = epi.cell_scale[cell];
    if (sc != 0.0) {
      const double ea1 = epi.a1 * sc, ea2 = epi.a2 * sc, ea3 = epi.a3 * sc;
      const double f2 = epi.o2_from_o1 ? ea1 + ea2 : ea2, f3 = epi.o2_from_o1 ? ea1 + ea3 : ea3;
      const double ij = 1.0 / jxw;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        const size_t idx = (size_t)cell * NF * ND + k * ND + q;
        const double dk = dr[k] * ij;
        epi.o1[idx] += ea1 * dk;
        if (epi.o2) {
          epi.o2[idx] += f2 * dk;
          if (epi.o3) epi.o3[idx] += f3 * dk;
        }
      }
    }
  }
}

}  // namespace exw
