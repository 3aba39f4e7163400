// dg_post.hpp -- post-processed pressure and its error.
//
// compute_post_pressure (source/wave_equation_operations.cc:786-848): per cell
// the local Neumann problem (grad phi, grad p*) = (grad phi, -rho v_t) on
// Q_{k+1} with the nodal basis of the k+2 Gauss-Lobatto points
// (fe_post_disp, source/explicit_wave.cc:165) and QGauss(k+2), v_t the
// velocity part of apply(solution); then the constant is set so that the cell
// average of p* equals that of p_h.
//
// Layout: with P = k+2 the Gauss-Lobatto nodes and the QGauss(P) points form
// grids of the same size P^dim, so thread t of a cell owns node t AND point t.
// Every CG vector (u, r, p, Ap) lives in registers, one value per thread; the
// symmetric geometry factor G = J^-1 J^-T det w of point t too.  The
// sum-factorised passes of the Laplacian exchange values through LDS:
//   nodes -> points: P^dim values, one pass per direction with the GL
//   interpolation matrix; gradient: collocation derivative of the QGauss(P)
//   points; flux G grad; the transposes back.
//
// Solver: the reference fixes the first unknown and runs CG to |r|/|r0| <
// 1e-11 (include/elementwise_cg.templates.h:66-145).  The result after the
// average adjustment is the same for every Neumann solution, so this kernel
// runs CG on the zero-sum subspace instead (the matrix's null space is the
// constant vector): condition numbers 20-500x smaller than with the fixed
// unknown, so the same iteration count reaches 1e-13.  The stopping rule is
// per cell (abs, rel, cap), the division guard that of the reference
// (adjust_division_by_zero, :44-61).
#pragma once
#include "dg_common.hpp"
#include "dg_misc.hpp"

namespace exw {

template <int DIM, int K>
struct PostShape {
  static constexpr int N = K + 1, P = K + 2, E = K + 3;
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;
  static constexpr int NP = DIM == 2 ? P * P : P * P * P;
  static constexpr int NE = DIM == 2 ? E * E : E * E * E;
  static constexpr int LANES = (NP + 63) / 64 * 64;  // whole waves per cell
  static constexpr int WAVES = LANES / 64;
  static constexpr int CPB = LANES >= 256 ? 1 : 256 / LANES;
  static constexpr int BLOCK = LANES * CPB;
  static_assert(BLOCK <= 1024, "cell does not fit a workgroup");
};

// Table (built on the host, per degree): offsets into one array
template <int K>
struct PostTab {
  static constexpr int N = K + 1, P = K + 2, E = K + 3;
  static constexpr int XQ = 0;             // QGauss(P) points on [0,1]
  static constexpr int WQ = XQ + P;        // and weights
  static constexpr int PV = WQ + P;        // PV[q*P+i] = l^GL_i(xq_q)
  static constexpr int DQ = PV + P * P;    // DQ[q*P+j] = l^Q_j'(xq_q) (collocation)
  static constexpr int IQ = DQ + P * P;    // IQ[q*N+i] = l^G(N)_i(xq_q)
  static constexpr int XE = IQ + P * N;    // QGauss(E) points (error rule)
  static constexpr int WE = XE + E;
  static constexpr int IE = WE + E;        // IE[q*P+i] = l^GL_i(xe_q)
  static constexpr int SIZE = IE + E * P;
  static constexpr int LDS = XE;           // what the solver kernel stages
};

struct PostSolver {
  double abs_tol, rel_tol;
  int max_iter;  // stop after max_iter + 1 iterations at most
};

namespace post {

// One pass along direction D of a square (P x P) matrix on a P^dim tensor at
// the thread's point ti: out = sum_j M(ti_D, j) in(ti with ti_D = j); TR uses
// M^T.
template <int DIM, int P, int D, bool TR>
__device__ __forceinline__ double sq(const double *M, const double *in, const int *ti) {
  constexpr int st = D == 0 ? 1 : (D == 1 ? P : P * P);
  const int base = ti[0] + P * ti[1] + (DIM == 3 ? P * P * ti[2] : 0) - ti[D] * st;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < P; ++j) s += (TR ? M[j * P + ti[D]] : M[ti[D] * P + j]) * in[base + j * st];
  return s;
}

// Rectangular pass D of an NO x NI matrix on the tensor whose directions < D
// already have NO points and directions >= D still NI: output entry o (of
// NO^(D+1) NI^(DIM-D-1)) or 0 when o is out of range.
template <int DIM, int NI, int NO, int D>
__device__ __forceinline__ double rect(const double *M, const double *in, int o) {
  constexpr int LO = D == 0 ? 1 : (D == 1 ? NO : NO * NO);
  constexpr int COUNT = LO * NO * (D == DIM - 1 ? 1 : (DIM - D - 1 == 1 ? NI : NI * NI));
  if (o >= COUNT) return 0.0;
  const int lower = o % LO, id = (o / LO) % NO, upper = o / (LO * NO);
  const double *src = in + lower + LO * NI * upper;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NI; ++j) s += M[id * NI + j] * src[LO * j];
  return s;
}

// Sum over the lanes of one cell (WAVES waves), every lane gets the total;
// fixed order (butterfly, then the waves in order): deterministic.
template <int WAVES, int CPB, int NV>
__device__ __forceinline__ void reduce(double *v, double (*red)[WAVES][3], int lc, int t) {
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if constexpr (WAVES > 1) {
    if ((t & 63) == 0)
#pragma unroll
      for (int k = 0; k < NV; ++k) red[lc][t >> 6][k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < WAVES; ++w) s += red[lc][w][k];
      v[k] = s;
    }
  }
}

}  // namespace post

// grid: ceil(n_cells / CPB) groups of BLOCK threads; dudt = apply(sol)
template <int DIM, int K>
__global__ void __launch_bounds__((PostShape<DIM, K>::BLOCK))
    post_pressure_kernel(const DevMesh mesh, const double *__restrict__ tab,
                         const double *__restrict__ sol, const double *__restrict__ dudt,
                         double *__restrict__ post, int *__restrict__ iterations,
                         const PostSolver cfg) {
  using S = PostShape<DIM, K>;
  using T = PostTab<K>;
  constexpr int N = S::N, P = S::P, NP = S::NP, ND = S::ND, NF = DIM + 1, LANES = S::LANES;
  constexpr int WAVES = S::WAVES, CPB = S::CPB, NG = DIM * (DIM + 1) / 2;
  __shared__ double s_tab[T::LDS];
  __shared__ double s_buf[CPB][DIM + 1][NP];
  __shared__ double s_red[2][CPB][WAVES][3];
  const int tid = threadIdx.x, lc = tid / LANES, t = tid % LANES;
  const int cell_raw = blockIdx.x * CPB + lc;
  const bool cell_ok = cell_raw < mesh.n_cells;
  const int cell = cell_ok ? cell_raw : mesh.n_cells - 1;  // spare groups redo the last cell
  const bool act = t < NP;
  const int tc = act ? t : NP - 1;
  const int ti[3] = {tc % P, (tc / P) % P, DIM == 3 ? tc / (P * P) : 0};
  for (int i = tid; i < T::LDS; i += S::BLOCK) s_tab[i] = tab[i];
  const double *xq = s_tab + T::XQ, *wq = s_tab + T::WQ, *PV = s_tab + T::PV,
               *DQ = s_tab + T::DQ, *IQ = s_tab + T::IQ;
  double *B[DIM + 1];
#pragma unroll
  for (int b = 0; b <= DIM; ++b) B[b] = s_buf[lc][b];
  int rb = 0;  // alternating reduction buffer
  auto reduce1 = [&](double x) {
    double v[1] = {x};
    post::reduce<WAVES, CPB, 1>(v, s_red[rb], lc, t);
    rb ^= 1;
    return v[0];
  };

  __syncthreads();

  // geometry at QGauss(P) point ti
  double G[NG], Ji[DIM * DIM], jxw;
  {
    double xi[3], J[9], w = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = xq[ti[d]];
      w *= wq[ti[d]];
    }
    const double det =
        multilinear_jacobian<DIM>(mesh.vertices + (size_t)cell * (1 << DIM) * DIM, xi, J, Ji);
    jxw = det * w;
    int g = 0;
#pragma unroll
    for (int e = 0; e < DIM; ++e)
#pragma unroll
      for (int e2 = e; e2 < DIM; ++e2) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) s += Ji[e * DIM + i] * Ji[e2 * DIM + i];
        G[g++] = s * jxw;
      }
  }

  // v_t (velocity of dudt) and p_h at the QGauss(P) points: N^dim -> P^dim
  double vt[DIM], ph = 0.0;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const double *src = (f < DIM ? dudt : sol) + ((size_t)cell * NF + f) * ND;
    __syncthreads();  // previous readers of B[0], B[1] are done
    if (t < ND) B[0][t] = src[t];
    __syncthreads();
    double val;
    if constexpr (DIM == 2) {
      const double a = post::rect<2, N, P, 0>(IQ, B[0], t);
      if (t < P * N) B[1][t] = a;
      __syncthreads();
      val = post::rect<2, N, P, 1>(IQ, B[1], tc);
    } else {
      const double a = post::rect<3, N, P, 0>(IQ, B[0], t);
      if (t < P * N * N) B[1][t] = a;
      __syncthreads();
      const double b = post::rect<3, N, P, 1>(IQ, B[1], t);
      if (t < P * P * N) B[2][t] = b;
      __syncthreads();
      val = post::rect<3, N, P, 2>(IQ, B[2], tc);
    }
    if (f < DIM)
      vt[f] = val;
    else
      ph = val;
  }
  __syncthreads();

  // The flux part of the operator: given F_e (the reference-space flux at the
  // thread's point, already in B[0..DIM-1]), sum_e grad_e^T F_e back to the
  // nodes, via B[DIM] and the transposed interpolation passes.
  auto grad_transpose = [&]() -> double {
    __syncthreads();
    double y = 0.0;
    y += post::sq<DIM, P, 0, true>(DQ, B[0], ti);
    y += post::sq<DIM, P, 1, true>(DQ, B[1], ti);
    if constexpr (DIM == 3) y += post::sq<DIM, P, 2, true>(DQ, B[2], ti);
    if (act) B[DIM][t] = y;
    __syncthreads();
    if constexpr (DIM == 2) {
      const double a = post::sq<2, P, 1, true>(PV, B[2], ti);
      if (act) B[0][t] = a;
      __syncthreads();
      return post::sq<2, P, 0, true>(PV, B[0], ti);
    } else {
      const double a = post::sq<3, P, 2, true>(PV, B[3], ti);
      if (act) B[0][t] = a;
      __syncthreads();
      const double b = post::sq<3, P, 1, true>(PV, B[0], ti);
      if (act) B[1][t] = b;
      __syncthreads();
      return post::sq<3, P, 0, true>(PV, B[1], ti);
    }
  };
  // nodal values -> values at the QGauss(P) points (the thread's own point).
  // 2D stages in B[1] (B[0] may still be read by the last transposed pass),
  // 3D in B[0].
  auto to_points = [&](double x) -> double {
    constexpr int st = DIM == 2 ? 1 : 0;
    if (act) B[st][t] = x;
    __syncthreads();
    if constexpr (DIM == 2) {
      const double a = post::sq<2, P, 0, false>(PV, B[1], ti);
      if (act) B[0][t] = a;
      __syncthreads();
      const double b = post::sq<2, P, 1, false>(PV, B[0], ti);
      if (act) B[2][t] = b;
      __syncthreads();
      return b;
    } else {
      const double a = post::sq<3, P, 0, false>(PV, B[0], ti);
      if (act) B[1][t] = a;
      __syncthreads();
      const double b = post::sq<3, P, 1, false>(PV, B[1], ti);
      if (act) B[2][t] = b;
      __syncthreads();
      const double c = post::sq<3, P, 2, false>(PV, B[2], ti);
      if (act) B[3][t] = c;
      __syncthreads();
      return c;
    }
  };
  // Neumann Laplacian of the nodal vector x at node t
  auto laplace = [&](double x) -> double {
    to_points(x);  // values at the points now in B[DIM]
    double g[DIM];
    g[0] = post::sq<DIM, P, 0, false>(DQ, B[DIM], ti);
    g[1] = post::sq<DIM, P, 1, false>(DQ, B[DIM], ti);
    if constexpr (DIM == 3) g[2] = post::sq<DIM, P, 2, false>(DQ, B[DIM], ti);
    double f[DIM];
    if constexpr (DIM == 2) {
      f[0] = G[0] * g[0] + G[1] * g[1];
      f[1] = G[1] * g[0] + G[2] * g[1];
    } else {
      f[0] = G[0] * g[0] + G[1] * g[1] + G[2] * g[2];
      f[1] = G[1] * g[0] + G[3] * g[1] + G[4] * g[2];
      f[2] = G[2] * g[0] + G[4] * g[1] + G[5] * g[2];
    }
    __syncthreads();  // everyone has read B[DIM]; B[0..DIM-1] are free
#pragma unroll
    for (int e = 0; e < DIM; ++e)
      if (act) B[e][t] = f[e];
    return grad_transpose();
  };

  // right-hand side (grad phi, -rho v_t): reference-space flux Ji (-rho v_t) jxw
  const double rho = mesh.rho[cell];
  {
    double f[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) s += Ji[e * DIM + i] * (-rho * vt[i]);
      f[e] = s * jxw;
    }
#pragma unroll
    for (int e = 0; e < DIM; ++e)
      if (act) B[e][t] = f[e];
  }
  double rhs = grad_transpose();
  if (!act) rhs = 0.0;
  const double mean = reduce1(rhs) / NP;  // project onto the zero-sum subspace
  if (act) rhs -= mean;

  // CG on the zero-sum subspace
  double u = 0.0, r = rhs, p = rhs;
  double rr = reduce1(r * r);
  const double nr0 = sqrt(rr);
  const double guard = fmax(1e-5 * cfg.abs_tol * cfg.abs_tol, 1e4 * 2.2250738585072014e-308);
  bool done = false;
  int it = 0;
  for (;;) {
    double Ap = laplace(p);
    if (!act) Ap = 0.0;
    double pAp = reduce1(p * Ap);
    if (!done) {
      if (pAp < guard) pAp = 1.0;
      const double alpha = rr / pAp;
      u += alpha * p;
      r -= alpha * Ap;
    }
    double v2[2] = {r, r * r};
    post::reduce<WAVES, CPB, 2>(v2, s_red[rb], lc, t);
    rb ^= 1;
    if (!done) {
      const double m = v2[0] / NP;
      if (act) r -= m;
      const double rn = fmax(v2[1] - v2[0] * m, 0.0);
      ++it;
      if (sqrt(rn) < cfg.abs_tol || sqrt(rn) < cfg.rel_tol * nr0 || it > cfg.max_iter) {
        done = true;
      } else {
        const double beta = rn / rr;
        rr = rn;
        p = act ? r + beta * p : 0.0;
      }
    }
    if (__syncthreads_and((int)done)) break;  // every cell of the group converged
  }

  // constant offset from the cell averages (:821-841)
  const double uq = to_points(u);
  double v3[3] = {act ? jxw : 0.0, act ? jxw * ph : 0.0, act ? jxw * uq : 0.0};
  post::reduce<WAVES, CPB, 3>(v3, s_red[rb], lc, t);
  const double adjust = (v3[1] - v3[2]) / v3[0];
  if (cell_ok && act) post[(size_t)cell * NP + t] = u + adjust;
  if (cell_ok && t == 0 && iterations) iterations[cell] = it;
}

// ||p* - p_exact||^2 partial sums with QGauss(k+3) (source/explicit_wave.cc:477-484)
template <int DIM, int K>
__global__ void __launch_bounds__(256)
    post_error_kernel(const DevMesh mesh, const double *__restrict__ tab, const DevModes modes,
                      double t, const double *__restrict__ post, double *__restrict__ partial) {
  using S = PostShape<DIM, K>;
  using T = PostTab<K>;
  constexpr int P = S::P, E = S::E, NP = S::NP, NE = S::NE;
  __shared__ double red[256];
  __shared__ double s_ie[E * P];
  for (int i = threadIdx.x; i < E * P; i += 256) s_ie[i] = tab[T::IE + i];
  __syncthreads();
  const double *xe = tab + T::XE, *we = tab + T::WE;
  double err = 0.0;
  const size_t total = (size_t)mesh.n_cells * NE;
  for (size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (size_t)gridDim.x * blockDim.x) {
    const int cell = int(gid / NE), q = int(gid % NE);
    const int qi[3] = {q % E, (q / E) % E, q / (E * E)};
    double xi[3], x[3], J[9], Ji[9], w = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = xe[qi[d]];
      w *= we[qi[d]];
    }
    const double *V = mesh.vertices + (size_t)cell * (1 << DIM) * DIM;
    multilinear_point<DIM>(V, xi, x);
    const double jxw = fabs(multilinear_jacobian<DIM>(V, xi, J, Ji)) * w;
    const double *pc = post + (size_t)cell * NP;
    double val = 0.0;
    for (int n = 0; n < NP; ++n) {
      const int ni[3] = {n % P, (n / P) % P, n / (P * P)};
      double s = s_ie[qi[0] * P + ni[0]] * s_ie[qi[1] * P + ni[1]];
      if (DIM == 3) s *= s_ie[qi[2] * P + ni[2]];
      val += s * pc[n];
    }
    const double d = val - modes_value<DIM>(modes, x, t, DIM);
    err += d * d * jxw;
  }
  red[threadIdx.x] = err;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace exw
