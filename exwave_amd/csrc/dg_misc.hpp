// dg_misc.hpp -- setup and norm kernels: initial projection, L2 errors, energy.
//
// project_initial_field (source/wave_equation_operations.cc:902-935): with the
// Gauss-collocated basis the QGauss(k+1) L2 projection is the value of the
// function at the mapped Gauss points.
// L2 errors (VectorTools::integrate_difference with QGauss(k+2),
// source/explicit_wave.cc:452-496): per cell, interpolate to the k+2 Gauss
// points, compare with the exact modes, reduce per workgroup; a second
// single-group pass sums the partials in a fixed order (deterministic).
#pragma once
#include "dg_common.hpp"

namespace exw {

struct DevModes {
  int n_modes;
  double amp[8];
  int m[8][3];
  double rho, c;
};

template <int DIM>
__device__ __forceinline__ double modes_value(const DevModes &e, const double *x, double t,
                                              int comp) {
  constexpr double PI = 3.14159265358979323846;
  double val = 0.0;
  for (int k = 0; k < e.n_modes; ++k) {
    double mm = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) mm += double(e.m[k][d]) * e.m[k][d];
    mm = sqrt(mm);
    const double omega = e.c * PI * mm;
    double v = comp == DIM ? e.amp[k] * cos(omega * t)
                           : -e.amp[k] * e.m[k][comp] / (e.rho * e.c * mm) * sin(omega * t);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      v *= d != comp ? sin(e.m[k][d] * PI * x[d]) : cos(e.m[k][d] * PI * x[d]);
    val += v;
  }
  return val;
}

template <int DIM, int K>
__global__ void project_kernel(const DevMesh mesh, const double *__restrict__ tab,
                               const DevModes modes, double t, double *__restrict__ dst) {
  using S = Shape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)mesh.n_cells * ND) return;
  const int cell = int(gid / ND), q = int(gid % ND);
  const int qi[3] = {q % N, (q / N) % N, q / (N * N)};
  double xi[3], x[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) xi[d] = tab[TabOff<N>::X + qi[d]];
  multilinear_point<DIM>(mesh.vertices + (size_t)cell * (1 << DIM) * DIM, xi, x);
#pragma unroll
  for (int f = 0; f < NF; ++f) dst[((size_t)cell * NF + f) * ND + q] = modes_value<DIM>(modes, x, t, f);
}

// errtab: xe[N+1], we[N+1], I[(N+1)*N] (I[q*N+i] = l_i(xe_q))
template <int DIM, int K>
__global__ void __launch_bounds__(256)
    l2_error_kernel(const DevMesh mesh, const double *__restrict__ errtab, const DevModes modes,
                    double t, const double *__restrict__ u, double *__restrict__ partial) {
  using S = Shape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF;
  constexpr int NE = N + 1;
  constexpr int NDE = DIM == 2 ? NE * NE : NE * NE * NE;
  __shared__ double red[3][256];
  double ep = 0.0, ev = 0.0, mp = 0.0;
  const double *xe = errtab, *we = errtab + NE, *I = errtab + 2 * NE;
  const size_t total = (size_t)mesh.n_cells * NDE;
  for (size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (size_t)gridDim.x * blockDim.x) {
    const int cell = int(gid / NDE), q = int(gid % NDE);
    const int qi[3] = {q % NE, (q / NE) % NE, q / (NE * NE)};
    double xi[3], x[3], J[9], Ji[9], w = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = xe[qi[d]];
      w *= we[qi[d]];
    }
    const double *V = mesh.vertices + (size_t)cell * (1 << DIM) * DIM;
    multilinear_point<DIM>(V, xi, x);
    const double jxw = fabs(multilinear_jacobian<DIM>(V, xi, J, Ji)) * w;
    const double *uc = u + (size_t)cell * NF * ND;
    double val[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) val[f] = 0.0;
    for (int n = 0; n < ND; ++n) {
      const int ni[3] = {n % N, (n / N) % N, n / (N * N)};
      double s = I[qi[0] * N + ni[0]] * I[qi[1] * N + ni[1]];
      if (DIM == 3) s *= I[qi[2] * N + ni[2]];
#pragma unroll
      for (int f = 0; f < NF; ++f) val[f] += s * uc[f * ND + n];
    }
    const double pex = modes_value<DIM>(modes, x, t, DIM);
    ep += (val[DIM] - pex) * (val[DIM] - pex) * jxw;
    mp += pex * pex * jxw;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double dv = val[d] - modes_value<DIM>(modes, x, t, d);
      ev += dv * dv * jxw;
    }
  }
  red[0][threadIdx.x] = ep;
  red[1][threadIdx.x] = ev;
  red[2][threadIdx.x] = mp;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) partial[blockIdx.x * 3 + k] = red[k][0];
}

template <int DIM, int K>
__global__ void __launch_bounds__(256)
    energy_kernel(const DevMesh mesh, const double *__restrict__ tab,
                  const double *__restrict__ u, double *__restrict__ partial) {
  using S = Shape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF;
  __shared__ double red[256];
  double E = 0.0;
  const size_t total = (size_t)mesh.n_cells * ND;
  for (size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (size_t)gridDim.x * blockDim.x) {
    const int cell = int(gid / ND), q = int(gid % ND);
    const int qi[3] = {q % N, (q / N) % N, q / (N * N)};
    double xi[3], J[9], Ji[9], w = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      xi[d] = tab[TabOff<N>::X + qi[d]];
      w *= tab[TabOff<N>::W + qi[d]];
    }
    const double *V = mesh.vertices + (size_t)cell * (1 << DIM) * DIM;
    const double jxw = fabs(multilinear_jacobian<DIM>(V, xi, J, Ji)) * w;
    const double *uc = u + (size_t)cell * NF * ND;
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) v2 += uc[d * ND + q] * uc[d * ND + q];
    const double p = uc[DIM * ND + q];
    const double rho = mesh.rho[cell], c = mesh.speed[cell];
    E += 0.5 * jxw * (rho * v2 + p * p / (rho * c * c));
  }
  red[threadIdx.x] = E;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// estimate_error (source/wave_equation_operations.cc:871-898): per cell
// err = sqrt(int_K |grad p_h - v_t|^2) at the QGauss(k+1) points (the nodes),
// dudt = apply(sol).  One lane per node; the cell's sum is taken by its first
// lane in node order (deterministic).
template <int DIM, int K>
__global__ void __launch_bounds__((Shape<DIM, K>::BLOCK))
    estimate_error_kernel(const DevMesh mesh, const double *__restrict__ tab,
                          const double *__restrict__ sol, const double *__restrict__ dudt,
                          double *__restrict__ err) {
  using S = Shape<DIM, K>;
  using TO = TabOff<S::N>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF, CPB = S::CPB;
  __shared__ double s_p[CPB][ND];
  __shared__ double s_e[CPB][ND];
  __shared__ double s_tab[TO::SIZE];
  const int tid = threadIdx.x, lc = tid / ND, q = tid - lc * ND;
  const int cell = blockIdx.x * CPB + lc;
  const bool active = lc < CPB && cell < mesh.n_cells;
  for (int i = tid; i < TO::SIZE; i += blockDim.x) s_tab[i] = tab[i];
  if (active) s_p[lc][q] = sol[((size_t)cell * NF + DIM) * ND + q];
  __syncthreads();
  if (active) {
    const int qi[3] = {q % N, (q / N) % N, DIM == 3 ? q / (N * N) : 0};
    double xi[DIM], J[DIM * DIM], Ji[DIM * DIM], w = 1.0, gref[DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      xi[e] = s_tab[TO::X + qi[e]];
      w *= s_tab[TO::W + qi[e]];
      const int se = pow_n<N>(e), q0 = q - qi[e] * se;
      double g = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) g += s_tab[TO::D + qi[e] * N + m] * s_p[lc][q0 + m * se];
      gref[e] = g;
    }
    const double det =
        multilinear_jacobian<DIM>(mesh.vertices + (size_t)cell * (1 << DIM) * DIM, xi, J, Ji);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      double gp = 0.0;
#pragma unroll
      for (int e = 0; e < DIM; ++e) gp += Ji[e * DIM + i] * gref[e];
      const double dl = gp - dudt[((size_t)cell * NF + i) * ND + q];
      s += dl * dl;
    }
    s_e[lc][q] = s * fabs(det) * w;
  }
  __syncthreads();
  if (active && q == 0) {
    double s = 0.0;
    for (int i = 0; i < ND; ++i) s += s_e[lc][i];
    err[cell] = sqrt(s);
  }
}

// u += I b on the pressure of the source cells selected by `sel` (bit s:
// source cell s): the time-integrated point source, one lane per node
static __global__ void __launch_bounds__(256)
    source_add_kernel(double *__restrict__ u, const int *__restrict__ cells,
                      const double *__restrict__ b, int n_src, int nd, int dim, unsigned sel,
                      double I) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = i / nd, q = i - s * nd;
  if (s >= n_src || !((sel >> s) & 1u)) return;
  u[((size_t)cells[s] * (dim + 1) + dim) * nd + q] += I * b[(size_t)s * nd + q];
}

// fixed-order final sum of `n` partial vectors of width `w`
static __global__ void __launch_bounds__(256)
    sum_partials_kernel(const double *__restrict__ partial, int n, int w, double *__restrict__ out) {
  __shared__ double red[256];
  for (int k = 0; k < w; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i * w + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

}  // namespace exw
