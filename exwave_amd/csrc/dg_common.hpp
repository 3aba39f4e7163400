// dg_common.hpp -- shared device-side definitions of the MI355X ExWave kernels.
//
// Polynomial space: Q_k per field on each multilinear cell, Lagrange basis
// collocated at the k+1 Gauss points per direction; all integrals use
// QGauss(k+1) like the reference (source/wave_equation_operations.cc:214-215),
// so the QGauss(k+1) mass matrix is diagonal (w_q * det J_q) and the inverse
// mass (source/wave_equation_operations.cc:633-661) becomes a pointwise
// scaling folded into the operator epilogue.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace exw {

constexpr int MAX_N = 14;  // points per direction (degree <= 12, +1 for the error rule)

template <int DIM, int K>
struct Shape {
  static constexpr int N = K + 1;                       // Gauss points per direction
  static constexpr int ND = DIM == 2 ? N * N : N * N * N;  // nodes per cell
  static constexpr int NT = DIM == 2 ? N : N * N;          // points per face
  static constexpr int NF = DIM + 1;                       // fields v_0..v_{dim-1}, p
  static constexpr int NFACE = 2 * DIM;
  // cells per workgroup: fill ~256 lanes; one cell per group beyond that
  static constexpr int CPB = ND >= 256 ? 1 : 256 / ND;
  static constexpr int BLOCK = ((CPB * ND + 63) / 64) * 64;
};

// Offsets (in doubles) into the per-operator basis table built on the host.
template <int N>
struct TabOff {
  static constexpr int D = 0;                 // D[q*N+i]  = l_i'(x_q)
  static constexpr int DT = D + N * N;        // DT[i*N+m] = D[m*N+i] * w_m / w_i
  static constexpr int L = DT + N * N;        // L[s*N+i]  = l_i(s), s = 0, 1
  static constexpr int LW = L + 2 * N;        // LW[s*N+i] = l_i(s) / w_i
  static constexpr int X = LW + 2 * N;        // Gauss points on [0,1]
  static constexpr int W = X + N;             // Gauss weights
  static constexpr int SIZE = W + N;
};

// Device view of the mesh and materials.
struct DevMesh {
  int n_cells;
  const int *__restrict__ neighbor;      // [n_cells][2*dim]
  const double *__restrict__ rho;        // [n_cells]
  const double *__restrict__ speed;      // [n_cells]
  const double *__restrict__ h;          // Cartesian: [n_cells][dim] cell extents
  const double *__restrict__ vertices;   // general: [n_cells][2^dim][dim]
  // Cartesian: per cell the face flux maps [2*dim][8] and the cell-term
  // factors [dim][2] (cart_coef_kernel); NULL where not set up
  const double *__restrict__ coef = nullptr;
  // optional: coef holds a table of distinct rows and cell c uses row
  // coef_class[c] (box meshes share a few dozen rows)
  const int *__restrict__ coef_class = nullptr;
};

// o1 = b + a1*s*k, o2 = (o1 or c) + a2*s*k, o3 = (o1 or c) + a3*s*k (if o3 is
// not NULL; needs o2) with s = cell_scale[cell] (1 if cell_scale is NULL);
// cells with s == 0 are inactive and skipped entirely.  b and c are read
// before any output of the same entry is written, so outputs may alias them.
// With cell_list the launch covers the cells cell_list[0..n_list) only (the
// active cells of a multi-instance pool of the serving loop; Cartesian kernel).
struct Epilogue {
  const double *b;
  double *o1;
  double a1;
  const double *c;
  double *o2;
  double a2;
  int o2_from_o1;
  const double *cell_scale;
  const int *cell_list;
  int n_list;
  double *o3;
  double a3;
  // set by the low-storage RK stage entry point: selects the pipelined
  // Cartesian kernel for its single-output last stage as for the others
  int rk_stage = 0;
};

// Optional cell selection of the local-time-stepping sweeps: the kernel runs
// over cells list[0..n) instead of 0..n_cells-1, and the expansion time of a
// cell is te[cluster[cell]] (its cluster's time level).
constexpr int MAX_CLUSTERS = 16;
struct CellSel {
  const int *list;
  int n;
  const int *cluster;
  double te[MAX_CLUSTERS];
};

// MI355X dispatches workgroup b to XCD b % 8.  Give every XCD one contiguous
// range of logical blocks (a bijection for any grid size), so that cells that
// are neighbours in the numbering share one L2.
constexpr int N_XCD = 8;

// calls f(integral_constant<int, I>) for I = 0 .. COUNT-1, at compile time
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int COUNT, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, COUNT>{});
}
__device__ __forceinline__ int xcd_remap(int b, int nblk) {
  const int q = nblk / N_XCD, r = nblk % N_XCD;
  const int x = b % N_XCD, i = b / N_XCD;
  return x * q + (x < r ? x : r) + i;
}

template <int N>
__device__ __forceinline__ constexpr int pow_n(int e) {
  return e == 0 ? 1 : (e == 1 ? N : N * N);
}

// J[i*DIM+e] = dx_i/dxi_e of the multilinear map with corners V (LDS or
// registers), and Jinv[e*DIM+i] = dxi_e/dx_i; returns det J.
template <int DIM, typename T, typename VT>
__device__ __forceinline__ T multilinear_jacobian(const VT *V, const T *xi, T *J, T *Jinv) {
#pragma unroll
  for (int k = 0; k < DIM * DIM; ++k) J[k] = T(0);
#pragma unroll
  for (int v = 0; v < (1 << DIM); ++v) {
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      T s = ((v >> e) & 1) ? T(1) : T(-1);
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        if (d != e) s *= ((v >> d) & 1) ? xi[d] : T(1) - xi[d];
#pragma unroll
      for (int i = 0; i < DIM; ++i) J[i * DIM + e] += s * T(V[v * DIM + i]);
    }
  }
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

template <int DIM, typename T, typename VT>
__device__ __forceinline__ void multilinear_point(const VT *V, const T *xi, T *x) {
#pragma unroll
  for (int i = 0; i < DIM; ++i) x[i] = T(0);
#pragma unroll
  for (int v = 0; v < (1 << DIM); ++v) {
    T s = T(1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) s *= ((v >> d) & 1) ? xi[d] : T(1) - xi[d];
#pragma unroll
    for (int i = 0; i < DIM; ++i) x[i] += s * T(V[v * DIM + i]);
  }
}

// Numerical flux at one face point, seen from the cell that owns the face
// (its outward normal nrm).  Restates evaluate_inner_face /
// evaluate_boundary_face (source/wave_equation_operations.cc:480-551, 569-629):
//   tau = 1/(rho c); inner: lambda = (v+.n - v-.n + tau+ p+ + tau- p-)/(tau+ + tau-)
//   boundary id 1: lambda = v+.n/tau + p+ ; 2: lambda = 0 ; 3: absorbing
//   test values: v: (p+ - lambda)/rho n ; p: rho c^2 (-v+.n + tau (lambda - p+))
// nv_m is the neighbour's v.n with the SAME normal (v-.n).
//
// lambda is linear in the traces: lambda = l[0] v+.n + l[1] p+ + l[2] v-.n + l[3] p-.
//
// Reciprocals: the hardware reciprocal refined by two Newton steps (within
// an ulp or two of the correctly rounded quotient, at half the instructions
// of an IEEE division).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ float recip(float x) { return 1.0f / x; }

template <typename T>
__device__ __forceinline__ void flux_lambda(T rho_p, T c_p, T rho_m, T c_m, int bid, T *l) {
  const T tau_p = recip(c_p * rho_p);
  l[0] = l[1] = l[2] = l[3] = T(0);
  if (bid == 0) {
    const T tau_m = recip(c_m * rho_m);
    const T inv = recip(tau_p + tau_m);
    l[0] = inv;
    l[1] = inv * tau_p;
    l[2] = -inv;
    l[3] = inv * tau_m;
  } else if (bid == 1) {
    l[0] = c_p * rho_p;  // 1/tau
    l[1] = T(1);
  } else if (bid == 3) {
    // first-order absorbing condition, the reference's expression with the
    // exterior impedance equal to the interior one
    const T tau_e = recip(c_p * rho_p);
    const T inv = recip(tau_p + tau_e);
    l[0] = inv;
    l[1] = tau_p * inv;
  }  // bid 2: lambda = 0
}

// Boundary code of a non-conforming (2:1 hanging) face (neighbour entry
// -NC_BID): the operator sweeps scale its contribution by zero (no branch in
// the flux, whose code the hot kernel is sensitive to); nc_face_kernel adds
// its subface integrals.
constexpr int NC_BID = 1000;

template <typename T>
__device__ __forceinline__ void flux(T nvp, T pp, T nv_m, T pm, T rho_p, T c_p, T rho_m, T c_m,
                                     int bid, T &val_pn, T &val_p) {
  T l[4];
  flux_lambda<T>(rho_p, c_p, rho_m, c_m, bid, l);
  const T tau_p = recip(c_p * rho_p);
  const T lambda = l[0] * nvp + l[1] * pp + l[2] * nv_m + l[3] * pm;
  val_pn = (pp - lambda) * recip(rho_p);  // multiplies n_d for the velocity test values
  val_p = c_p * c_p * rho_p * (-nvp + tau_p * (lambda - pp));
}

// The same flux as a linear map of (own v.n, own p, neighbour v.n, neighbour
// p) -> (v test value, p test value): a[j], b[j] multiply input j.
template <typename T>
__device__ __forceinline__ void flux_matrix(T rho_p, T c_p, T rho_m, T c_m, int bid, T *a, T *b) {
  T l[4];
  flux_lambda<T>(rho_p, c_p, rho_m, c_m, bid, l);
  const T irho = recip(rho_p), tau_p = recip(c_p * rho_p), z = c_p * c_p * rho_p;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const T epp = j == 1 ? T(1) : T(0), env = j == 0 ? T(1) : T(0);
    a[j] = (epp - l[j]) * irho;
    b[j] = z * (-env + tau_p * (l[j] - epp));
  }
}

}  // namespace exw
