// gauss_cx.hpp -- Gauss-Legendre rules and Lagrange tables as constexpr code.
//
// One implementation serves the host tables of every kernel (build_tables)
// and the compile-time 1D matrices of the Cartesian kernel, which the
// compiler folds into the fully unrolled line loops as immediate constants
// (no table in kernel arguments, registers or memory).
#pragma once

namespace exw {
namespace cx {

constexpr double fabs_(double x) { return x < 0 ? -x : x; }

// cos on [0, pi] from its Taylor series (the initial guesses of the roots)
constexpr double cos_(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 40; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}

// n-point Gauss-Legendre rule on [0,1], ascending (Newton iteration on P_n)
constexpr void gauss01(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = cos_(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 200; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 1 ? z : p1, pnm1 = n == 1 ? 1.0 : p0;
      dp = n * (z * pn - pnm1) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (fabs_(dz) < 1e-17) break;
    }
    double p0 = 1.0, p1 = z;
    for (int j = 2; j <= n; ++j) {
      const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    const double pn = n == 1 ? z : p1, pnm1 = n == 1 ? 1.0 : p0;
    dp = n * (z * pn - pnm1) / (z * z - 1.0);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);  // 2/((1-z^2)P'^2) scaled by 1/2
  }
}

// n-point Gauss-Lobatto points on [0,1], ascending: the ends and the roots of
// P'_{n-1} (Newton iteration on P'_{n-1} from the Chebyshev-Lobatto points)
constexpr void gauss_lobatto01(int n, double *x) {
  const int m = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  for (int i = 1; i < n - 1; ++i) {
    double z = -cos_(3.14159265358979323846 * i / m);
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= m; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      const double d1 = m * (z * p1 - p0) / (z * z - 1.0);              // P'_m
      const double d2 = (2.0 * z * d1 - m * (m + 1.0) * p1) / (1.0 - z * z);  // P''_m
      const double dz = d1 / d2;
      z -= dz;
      if (fabs_(dz) < 1e-17) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
}

// barycentric Lagrange basis on nodes[0..n): values l_i(x) and derivatives l_i'(x)
constexpr void lagrange(int n, const double *nodes, double x, double *val, double *der) {
  double bw[16] = {};
  for (int i = 0; i < n; ++i) {
    bw[i] = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != i) bw[i] /= (nodes[i] - nodes[j]);
  }
  int hit = -1;
  for (int i = 0; i < n; ++i)
    if (x == nodes[i]) hit = i;
  if (hit >= 0) {
    // at a node: l_i = delta, l_i'(x_k) = bw_i/bw_k/(x_k-x_i), l_k'(x_k) = -sum_{i!=k}
    for (int i = 0; i < n; ++i) {
      if (val) val[i] = i == hit ? 1.0 : 0.0;
      if (der && i != hit) der[i] = bw[i] / bw[hit] / (x - nodes[i]);
    }
    if (der) {
      double s = 0.0;
      for (int i = 0; i < n; ++i)
        if (i != hit) s += der[i];
      der[hit] = -s;
    }
    return;
  }
  double s = 0.0, sd = 0.0;
  for (int i = 0; i < n; ++i) {
    s += bw[i] / (x - nodes[i]);
    sd += bw[i] / ((x - nodes[i]) * (x - nodes[i]));
  }
  for (int i = 0; i < n; ++i) {
    const double li = bw[i] / (x - nodes[i]) / s;
    if (val) val[i] = li;
    if (der) der[i] = li * (-1.0 / (x - nodes[i]) + sd / s);
  }
}

}  // namespace cx

// The 1D matrices of N Gauss points (meaning as TabOff): D[q*N+i] = l_i'(x_q),
// DT[i*N+m] = D[m*N+i] w_m / w_i, L[s*N+i] = l_i(s), LW[s*N+i] = l_i(s) / w_i.
template <int N>
struct CartConst {
  double x[N], w[N], D[N * N], DT[N * N], L[2 * N], LW[2 * N];
  constexpr CartConst() : x(), w(), D(), DT(), L(), LW() {
    cx::gauss01(N, x, w);
    for (int q = 0; q < N; ++q) cx::lagrange(N, x, x[q], nullptr, D + q * N);
    for (int i = 0; i < N; ++i)
      for (int m = 0; m < N; ++m) DT[i * N + m] = D[m * N + i] * w[m] / w[i];
    cx::lagrange(N, x, 0.0, L, nullptr);
    cx::lagrange(N, x, 1.0, L + N, nullptr);
    for (int s = 0; s < 2; ++s)
      for (int i = 0; i < N; ++i) LW[s * N + i] = L[s * N + i] / w[i];
  }
};

template <int N>
struct CartTable {
  static constexpr CartConst<N> value{};
};

}  // namespace exw
