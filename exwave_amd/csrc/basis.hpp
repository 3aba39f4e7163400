// basis.hpp -- 1D Gauss rules and Lagrange tables built on the host.
#pragma once
#include <cmath>
#include <vector>

#include "dg_ader.hpp"
#include "gauss_cx.hpp"

namespace exw {

// The 1D rules are computed by constexpr code so that the same arithmetic
// yields the host tables (build_tables) and the compile-time tables of the
// Cartesian kernel (CartConst in gauss_cx.hpp).
inline void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  cx::gauss01(n, x.data(), w.data());
}

// Barycentric Lagrange basis on `nodes`: values l_i(x) and derivatives l_i'(x).
inline void lagrange_barycentric(const std::vector<double> &nodes, double x, double *val,
                                 double *der) {
  cx::lagrange(int(nodes.size()), nodes.data(), x, val, der);
}


// Host tables of one degree: the operator table (layout TabOff), the QGauss(k+2)
// error table (xe, we, I[q*N+i]) and the CK level tables (layout CkTab).
inline void build_tables(int degree, std::vector<double> &tab, std::vector<double> &etab,
                         std::vector<double> &ckt, CkTab &ck) {
  const int N = degree + 1;
  std::vector<double> x, w;
  gauss_legendre01(N, x, w);
  tab.assign(2 * N * N + 6 * N, 0.0);
  double *D = tab.data(), *DT = D + N * N, *L = DT + N * N, *LW = L + 2 * N, *X = LW + 2 * N,
         *W = X + N;
  for (int q = 0; q < N; ++q) lagrange_barycentric(x, x[q], nullptr, D + q * N);
  for (int i = 0; i < N; ++i)
    for (int m = 0; m < N; ++m) DT[i * N + m] = D[m * N + i] * w[m] / w[i];
  lagrange_barycentric(x, 0.0, L, nullptr);
  lagrange_barycentric(x, 1.0, L + N, nullptr);
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < N; ++i) LW[s * N + i] = L[s * N + i] / w[i];
  for (int i = 0; i < N; ++i) {
    X[i] = x[i];
    W[i] = w[i];
  }
  // error rule QGauss(k+2)
  std::vector<double> xe, we;
  gauss_legendre01(N + 1, xe, we);
  etab.assign(2 * (N + 1) + (N + 1) * N, 0.0);
  for (int i = 0; i <= N; ++i) {
    etab[i] = xe[i];
    etab[N + 1 + i] = we[i];
    lagrange_barycentric(x, xe[i], etab.data() + 2 * (N + 1) + i * N, nullptr);
  }
  // CK level tables
  ckt.clear();
  ck = CkTab{};
  const int nsteps = degree >= 2 ? degree - 1 : 1;
  ck.n_levels = (nsteps - 1) / 2 + 1;
  std::vector<double> xprev, wprev;
  for (int l = 0; l < ck.n_levels; ++l) {
    const int n = N - 2 * l;
    std::vector<double> xl, wl;
    gauss_legendre01(n, xl, wl);
    ck.off_D[l] = (int)ckt.size();
    ckt.resize(ckt.size() + n * n);
    for (int q = 0; q < n; ++q) lagrange_barycentric(xl, xl[q], nullptr, &ckt[ck.off_D[l] + q * n]);
    ck.off_X[l] = (int)ckt.size();
    for (int i = 0; i < n; ++i) ckt.push_back(xl[i]);
    if (l > 0) {
      const int nh = n + 2;
      std::vector<double> vl(n);
      ck.off_P[l] = (int)ckt.size();
      ckt.resize(ckt.size() + n * nh);
      ck.off_V[l] = (int)ckt.size();
      ckt.resize(ckt.size() + nh * n);
      for (int q = 0; q < nh; ++q) {
        lagrange_barycentric(xl, xprev[q], vl.data(), nullptr);
        for (int i = 0; i < n; ++i) {
          ckt[ck.off_P[l] + i * nh + q] = vl[i] * wprev[q] / wl[i];
          ckt[ck.off_V[l] + q * n + i] = vl[i];
        }
      }
    }
    xprev = xl;
    wprev = wl;
  }
  ck.size = (int)ckt.size();
}

}  // namespace exw
