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
// contributions live in LDS.  On Cartesian cells each derivative runs in line
// form (one lane per line of the level, then one per node) where that was
// measured faster.
#pragma once
#include <type_traits>
#include <utility>

#include "dg_common.hpp"
#include "gauss_cx.hpp"

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
  // points per direction of level l (at least 1), nodes in DIM dimensions, and
  // the offset of level l >= 1 among the lower levels' contributions
  static constexpr int nl(int l) { return K + 1 - 2 * l > 1 ? K + 1 - 2 * l : 1; }
  static constexpr int nd(int dim, int l) { return dim == 2 ? nl(l) * nl(l) : nl(l) * nl(l) * nl(l); }
  static constexpr int low_off(int dim, int l) {
    int s = 0;
    for (int j = 1; j < l; ++j) s += nd(dim, j);
    return s;
  }
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

// Cells per workgroup of the CK kernel.  Its 18 barrier-separated passes per
// cell wait on every wave of the group, so fewer cells per group pay: in 3D
// one cell per group from 100 nodes on (k=4, 64^3: 1.041 against 1.117 ms
// with two), two below (k=3: 0.315 against 0.323 ms with four); 2D keeps the
// operator's grouping.
template <int DIM, int K>
struct CkGroup {
  static constexpr int ND = Shape<DIM, K>::ND;
  static constexpr int MAXC = DIM == 3 ? (ND >= 100 ? 1 : 2) : 16;
  static constexpr int CPB = Shape<DIM, K>::CPB < MAXC ? Shape<DIM, K>::CPB : MAXC;
  static constexpr int BLOCK = ((CPB * ND + 63) / 64) * 64;
  // 3D k=6 sits one register above 6 waves per SIMD (81 VGPRs): held to 6 it
  // spills 2 and runs 0.633 against 0.735 ms at 32^3 (FP64); elsewhere no hint
  static constexpr int MIN_WAVES = (DIM == 3 && K == 6) ? 6 : 1;
};

template <int DIM, int N>
__device__ __forceinline__ constexpr int ipow_c() {
  return DIM == 2 ? N * N : N * N * N;
}

namespace ck {

using exw::static_for;

// Physical derivative at node qq of a level field of NN points per direction
// (in[f][...] in LDS): outv_d = rho^-1 d_d p, outp = rho c^2 div v (the
// time derivative is minus this).  Cartesian: d_d = (1/h_d) d/dxi_d.
template <int DIM, int NN, int LD, bool CART, typename Number>
__device__ __forceinline__ void derive(const Number (*in)[LD], const Number *Dl, const Number *Xl,
                                       int qq, const Number *ih, Number rho_inv, Number rc2,
                                       const double *V, Number *outv, Number &outp) {
  const int qd[3] = {qq % NN, (qq / NN) % NN, DIM == 3 ? qq / (NN * NN) : 0};
  constexpr int st[3] = {1, NN, NN * NN};
  if constexpr (CART) {
    Number div = Number(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int q0 = qq - qd[d] * st[d];
      const Number *Dr = Dl + qd[d] * NN;
      Number gp = Number(0), gv = Number(0);
#pragma unroll
      for (int m = 0; m < NN; ++m) {
        const Number dm = Dr[m];
        gp += dm * in[DIM][q0 + m * st[d]];
        gv += dm * in[d][q0 + m * st[d]];
      }
      outv[d] = rho_inv * gp * ih[d];
      div += gv * ih[d];
    }
    outp = rc2 * div;
  } else {
    Number gr[DIM + 1][DIM];
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      const int q0 = qq - qd[e] * st[e];
      Number dr[NN];
#pragma unroll
      for (int m = 0; m < NN; ++m) dr[m] = Dl[qd[e] * NN + m];
#pragma unroll
      for (int f = 0; f <= DIM; ++f) {
        Number g = Number(0);
#pragma unroll
        for (int m = 0; m < NN; ++m) g += dr[m] * in[f][q0 + m * st[e]];
        gr[f][e] = g;
      }
    }
    Number xi[DIM], J[DIM * DIM], Ji[DIM * DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xi[d] = Xl[qd[d]];
    multilinear_jacobian<DIM>(V, xi, J, Ji);
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
}

// The same Cartesian derivative in line form, for lane `line` < DIM * NN^(DIM-1)
// of a cell: along its line (direction d) the lane reads p and v_d of `in` into
// registers once and, at the line's nodes, writes rho^-1 (d/dxi_d p) into
// out[d] and d/dxi_d v_d over in[d] (only this lane reads those entries of
// in[d]); derive_nodes then applies 1/h and forms the divergence.  The 1D
// matrix is the compile-time one of NN Gauss points (CartTable, built by the
// same routines as the level's table), and every sum and product runs in
// derive's order: N reads per line instead of N per node and direction.
template <int DIM, int NN, int LD, typename Number>
__device__ __forceinline__ void derive_line(Number (*in)[LD], Number (*out)[LD], int line,
                                            Number rho_inv) {
  constexpr int NT = DIM == 2 ? NN : NN * NN;
  constexpr const CartConst<NN> &T = CartTable<NN>::value;
  const int d = line / NT, t = line - d * NT;
  int base, sd;
  if constexpr (DIM == 2) {
    base = d == 0 ? t * NN : t;
    sd = d == 0 ? 1 : NN;
  } else {
    const int t0 = t % NN, t1 = t / NN;
    base = d == 0 ? t0 * NN + t1 * NN * NN : (d == 1 ? t0 + t1 * NN * NN : t0 + t1 * NN);
    sd = d == 0 ? 1 : (d == 1 ? NN : NN * NN);
  }
  Number pl[NN], vl[NN];
#pragma unroll
  for (int m = 0; m < NN; ++m) {
    pl[m] = in[DIM][base + m * sd];
    vl[m] = in[d][base + m * sd];
  }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    Number gp = Number(0), gv = Number(0);
#pragma unroll
    for (int m = 0; m < NN; ++m) {
      const Number dm = Number(T.D[i * NN + m]);
      gp += dm * pl[m];
      gv += dm * vl[m];
    }
    out[d][base + i * sd] = rho_inv * gp;
    in[d][base + i * sd] = gv;
  }
}

// The node half of the line form (after a barrier): outv_d = (rho^-1 d/dxi_d p)
// / h_d from out[d], and outp = rho c^2 sum_d (d/dxi_d v_d) / h_d from in[d].
template <int DIM, int LD, typename Number>
__device__ __forceinline__ void derive_nodes(const Number (*in)[LD], const Number (*out)[LD],
                                             int qq, const Number *ih, Number rc2, Number *outv,
                                             Number &outp) {
  Number div = Number(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    outv[d] = out[d][qq] * ih[d];
    div += in[d][qq] * ih[d];
  }
  outp = rc2 * div;
}

// whether the lines of a level of NN points fit the cell's lanes
template <int DIM, int NN, int ND>
constexpr bool lines_fit() {
  return DIM * (DIM == 2 ? NN : NN * NN) <= ND;
}

// where the line form was measured faster (CK kernel alone): FP64 everywhere
// (3D k=4: 1.11 against 1.31 ms at 64^3, k=6: 0.66 against 0.93 ms at 32^3,
// 3D k=3 equal); FP32 only in 3D from k=4 (at 3D k=3 and 2D it lost 8-10 %)
template <int DIM, int K, typename Number>
constexpr bool lines_pay() {
  return std::is_same_v<Number, double> || (DIM == 3 && K >= 4);
}

// One tensor pass along direction DIR of the matrix M (NO x NI, row-major) on
// every field: the input has sizes (A0, A1, A2) with A_DIR = NI, the output
// the same with NO in direction DIR.  Lane o computes output entry o.
template <int DIM, int DIR, int NI, int NO, int A0, int A1, int A2, int LD, typename Number,
          int LDI = LD>
__device__ __forceinline__ void pass(const Number *M, const Number (*in)[LDI], Number (*out)[LD],
                                     int o) {
  constexpr int O0 = DIR == 0 ? NO : A0, O1 = DIR == 1 ? NO : A1, O2 = DIR == 2 ? NO : A2;
  constexpr int TOTAL = O0 * O1 * (DIM == 3 ? O2 : 1);
  constexpr int SI[3] = {1, A0, A0 * A1};
  if (o >= TOTAL) return;
  int idx[3] = {o % O0, (o / O0) % O1, DIM == 3 ? o / (O0 * O1) : 0};
  const int r = idx[DIR];
  idx[DIR] = 0;
  const int b = idx[0] * SI[0] + idx[1] * SI[1] + (DIM == 3 ? idx[2] * SI[2] : 0);
  // the lane's matrix row once in registers (the output may alias nothing
  // the compiler can prove, so it would reload the row for every field)
  Number mr[NI];
#pragma unroll
  for (int c = 0; c < NI; ++c) mr[c] = M[r * NI + c];
#pragma unroll
  for (int f = 0; f <= DIM; ++f) {
    Number acc = Number(0);
#pragma unroll
    for (int c = 0; c < NI; ++c) acc += mr[c] * in[f][b + c * SI[DIR]];
    out[f][o] = acc;
  }
}

// The same pass, its output entry o returned in res (all fields) instead of
// stored: the last interpolation pass, whose entry o is lane o's own node.
template <int DIM, int DIR, int NI, int NO, int A0, int A1, int A2, int LDI, typename Number>
__device__ __forceinline__ bool pass_reg(const Number *M, const Number (*in)[LDI], int o,
                                         Number *res) {
  constexpr int O0 = DIR == 0 ? NO : A0, O1 = DIR == 1 ? NO : A1, O2 = DIR == 2 ? NO : A2;
  constexpr int TOTAL = O0 * O1 * (DIM == 3 ? O2 : 1);
  constexpr int SI[3] = {1, A0, A0 * A1};
  if (o >= TOTAL) return false;
  int idx[3] = {o % O0, (o / O0) % O1, DIM == 3 ? o / (O0 * O1) : 0};
  const int r = idx[DIR];
  idx[DIR] = 0;
  const int b = idx[0] * SI[0] + idx[1] * SI[1] + (DIM == 3 ? idx[2] * SI[2] : 0);
  Number mr[NI];
#pragma unroll
  for (int c = 0; c < NI; ++c) mr[c] = M[r * NI + c];
#pragma unroll
  for (int f = 0; f <= DIM; ++f) {
    Number acc = Number(0);
#pragma unroll
    for (int c = 0; c < NI; ++c) acc += mr[c] * in[f][b + c * SI[DIR]];
    res[f] = acc;
  }
  return true;
}

}  // namespace ck

// The cell's factors of the CK derivatives: 1/rho, rho c^2 and (Cartesian)
// 1/h_d; 1 for lanes without a cell.  Reciprocals by the hardware estimate
// and two Newton steps (recip), not IEEE divisions: every lane computes them.
template <int DIM, bool CART, typename Number>
__device__ __forceinline__ void ck_factors(const DevMesh &mesh, int cell, bool active,
                                           Number &rho_inv, Number &rc2, Number (&ih)[DIM]) {
  Number rho = Number(1), cc = Number(1);
#pragma unroll
  for (int d = 0; d < DIM; ++d) ih[d] = Number(1);
  if (active) {
    rho = Number(mesh.rho[cell]);
    cc = Number(mesh.speed[cell]);
    if constexpr (CART) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) ih[d] = recip(Number(mesh.h[(size_t)cell * DIM + d]));
    }
  }
  rho_inv = recip(rho);
  rc2 = rho * cc * cc;
}

// The Cauchy-Kowalevski work of the group's cells once the lane's node value
// u of its cell is known (in registers): the Taylor levels in LDS, the time
// integral over [t1, t2] about te written to tempsrc.  Shared by
// ader_ck_kernel and the fused ADER sweep, so both give the same bits.
template <int DIM, int K, typename Number, bool CART>
__device__ __forceinline__ void ck_cell(
    const DevMesh &mesh, const double *__restrict__ recon, const Number *s_ck, const CkTab &ckt,
    double t2, double t1, double te, Number (*s_a)[Shape<DIM, K>::NF][Shape<DIM, K>::ND],
    Number (*s_b)[Shape<DIM, K>::NF][Shape<DIM, K>::ND],
    Number (*s_c)[Shape<DIM, K>::NF]
                 [CkShape<K>::low_off(DIM, CkShape<K>::N_LEVELS) > 0
                      ? CkShape<K>::low_off(DIM, CkShape<K>::N_LEVELS)
                      : 1],
    const double *Vc, int lc, int lcc, int q, bool lane_ok, bool active, size_t cbase,
    Number rho_inv, Number rc2, const Number (&ih)[DIM], const Number (&u)[Shape<DIM, K>::NF],
    double *__restrict__ tempsrc) {
  using S = Shape<DIM, K>;
  using CS = CkShape<K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF;
  constexpr int NL = CS::N_LEVELS, NSTEPS = CS::N_STEPS;
  constexpr int LOW = CS::low_off(DIM, NL) > 0 ? CS::low_off(DIM, NL) : 1;
  if (lane_ok) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_a[lc][f][q] = u[f];
  }
  __syncthreads();

  // contrib (level 0) and the first derivative
  const double a2 = t2 - te, a1 = t1 - te;
  const Number c1 = Number(0.5 * (a2 * a2 - a1 * a1));
  Number contrib[NF], up[NF];
  if (recon) {
#pragma unroll
    for (int f = 0; f < NF; ++f) up[f] = active ? Number(recon[cbase + f * ND + q]) : Number(0);
  } else {
    Number ov[DIM], op;
    if constexpr (CART && ck::lines_fit<DIM, N, ND>() && ck::lines_pay<DIM, K, Number>()) {
      if (lane_ok && q < DIM * (DIM == 2 ? N : N * N))
        ck::derive_line<DIM, N, ND>(s_a[lc], s_b[lc], q, rho_inv);
      __syncthreads();
      ck::derive_nodes<DIM, ND>(s_a[lcc], s_b[lcc], q, ih, rc2, ov, op);
    } else {
      ck::derive<DIM, N, ND, CART>(s_a[lcc], s_ck + ckt.off_D[0], s_ck + ckt.off_X[0], q, ih,
                                   rho_inv, rc2, Vc, ov, op);
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) up[d] = -ov[d];
    up[DIM] = -op;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) contrib[f] = Number(a2 - a1) * u[f] + c1 * up[f];
  // the node form of the derivative reads other nodes' entries of s_a: wait
  // for every lane before overwriting; the line form's node half (and the
  // reconstruction input) read only the lane's own entries
  if constexpr (!(CART && ck::lines_fit<DIM, N, ND>() && ck::lines_pay<DIM, K, Number>()))
    __syncthreads();
  if (lane_ok) {
#pragma unroll
    for (int f = 0; f < NF; ++f) s_a[lc][f][q] = -up[f];
  }
  __syncthreads();

  // the derivative steps; cur/nxt alternate between s_a and s_b at every step
  // and every projection pass
  double fac = -0.5, p2 = a2 * a2, p1 = a1 * a1;
  int parity = 0;  // 0: current field in s_a
  ck::static_for<NSTEPS>([&](auto S_) {
    constexpr int STEP = decltype(S_)::value;
    constexpr int LEVEL = STEP / 2;
    constexpr int n = CS::nl(LEVEL), nd = CS::nd(DIM, LEVEL);
    fac /= -double(STEP + 3);
    p2 *= a2;
    p1 *= a1;
    const Number fac_t = Number(fac * (p2 - p1));
    Number(*cur)[ND] = parity ? s_b[lcc] : s_a[lcc];
    Number(*nxt)[ND] = parity ? s_a[lcc] : s_b[lcc];
    constexpr bool LINES = CART && ck::lines_fit<DIM, n, ND>() && ck::lines_pay<DIM, K, Number>();
    if constexpr (LINES) {
      if (lane_ok && q < DIM * (DIM == 2 ? n : n * n))
        ck::derive_line<DIM, n, ND>(cur, nxt, q, rho_inv);
      __syncthreads();
    }
    if (q < nd) {
      Number ov[DIM], op;
      if constexpr (LINES)
        ck::derive_nodes<DIM, ND>(cur, nxt, q, ih, rc2, ov, op);
      else
        ck::derive<DIM, n, ND, CART>(cur, s_ck + ckt.off_D[LEVEL], s_ck + ckt.off_X[LEVEL], q,
                                     ih, rho_inv, rc2, Vc, ov, op);
      if (lane_ok) {
        // (the line phase left rho^-1 d/dxi_d p there: this lane's own entries)
#pragma unroll
        for (int d = 0; d < DIM; ++d) nxt[d][q] = ov[d];
        nxt[DIM][q] = op;
      }
      if constexpr (LEVEL == 0) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) contrib[d] += fac_t * ov[d];
        contrib[DIM] += fac_t * op;
      } else if (lane_ok) {
        constexpr bool FIRST = (STEP & 1) == 0;  // the first step of a level sets
        constexpr int OFF = CS::low_off(DIM, LEVEL);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          s_c[lc][d][OFF + q] = (FIRST ? Number(0) : s_c[lc][d][OFF + q]) + fac_t * ov[d];
        s_c[lc][DIM][OFF + q] = (FIRST ? Number(0) : s_c[lc][DIM][OFF + q]) + fac_t * op;
      }
    }
    __syncthreads();
    parity ^= 1;
    // project to the next level after an odd step that is not the last one
    if constexpr ((STEP & 1) == 1 && STEP < K - 2) {
      constexpr int NLO = n - 2;
      const Number *P = s_ck + ckt.off_P[LEVEL + 1];
      ck::static_for<DIM>([&](auto D_) {
        constexpr int DIR = decltype(D_)::value;
        // sizes before this pass: directions < DIR already reduced
        constexpr int A0 = DIR > 0 ? NLO : n, A1 = DIR > 1 ? NLO : n, A2 = n;
        Number(*in)[ND] = parity ? s_b[lcc] : s_a[lcc];
        Number(*out)[ND] = parity ? s_a[lcc] : s_b[lcc];
        if (lane_ok) ck::pass<DIM, DIR, n, NLO, A0, A1, A2, ND>(P, in, out, q);
        __syncthreads();
        parity ^= 1;
      });
    }
  });

  // interpolate the lower levels' contributions back up and add: the first
  // pass of a level reads its contribution where it lies in s_c, the last one
  // hands lane o its own output entry (level 0: the lane's node)
  if constexpr (NL > 1) {
    ck::static_for<NL - 1>([&](auto L_) {
      constexpr int LV = NL - 1 - decltype(L_)::value;  // NL-1 .. 1
      constexpr int NLO = CS::nl(LV), NHI = NLO + 2;
      constexpr int OFF = CS::low_off(DIM, LV);
      const Number *Vt = s_ck + ckt.off_V[LV];
      const Number(*lvl)[LOW] = reinterpret_cast<const Number(*)[LOW]>(&s_c[lcc][0][OFF]);
      bool in_b = true;  // the latest pass output is in s_b (else s_a)
      ck::static_for<DIM>([&](auto D_) {
        constexpr int DIR = decltype(D_)::value;
        constexpr int A0 = DIR > 0 ? NHI : NLO, A1 = DIR > 1 ? NHI : NLO, A2 = NLO;
        if constexpr (DIR < DIM - 1) {
          if constexpr (DIR == 0) {
            if (lane_ok) ck::pass<DIM, DIR, NLO, NHI, A0, A1, A2, ND, Number, LOW>(Vt, lvl, s_b[lcc], q);
            in_b = true;
          } else {
            if (lane_ok)
              ck::pass<DIM, DIR, NLO, NHI, A0, A1, A2, ND>(Vt, in_b ? s_b[lcc] : s_a[lcc],
                                                         in_b ? s_a[lcc] : s_b[lcc], q);
            in_b = !in_b;
          }
          __syncthreads();
        } else {
          Number res[NF];
          bool have = false;
          if (lane_ok)
            have = ck::pass_reg<DIM, DIR, NLO, NHI, A0, A1, A2, ND>(
                Vt, in_b ? s_b[lcc] : s_a[lcc], q, res);
          if (have) {
            if constexpr (LV - 1 == 0) {
#pragma unroll
              for (int f = 0; f < NF; ++f) contrib[f] += res[f];
            } else {
#pragma unroll
              for (int f = 0; f < NF; ++f) s_c[lc][f][CS::low_off(DIM, LV - 1) + q] += res[f];
            }
          }
          // the next level's first pass reads s_c; the last level needs no wait
          if constexpr (LV - 1 > 0) __syncthreads();
        }
      });
    });
  }

  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) tempsrc[cbase + f * ND + q] = double(contrib[f]);
  }
}

template <int DIM, int K, typename Number, bool CART>
__global__ void __launch_bounds__((CkGroup<DIM, K>::BLOCK))
    __attribute__((amdgpu_waves_per_eu(CkGroup<DIM, K>::MIN_WAVES)))
    ader_ck_kernel(const DevMesh mesh, const double *__restrict__ src,
                   const double *__restrict__ recon, const double *__restrict__ cktab,
                   const CkTab ckt, double t2, double t1, double te_all, const CellSel sel,
                   double *__restrict__ tempsrc) {
  using S = Shape<DIM, K>;
  using CS = CkShape<K>;
  constexpr int N = S::N, ND = S::ND, NF = S::NF, CPB = CkGroup<DIM, K>::CPB;
  constexpr int NL = CS::N_LEVELS, NSTEPS = CS::N_STEPS;
  // the lower levels' contributions, one after the other in s_c
  constexpr int LOW = CS::low_off(DIM, NL) > 0 ? CS::low_off(DIM, NL) : 1;

  __shared__ Number s_a[CPB][NF][ND];
  __shared__ Number s_b[CPB][NF][ND];
  __shared__ Number s_c[CPB][NF][LOW];
  __shared__ double s_V[CART ? 1 : CPB][CART ? 1 : (1 << DIM) * DIM];
  __shared__ Number s_ck[CS::TAB_SIZE];

  const int tid = threadIdx.x;
  for (int i = tid; i < CS::TAB_SIZE && i < ckt.size; i += blockDim.x) s_ck[i] = Number(cktab[i]);
  const int lc = tid / ND;
  const int q = tid - lc * ND;
  const bool lane_ok = lc < CPB;
  const int lcc = lane_ok ? lc : 0;  // LDS slot used by spare lanes (they store nothing)
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
  Number rho_inv, rc2, ih[DIM];
  ck_factors<DIM, CART>(mesh, cell, active, rho_inv, rc2, ih);
  const double *Vc = CART ? nullptr : s_V[CART ? 0 : lcc];

  const size_t cbase = (size_t)cell * NF * ND;
  Number u[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) u[f] = Number(0);
  if (active) {
#pragma unroll
    for (int f = 0; f < NF; ++f) u[f] = Number(src[cbase + f * ND + q]);
  }
  ck_cell<DIM, K, Number, CART>(mesh, recon, s_ck, ckt, t2, t1, te, s_a, s_b, s_c, Vc, lc, lcc,
                                q, lane_ok, active, cbase, rho_inv, rc2, ih, u, tempsrc);
}

}  // namespace exw
