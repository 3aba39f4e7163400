// dg_operator_cartg.hpp -- the Cartesian operator + RK epilogue for large
// cells (3D k >= 6) with the neighbours' face traces gathered from memory.
//
// dg_operator_cart_kernel stages the 2*dim neighbours' (v_n, p) nodes in LDS
// so that a face trace is a line sum there: 12 fields of the cell's size, 33
// of its 44 KB of LDS at 3D k=6, which leaves 3 cells per CU.  Here the lane
// that owns line t of direction d in the line phase sums the neighbour traces
// of faces 2d and 2d+1 at point t itself, straight from the neighbours' nodes
// in memory (L2 hits: the neighbours are cells of the same tile), and keeps
// the four sums in registers.  The LDS holds the cell's fields and the lines'
// dp/dt shares, 7 fields (19.5 KB at k=6: 8 cells per CU by LDS).
//
// Lanes: wave w works on direction d = w / WPD in the line phase (WPD waves
// of 64 lanes per direction, line t = 64 (w - d WPD) + lane), so the stride
// of a line and the neighbour across each of its faces are wave-uniform and
// every gathered load is a scalar base plus one per-lane offset; x lines are
// contiguous and read in 16-byte loads.  In the node phases each lane takes
// the nodes lane + j*BLOCK, j < NPL.  The cell's own fields reach the LDS by
// LDS-DMA (FP64, 16-byte aligned vector), so that they hold no registers
// while the traces are gathered.
//
// The arithmetic is dg_operator_cart_kernel's, expression for expression:
// every trace is the same chain of FMAs over the same values, the line phase
// evaluates the same chains per value, the pressure shares are summed in the
// same order, and the epilogue is the same; both kernels give bit-identical
// results (tested).  At 3D k = 4, 5 it was measured slower than the staged
// kernel (more lanes idle in the line phase, 3 waves for 125 / 216 nodes), so
// it is compiled from k = 6 on.
#pragma once
#include "dg_operator_pipe.hpp"

namespace exw {

// x[0..N) = p[0..N) of an 8-byte aligned line, in 16-byte loads where they
// fit (the hardware and the compiler's target allow unaligned dword-multiple
// global accesses)
template <int N>
__device__ __forceinline__ void line_load(const double *p, double (&x)[N]) {
  typedef double dbl2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int i = 0; i + 1 < N; i += 2) {
    dbl2 v;
    __builtin_memcpy(&v, p + i, 16);
    x[i] = v.x;
    x[i + 1] = v.y;
  }
  if constexpr (N % 2) x[N - 1] = p[N - 1];
}

#ifndef EXW_CARTG_MIN_K
#define EXW_CARTG_MIN_K 6
#endif

template <int DIM, int K>
struct CartGShape {
  using C = CartShape<DIM, K>;
  static constexpr int N = C::N, ND = C::ND, NT = C::NT, NF = C::NF, NFACE = 2 * DIM;
  static constexpr int WPD = (NT + 63) / 64;  // waves per direction
  static constexpr int BLOCK = DIM * WPD * 64;
  static constexpr int NPL = (ND + BLOCK - 1) / BLOCK;  // nodes per lane
  static constexpr bool ENABLED = DIM == 3 && K >= EXW_CARTG_MIN_K && BLOCK <= 1024;
};

// One cell per group.  The same epilogue as dg_operator_cart_kernel (b, c,
// o1..o3, cell scales, cell lists).

template <int DIM, int K, typename Number>
__global__ void __launch_bounds__((CartGShape<DIM, K>::BLOCK))
    dg_operator_cartg_kernel(const DevMesh mesh, const double *__restrict__ src,
                             const Epilogue epi) {
  using S = CartGShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT, NF = S::NF, NFACE = S::NFACE;
  constexpr int BLOCK = S::BLOCK, NPL = S::NPL, WPD = S::WPD;
  constexpr const CartConst<N> &T = CartTable<N>::value;
  // the cell's fields; the line phase writes each line's dv_d/dt over its v_d
  __shared__ __align__(16) Number s_u[NF][ND];
  // the lines' dp/dt shares, by direction
  __shared__ Number s_kp[DIM][ND];
  // per face the numerical flux as a linear map (as in dg_operator_cart_kernel)
  __shared__ Number s_fc[NFACE][8];
  __shared__ Number s_vc[DIM][2];

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int d = wave / WPD;                       // the line direction of this wave
  const int t = (wave - d * WPD) * 64 + (tid & 63);  // its line within direction d
  const int e = xcd_remap(blockIdx.x, gridDim.x);
  const int cell = epi.cell_list ? (e < epi.n_list ? pipe::sld(epi.cell_list, e) : -1) : e;
  if (cell < 0 || cell >= mesh.n_cells) return;  // uniform over the group
  double cscale = 1.0;
  if (epi.cell_scale) {
    cscale = pipe::sld(epi.cell_scale, cell);
    if (cscale == 0.0) return;  // inactive cell: nothing computed or stored
  }
  int nbr[NFACE];
#pragma unroll
  for (int f = 0; f < NFACE; ++f) nbr[f] = pipe::sld(mesh.neighbor, (size_t)cell * NFACE + f);
  const size_t cbase = (size_t)cell * NF * ND;

  // ---------------- A: the cell's nodes, the lines' neighbour traces ----------------
  // line t of direction d: first node and stride (uniform), as in the line phase
  const int tc = t < NT ? t : NT - 1;
  int base, sd;
  {
    const int t0 = tc % N, t1 = tc / N;
    base = d == 0 ? t0 * N + t1 * N * N : (d == 1 ? t0 + t1 * N * N : t0 + t1 * N);
    sd = d == 0 ? 1 : (d == 1 ? N : N * N);
  }
  // the cell's fields into LDS: for FP64 with a 16-byte aligned vector through
  // LDS-DMA (no registers; waited for before the barrier), else through
  // registers, stored before the traces are gathered
  constexpr int OWN_BYTES = NF * ND * 8, NDMA = (OWN_BYTES + 1023) / 1024;
  const bool dma = std::is_same_v<Number, double> && (uintptr_t)src % 16 == 0;
  if (dma) {
    const uint32_t lbase = pipe::lds_addr(&s_u[0][0]);
    const uint32_t lane16 = uint32_t(tid & 63) * 16u;
    const double *own = src + cbase;  // wave-uniform, 16-byte aligned (OWN_BYTES = 32 ND)
    static_for<NDMA>([&](auto I_) __attribute__((always_inline)) {
      constexpr int I = decltype(I_)::value;
      if (wave == I % (BLOCK / 64) && int(I * 1024 + lane16) < OWN_BYTES)
        pipe::dma_dwordx4(lbase + I * 1024, own, lane16 + I * 1024);
    });
  } else {
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
      const int n = tid + j * BLOCK;
      if (n < ND) {
#pragma unroll
        for (int f = 0; f < NF; ++f) s_u[f][n] = Number(src[cbase + f * ND + n]);
      }
    }
  }
  // traces of the neighbours across faces 2d (s = 0) and 2d+1 (s = 1) at this
  // line's face point: neighbour side 1-s (boundary and hanging faces read the
  // cell itself; their flux coefficients are zero).  Summed here, before the
  // barrier, so that the loads' registers are free again in the line phase.
  Number tv[2], tp[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int nb = d == 0 ? nbr[s] : (d == 1 ? nbr[2 + s] : nbr[4 + s]);
    const double *un = src + (size_t)(nb >= 0 ? nb : cell) * NF * ND;
    const double *uv = un + d * ND + base, *up = un + DIM * ND + base;
    double xv[N], xp[N];
    if (d == 0) {
      // x lines are contiguous: 16-byte loads (8-byte aligned), a third of the
      // instructions of one load per node
      line_load<N>(uv, xv);
      line_load<N>(up, xp);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        xv[i] = uv[i * sd];
        xp[i] = up[i * sd];
      }
    }
    Number vn = Number(0), pn = Number(0);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const Number lo = Number(T.L[(1 - s) * N + i]);
      vn += lo * Number(xv[i]);
      pn += lo * Number(xp[i]);
    }
    tv[s] = vn;
    tp[s] = pn;
  }
  // (an empty asm that takes the sums: the compiler would otherwise sink the
  // sums, and with them every gathered load's registers, past the barrier)
  asm volatile("" : "+v"(tv[0]), "+v"(tv[1]), "+v"(tp[0]), "+v"(tp[1]));
  if (tid < NFACE) {
    // face lane: columns of the linear flux map (evaluate_inner_face /
    // evaluate_boundary_face, source/wave_equation_operations.cc:480-629),
    // dg_operator_cart_kernel's face-lane code
    const int f = tid, fd = f >> 1;
    int nb = nbr[0];
#pragma unroll
    for (int g = 1; g < NFACE; ++g)
      if (tid == g) nb = nbr[g];
    const double rho = mesh.rho[cell], cc = mesh.speed[cell];
    const double rho_m = nb >= 0 ? mesh.rho[nb] : rho, c_m = nb >= 0 ? mesh.speed[nb] : cc;
    const double ih = recip(mesh.h[(size_t)cell * DIM + fd]);
    const double sg = (f & 1) ? 1.0 : -1.0;
    const double ihf = nb == -NC_BID ? 0.0 : ih;
    double a[4], b[4];
    flux_matrix<double>(rho, cc, rho_m, c_m, nb >= 0 ? 0 : -nb, a, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double in = (j == 0 || j == 2) ? sg : 1.0;
      s_fc[f][j] = Number(a[j] * in * sg * ihf);
      s_fc[f][4 + j] = Number(b[j] * in * ihf);
    }
    if ((f & 1) == 0) {
      s_vc[fd][0] = Number(-ih * recip(rho));
      s_vc[fd][1] = Number(rho * cc * cc * ih);
    }
  }
  if (dma) pipe::wait_vm<0>();
  __syncthreads();

  // the accumulator, in flight during the line phase
  double bv[NPL][NF];
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int n = tid + j * BLOCK, nc = n < ND ? n : ND - 1;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      bv[j][f] = epi.b ? __builtin_nontemporal_load(epi.b + cbase + f * ND + nc) : 0.0;
  }

  // ---------------- B: one lane per line ----------------
  // dg_operator_cart_kernel's line phase with the same chains of FMAs per
  // value, in three passes over the line so that at most one line of values is
  // live at a time: the traces of v_d and p and the fluxes; v_d again (the weak
  // derivative plus the lifted fluxes: the dp/dt share); p again (the gradient
  // plus the lifted fluxes: dv_d/dt).  The empty asm with a memory clobber
  // between the passes makes the compiler read the line from LDS again instead
  // of keeping the first pass's values (3D k=6: 92 -> 71 VGPRs; k=8: 105 -> 84
  // and 121 -> 127 GDoF/s)
  if (t < NT) {
    const Number cvv = s_vc[d][0], cvp = s_vc[d][1];
    Number fv[2], fp[2];
    {
      Number vo[2], po[2];
      {
        Number vl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) vl[i] = s_u[d][base + i * sd];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          Number a = Number(0);
#pragma unroll
          for (int i = 0; i < N; ++i) a += Number(T.L[s * N + i]) * vl[i];
          vo[s] = a;
        }
      }
      {
        Number pl[N];
#pragma unroll
        for (int i = 0; i < N; ++i) pl[i] = s_u[DIM][base + i * sd];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          Number a = Number(0);
#pragma unroll
          for (int i = 0; i < N; ++i) a += Number(T.L[s * N + i]) * pl[i];
          po[s] = a;
        }
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const Number *c = s_fc[2 * d + s];
        fv[s] = c[0] * vo[s] + c[1] * po[s] + c[2] * tv[s] + c[3] * tp[s];
        fp[s] = c[4] * vo[s] + c[5] * po[s] + c[6] * tv[s] + c[7] * tp[s];
      }
    }
    asm volatile("" ::: "memory");  // reload, do not keep the first loads
    {
      Number vl[N];
#pragma unroll
      for (int i = 0; i < N; ++i) vl[i] = s_u[d][base + i * sd];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Number dv = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) dv += Number(T.DT[i * N + m]) * vl[m];
        Number k = cvp * dv;
        k += Number(T.LW[i]) * fp[0];
        k += Number(T.LW[N + i]) * fp[1];
        s_kp[d][base + i * sd] = k;
      }
    }
    asm volatile("" ::: "memory");  // reload, do not keep the first loads
    {
      Number pl[N];
#pragma unroll
      for (int i = 0; i < N; ++i) pl[i] = s_u[DIM][base + i * sd];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Number gp = Number(0);
#pragma unroll
        for (int m = 0; m < N; ++m) gp += Number(T.D[i * N + m]) * pl[m];
        Number k = cvv * gp;
        k += Number(T.LW[i]) * fv[0];
        k += Number(T.LW[N + i]) * fv[1];
        s_u[d][base + i * sd] = k;
      }
    }
  }
  __syncthreads();

  // ---------------- C: the nodes, fused epilogue ----------------
  const double ea1 = epi.a1 * cscale, ea2 = epi.a2 * cscale, ea3 = epi.a3 * cscale;
#pragma unroll
  for (int j = 0; j < NPL; ++j) {
    const int n = tid + j * BLOCK;
    if (n >= ND) continue;
    Number k[NF];
    Number kp = s_kp[0][n];
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd) {
      k[dd] = s_u[dd][n];
      if (dd > 0) kp += s_kp[dd][n];
    }
    k[DIM] = kp;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const size_t idx = cbase + f * ND + n;
      const double kk = double(k[f]);
      const double o1 = bv[j][f] + ea1 * kk;
      __builtin_nontemporal_store(o1, epi.o1 + idx);
      if (epi.o2) {
        const double cv = epi.o2_from_o1 ? o1 : epi.c[idx];
        __builtin_nontemporal_store(cv + ea2 * kk, epi.o2 + idx);
        if (epi.o3) __builtin_nontemporal_store(cv + ea3 * kk, epi.o3 + idx);
      }
    }
  }
}

}  // namespace exw
