// dg_operator_pipe.hpp -- the Cartesian operator + RK epilogue as a
// software-pipelined FP64 kernel for cells of >= 48 nodes whose two LDS
// buffers fit the workgroup (3D k = 3..6, 2D k >= 6; the host picks it where
// it was measured faster than dg_operator_cart_kernel).
//
// The arithmetic is dg_operator_cart_kernel's, expression for expression (its
// face-lane flux maps are precomputed per cell by cart_coef_kernel from the
// same code; its line phase and epilogue are restated below), so both kernels
// give bit-identical results; what differs is how the inputs reach the LDS.
// Each workgroup computes two consecutive cells one after the other.  While
// it computes cell t, the inputs of cell t+1 -- its nodes, the RK accumulator,
// the 2*dim neighbours' (v_n, p) fields and its flux coefficients -- already
// stream into the second of two LDS buffers through gfx950's LDS-DMA loads
// (global_load_lds_dword / _dwordx4: lane l's 4 or 16 bytes land at
// M0 + 4l or M0 + 16l), which occupy no registers.
//
// The DMA is issued from inline assembly and waited for with explicit vmcnt
// counts: the compiler treats every LDS-DMA as one alias class and would wait
// for the prefetch together with the current cell.  Vector memory operations
// of one wave complete in issue order, so "at most n outstanding" means all but
// the last n issued are done.  For the count to be exact the kernel issues no
// vector loads the compiler knows of: the neighbour rows are wave-uniform
// scalar loads through the constant address space, and the per-cell LDS
// layout comes from a compile-time plan (PipeShape::PLAN) whose every DMA
// instruction has at least lane 0 active, so that each wave issues a fixed
// number of them.
#pragma once
#include "dg_operator_cart.hpp"

// LDS of one workgroup's two buffers at most: gfx950's 160 KB (3D k=5 runs
// 2 groups per CU, k=6 one; the host launches the latter nowhere, as the
// one-cell kernel was measured faster there)
// the accumulator's DMA with the non-temporal cache policy (1) or not (0)
#ifndef EXW_PIPE_ACC_NT
#define EXW_PIPE_ACC_NT 1
#endif

#ifndef EXW_PIPE_LDS_MAX
#define EXW_PIPE_LDS_MAX 163840
#endif

namespace exw {
namespace pipe {

__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

// 64 lanes x 4 bytes: global base + voff -> LDS m0 + 4 * lane.  M0 is a
// reserved register that the compiler sets right before each instruction of
// its own that reads it, so it holds nothing of the compiler's across this.
__device__ __forceinline__ void dma_dword(uint32_t m0, const void *base, uint32_t voff) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, %2" ::"s"(m0), "v"(voff),
               "s"(base)
               : "memory");
}

// 64 lanes x 16 bytes: global base + voff -> LDS m0 + 16 * lane (gfx950)
__device__ __forceinline__ void dma_dwordx4(uint32_t m0, const void *base, uint32_t voff) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(m0), "v"(voff),
               "s"(base)
               : "memory");
}

// the same with the non-temporal cache policy: data read once (the RK
// accumulator) is marked to leave the L2 to the state vector, whose nodes
// the neighbouring cells read again
__device__ __forceinline__ void dma_dwordx4_nt(uint32_t m0, const void *base, uint32_t voff) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2 nt" ::"s"(m0), "v"(voff),
               "s"(base)
               : "memory");
}
__device__ __forceinline__ void dma_dword_nt(uint32_t m0, const void *base, uint32_t voff) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, %2 nt" ::"s"(m0), "v"(voff),
               "s"(base)
               : "memory");
}

// wait until at most N vector memory operations of this wave are outstanding
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N <= 63, "vmcnt is a 6-bit counter");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// workgroup barrier ordering LDS only (no wait on global stores or DMA)
__device__ __forceinline__ void barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// wave-uniform load through the constant address space (a scalar load); only
// for data the kernel does not write
template <typename T>
__device__ __forceinline__ T sld(const T *p, size_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

}  // namespace pipe

// Per-cell flux coefficients of a Cartesian mesh for the cells first ..
// first+count-1, one thread per (cell, face):
// exactly the values dg_operator_cart_kernel's face lanes compute (the same
// expressions; evaluate_inner_face / evaluate_boundary_face,
// source/wave_equation_operations.cc:480-629).  coef[cell]: the face flux maps
// [2*dim][8] (the numerical flux as a linear map of own v_n, own p, neighbour
// v_n, neighbour p -> v and p lifting, with 1/h and the normal's sign folded
// in; zero on hanging faces, which nc_face_kernel handles), then the cell-term
// factors [dim][2] = (-1/(rho h_d), rho c^2/h_d).
template <int DIM>
__global__ void cart_coef_kernel(const DevMesh mesh, int first, int count,
                                 double *__restrict__ coef) {
  constexpr int NFACE = 2 * DIM, NCOEF = NFACE * 8 + DIM * 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)count * NFACE) return;
  const int cell = first + int(i / NFACE), f = int(i % NFACE), d = f >> 1;
  const int nb = mesh.neighbor[(size_t)cell * NFACE + f];
  const double rho = mesh.rho[cell], cc = mesh.speed[cell];
  const double rho_m = nb >= 0 ? mesh.rho[nb] : rho, c_m = nb >= 0 ? mesh.speed[nb] : cc;
  const double ih = recip(mesh.h[(size_t)cell * DIM + d]);
  const double sg = (f & 1) ? 1.0 : -1.0;
  const double ihf = nb == -NC_BID ? 0.0 : ih;
  double a[4], b[4];
  flux_matrix<double>(rho, cc, rho_m, c_m, nb >= 0 ? 0 : -nb, a, b);
  double *o = coef + (size_t)cell * NCOEF;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double in = (j == 0 || j == 2) ? sg : 1.0;
    o[f * 8 + j] = a[j] * in * sg * ihf;
    o[f * 8 + 4 + j] = b[j] * in * ihf;
  }
  if ((f & 1) == 0) {
    o[NFACE * 8 + 2 * d] = -ih * recip(rho);
    o[NFACE * 8 + 2 * d + 1] = rho * cc * cc * ih;
  }
}

// One DMA instruction of a cell's inputs: 64 lanes of 4 or 16 bytes (the
// first `lanes` active) from byte goff of a source block to byte loff of the
// cell's LDS buffer.  Sources: the cell's fields, the accumulator's, the
// neighbour field fld = 2*face + component (0: v_n, 1: p), the coefficients.
enum DmaSrc { DMA_OWN, DMA_ACC, DMA_NB, DMA_COEF };
struct DmaOp {
  int src, fld, x4, lanes, goff, loff;
};
constexpr int MAX_DMA_OPS = 128;
struct DmaPlan {
  DmaOp op[MAX_DMA_OPS];
  int n;
  int slot[12];  // LDS byte offset of neighbour field 2*face + component
  int coef;      // LDS byte offset of the coefficients
  int bytes;     // LDS bytes of one buffer, a multiple of 16
};

// Appends the instructions that cover `len` bytes at global byte g / LDS
// byte l (equal modulo 16; every instruction stays inside the piece).  x4:
// an 8-byte head in dwords where they are 8 mod 16, 16-byte lanes, an 8-byte
// tail in dwords; else dword lanes.  Runs of >= 64 lanes are covered by full
// instructions at lane offsets 0, 64, ... and a last one ending exactly at
// the run's end (overlapping the one before with the same data).
// (Plans longer than MAX_DMA_OPS are counted but not stored; such shapes are
// not enabled.)
constexpr void dma_cover(DmaPlan &p, int src, int fld, bool x4, int g, int l, int len) {
  auto emit = [&](const DmaOp &o) {
    if (p.n < MAX_DMA_OPS) p.op[p.n] = o;
    ++p.n;
  };
  auto run = [&](bool wide, int off, int units) {  // units of 4 or 16 bytes from off
    const int u = wide ? 16 : 4;
    if (units >= 64) {
      for (int q = 0; 64 * q < units; ++q) {
        const int o = off + u * (64 * q < units - 64 ? 64 * q : units - 64);
        emit(DmaOp{src, fld, wide ? 1 : 0, 64, g + o, l + o});
      }
    } else if (units > 0) {
      emit(DmaOp{src, fld, wide ? 1 : 0, units, g + off, l + off});
    }
  };
  if (!x4) {
    run(false, 0, len / 4);
    return;
  }
  const int head = (g % 16) ? 16 - g % 16 : 0;
  run(false, 0, head / 4);
  const int q = (len - head) / 16;
  run(true, head, q);
  run(false, head + 16 * q, (len - head - 16 * q) / 4);
}

template <int DIM, int K>
struct PipeShape {
  using C = CartShape<DIM, K>;
  static constexpr int ND = C::ND, NF = C::NF, NFACE = 2 * DIM, NCOEF = NFACE * 8 + DIM * 2;
  static constexpr int WV = (ND + 63) / 64;  // waves per cell
  static constexpr int BLOCK = 64 * WV;
  // 16-byte lanes in 3D, where every cell block (NF*ND*8 = 32*ND bytes) and
  // coefficient block (NCOEF*8 = 432 bytes) starts 16-byte aligned in a
  // 16-byte aligned vector (the host checks the vectors); a neighbour field
  // then starts at 8*(field*ND % 2) mod 16.  2D: dword lanes throughout.
  static constexpr bool X4 = DIM == 3;
  static constexpr int OWN_BYTES = NF * ND * 8, FLD_BYTES = ND * 8;
  static constexpr int field_of(int j) { return (j & 1) ? DIM : (j >> 2); }

  static constexpr DmaPlan make_plan() {
    DmaPlan p{};
    p.n = 0;
    dma_cover(p, DMA_OWN, 0, X4, 0, 0, OWN_BYTES);
    dma_cover(p, DMA_ACC, 0, X4, 0, OWN_BYTES, OWN_BYTES);
    // neighbour fields: each placed in LDS at its global alignment modulo 16,
    // the next field taken among those that need no pad where there is one
    int o = 2 * OWN_BYTES;
    bool placed[12] = {};
    for (int k = 0; k < 2 * NFACE; ++k) {
      int pick = -1;
      for (int pass = 0; pass < 2 && pick < 0; ++pass)
        for (int j = 0; j < 2 * NFACE && pick < 0; ++j)
          if (!placed[j] && (pass == 1 || !X4 || (field_of(j) * FLD_BYTES) % 16 == o % 16))
            pick = j;
      const int g = field_of(pick) * FLD_BYTES;
      if (X4) o += (g % 16 - o % 16 + 16) % 16;
      p.slot[pick] = o;
      placed[pick] = true;
      dma_cover(p, DMA_NB, pick, X4, g, o, FLD_BYTES);
      o += FLD_BYTES;
    }
    o = (o + 15) / 16 * 16;
    p.coef = o;
    dma_cover(p, DMA_COEF, 0, X4, 0, o, NCOEF * 8);
    p.bytes = (o + NCOEF * 8 + 15) / 16 * 16;
    return p;
  }
  static constexpr DmaPlan PLAN = make_plan();
  static constexpr int TOTAL = PLAN.n;
  static constexpr int BUF = PLAN.bytes / 8;  // doubles of one LDS buffer
  static constexpr int OFF_B = NF * ND;       // the accumulator, in doubles
  static constexpr int OFF_C = PLAN.coef / 8;
  // the slot (in doubles) of field c of face f
  static constexpr int slot(int f, int c) { return PLAN.slot[2 * f + c] / 8; }

  // wave w issues the instructions i with i % WV == w
  static constexpr int count(int w, bool has_b) {
    int n = 0;
    for (int i = w; i < TOTAL && i < MAX_DMA_OPS; i += WV)
      if (has_b || PLAN.op[i].src != DMA_ACC) ++n;
    return n;
  }
  static constexpr int min_count(bool has_b) {
    int m = count(0, has_b);
    for (int w = 1; w < WV; ++w) m = count(w, has_b) < m ? count(w, has_b) : m;
    return m;
  }
  static constexpr int max_count(bool has_b) {
    int m = count(0, has_b);
    for (int w = 1; w < WV; ++w) m = count(w, has_b) > m ? count(w, has_b) : m;
    return m;
  }
  static constexpr int LDS_BYTES = 2 * PLAN.bytes;
  // cells that fill whole waves, and two buffers within EXW_PIPE_LDS_MAX bytes
  static constexpr bool ENABLED = ND >= 48 && LDS_BYTES <= EXW_PIPE_LDS_MAX && TOTAL <= MAX_DMA_OPS;
  static_assert(!ENABLED || max_count(true) + NF <= 63, "outstanding operations exceed vmcnt");
};

// The line phase of dg_operator_cart_kernel (phase B) on one LDS buffer, for
// the line of direction d = line / NT: buf holds the cell's fields [NF][ND]
// and the neighbour fields at their plan slots; the line's dv_d/dt and dp/dt
// share are written over the fields of face 2d at its own nodes; fc the face
// flux maps [2*DIM][8]; vc the cell-term factors [DIM][2].
template <int DIM, int K>
__device__ __forceinline__ void pipe_line(int line, double *buf, const double *fc,
                                          const double *vc) {
  using S = CartShape<DIM, K>;
  using P = PipeShape<DIM, K>;
  constexpr int N = S::N, ND = S::ND, NT = S::NT;
  constexpr const CartConst<N> &T = CartTable<N>::value;
  const int d = line / NT, t = line - d * NT;
  int base, sd;
  if constexpr (DIM == 2) {
    base = d == 0 ? t * N : t;
    sd = d == 0 ? 1 : N;
  } else {
    const int t0 = t % N, t1 = t / N;
    base = d == 0 ? t0 * N + t1 * N * N : (d == 1 ? t0 + t1 * N * N : t0 + t1 * N);
    sd = d == 0 ? 1 : (d == 1 ? N : N * N);
  }
  // slot of field C of face 2d + S: a select between compile-time constants
  // (never an index into the plan, which would become a memory load)
  auto slot = [&](auto S_, auto C_) {
    constexpr int S = decltype(S_)::value, C = decltype(C_)::value;
    constexpr int a = P::slot(S, C), b = P::slot(2 + S, C), e = P::slot(DIM == 3 ? 4 + S : S, C);
    if constexpr (DIM == 2) return d == 0 ? a : b;
    else return d == 0 ? a : (d == 1 ? b : e);
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  const double cvv = vc[2 * d], cvp = vc[2 * d + 1];
  double vl[N], pl[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int n = base + i * sd;
    vl[i] = buf[d * ND + n];
    pl[i] = buf[DIM * ND + n];
  }
  double kv[N], kp[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double gp = 0.0, dv = 0.0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      gp += T.D[i * N + m] * pl[m];
      dv += T.DT[i * N + m] * vl[m];
    }
    kv[i] = cvv * gp;
    kp[i] = cvp * dv;
  }
  static_for<2>([&](auto S_) __attribute__((always_inline)) {
    constexpr int s = decltype(S_)::value;
    constexpr const CartConst<N> &T = CartTable<N>::value;
    const int f = 2 * d + s;
    const double *nv = buf + slot(S_, I0{}), *np = buf + slot(S_, I1{});
    double vo = 0.0, po = 0.0, vn = 0.0, pn = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int n = base + i * sd;
      const double ls = T.L[s * N + i], lo = T.L[(1 - s) * N + i];
      vo += ls * vl[i];
      po += ls * pl[i];
      vn += lo * nv[n];
      pn += lo * np[n];
    }
    const double *c = fc + 8 * f;
    const double fv = c[0] * vo + c[1] * po + c[2] * vn + c[3] * pn;
    const double fp = c[4] * vo + c[5] * po + c[6] * vn + c[7] * pn;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double lw = T.LW[s * N + i];
      kv[i] += lw * fv;
      kp[i] += lw * fp;
    }
  });
  double *ov = buf + slot(I0{}, I0{}), *op = buf + slot(I0{}, I1{});
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int n = base + i * sd;
    ov[n] = kv[i];
    op[n] = kp[i];
  }
}

// TPG consecutive cells (or cell-list entries) per group, one after the
// other, the loop unrolled at compile time (a run-time loop would have its
// 1D-matrix constants hoisted out of it and spilled).  Needs mesh.coef.
// Epilogue: b optional, o2 only with o2_from_o1, cell list and cell scales
// with SEL (a separate instantiation, so that the plain sweep keeps its
// code); no c (the host uses dg_operator_cart_kernel for that).  The vectors must be below 4 GiB
// (n_cells * NF * ND * 8 < 2^32), and 16-byte aligned where PipeShape::X4.
//
// The two buffers admit about 2 waves per SIMD (3D k=4: 4 groups of 2 waves
// per CU); the register budget is capped to match.
template <int DIM, int K, int TPG, bool SEL>
__global__ void __launch_bounds__((PipeShape<DIM, K>::BLOCK)) __attribute__((amdgpu_waves_per_eu(2)))
    dg_operator_pipe_kernel(const DevMesh mesh, const double *__restrict__ src,
                            const Epilogue epi) {
  using P = PipeShape<DIM, K>;
  constexpr int ND = P::ND, NF = P::NF, NFACE = P::NFACE, WV = P::WV;
  constexpr int NLINE = CartShape<DIM, K>::NLINE;
  __shared__ __align__(16) double s_buf[2][P::BUF];

  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const uint32_t lane4 = uint32_t(lane) * 4u, lane16 = lane4 * 4u;
  // entries e = c0 .. c0+nt-1 of the cell sequence (all cells, or the list)
  const int n_entries = SEL && epi.cell_list ? epi.n_list : mesh.n_cells;
  const int c0 = xcd_remap(blockIdx.x, gridDim.x) * TPG;
  const int nt = min(TPG, n_entries - c0);  // >= 1: the grid is ceil(n_entries / TPG)
  const bool has_b = epi.b != nullptr;

  // the group's cells (entries past the end repeat the first), their scale
  // (a cell of scale 0 is inactive: loaded, not computed, not stored) and
  // neighbour rows
  int cells[TPG], rows[TPG][NFACE];
  double cs[TPG];
#pragma unroll
  for (int t = 0; t < TPG; ++t) {
    const int e = t < nt ? c0 + t : c0;
    cells[t] = SEL && epi.cell_list ? pipe::sld(epi.cell_list, e) : e;
    cs[t] = SEL && epi.cell_scale ? pipe::sld(epi.cell_scale, cells[t]) : 1.0;
#pragma unroll
    for (int f = 0; f < NFACE; ++f)
      rows[t][f] = pipe::sld(mesh.neighbor, (size_t)cells[t] * NFACE + f);
  }

  // this wave's share of cell c's DMA into buffer buf; global addresses are
  // the array's base (scalar) + a 32-bit byte offset (vectors below 4 GiB)
  auto issue = [&](int c, const int (&nb)[NFACE], int buf) __attribute__((always_inline)) {
    const uint32_t lbase = pipe::lds_addr(&s_buf[buf][0]);
    // the cell's row of flux coefficients (its class's row in a shared table)
    const int cls = mesh.coef_class ? pipe::sld(mesh.coef_class, c) : c;
    const uint32_t own = uint32_t(c) * uint32_t(P::OWN_BYTES);
    static_for<WV>([&](auto W_) __attribute__((always_inline)) {
      constexpr int W = decltype(W_)::value;
      if (wave != W) return;
      static_for<(P::TOTAL - W + WV - 1) / WV>([&](auto J_) __attribute__((always_inline)) {
        constexpr DmaOp op = P::PLAN.op[W + decltype(J_)::value * WV];
        if (op.src == DMA_ACC && !has_b) return;
        const double *base = op.src == DMA_ACC ? epi.b : (op.src == DMA_COEF ? mesh.coef : src);
        uint32_t g;
        if constexpr (op.src == DMA_NB) {
          // boundary and hanging faces read the cell itself (zero flux coefficients)
          const int f = op.fld >> 1, cn = nb[f] >= 0 ? nb[f] : c;
          g = uint32_t(cn) * uint32_t(P::OWN_BYTES);
        } else if constexpr (op.src == DMA_COEF) {
          g = uint32_t(cls) * uint32_t(P::NCOEF * 8);
        } else {
          g = own;
        }
        if (lane < op.lanes) {
          if constexpr (op.src == DMA_ACC && EXW_PIPE_ACC_NT) {
            if constexpr (op.x4)
              pipe::dma_dwordx4_nt(lbase + op.loff, base, lane16 + (g + op.goff));
            else
              pipe::dma_dword_nt(lbase + op.loff, base, lane4 + (g + op.goff));
          } else {
            if constexpr (op.x4)
              pipe::dma_dwordx4(lbase + op.loff, base, lane16 + (g + op.goff));
            else
              pipe::dma_dword(lbase + op.loff, base, lane4 + (g + op.goff));
          }
        }
      });
    });
  };

  issue(cells[0], rows[0], 0);
  static_for<TPG>([&](auto T_) __attribute__((always_inline)) {
    constexpr int t = decltype(T_)::value;
    if (t >= nt) return;  // uniform over the group
    const int c = cells[t];
    constexpr int tn = t + 1 < TPG ? t + 1 : t;
    // prefetch cell t+1, then wait for cell t.  Outstanding, oldest first:
    // DMA(t), stores(t-1) (>= NF: o1 of every field, if cell t-1 was
    // active; none otherwise), DMA(t+1).
    const bool stored = t > 0 && (!SEL || cs[t > 0 ? t - 1 : 0] != 0.0);
    if (t + 1 < nt) {
      if (has_b)
        pipe::wait_vm<63 - P::max_count(true)>();
      else
        pipe::wait_vm<63 - P::max_count(false)>();
      issue(cells[tn], rows[tn], (t + 1) & 1);
      if (stored) {
        if (has_b)
          pipe::wait_vm<P::min_count(true) + NF>();
        else
          pipe::wait_vm<P::min_count(false) + NF>();
      } else {
        if (has_b)
          pipe::wait_vm<P::min_count(true)>();
        else
          pipe::wait_vm<P::min_count(false)>();
      }
    } else {
      if (stored)
        pipe::wait_vm<NF>();
      else
        pipe::wait_vm<0>();
    }
    pipe::barrier();
    const bool active = !SEL || cs[t] != 0.0;  // uniform over the group

    double *buf = s_buf[t & 1];
    if (active && tid < NLINE)
      pipe_line<DIM, K>(tid, buf, buf + P::OFF_C, buf + P::OFF_C + NFACE * 8);
    pipe::barrier();

    // fused epilogue, RKVectorUpdater semantics (source/time_integrators.cc:128-155)
    if (active && tid < ND) {
      // (without SEL the scale is 1 and the products are exact)
      const double ea1 = SEL ? epi.a1 * cs[t] : epi.a1, ea2 = SEL ? epi.a2 * cs[t] : epi.a2,
                   ea3 = SEL ? epi.a3 * cs[t] : epi.a3;
      double k[NF];
      double kp = 0.0;
      static_for<DIM>([&](auto D_) __attribute__((always_inline)) {
        constexpr int d = decltype(D_)::value;
        constexpr int sv = P::slot(2 * d, 0), sp = P::slot(2 * d, 1);
        k[d] = buf[sv + tid];
        kp = d == 0 ? buf[sp + tid] : kp + buf[sp + tid];
      });
      k[DIM] = kp;
      const size_t cbase = (size_t)c * NF * ND;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const size_t idx = cbase + f * ND + tid;
        const double bv = has_b ? buf[P::OFF_B + f * ND + tid] : 0.0;
        const double o1 = bv + ea1 * k[f];
        __builtin_nontemporal_store(o1, epi.o1 + idx);
        if (epi.o2) {
          __builtin_nontemporal_store(o1 + ea2 * k[f], epi.o2 + idx);
          if (epi.o3) __builtin_nontemporal_store(o1 + ea3 * k[f], epi.o3 + idx);
        }
      }
    }
    // the next cell's prefetch reuses this buffer
    if (t + 1 < TPG) pipe::barrier();
  });
}

}  // namespace exw
