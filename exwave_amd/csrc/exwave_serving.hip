// exwave_serving.hip -- continuous batching of independent wave solves.
//
// Pools: one per (dim, degree, precision).  A pool owns `cap` cells of
// device memory: the connectivity/material/geometry arrays of the operator,
// a per-cell scale (the solve's dt; 0 = free cell) and three RK registers.
// A solve occupies a contiguous range of cells; its neighbour indices point
// inside its own range, so solves never interact.  One LSRK45 step of every
// solve of a pool is five launches of the fused operator over the pool's
// cells (epilogue coefficients a_s*dt_K per cell), captured once in a
// hipGraph per register parity and replayed.  Admission, deactivation and the
// final error evaluation are enqueued on the pool's stream in order with the
// steps; completion is detected with events, never by blocking the loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "basis.hpp"
#include "launchers.hpp"
#include "lsrk.hpp"

using namespace exw;

namespace exw {
void set_last_error(const std::string &msg);
}

namespace {

int sfail(int code, const std::string &msg) {
  set_last_error(msg);
  return code;
}
#define SHIP(call)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return sfail(EXW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

const DimKernels &dim_kernels(int dim) { return dim == 2 ? kernels_2d() : kernels_3d(); }

__global__ void fill_kernel(double *p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int solve_steps(const exw_request &r, double &dt, double &t_end) {
  const double h = 1.0 / double(1 << r.n_refinements);
  dt = r.cfl_number / std::pow((double)r.degree, 1.5) * h / r.c;
  double t = 0.0;
  int s = 0;
  while (t < r.final_time) {  // TimeControl::done (source/utilities.cc:84-87)
    t += dt;
    ++s;
  }
  t_end = t;
  return s;
}

int check_request(const exw_request &r) {
  // the pools sweep cell lists, which the kernels of cells with more nodes than
  // a workgroup has lanes (3D k >= 10) do not take
  if ((r.dim != 2 && r.dim != 3) || r.degree < 1 || r.degree > max_degree(r.dim) ||
      (r.dim == 3 && r.degree > 9) ||
      r.n_refinements < 0 || r.n_refinements > 10 || !(r.rho > 0) || !(r.c > 0) ||
      !(r.final_time > 0) || !(r.cfl_number > 0) || r.modes.n_modes < 0 || r.modes.n_modes > 8 ||
      (r.precision != EXW_FP64 && r.precision != EXW_FP32 && r.precision != EXW_PRECISION_AUTO))
    return sfail(EXW_EINVAL, "invalid serving request");
  for (int j = 0; j < r.modes.n_modes; ++j)
    for (int d = 0; d < r.dim; ++d)
      if (r.modes.m[j][d] < 0 || r.modes.m[j][d] > 64)
        return sfail(EXW_EINVAL, "invalid serving request: mode numbers must be 0..64");
  return EXW_OK;
}

// ||p0 - I p0|| with QGauss(k+2) on the 2^r-cell unit hypercube, where p0 =
// sum_j a_j prod_d sin(m_jd pi x_d) and I is the interpolant at the Gauss
// points (project_initial_field in the collocated basis).  With tensor
// quadrature, prod f - prod If = sum_d (prod_{e<d} If_e)(f_d - If_d)(prod_{e>d} f_e),
// so the squared norm is a sum of products of 1D inner products, each of
// which involves f - If itself: no cancellation between large terms.
void initial_error(const exw_request &r, double &e0, double &norm0) {
  const int dim = r.dim, k = r.degree, n = k + 1, ne = k + 2, nc = 1 << r.n_refinements;
  const int M = r.modes.n_modes;
  const double h = 1.0 / nc, pi = 3.14159265358979323846;
  std::vector<double> xg, wg, xe, we;
  gauss_legendre01(n, xg, wg);
  gauss_legendre01(ne, xe, we);
  std::vector<double> I((size_t)ne * n);
  for (int q = 0; q < ne; ++q) lagrange_barycentric(xg, xe[q], &I[(size_t)q * n], nullptr);
  const size_t L = (size_t)nc * ne;
  // per mode j and direction d: f, If and f - If at every error point
  std::vector<double> F((size_t)M * dim * L), G(F.size()), E(F.size());
  std::vector<double> fg(n);
  for (int j = 0; j < M; ++j)
    for (int d = 0; d < dim; ++d) {
      const double m = r.modes.m[j][d];
      const size_t o = ((size_t)j * dim + d) * L;
      for (int c = 0; c < nc; ++c) {
        for (int i = 0; i < n; ++i) fg[i] = std::sin(m * pi * (c + xg[i]) * h);
        for (int q = 0; q < ne; ++q) {
          const double f = std::sin(m * pi * (c + xe[q]) * h);
          double g = 0.0;
          for (int i = 0; i < n; ++i) g += I[(size_t)q * n + i] * fg[i];
          F[o + (size_t)c * ne + q] = f;
          G[o + (size_t)c * ne + q] = g;
          E[o + (size_t)c * ne + q] = f - g;
        }
      }
    }
  auto inner = [&](const double *a, const double *b) {
    double s = 0.0;
    for (int c = 0; c < nc; ++c)
      for (int q = 0; q < ne; ++q) s += h * we[q] * a[(size_t)c * ne + q] * b[(size_t)c * ne + q];
    return s;
  };
  // factor e of the telescoping term d of mode j
  auto factor = [&](int j, int d, int e) {
    const size_t o = ((size_t)j * dim + e) * L;
    return e < d ? &G[o] : (e == d ? &E[o] : &F[o]);
  };
  double err2 = 0.0, nrm2 = 0.0;
  for (int j = 0; j < M; ++j)
    for (int l = 0; l < M; ++l) {
      const double a = r.modes.amp[j] * r.modes.amp[l];
      double pf = 1.0;
      for (int e = 0; e < dim; ++e)
        pf *= inner(&F[((size_t)j * dim + e) * L], &F[((size_t)l * dim + e) * L]);
      nrm2 += a * pf;
      for (int d = 0; d < dim; ++d)
        for (int d2 = 0; d2 < dim; ++d2) {
          double t = 1.0;
          for (int e = 0; e < dim; ++e) t *= inner(factor(j, d, e), factor(l, d2, e));
          err2 += a * t;
        }
    }
  e0 = std::sqrt(std::max(err2, 0.0));
  norm0 = std::sqrt(std::max(nrm2, 0.0));
}

// EXW_PRECISION_AUTO: FP32 when the FP32 score clears EXW_FP32_SCORE_MIN
int auto_precision(const exw_request &r, int n_steps, double *e0_out, double *norm_out,
                   double *score_out) {
  double e0, norm0;
  initial_error(r, e0, norm0);
  const double floor32 = std::ldexp(1.0, -24) * std::sqrt((double)n_steps) * norm0;
  const double score = floor32 > 0 ? e0 / floor32 : 0.0;
  if (e0_out) *e0_out = e0;
  if (norm_out) *norm_out = norm0;
  if (score_out) *score_out = score;
  return score >= EXW_FP32_SCORE_MIN ? EXW_FP32 : EXW_FP64;
}

DevModes modes_of(const exw_request &r) {
  DevModes d{};
  d.n_modes = r.modes.n_modes;
  for (int k = 0; k < 8; ++k) {
    d.amp[k] = r.modes.amp[k];
    for (int j = 0; j < 3; ++j) d.m[k][j] = r.modes.m[k][j];
  }
  d.rho = r.modes.rho;
  d.c = r.modes.c;
  return d;
}

// Per-cell flux coefficients of a pool: only the FP64 sweeps that may run the
// pipelined kernel read them (DimKernels::pipe); other pools hold none.
size_t pool_ncoef(int dim, int k, int prec) {
  return (prec == EXW_FP64 && dim_kernels(dim).pipe[k]) ? (size_t)2 * dim * 8 + 2 * (size_t)dim
                                                         : 0;
}

// Device bytes per pool cell: three RK registers, rho, speed, scale, h[dim],
// the vertices, the flux coefficients and the neighbour indices.  Admission
// (Pool growth) and the submit-time size check both use this figure; the
// int share covers the neighbour indices and the cell's active-list entry.
size_t pool_cell_bytes(int dim, int k, int prec) {
  const size_t nv = (size_t)1 << dim;
  size_t nd = 1;
  for (int d = 0; d < dim; ++d) nd *= (size_t)(k + 1);
  const size_t dpc = (size_t)(dim + 1) * nd;
  return sizeof(double) * (3 * dpc + 3 + dim + nv * dim + pool_ncoef(dim, k, prec)) +
         sizeof(int) * (2 * (size_t)dim + 1);
}

struct Instance {
  int64_t id;
  exw_request req;
  int n_cells = 0, start = -1, n_steps = 0, steps_enq = 0;
  double dt = 0.0, t_end = 0.0;
  // host staging of the uploads, kept alive until the solve leaves
  std::vector<double> vert, h, rho, speed, scale;
  std::vector<int> nbr;
  double *d_partial = nullptr, *h_sum = nullptr;
  hipEvent_t done = nullptr;
  struct Pool *pool = nullptr;
};

struct Pool {
  int dim, k, prec, dpc, nv, nface;
  int cap = 0;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  int *nbr = nullptr;
  double *rho = nullptr, *speed = nullptr, *h = nullptr, *vert = nullptr, *scale = nullptr;
  double *coef = nullptr;  // per-cell flux coefficients (DevMesh::coef), set at admission
  double *reg[3] = {nullptr, nullptr, nullptr};
  std::vector<double> tab, etab, ckt;
  CkTab ck{};
  double *d_tab = nullptr, *d_etab = nullptr;
  std::map<int, int> free_ranges;  // start -> length
  int parity = 0;                  // the current solution lives in reg[parity]
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool graph_ok = false;
  std::vector<Instance *> active;
  int finishing = 0;  // solves whose error evaluation has not been polled yet
  // the launches cover the active cells only: d_list holds them in ascending
  // order, uploaded in stream order from a ring of pinned staging buffers
  static constexpr int RING = 4;
  int *d_list = nullptr;
  int list_cap = 0, n_list = 0;
  bool list_dirty = true;
  int *ring[RING] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ring_ev[RING] = {nullptr, nullptr, nullptr, nullptr};
  int ring_cap = 0, ring_next = 0;
  std::deque<hipEvent_t> inflight;
  std::vector<hipEvent_t> event_pool;

  size_t ncoef() const { return pool_ncoef(dim, k, prec); }
  size_t bytes_per_cell() const { return pool_cell_bytes(dim, k, prec); }
  DevMesh view(int start, int n) const {
    DevMesh m;
    m.n_cells = n;
    m.neighbor = nbr + (size_t)start * nface;
    m.rho = rho + start;
    m.speed = speed + start;
    m.h = h + (size_t)start * dim;
    m.vertices = vert + (size_t)start * nv * dim;
    m.coef = coef ? coef + (size_t)start * ncoef() : nullptr;
    return m;
  }
};

}  // namespace

struct exw_server_s {
  exw_ctx_s *ctx = nullptr;
  int device = 0;
  size_t budget = 0, used = 0;
  std::map<int, Pool *> pools;
  std::deque<Instance *> queue;
  std::vector<Instance *> finishing;
  int64_t next_id = 0;
  int lookahead = 32;
  // error-evaluation buffers of polled results (device partials, pinned sums),
  // reused by later finalisations: hipFree would wait for every pool's stream
  std::vector<std::pair<double *, double *>> spare_bufs;
};

// exw_ctx is opaque here; its device is set by the caller before any call
struct exw_ctx_s;

namespace {

int pool_key(int dim, int k, int prec) { return (dim * 16 + k) * 2 + prec; }

void free_pool_arrays(Pool *p) {
  hipFree(p->nbr);
  hipFree(p->rho);
  hipFree(p->speed);
  hipFree(p->h);
  hipFree(p->vert);
  hipFree(p->scale);
  hipFree(p->coef);
  for (int i = 0; i < 3; ++i) hipFree(p->reg[i]);
  p->nbr = nullptr;
  p->rho = p->speed = p->h = p->vert = p->scale = p->coef = nullptr;
  for (int i = 0; i < 3; ++i) p->reg[i] = nullptr;
}

void drop_graphs(Pool *p) {
  for (int i = 0; i < 2; ++i)
    if (p->exec[i]) hipGraphExecDestroy(p->exec[i]);
  p->exec[0] = p->exec[1] = nullptr;
  p->graph_ok = false;
}

// rebuild the list of active cells (ascending) and upload it in stream order
int update_list(Pool *p) {
  std::vector<std::pair<int, int>> ranges;
  int n = 0;
  for (Instance *in : p->active) {
    ranges.emplace_back(in->start, in->n_cells);
    n += in->n_cells;
  }
  std::sort(ranges.begin(), ranges.end());
  if (n > p->list_cap) {
    // the in-flight steps read the old list
    SHIP(hipStreamSynchronize(p->stream));
    hipFree(p->d_list);
    p->d_list = nullptr;
    p->list_cap = std::max(n, 2 * p->list_cap);
    if (hipMalloc((void **)&p->d_list, sizeof(int) * p->list_cap) != hipSuccess)
      return sfail(EXW_ENOMEM, "serving: cell list allocation failed");
  }
  if (n > p->ring_cap) {
    for (int i = 0; i < Pool::RING; ++i) {
      if (p->ring_ev[i]) SHIP(hipEventSynchronize(p->ring_ev[i]));
      if (p->ring[i]) hipHostFree(p->ring[i]);
      p->ring[i] = nullptr;
    }
    p->ring_cap = std::max(n, 2 * p->ring_cap);
    for (int i = 0; i < Pool::RING; ++i) {
      if (hipHostMalloc((void **)&p->ring[i], sizeof(int) * p->ring_cap, hipHostMallocDefault) !=
          hipSuccess)
        return sfail(EXW_ENOMEM, "serving: pinned staging allocation failed");
      if (!p->ring_ev[i]) SHIP(hipEventCreateWithFlags(&p->ring_ev[i], hipEventDisableTiming));
    }
  }
  const int slot = p->ring_next;
  p->ring_next = (slot + 1) % Pool::RING;
  SHIP(hipEventSynchronize(p->ring_ev[slot]));  // the slot's previous upload has been read
  int *h = p->ring[slot], k = 0;
  for (auto &r : ranges)
    for (int c = 0; c < r.second; ++c) h[k++] = r.first + c;
  if (n > 0) {
    SHIP(hipMemcpyAsync(p->d_list, h, sizeof(int) * n, hipMemcpyHostToDevice, p->stream));
  }
  SHIP(hipEventRecord(p->ring_ev[slot], p->stream));
  p->n_list = n;
  p->list_dirty = false;
  drop_graphs(p);
  return EXW_OK;
}

// grow the pool to new_cap cells (copies the live contents)
int grow_pool(exw_server s, Pool *p, int new_cap) {
  const size_t bpc = p->bytes_per_cell();
  const size_t extra = bpc * (size_t)(new_cap - p->cap);
  if (s->used + extra > s->budget) return sfail(EXW_ENOMEM, "serving HBM budget exhausted");
  int *nbr = nullptr;
  double *rho = nullptr, *speed = nullptr, *h = nullptr, *vert = nullptr, *scale = nullptr,
         *coef = nullptr, *reg[3] = {nullptr, nullptr, nullptr};
  const size_t nc = (size_t)new_cap;
  bool ok = hipMalloc((void **)&nbr, sizeof(int) * nc * p->nface) == hipSuccess &&
            hipMalloc((void **)&rho, sizeof(double) * nc) == hipSuccess &&
            hipMalloc((void **)&speed, sizeof(double) * nc) == hipSuccess &&
            hipMalloc((void **)&h, sizeof(double) * nc * p->dim) == hipSuccess &&
            hipMalloc((void **)&vert, sizeof(double) * nc * p->nv * p->dim) == hipSuccess &&
            hipMalloc((void **)&scale, sizeof(double) * nc) == hipSuccess &&
            (p->ncoef() == 0 ||
             hipMalloc((void **)&coef, sizeof(double) * nc * p->ncoef()) == hipSuccess);
  for (int i = 0; i < 3 && ok; ++i)
    ok = hipMalloc((void **)&reg[i], sizeof(double) * nc * p->dpc) == hipSuccess;
  if (!ok) {
    hipFree(nbr);
    hipFree(rho);
    hipFree(speed);
    hipFree(h);
    hipFree(vert);
    hipFree(scale);
    hipFree(coef);
    for (int i = 0; i < 3; ++i) hipFree(reg[i]);
    return sfail(EXW_ENOMEM, "serving pool allocation failed");
  }
  const size_t oc = (size_t)p->cap;
  hipStream_t st = p->stream;
  if (oc) {
    SHIP(hipMemcpyAsync(nbr, p->nbr, sizeof(int) * oc * p->nface, hipMemcpyDeviceToDevice, st));
    SHIP(hipMemcpyAsync(rho, p->rho, sizeof(double) * oc, hipMemcpyDeviceToDevice, st));
    SHIP(hipMemcpyAsync(speed, p->speed, sizeof(double) * oc, hipMemcpyDeviceToDevice, st));
    SHIP(hipMemcpyAsync(h, p->h, sizeof(double) * oc * p->dim, hipMemcpyDeviceToDevice, st));
    SHIP(hipMemcpyAsync(vert, p->vert, sizeof(double) * oc * p->nv * p->dim,
                        hipMemcpyDeviceToDevice, st));
    SHIP(hipMemcpyAsync(scale, p->scale, sizeof(double) * oc, hipMemcpyDeviceToDevice, st));
    if (coef)
      SHIP(hipMemcpyAsync(coef, p->coef, sizeof(double) * oc * p->ncoef(),
                          hipMemcpyDeviceToDevice, st));
    for (int i = 0; i < 3; ++i)
      SHIP(hipMemcpyAsync(reg[i], p->reg[i], sizeof(double) * oc * p->dpc,
                          hipMemcpyDeviceToDevice, st));
  }
  // new cells are free: scale 0, harmless connectivity
  const int64_t nn = (int64_t)(nc - oc);
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                     scale + oc, nn, 0.0);
  SHIP(hipGetLastError());
  SHIP(hipMemsetAsync(nbr + oc * p->nface, 0xff, sizeof(int) * nn * p->nface, st));
  SHIP(hipStreamSynchronize(st));
  free_pool_arrays(p);
  p->nbr = nbr;
  p->rho = rho;
  p->speed = speed;
  p->h = h;
  p->vert = vert;
  p->scale = scale;
  p->coef = coef;
  for (int i = 0; i < 3; ++i) p->reg[i] = reg[i];
  // the new tail joins the free list (merged with a free range ending at cap)
  int start = (int)oc, len = (int)nn;
  if (!p->free_ranges.empty()) {
    auto it = std::prev(p->free_ranges.end());
    if (it->first + it->second == (int)oc) {
      start = it->first;
      len += it->second;
      p->free_ranges.erase(it);
    }
  }
  p->free_ranges[start] = len;
  s->used += extra;
  p->bytes += extra;
  p->cap = new_cap;
  p->list_dirty = true;
  drop_graphs(p);
  return EXW_OK;
}

int alloc_range(Pool *p, int n) {
  for (auto it = p->free_ranges.begin(); it != p->free_ranges.end(); ++it) {
    if (it->second >= n) {
      const int start = it->first, len = it->second;
      p->free_ranges.erase(it);
      if (len > n) p->free_ranges[start + n] = len - n;
      return start;
    }
  }
  return -1;
}

void release_range(Pool *p, int start, int n) {
  auto it = p->free_ranges.emplace(start, n).first;
  // merge with the next and previous ranges
  auto nx = std::next(it);
  if (nx != p->free_ranges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    p->free_ranges.erase(nx);
  }
  if (it != p->free_ranges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      p->free_ranges.erase(it);
    }
  }
}

int get_pool(exw_server s, int dim, int k, int prec, Pool **out) {
  const int key = pool_key(dim, k, prec);
  auto it = s->pools.find(key);
  if (it != s->pools.end()) {
    *out = it->second;
    return EXW_OK;
  }
  Pool *p = new Pool;
  p->dim = dim;
  p->k = k;
  p->prec = prec;
  p->nv = 1 << dim;
  p->nface = 2 * dim;
  int nd = 1;
  for (int d = 0; d < dim; ++d) nd *= k + 1;
  p->dpc = (dim + 1) * nd;
  build_tables(k, p->tab, p->etab, p->ckt, p->ck);
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&p->d_tab, sizeof(double) * p->tab.size()) != hipSuccess ||
      hipMalloc((void **)&p->d_etab, sizeof(double) * p->etab.size()) != hipSuccess) {
    delete p;
    return sfail(EXW_EDEVICE, "serving pool setup failed");
  }
  SHIP(hipMemcpy(p->d_tab, p->tab.data(), sizeof(double) * p->tab.size(), hipMemcpyHostToDevice));
  SHIP(hipMemcpy(p->d_etab, p->etab.data(), sizeof(double) * p->etab.size(),
                 hipMemcpyHostToDevice));
  s->pools[key] = p;
  *out = p;
  return EXW_OK;
}

// one LSRK45 step of every cell of the pool, n = reg[from], np = reg[1-from]
int enqueue_step_launches(Pool *p, int from) {
  OpLaunch l = dim_kernels(p->dim).op[p->k][p->prec][1];
  DevMesh m = p->view(0, p->cap);
  const double *host_tab = p->tab.data();
  const int rc = lsrk_fused_schedule(
      Lsrk45::S, Lsrk45::a, Lsrk45::b, 1.0, p->reg[from], p->reg[1 - from], p->reg[2],
      [&](const double *src, const double *acc, double *y_out, double *b_out, double f1,
          double f2, bool last) {
        Epilogue e{acc,     y_out,   f1, nullptr, last ? nullptr : b_out, f2, 1, p->scale,
                   p->d_list, p->n_list};
        e.rk_stage = 1;
        return l(m, src, p->d_tab, host_tab, e, p->stream) == hipSuccess ? 0 : EXW_EDEVICE;
      });
  if (rc) return sfail(EXW_EDEVICE, "serving step launch failed");
  return EXW_OK;
}

int capture_graphs(Pool *p) {
  drop_graphs(p);
  for (int par = 0; par < 2; ++par) {
    hipGraph_t g = nullptr;
    SHIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step_launches(p, par);
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return sfail(EXW_EDEVICE, "hipStreamEndCapture failed");
    e = hipGraphInstantiate(&p->exec[par], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return sfail(EXW_EDEVICE, "hipGraphInstantiate failed");
  }
  p->graph_ok = true;
  return EXW_OK;
}

int admit(exw_server s, Instance *in) {
  const exw_request &r = in->req;
  Pool *p;
  if (int rc = get_pool(s, r.dim, r.degree, r.precision, &p)) return rc;
  int start = alloc_range(p, in->n_cells);
  if (start < 0) {
    const int want = std::max(p->cap * 2, p->cap + in->n_cells);
    int new_cap = want;
    // shrink the growth to what the budget allows
    const size_t bpc = p->bytes_per_cell();
    if (s->used + bpc * (size_t)(new_cap - p->cap) > s->budget)
      new_cap = p->cap + (int)std::min<size_t>((s->budget - s->used) / bpc, (size_t)INT32_MAX / 2);
    if (new_cap < p->cap + in->n_cells) return 1;  // does not fit now: stays queued
    if (int rc = grow_pool(s, p, new_cap)) return rc;
    start = alloc_range(p, in->n_cells);
    if (start < 0) return 1;
  }
  in->start = start;
  // host arrays of this solve (Cartesian unit hypercube)
  const int dim = r.dim, nv = 1 << dim, nface = 2 * dim, n = in->n_cells;
  int nc = 0;
  exw_mesh_hypercube(dim, 1 << r.n_refinements, 0, EXW_TRANSFORM_NONE, 0.0, 2, &nc, nullptr,
                     nullptr);
  in->vert.resize((size_t)n * nv * dim);
  in->nbr.resize((size_t)n * nface);
  exw_mesh_hypercube(dim, 1 << r.n_refinements, 0, EXW_TRANSFORM_NONE, 0.0, 2, &nc,
                     in->vert.data(), in->nbr.data());
  for (auto &v : in->nbr)
    if (v >= 0) v += start;
  in->rho.assign(n, r.rho);
  in->speed.assign(n, r.c);
  in->h.resize((size_t)n * dim);
  for (int c = 0; c < n; ++c)
    for (int d = 0; d < dim; ++d)
      in->h[(size_t)c * dim + d] = in->vert[((size_t)c * nv + (1 << d)) * dim + d] -
                                   in->vert[(size_t)c * nv * dim + d];
  in->scale.assign(n, in->dt);
  hipStream_t st = p->stream;
  SHIP(hipMemcpyAsync(p->nbr + (size_t)start * nface, in->nbr.data(), sizeof(int) * in->nbr.size(),
                      hipMemcpyHostToDevice, st));
  SHIP(hipMemcpyAsync(p->rho + start, in->rho.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  SHIP(hipMemcpyAsync(p->speed + start, in->speed.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                      st));
  SHIP(hipMemcpyAsync(p->h + (size_t)start * dim, in->h.data(), sizeof(double) * in->h.size(),
                      hipMemcpyHostToDevice, st));
  SHIP(hipMemcpyAsync(p->vert + (size_t)start * nv * dim, in->vert.data(),
                      sizeof(double) * in->vert.size(), hipMemcpyHostToDevice, st));
  SHIP(hipMemcpyAsync(p->scale + start, in->scale.data(), sizeof(double) * n,
                      hipMemcpyHostToDevice, st));
  // the flux coefficients of the new cells (neighbour indices are pool-wide)
  if (p->coef) SHIP(dim_kernels(dim).coef(p->view(0, p->cap), start, n, p->coef, st));
  // initial field into the pool's current register (project_initial_field)
  SHIP(dim_kernels(dim).project[r.degree](p->view(start, n), p->d_tab, modes_of(r), 0.0,
                                           p->reg[p->parity] + (size_t)start * p->dpc, st));
  p->active.push_back(in);
  p->list_dirty = true;
  return EXW_OK;
}

// the solve has had its last step enqueued: deactivate, evaluate, release
int finalize(exw_server s, Pool *p, Instance *in) {
  hipStream_t st = p->stream;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((in->n_cells + 255) / 256)), dim3(256), 0, st,
                     p->scale + in->start, (int64_t)in->n_cells, 0.0);
  SHIP(hipGetLastError());
  const int nb = 1024;  // as exw_l2_errors: same summation order, same result
  if (!s->spare_bufs.empty()) {
    in->d_partial = s->spare_bufs.back().first;
    in->h_sum = s->spare_bufs.back().second;
    s->spare_bufs.pop_back();
  } else {
    SHIP(hipMalloc((void **)&in->d_partial, sizeof(double) * (3 * nb + 4)));
    SHIP(hipHostMalloc((void **)&in->h_sum, sizeof(double) * 4, hipHostMallocDefault));
  }
  SHIP(dim_kernels(p->dim).l2err[p->k](p->view(in->start, in->n_cells), p->d_etab,
                                        modes_of(in->req), in->t_end,
                                        p->reg[p->parity] + (size_t)in->start * p->dpc,
                                        in->d_partial, nb, st));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, in->d_partial, nb, 3,
                     in->d_partial + 3 * nb);
  SHIP(hipGetLastError());
  SHIP(hipMemcpyAsync(in->h_sum, in->d_partial + 3 * nb, sizeof(double) * 3,
                      hipMemcpyDeviceToHost, st));
  SHIP(hipEventCreateWithFlags(&in->done, hipEventDisableTiming));
  SHIP(hipEventRecord(in->done, st));
  release_range(p, in->start, in->n_cells);
  p->list_dirty = true;
  in->pool = p;
  ++p->finishing;
  s->finishing.push_back(in);
  return EXW_OK;
}

// give the memory of pools without work back to the budget
int release_idle_pools(exw_server s) {
  for (auto &kv : s->pools) {
    Pool *p = kv.second;
    if (p->cap == 0 || !p->active.empty() || p->finishing > 0) continue;
    SHIP(hipStreamSynchronize(p->stream));
    while (!p->inflight.empty()) {
      p->event_pool.push_back(p->inflight.front());
      p->inflight.pop_front();
    }
    drop_graphs(p);
    free_pool_arrays(p);
    hipFree(p->d_list);
    p->d_list = nullptr;
    p->list_cap = 0;
    p->list_dirty = true;
    s->used -= p->bytes;
    p->bytes = 0;
    p->cap = 0;
    p->free_ranges.clear();
  }
  return EXW_OK;
}

// the pool bytes a request occupies once admitted (its precision resolved)
size_t footprint(const exw_request &r, int n_cells) {
  return (size_t)n_cells * pool_cell_bytes(r.dim, r.degree, r.precision);
}

}  // namespace

extern "C" {

int exw_server_create(exw_ctx ctx, size_t hbm_budget_bytes, exw_server *out) {
  if (!ctx || !out) return sfail(EXW_EINVAL, "exw_server_create: null argument");
  exw_server s = new exw_server_s;
  s->ctx = ctx;
  s->budget = hbm_budget_bytes;
  *out = s;
  return EXW_OK;
}

int exw_server_destroy(exw_server s) {
  if (!s) return EXW_OK;
  for (auto &kv : s->pools) {
    Pool *p = kv.second;
    hipStreamSynchronize(p->stream);
    drop_graphs(p);
    free_pool_arrays(p);
    hipFree(p->d_tab);
    hipFree(p->d_etab);
    for (auto e : p->event_pool) hipEventDestroy(e);
    for (auto e : p->inflight) hipEventDestroy(e);
    hipFree(p->d_list);
    for (int i = 0; i < Pool::RING; ++i) {
      if (p->ring[i]) hipHostFree(p->ring[i]);
      if (p->ring_ev[i]) hipEventDestroy(p->ring_ev[i]);
    }
    for (auto *in : p->active) delete in;
    hipStreamDestroy(p->stream);
    delete p;
  }
  for (auto *in : s->finishing) {
    hipFree(in->d_partial);
    hipHostFree(in->h_sum);
    if (in->done) hipEventDestroy(in->done);
    delete in;
  }
  for (auto *in : s->queue) delete in;
  for (auto &b : s->spare_bufs) {
    hipFree(b.first);
    hipHostFree(b.second);
  }
  delete s;
  return EXW_OK;
}

int exw_server_submit(exw_server s, const exw_request *r, int64_t *id) {
  if (!s || !r) return sfail(EXW_EINVAL, "exw_server_submit: null argument");
  if (int rc = check_request(*r)) return rc;
  Instance *in = new Instance;
  in->id = s->next_id++;
  in->req = *r;
  in->n_steps = solve_steps(*r, in->dt, in->t_end);
  if (r->precision == EXW_PRECISION_AUTO)
    in->req.precision = auto_precision(*r, in->n_steps, nullptr, nullptr, nullptr);
  int nc = 0;
  exw_mesh_hypercube(r->dim, 1 << r->n_refinements, 0, EXW_TRANSFORM_NONE, 0.0, 2, &nc, nullptr,
                     nullptr);
  in->n_cells = nc;
  if (footprint(in->req, nc) > s->budget) {
    delete in;
    return sfail(EXW_ENOMEM, "request larger than the serving HBM budget");
  }
  s->queue.push_back(in);
  if (id) *id = in->id;
  return EXW_OK;
}

int exw_server_progress(exw_server s, int max_steps) {
  if (!s) return sfail(EXW_EINVAL, "exw_server_progress: null server");
  // admission in arrival order while the budget allows; when the head does
  // not fit, idle pools give their memory back and the head is tried again
  bool released = false;
  while (!s->queue.empty()) {
    Instance *in = s->queue.front();
    const int rc = admit(s, in);
    if (rc < 0) return rc;
    if (rc == 1) {
      if (released) break;
      if (int r2 = release_idle_pools(s)) return r2;
      released = true;
      continue;
    }
    s->queue.pop_front();
  }
  int n_active = 0;
  for (int it = 0; it < max_steps; ++it) {
    n_active = 0;
    int launched = 0, waiting = 0;
    for (auto &kv : s->pools) {
      Pool *p = kv.second;
      if (p->active.empty()) continue;
      n_active += (int)p->active.size();
      // retire finished look-ahead steps; a pool with a full look-ahead waits
      while (!p->inflight.empty() && hipEventQuery(p->inflight.front()) == hipSuccess) {
        p->event_pool.push_back(p->inflight.front());
        p->inflight.pop_front();
      }
      if ((int)p->inflight.size() >= s->lookahead) {
        ++waiting;
        continue;
      }
      if (p->list_dirty)
        if (int rc = update_list(p)) return rc;
      if (!p->graph_ok)
        if (int rc = capture_graphs(p)) return rc;
      SHIP(hipGraphLaunch(p->exec[p->parity], p->stream));
      ++launched;
      p->parity ^= 1;
      hipEvent_t e;
      if (!p->event_pool.empty()) {
        e = p->event_pool.back();
        p->event_pool.pop_back();
      } else {
        SHIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      }
      SHIP(hipEventRecord(e, p->stream));
      p->inflight.push_back(e);
      for (size_t i = 0; i < p->active.size();) {
        Instance *in = p->active[i];
        if (++in->steps_enq >= in->n_steps) {
          if (int rc = finalize(s, p, in)) return rc;
          p->active[i] = p->active.back();
          p->active.pop_back();
        } else {
          ++i;
        }
      }
    }
    if (n_active == 0) break;
    if (!launched && waiting) {
      // every active pool has a full look-ahead: return to the caller, which
      // polls and calls again (no pool is waited for while others could step)
      break;
    }
  }
  // retire completed look-ahead events without blocking
  for (auto &kv : s->pools) {
    Pool *p = kv.second;
    while (!p->inflight.empty() && hipEventQuery(p->inflight.front()) == hipSuccess) {
      p->event_pool.push_back(p->inflight.front());
      p->inflight.pop_front();
    }
  }
  return n_active;
}

int exw_server_poll(exw_server s, exw_result *out, int max_results) {
  if (!s || (!out && max_results > 0)) return sfail(EXW_EINVAL, "exw_server_poll: null argument");
  int n = 0;
  for (size_t i = 0; i < s->finishing.size() && n < max_results;) {
    Instance *in = s->finishing[i];
    if (hipEventQuery(in->done) == hipSuccess) {
      exw_result &r = out[n++];
      r.id = in->id;
      r.n_steps = in->n_steps;
      r.time = in->t_end;
      r.err_p = std::sqrt(in->h_sum[0]);
      r.err_v = std::sqrt(in->h_sum[1]);
      r.mag_p = std::sqrt(in->h_sum[2]);
      r.precision = in->req.precision;
      // the copy into h_sum has completed (in->done): both buffers are free
      s->spare_bufs.emplace_back(in->d_partial, in->h_sum);
      hipEventDestroy(in->done);
      if (in->pool) --in->pool->finishing;
      delete in;
      s->finishing[i] = s->finishing.back();
      s->finishing.pop_back();
    } else {
      ++i;
    }
  }
  return n;
}

int exw_server_status(exw_server s, int *n_queued, int *n_active, int *n_finishing,
                      size_t *hbm_bytes) {
  if (!s) return sfail(EXW_EINVAL, "exw_server_status: null server");
  int na = 0;
  for (auto &kv : s->pools) na += (int)kv.second->active.size();
  if (n_queued) *n_queued = (int)s->queue.size();
  if (n_active) *n_active = na;
  if (n_finishing) *n_finishing = (int)s->finishing.size();
  if (hbm_bytes) *hbm_bytes = s->used;
  return EXW_OK;
}

int exw_solve(exw_ctx ctx, const exw_request *r, exw_result *out) {
  if (!ctx || !r || !out) return sfail(EXW_EINVAL, "exw_solve: null argument");
  if (int rc = check_request(*r)) return rc;
  const int dim = r->dim, nv = 1 << dim;
  int nc = 0, rc;
  exw_mesh_hypercube(dim, 1 << r->n_refinements, 0, EXW_TRANSFORM_NONE, 0.0, 2, &nc, nullptr,
                     nullptr);
  std::vector<double> V((size_t)nc * nv * dim);
  std::vector<int> nb((size_t)nc * 2 * dim);
  exw_mesh_hypercube(dim, 1 << r->n_refinements, 0, EXW_TRANSFORM_NONE, 0.0, 2, &nc, V.data(),
                     nb.data());
  std::vector<double> rho(nc, r->rho), speed(nc, r->c);
  exw_op op = nullptr;
  if ((rc = exw_op_create(ctx, dim, r->degree, nc, V.data(), nb.data(), rho.data(), speed.data(),
                          &op)))
    return rc;
  const int64_t N = exw_op_n_dofs(op);
  double *buf = nullptr;
  if ((rc = exw_alloc(ctx, sizeof(double) * N * 3, (void **)&buf))) {
    exw_op_destroy(op);
    return rc;
  }
  double *sol = buf, *tmp = buf + N, *work = buf + 2 * N;
  double dt, t_end;
  const int steps = solve_steps(*r, dt, t_end);
  const int prec = r->precision == EXW_PRECISION_AUTO
                       ? auto_precision(*r, steps, nullptr, nullptr, nullptr)
                       : r->precision;
  rc = exw_project(op, &r->modes, 0.0, sol);
  for (int s = 0; s < steps && !rc; ++s) {
    std::swap(sol, tmp);
    rc = exw_time_step(op, EXW_LSRK45R2, dt, tmp, sol, work, 0, prec);
  }
  double err[3] = {0, 0, 0};
  if (!rc) rc = exw_l2_errors(op, &r->modes, t_end, sol, err);
  exw_free(ctx, buf);
  exw_op_destroy(op);
  if (rc) return rc;
  out->id = -1;
  out->n_steps = steps;
  out->time = t_end;
  out->err_p = err[0];
  out->err_v = err[1];
  out->mag_p = err[2];
  out->precision = prec;
  return EXW_OK;
}

int exw_request_initial_error(const exw_request *r, double *e0, double *norm_p0, double *score,
                              int *precision) {
  if (!r) return sfail(EXW_EINVAL, "exw_request_initial_error: null request");
  if (int rc = check_request(*r)) return rc;
  double dt, t_end;
  const int steps = solve_steps(*r, dt, t_end);
  const int p = auto_precision(*r, steps, e0, norm_p0, score);
  if (precision) *precision = p;
  return EXW_OK;
}

}  // extern "C"
