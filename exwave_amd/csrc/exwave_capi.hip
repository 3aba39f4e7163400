// exwave_capi.hip -- C-ABI of the MI355X ExWave hot path (include/exwave_amd.h).
//
// Host side of the library: operator setup (replaces WaveEquationOperation::
// setup / reset_data_vectors), the fused operator + RK stage launches, the
// ADER step and the explicit integrators (ExplicitIntegrator::perform_time_step
// of source/time_integrators.cc), projection and norms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "basis.hpp"
#include "launchers.hpp"
#include "lsrk.hpp"
#include "capi_internal.hpp"

using namespace exw;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define EXW_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(EXW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

const DimKernels &kernels(int dim) { return dim == 2 ? kernels_2d() : kernels_3d(); }
}  // namespace

namespace exw {
// error message of the calling thread, shared by every translation unit
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace exw

// Scratch of exw_l2_errors_noalloc: the per-block partial sums (3 per block)
// and the three totals.  One copy per GPU, held until the process ends; host
// code never uses the array's name as an address, only the per-device
// pointer from hipGetSymbolAddress.
#define EXW_NOALLOC_MAX_PARTIALS 1024
static __device__ __align__(256) double scratch[3 * EXW_NOALLOC_MAX_PARTIALS + 4];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once per
// device, with that device current)
int scratch_on_device(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return fail(EXW_EINVAL, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(EXW_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
  }
  *out = g_scratch_ptr[device];
  return EXW_OK;
}

__global__ void sqrt3_kernel(const double *__restrict__ sums, double *__restrict__ out) {
  if (threadIdx.x < 3) out[threadIdx.x] = sqrt(sums[threadIdx.x]);
}
}  // namespace

extern "C" {

const char *exw_last_error(void) { return g_err.c_str(); }
const char *exw_version(void) { return "exwave_amd 0.1 (gfx950)"; }

int exw_ctx_create(int device, void *stream, exw_ctx *out) {
  if (!out) return fail(EXW_EINVAL, "exw_ctx_create: out is NULL");
  int n = 0;
  EXW_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(EXW_EINVAL, "exw_ctx_create: bad device id");
  EXW_HIP(hipSetDevice(device));
  double *scr = nullptr;
  if (int rc = scratch_on_device(device, &scr)) return rc;
  exw_ctx c = new exw_ctx_s;
  c->device = device;
  c->stream = (hipStream_t)stream;
  c->scratch = scr;
  *out = c;
  return EXW_OK;
}

int exw_ctx_set_stream(exw_ctx ctx, void *stream) {
  if (!ctx) return fail(EXW_EINVAL, "null context");
  ctx->stream = (hipStream_t)stream;
  return EXW_OK;
}

int exw_ctx_destroy(exw_ctx ctx) {
  delete ctx;
  return EXW_OK;
}

int exw_sync(exw_ctx ctx) {
  if (!ctx) return fail(EXW_EINVAL, "null context");
  EXW_HIP(hipStreamSynchronize(ctx->stream));
  return EXW_OK;
}

int exw_alloc(exw_ctx ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return fail(EXW_EINVAL, "exw_alloc: null argument");
  EXW_HIP(hipSetDevice(ctx->device));
  if (hipMalloc(dptr, bytes) != hipSuccess) return fail(EXW_ENOMEM, "exw_alloc: hipMalloc failed");
  return EXW_OK;
}

int exw_free(exw_ctx ctx, void *dptr) {
  if (!ctx) return fail(EXW_EINVAL, "exw_free: null context");
  EXW_HIP(hipFree(dptr));
  return EXW_OK;
}

int exw_memcpy_h2d(exw_ctx ctx, void *dst, const void *src, size_t bytes) {
  EXW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  EXW_HIP(hipStreamSynchronize(ctx->stream));
  return EXW_OK;
}

int exw_memcpy_d2h(exw_ctx ctx, void *dst, const void *src, size_t bytes) {
  EXW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  EXW_HIP(hipStreamSynchronize(ctx->stream));
  return EXW_OK;
}

static void free_op(exw_op op) {
  if (!op) return;
  hipFree(op->d_neighbor);
  hipFree(op->d_rho);
  hipFree(op->d_speed);
  hipFree(op->d_h);
  hipFree(op->d_vertices);
  hipFree(op->d_coef);
  hipFree(op->d_coef_class);
  hipFree(op->d_tab);
  hipFree(op->d_errtab);
  hipFree(op->d_cktab);
  hipFree(op->d_posttab);
  hipFree(op->d_nc_cells);
  hipFree(op->d_nc_kind);
  hipFree(op->d_nc_other);
  hipFree(op->d_nctab);
  hipFree(op->d_partial);
  hipFree(op->d_sum);
  hipFree(op->d_src_cells);
  hipFree(op->d_src_b);
  delete op;
}

// 1D matrices of the post-processing (layout PostTab<k>): QGauss(k+2) points
// and weights, the Gauss-Lobatto(k+2) Lagrange basis and the QGauss(k+2)
// collocation derivative at those points, the QGauss(k+1) (solution) basis at
// them, and the QGauss(k+3) error rule with the Gauss-Lobatto basis there.
static std::vector<double> build_post_table(int k) {
  const int N = k + 1, P = k + 2, E = k + 3;
  std::vector<double> t(2 * P + 2 * P * P + P * N + 2 * E + E * P);
  double xg[16], wg[16], xl[16];
  double *xq = t.data(), *wq = xq + P, *PV = wq + P, *DQ = PV + P * P, *IQ = DQ + P * P,
         *xe = IQ + P * N, *we = xe + E, *IE = we + E;
  cx::gauss01(N, xg, wg);
  cx::gauss01(P, xq, wq);
  cx::gauss_lobatto01(P, xl);
  cx::gauss01(E, xe, we);
  for (int q = 0; q < P; ++q) {
    cx::lagrange(P, xl, xq[q], PV + q * P, nullptr);
    cx::lagrange(P, xq, xq[q], nullptr, DQ + q * P);
    cx::lagrange(N, xg, xq[q], IQ + q * N, nullptr);
  }
  for (int q = 0; q < E; ++q) cx::lagrange(P, xl, xe[q], IE + q * P, nullptr);
  return t;
}

extern "C++" {
template <typename T>
static int upload(exw_op op, T **dptr, const T *host, size_t n) {
  if (hipMalloc((void **)dptr, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return fail(EXW_ENOMEM, "exw_op_create: hipMalloc failed");
  if (n) EXW_HIP(hipMemcpy(*dptr, host, n * sizeof(T), hipMemcpyHostToDevice));
  return EXW_OK;
}
}  // extern "C++"

// the hanging-face records of exw_op_create_nc against the connectivity and
// the geometry; fills the per-cell face kinds (0 conforming, 1 fine side,
// 2 coarse side) and partners of the NcArgs layout
static int check_hanging(int dim, int n_cells, const double *vertices, const int *neighbor,
                         int n_hanging, const int *hanging, std::vector<int> &kind,
                         std::vector<int> &other) {
  const int nv = 1 << dim, nface = 2 * dim, nsub = 1 << (dim - 1), rec = 2 + nsub;
  kind.assign((size_t)n_cells * nface, 0);
  other.assign((size_t)n_cells * nface * 4, -1);
  auto corner = [&](int c, const double *xi, double *x) {
    const double *V = vertices + (size_t)c * nv * dim;
    for (int i = 0; i < dim; ++i) x[i] = 0.0;
    for (int u = 0; u < nv; ++u) {
      double w = 1.0;
      for (int d = 0; d < dim; ++d) w *= ((u >> d) & 1) ? xi[d] : 1.0 - xi[d];
      for (int i = 0; i < dim; ++i) x[i] += w * V[u * dim + i];
    }
  };
  for (int h = 0; h < n_hanging; ++h) {
    const int *r = hanging + (size_t)h * rec;
    const int c = r[0], f = r[1];
    if (c < 0 || c >= n_cells || f < 0 || f >= nface || neighbor[(size_t)c * nface + f] != EXW_NC_FACE ||
        kind[(size_t)c * nface + f] != 0)
      return fail(EXW_EINVAL, "hanging record " + std::to_string(h) + ": bad coarse cell/face");
    kind[(size_t)c * nface + f] = 2;
    const int d = f >> 1, s = f & 1;
    // size of the coarse face, for the geometric tolerance
    double lo[3], hi[3], x[3], y[3], xi[3];
    for (int e = 0; e < dim; ++e) xi[e] = e == d ? s : 0.0;
    corner(c, xi, lo);
    for (int e = 0; e < dim; ++e) xi[e] = e == d ? s : 1.0;
    corner(c, xi, hi);
    double diam = 0.0;
    for (int i = 0; i < dim; ++i) diam = std::max(diam, std::fabs(hi[i] - lo[i]));
    for (int b = 0; b < nsub; ++b) {
      const int fc = r[2 + b];
      if (fc < 0 || fc >= n_cells || neighbor[(size_t)fc * nface + (f ^ 1)] != EXW_NC_FACE ||
          kind[(size_t)fc * nface + (f ^ 1)] != 0)
        return fail(EXW_EINVAL, "hanging record " + std::to_string(h) + ": bad fine cell");
      kind[(size_t)fc * nface + (f ^ 1)] = 1;
      other[((size_t)fc * nface + (f ^ 1)) * 4 + 0] = c;
      other[((size_t)fc * nface + (f ^ 1)) * 4 + 1] = b;
      other[((size_t)c * nface + f) * 4 + b] = fc;
      // the fine face's corners are the coarse face's map at the subface corners
      for (int v = 0; v < nsub; ++v) {
        double xf[3], xc[3];
        int j = 0;
        for (int e = 0; e < dim; ++e) {
          if (e == d) {
            xf[e] = 1 - s;
            xc[e] = s;
          } else {
            const int bit = (v >> j) & 1;
            xf[e] = bit;
            xc[e] = 0.5 * (bit + ((b >> j) & 1));
            ++j;
          }
        }
        corner(fc, xf, x);
        corner(c, xc, y);
        for (int i = 0; i < dim; ++i)
          if (std::fabs(x[i] - y[i]) > 1e-12 * diam)
            return fail(EXW_EINVAL, "hanging record " + std::to_string(h) +
                                        ": fine face is not the subface's image");
      }
    }
  }
  for (size_t i = 0; i < kind.size(); ++i)
    if (neighbor[i] == EXW_NC_FACE && kind[i] == 0)
      return fail(EXW_EINVAL, "EXW_NC_FACE in neighbor[] without a hanging record");
  return EXW_OK;
}

static int op_create(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                     const int *neighbor, int n_hanging, const int *hanging, const double *rho,
                     const double *speed, exw_op *out);

int exw_mesh_check(int dim, int n_cells, const double *vertices, const int *neighbor,
                   int n_hanging, const int *hanging) {
  if ((dim != 2 && dim != 3) || n_cells < 1 || !vertices || !neighbor ||
      n_hanging < 0 || (n_hanging > 0 && !hanging))
    return fail(EXW_EINVAL, "exw_mesh_check: invalid argument");
  const int nv = 1 << dim, nface = 2 * dim;
  for (int c = 0; c < n_cells; ++c) {
    const double *V = vertices + (size_t)c * nv * dim;
    double h = 0.0;
    for (int d = 0; d < dim; ++d) h = std::max(h, std::fabs(V[(nv - 1) * dim + d] - V[d]));
    for (int f = 0; f < nface; ++f) {
      const int nb = neighbor[(size_t)c * nface + f];
      if (nb == EXW_NC_FACE && n_hanging > 0) continue;  // checked with its record
      if (nb >= n_cells || (nb < 0 && (nb < -3 || nb > -1)))
        return fail(EXW_EINVAL, "set your boundary ids correctly: 1 - soft wall, 2 - hard wall, "
                                "3 - first order ABC (or neighbour index out of range)");
      if (nb < 0) continue;
      if (neighbor[(size_t)nb * nface + (f ^ 1)] != c)
        return fail(EXW_EINVAL, "cell " + std::to_string(c) + " face " + std::to_string(f) +
                                    ": the neighbour does not see this cell across face " +
                                    std::to_string(f ^ 1) +
                                    " (faces pair as f and f^1 in standard orientation; other "
                                    "pairings are not supported)");
      // standard orientation: corner v of face f (bit d of v = s) is corner
      // v with bit d flipped of the neighbour's face f^1, at the same point
      const int d = f >> 1, s = f & 1;
      const double *W = vertices + (size_t)nb * nv * dim;
      for (int v = 0; v < nv; ++v) {
        if (((v >> d) & 1) != s) continue;
        const int w = v ^ (1 << d);
        for (int i = 0; i < dim; ++i)
          if (std::fabs(V[v * dim + i] - W[w * dim + i]) > 1e-10 * h)
            return fail(EXW_EINVAL,
                        "cell " + std::to_string(c) + " face " + std::to_string(f) +
                            ": face not in standard orientation (its corners are not the "
                            "neighbour's face corners in the same order; rotated or flipped "
                            "faces are not supported)");
      }
    }
  }
  return EXW_OK;
}

int exw_op_create(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                  const int *neighbor, const double *rho, const double *speed, exw_op *out) {
  return op_create(ctx, dim, degree, n_cells, vertices, neighbor, 0, nullptr, rho, speed, out);
}

int exw_op_create_nc(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                     const int *neighbor, int n_hanging, const int *hanging, const double *rho,
                     const double *speed, exw_op *out) {
  if (n_hanging < 0 || (n_hanging > 0 && !hanging))
    return fail(EXW_EINVAL, "exw_op_create_nc: bad hanging records");
  if (n_hanging > 0 && (dim == 2 || dim == 3) && degree >= 1 && degree <= max_degree(dim) &&
      !kernels(dim).nc[degree])
    return fail(EXW_EUNSUP, "exw_op_create_nc: hanging faces not supported at this degree");
  return op_create(ctx, dim, degree, n_cells, vertices, neighbor, n_hanging, hanging, rho, speed,
                   out);
}

static int op_create(exw_ctx ctx, int dim, int degree, int n_cells, const double *vertices,
                     const int *neighbor, int n_hanging, const int *hanging, const double *rho,
                     const double *speed, exw_op *out) {
  if (!ctx || !out || !vertices || !neighbor || !rho || !speed)
    return fail(EXW_EINVAL, "exw_op_create: null argument");
  if (dim != 2 && dim != 3) return fail(EXW_EINVAL, "exw_op_create: dim must be 2 or 3");
  if (degree < 1 || degree > max_degree(dim))
    return fail(EXW_EUNSUP, "exw_op_create: degree " + std::to_string(degree) +
                                " not compiled (1.." + std::to_string(max_degree(dim)) + " in " +
                                std::to_string(dim) + "D)");
  if (n_cells < 1) return fail(EXW_EINVAL, "exw_op_create: n_cells < 1");
  const int nv = 1 << dim, nface = 2 * dim, N = degree + 1;
  // validate connectivity, face orientation and materials
  // (source/wave_equation_operations.cc:615-616)
  if (int rc = exw_mesh_check(dim, n_cells, vertices, neighbor, n_hanging, hanging)) return rc;
  std::vector<int> nc_kind, nc_other;
  if (n_hanging > 0) {
    if (int rc = check_hanging(dim, n_cells, vertices, neighbor, n_hanging, hanging, nc_kind,
                               nc_other))
      return rc;
  }
  for (int c = 0; c < n_cells; ++c)
    if (!(rho[c] > 0.0) || !(speed[c] > 0.0))
      return fail(EXW_EINVAL, "exw_op_create: density and speed must be positive");
  EXW_HIP(hipSetDevice(ctx->device));
  exw_op op = new exw_op_s;
  op->ctx = ctx;
  op->dim = dim;
  op->degree = degree;
  op->n_cells = n_cells;
  // geometry classification: axis-aligned boxes -> Cartesian kernels
  bool cart = true;
  std::vector<double> h((size_t)n_cells * dim);
  double minvd = 1e300;
  for (int c = 0; c < n_cells; ++c) {
    const double *V = vertices + (size_t)c * nv * dim;
    for (int d = 0; d < dim; ++d) h[(size_t)c * dim + d] = V[(1 << d) * dim + d] - V[d];
    for (int v = 0; v < nv && cart; ++v)
      for (int i = 0; i < dim; ++i) {
        const double want = V[i] + (((v >> i) & 1) ? h[(size_t)c * dim + i] : 0.0);
        if (std::fabs(V[v * dim + i] - want) > 1e-13 * std::fabs(h[(size_t)c * dim + i]))
          cart = false;
      }
    // minimum vertex distance over all vertex pairs / c (source/explicit_wave.cc:529-545)
    for (int a = 0; a < nv; ++a)
      for (int b = a + 1; b < nv; ++b) {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) {
          const double df = V[a * dim + d] - V[b * dim + d];
          s += df * df;
        }
        minvd = std::min(minvd, std::sqrt(s) / speed[c]);
      }
  }
  op->cart = cart;
  op->minvd = minvd;
  int rc;
  if ((rc = upload(op, &op->d_neighbor, neighbor, (size_t)n_cells * nface)) ||
      (rc = upload(op, &op->d_rho, rho, (size_t)n_cells)) ||
      (rc = upload(op, &op->d_speed, speed, (size_t)n_cells)) ||
      (rc = upload(op, &op->d_h, h.data(), h.size())) ||
      (rc = upload(op, &op->d_vertices, vertices, (size_t)n_cells * nv * dim))) {
    free_op(op);
    return rc;
  }
  if (cart && n_cells > 0) {
    // the per-cell flux coefficients of the pipelined Cartesian kernel
    const size_t n = (size_t)n_cells * (2 * dim * 8 + 2 * dim);
    if (hipMalloc((void **)&op->d_coef, n * sizeof(double)) != hipSuccess) {
      free_op(op);
      return fail(EXW_ENOMEM, "exw_op_create: hipMalloc failed");
    }
    if (kernels(dim).coef(op->view(), 0, n_cells, op->d_coef, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      free_op(op);
      return fail(EXW_EDEVICE, "exw_op_create: flux coefficient set-up failed");
    }
    // rows shared by many cells (same extents, materials on both sides and
    // boundary pattern) are kept once: the sweeps then read a small table that
    // stays in L2 instead of a per-cell array from HBM (the same values)
    const int nco = 2 * dim * 8 + 2 * dim;
    std::vector<double> rows(n);
    EXW_HIP(hipMemcpy(rows.data(), op->d_coef, n * sizeof(double), hipMemcpyDeviceToHost));
    std::map<std::vector<uint64_t>, int> seen;
    std::vector<int> cls(n_cells);
    std::vector<double> table;
    for (int c = 0; c < n_cells && (int)seen.size() <= n_cells / 8; ++c) {
      std::vector<uint64_t> key(nco);
      std::memcpy(key.data(), rows.data() + (size_t)c * nco, nco * sizeof(double));
      auto it = seen.find(key);
      if (it == seen.end()) {
        it = seen.emplace(key, (int)seen.size()).first;
        table.insert(table.end(), rows.begin() + (size_t)c * nco, rows.begin() + (size_t)(c + 1) * nco);
      }
      cls[c] = it->second;
    }
    if ((int)seen.size() <= n_cells / 8) {
      double *d_table = nullptr;
      if ((rc = upload(op, &d_table, table.data(), table.size())) ||
          (rc = upload(op, &op->d_coef_class, cls.data(), cls.size()))) {
        hipFree(d_table);
        free_op(op);
        return rc;
      }
      hipFree(op->d_coef);
      op->d_coef = d_table;
      op->n_coef_rows = (int)seen.size();
    }
  }
  if (n_hanging > 0) {
    // the cells that touch a hanging face, in cell order, with their face kinds
    std::vector<int> cells, kind, other;
    for (int c = 0; c < n_cells; ++c) {
      bool any = false;
      for (int f = 0; f < nface; ++f) any |= nc_kind[(size_t)c * nface + f] != 0;
      if (!any) continue;
      cells.push_back(c);
      kind.insert(kind.end(), nc_kind.begin() + (size_t)c * nface,
                  nc_kind.begin() + (size_t)(c + 1) * nface);
      other.insert(other.end(), nc_other.begin() + (size_t)c * nface * 4,
                   nc_other.begin() + (size_t)(c + 1) * nface * 4);
    }
    std::vector<double> nt(4 * N + 2 * N * N);
    double *X = nt.data(), *W = X + N, *L = W + N, *H = L + 2 * N;
    cx::gauss01(N, X, W);
    cx::lagrange(N, X, 0.0, L, nullptr);
    cx::lagrange(N, X, 1.0, L + N, nullptr);
    for (int b = 0; b < 2; ++b)
      for (int t = 0; t < N; ++t) cx::lagrange(N, X, 0.5 * (X[t] + b), H + (b * N + t) * N, nullptr);
    op->n_nc = (int)cells.size();
    op->n_hanging = n_hanging;
    if ((rc = upload(op, &op->d_nc_cells, cells.data(), cells.size())) ||
        (rc = upload(op, &op->d_nc_kind, kind.data(), kind.size())) ||
        (rc = upload(op, &op->d_nc_other, other.data(), other.size())) ||
        (rc = upload(op, &op->d_nctab, nt.data(), nt.size()))) {
      free_op(op);
      return rc;
    }
  }
  std::vector<double> tab, etab, ckt;
  CkTab ck{};
  build_tables(degree, tab, etab, ckt, ck);
  op->ck = ck;
  op->host_tab = tab;
  op->h_vertices.assign(vertices, vertices + (size_t)n_cells * nv * dim);
  op->h_neighbor.assign(neighbor, neighbor + (size_t)n_cells * nface);
  op->h_speed.assign(speed, speed + n_cells);
  op->h_nc_kind = std::move(nc_kind);
  op->h_nc_other = std::move(nc_other);
  if ((rc = upload(op, &op->d_tab, tab.data(), tab.size())) ||
      (rc = upload(op, &op->d_errtab, etab.data(), etab.size())) ||
      (rc = upload(op, &op->d_cktab, ckt.data(), ckt.size()))) {
    free_op(op);
    return rc;
  }
  {
    const std::vector<double> pt = build_post_table(degree);
    if ((rc = upload(op, &op->d_posttab, pt.data(), pt.size()))) {
      free_op(op);
      return rc;
    }
  }
  op->n_partial = 1024;
  if (hipMalloc((void **)&op->d_partial, sizeof(double) * 3 * op->n_partial) != hipSuccess ||
      hipMalloc((void **)&op->d_sum, sizeof(double) * 4) != hipSuccess) {
    free_op(op);
    return fail(EXW_ENOMEM, "exw_op_create: hipMalloc failed");
  }
  *out = op;
  return EXW_OK;
}

int exw_op_destroy(exw_op op) {
  free_op(op);
  return EXW_OK;
}

int64_t exw_op_n_dofs(exw_op op) { return op ? op->n_dofs() : -1; }
int exw_op_n_cells(exw_op op) { return op ? op->n_cells : -1; }
int exw_op_is_cartesian(exw_op op) { return op ? (int)op->cart : -1; }
double exw_op_min_vertex_distance_over_c(exw_op op) { return op ? op->minvd : -1.0; }

// one operator sweep with the library-internal epilogue (three outputs)
static int launch_epilogue(exw_op op, const double *src, const Epilogue &ep, int precision) {
  OpLaunch l = kernels(op->dim).op[op->degree][precision][op->cart ? 1 : 0];
  if (!l) return fail(EXW_EUNSUP, "operator kernel not compiled");
  EXW_HIP(l(op->view(), src, op->d_tab, op->host_tab.data(), ep, op->ctx->stream));
  if (op->n_nc > 0) {  // the hanging faces' subface integrals, added to the outputs
    NcLaunch ln = kernels(op->dim).nc[op->degree];
    if (!ln) return fail(EXW_EUNSUP, "hanging-face kernel not compiled");
    EXW_HIP(ln(op->view(), op->d_nctab, op->nc_args(), src, ep, op->ctx->stream));
  }
  return EXW_OK;
}

static int check_prec(int precision) {
  if (precision != EXW_FP64 && precision != EXW_FP32)
    return fail(EXW_EINVAL, "precision must be EXW_FP64 or EXW_FP32");
  return EXW_OK;
}

static int apply_fused(exw_op op, const double *src, const exw_epilogue *e, int precision,
                       bool rk_stage) {
  if (!op || !src || !e || !e->o1) return fail(EXW_EINVAL, "exw_op_apply_fused: null argument");
  if (int rc = check_prec(precision)) return rc;
  if ((const double *)e->o1 == src || (const double *)e->o2 == src)
    return fail(EXW_EINVAL, "exw_op_apply_fused: an output aliases src (neighbour reads race)");
  if (e->o2 && !e->o2_from_o1 && !e->c)
    return fail(EXW_EINVAL, "exw_op_apply_fused: o2 needs c or o2_from_o1");
  Epilogue ep{e->b, e->o1, e->a1, e->c, e->o2, e->a2, e->o2_from_o1, e->cell_scale};
  ep.rk_stage = rk_stage ? 1 : 0;
  return launch_epilogue(op, src, ep, precision);
}

int exw_op_apply_fused(exw_op op, const double *src, const exw_epilogue *e, int precision) {
  return apply_fused(op, src, e, precision, false);
}

int exw_op_apply(exw_op op, const double *src, double *dst, int precision) {
  exw_epilogue e{nullptr, dst, 1.0, nullptr, nullptr, 0.0, 0, nullptr};
  return exw_op_apply_fused(op, src, &e, precision);
}

int exw_rk_stage(exw_op op, const double *src, const double *b, double *y_out, double *b_out,
                 double f1, double f2, int is_last, int precision) {
  if (!b) return fail(EXW_EINVAL, "exw_rk_stage: b is NULL");
  if (!is_last && !b_out) return fail(EXW_EINVAL, "exw_rk_stage: b_out is NULL");
  exw_epilogue e{b, y_out, f1, nullptr, is_last ? nullptr : b_out, f2, 1, nullptr};
  return apply_fused(op, src, &e, precision, true);
}

}  // extern "C"

namespace {
// two entries per lane (16-byte accesses) where all three vectors are 16-byte
// aligned, streaming (non-temporal) loads and stores: each entry is touched once
__device__ __forceinline__ void rk_update_entry(double up, double f1, double f2, int is_last,
                                                double *v2, double *v3) {
  if (is_last) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(v3) + f1 * up, v2);
  } else {
    const double y = __builtin_nontemporal_load(v2) + f1 * up;
    __builtin_nontemporal_store(y, v2);
    __builtin_nontemporal_store(y + f2 * up, v3);
  }
}
typedef double dbl2 __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256)
    rk_update_kernel(int64_t n, double f1, double f2, int is_last, const double *__restrict__ v1,
                     double *__restrict__ v2, double *__restrict__ v3) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const bool pairs = ((uintptr_t)v1 | (uintptr_t)v2 | (uintptr_t)v3) % 16 == 0;
  const int64_t np = pairs ? n / 2 : 0;
  const dbl2 *p1 = reinterpret_cast<const dbl2 *>(v1);
  dbl2 *p2 = reinterpret_cast<dbl2 *>(v2), *p3 = reinterpret_cast<dbl2 *>(v3);
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < np; j += stride) {
    const dbl2 up = __builtin_nontemporal_load(p1 + j);
    if (is_last) {
      __builtin_nontemporal_store(__builtin_nontemporal_load(p3 + j) + f1 * up, p2 + j);
    } else {
      const dbl2 y = __builtin_nontemporal_load(p2 + j) + f1 * up;
      __builtin_nontemporal_store(y, p2 + j);
      __builtin_nontemporal_store(y + f2 * up, p3 + j);
    }
  }
  for (int64_t i = 2 * np + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    rk_update_entry(__builtin_nontemporal_load(v1 + i), f1, f2, is_last, v2 + i, v3 + i);
}
// out = sum_i coef[i] * in[i] over n_in <= 16 vectors (out may alias an input)
struct LinComb {
  const double *in[16];
  double coef[16];
  int n_in;
};
__global__ void __launch_bounds__(256)
    lincomb_kernel(int64_t n, const LinComb lc, double *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < lc.n_in; ++j) s += lc.coef[j] * __builtin_nontemporal_load(lc.in[j] + i);
    __builtin_nontemporal_store(s, out + i);
  }
}
}  // namespace

static int lincomb(exw_op op, const LinComb &lc, double *out) {
  const int64_t n = op->n_dofs();
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)blocks), dim3(256), 0, op->ctx->stream, n, lc,
                     out);
  EXW_HIP(hipGetLastError());
  return EXW_OK;
}

extern "C" {

int exw_rk_update(exw_ctx ctx, int64_t n, double factor1, double factor2, int is_last,
                  const double *vector1, double *vector2, double *vector3) {
  if (!ctx || !vector1 || !vector2 || !vector3 || n < 0)
    return fail(EXW_EINVAL, "exw_rk_update: null argument");
  if (n == 0) return EXW_OK;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(rk_update_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, n,
                     factor1, factor2, is_last, vector1, vector2, vector3);
  EXW_HIP(hipGetLastError());
  return EXW_OK;
}

int exw_ader_ck(exw_op op, const double *src, const double *recon, double t2, double t1,
                double te, double *tempsrc, int precision) {
  if (!op || !src || !tempsrc) return fail(EXW_EINVAL, "exw_ader_ck: null argument");
  if (int rc = check_prec(precision)) return rc;
  if (tempsrc == src || tempsrc == recon)
    return fail(EXW_EINVAL, "exw_ader_ck: tempsrc aliases an input");
  CkLaunch l = kernels(op->dim).ck[op->degree][precision][op->cart ? 1 : 0];
  if (!l) return fail(EXW_EUNSUP, "CK kernel not compiled");
  EXW_HIP(l(op->view(), src, recon, op->d_cktab, op->ck, t2, t1, te, nullptr, tempsrc,
             op->ctx->stream));
  return EXW_OK;
}

int exw_ader_step(exw_op op, const double *vec_n, double *vec_np, double *work, double dt,
                  int recon, int precision) {
  if (!op || !vec_n || !vec_np || !work) return fail(EXW_EINVAL, "exw_ader_step: null argument");
  const int64_t N = op->n_dofs();
  double *tempsrc = work, *rec = work + N;
  int rc;
  if (recon && (rc = exw_op_apply(op, vec_n, rec, precision))) return rc;
  if ((rc = exw_ader_ck(op, vec_n, recon ? rec : nullptr, dt, 0.0, 0.0, tempsrc, precision)))
    return rc;
  // vec_np = vec_n - (-M^-1 A tempsrc)  (source/time_integrators.cc:44-47 with the sign
  // flipped second sweep of source/wave_equation_operations.cc:1301-1342)
  exw_epilogue e{vec_n, vec_np, 1.0, nullptr, nullptr, 0.0, 0, nullptr};
  return exw_op_apply_fused(op, tempsrc, &e, precision);
}

int exw_ader_steps(exw_op op, int n_steps, double dt, double *vec, double *work,
                   int precision) {
  if (!op || !vec || !work) return fail(EXW_EINVAL, "exw_ader_steps: null argument");
  if (n_steps < 0) return fail(EXW_EINVAL, "exw_ader_steps: negative step count");
  if (int rc = check_prec(precision)) return rc;
  if (n_steps == 0) return EXW_OK;
  const int64_t N = op->n_dofs();
  double *w[2] = {work, work + N};
  // the update sweep vec += M^-1 A(w) in place (its epilogue reads b before
  // writing o1 at the same entry; the neighbours' w is read, never vec)
  auto sweep = [&](const double *wi) {
    exw_epilogue e{vec, vec, 1.0, nullptr, nullptr, 0.0, 0, nullptr};
    return exw_op_apply_fused(op, wi, &e, precision);
  };
  int rc;
  AderFusedLaunch fused = op->cart && op->n_nc == 0
                              ? kernels(op->dim).ader_fused[op->degree][precision]
                              : nullptr;
  if (!fused) {
    // one exw_ader_step per step, in place
    for (int s = 0; s < n_steps; ++s)
      if ((rc = exw_ader_ck(op, vec, nullptr, dt, 0.0, 0.0, w[0], precision)) ||
          (rc = sweep(w[0])))
        return rc;
    return EXW_OK;
  }
  if ((rc = exw_ader_ck(op, vec, nullptr, dt, 0.0, 0.0, w[0], precision))) return rc;
  for (int s = 1; s < n_steps; ++s)
    EXW_HIP(fused(op->view(), w[(s - 1) & 1], vec, op->d_cktab, op->ck, dt, w[s & 1],
                  op->ctx->stream));
  return sweep(w[(n_steps - 1) & 1]);
}

int exw_time_step_work_vectors(int integrator) {
  switch (integrator) {
    case EXW_EXPL_EULER: return 0;
    case EXW_CLRK4: return 2;
    case EXW_LSRK45R2:
    case EXW_LSRK33R2:
    case EXW_LSRK59R2: return 1;
    case EXW_ADER: return 2;
    case EXW_LSRK45R3: return 2;
    case EXW_SSPRK: return 14;
    case EXW_ADER_ADCONFULL: return 2;
    default: return -1;
  }
}

// Shu-Osher coefficients of the optimal SSPRK with 8 stages and order 4
// (S. J. Ruuth, Math. Comp. 75 (2006) 183-207) as the reference tabulates them
// (SSPRK::initialize_coeffs, source/time_integrators.cc:726-858): stage s is
// sum_{l<s} A[s-1][l] u_l + dt B[s-1][l] F(u_l).
static const double kSspA[8][8] = {
    {1.},
    {0.538569155333175, 0.461430844666825},
    {0.004485387460763, 0., 0.995514612539237},
    {0.164495299288580, 0.016875060685979, 0., 0.818629640025440},
    {0.426933682982668, 0.157047028197878, 0.023164224070770, 0., 0.392855064748685},
    {0.082083400476958, 0.000000039091042, 0.033974171137350, 0.005505195713107, 0.,
     0.878437193581543},
    {0.006736365648625, 0.010581829625529, 0.009353386191951, 0.101886062556838,
     0.000023428364930, 0., 0.871418927612128},
    {0.071115287415749, 0.018677648343953, 0.007902408660034, 0.319384027162348,
     0.007121989995845, 0.001631615692736, 0., 0.574167022729334}};
static const double kSspB[8][8] = {
    {0.282318339066479},
    {0., 0.130270389660380},
    {0.003963092203460, 0., 0.281052031928487},
    {0.000038019518678, 0.004764139104512, 0., 0.231114160282572},
    {0.000019921336144, 0.044337256156151, 0.006539685265423, 0., 0.110910189373703},
    {0.000000034006679, 0.000000011036118, 0.009591531566657, 0.001554217709960, 0.,
     0.247998929466160},
    {0.013159891155054, 0.002987444564164, 0.002640632454359, 0.028764303955070,
     0.000006614257074, 0., 0.246017544274548},
    {0.000000010647874, 0.005273042658132, 0.002230994887525, 0.090167968072837,
     0.002010668386475, 0.000460635032368, 0., 0.162097880203691}};

// SSPRK(8,4): stage s first combines every known term (u_0..u_{s-1},
// F_0..F_{s-2}) into its output vector, then the sweep over u_{s-1} stores
// F_{s-1} and adds dt B[s-1][s-1] F_{s-1} to that output in its epilogue; the
// last stage adds into vec_np and stores no F.  work: u_1..u_7, F_0..F_6.
static int ssprk84(exw_op op, double dt, const double *vec_n, double *vec_np, double *work,
                   int precision) {
  const int64_t N = op->n_dofs();
  const double *u[8];
  double *F[7];
  u[0] = vec_n;
  for (int l = 1; l < 8; ++l) u[l] = work + (l - 1) * N;
  for (int l = 0; l < 7; ++l) F[l] = work + (7 + l) * N;
  for (int st = 1; st <= 8; ++st) {
    double *out = st < 8 ? const_cast<double *>(u[st]) : vec_np;
    const double *c = out;
    if (st == 1) {
      c = vec_n;  // u_1 = u_0 + dt B[0][0] F_0
    } else {
      LinComb lc{};
      for (int l = 0; l < st; ++l) {
        if (kSspA[st - 1][l] != 0.0) {
          lc.in[lc.n_in] = u[l];
          lc.coef[lc.n_in++] = kSspA[st - 1][l];
        }
        if (l < st - 1 && kSspB[st - 1][l] != 0.0) {
          lc.in[lc.n_in] = F[l];
          lc.coef[lc.n_in++] = dt * kSspB[st - 1][l];
        }
      }
      if (int rc = lincomb(op, lc, out)) return rc;
    }
    Epilogue ep{};
    if (st < 8) {
      ep.o1 = F[st - 1];  // F_{s-1} = k
      ep.a1 = 1.0;
      ep.c = c;
      ep.o2 = out;
      ep.a2 = dt * kSspB[st - 1][st - 1];
    } else {
      ep.b = out;
      ep.o1 = out;
      ep.a1 = dt * kSspB[7][7];
    }
    if (int rc = launch_epilogue(op, u[st - 1], ep, precision)) return rc;
  }
  return EXW_OK;
}

// Low-storage RK with the RKVectorUpdater fused into each operator sweep
// (register schedule in lsrk.hpp).
static int lsrk_fused(exw_op op, int S, const double *a, const double *b, double dt,
                      double *vec_n, double *vec_np, double *work, int precision) {
  const int rc = lsrk_fused_schedule(
      S, a, b, dt, vec_n, vec_np, work,
      [&](const double *src, const double *acc, double *y_out, double *b_out, double f1,
          double f2, bool last) {
        return exw_rk_stage(op, src, acc, y_out, b_out, f1, f2, last ? 1 : 0, precision);
      });
  if (rc == -1) return fail(EXW_EINVAL, "internal: LSRK register plan failed");
  return rc;
}

int exw_time_step(exw_op op, int integrator, double dt, double *vec_n, double *vec_np,
                  double *work, int ader_recon, int precision) {
  if (!op || !vec_n || !vec_np) return fail(EXW_EINVAL, "exw_time_step: null argument");
  if (vec_n == vec_np) return fail(EXW_EINVAL, "exw_time_step: vec_n aliases vec_np");
  const int64_t N = op->n_dofs();
  switch (integrator) {
    case EXW_EXPL_EULER: {
      // source/time_integrators.cc:24-34: vec_np = vec_n + dt*M^-1 A vec_n
      exw_epilogue e{vec_n, vec_np, dt, nullptr, nullptr, 0.0, 0, nullptr};
      return exw_op_apply_fused(op, vec_n, &e, precision);
    }
    case EXW_CLRK4: {
      // source/time_integrators.cc:66-102 with the vector updates fused
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      double *ta = work, *tb = work + N;
      const double wgt[4] = {dt / 6., dt / 3., dt / 3., dt / 6.};
      const double cf[3] = {0.5 * dt, 0.5 * dt, dt};
      const double *src = vec_n;
      double *nxt = ta;
      for (int s = 0; s < 4; ++s) {
        exw_epilogue e{s == 0 ? vec_n : vec_np, vec_np, wgt[s], vec_n, s < 3 ? nxt : nullptr,
                       s < 3 ? cf[s] : 0.0, 0, nullptr};
        if (int rc = exw_op_apply_fused(op, src, &e, precision)) return rc;
        src = nxt;
        nxt = nxt == ta ? tb : ta;
      }
      return EXW_OK;
    }
    case EXW_LSRK45R2: {
      // source/time_integrators.cc:235-283
      const double a[5] = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
                           2251764453980. / 15575788980749., 26877169314380. / 34165994151039.,
                           0.0};
      const double b[5] = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
                           -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
                           5198255086312. / 14908931495163.};
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      return lsrk_fused(op, 5, a, b, dt, vec_n, vec_np, work, precision);
    }
    case EXW_LSRK33R2: {
      // source/time_integrators.cc:200-231
      const double a[3] = {0.755726351946097, 0.386954477304099, 0.0};
      const double b[3] = {0.245170287303492, 0.184896052186740, 0.569933660509768};
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      return lsrk_fused(op, 3, a, b, dt, vec_n, vec_np, work, precision);
    }
    case EXW_LSRK59R2: {
      // source/time_integrators.cc:286-363
      const double a[9] = {1107026461565. / 5417078080134.,   38141181049399. / 41724347789894.,
                           493273079041. / 11940823631197.,   1851571280403. / 6147804934346.,
                           11782306865191. / 62590030070788., 9452544825720. / 13648368537481.,
                           4435885630781. / 26285702406235.,  2357909744247. / 11371140753790.,
                           0.0};
      const double b[9] = {2274579626619. / 23610510767302., 693987741272. / 12394497460941.,
                           -347131529483. / 15096185902911., 1144057200723. / 32081666971178.,
                           1562491064753. / 11797114684756., 13113619727965. / 44346030145118.,
                           393957816125. / 7825732611452.,   720647959663. / 6565743875477.,
                           3559252274877. / 14424734981077.};
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      return lsrk_fused(op, 9, a, b, dt, vec_n, vec_np, work, precision);
    }
    case EXW_ADER:
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      return exw_ader_step(op, vec_n, vec_np, work, dt, ader_recon, precision);
    case EXW_LSRK45R3: {
      // LowStorageRK45Reg3 (source/time_integrators.cc:366-425).  With
      // y_s = u + dt sum_{i<=s} b_i k_i the stage inputs are
      // x_{s+1} = y_{s-2} + dt (a_{s+1,s-1} k_{s-1} + a_{s+1,s} k_s); each sweep
      // (k_s in registers) writes x_{s+1} = z_{s-1} + dt a_{s+1,s} k_s, the
      // known part of the stage after next z_s = y_{s-1} + dt a_{s+2,s} k_s,
      // and y_s = y_{s-1} + dt b_s k_s (kept in vec_np).  Registers: vec_n
      // (u, later a stage input), work[0] (stage inputs), work[1] (z).
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      const double a21 = 2365592473904. / 8146167614645., a32 = 4278267785271. / 6823155464066.,
                   a43 = 2789585899612. / 8986505720531., a54 = 15310836689591. / 24358012670437.,
                   a31 = -722262345248. / 10870640012513., a42 = 1365858020701. / 8494387045469.,
                   a53 = 3819021186. / 2763618202291.;
      const double b1 = 846876320697. / 6523801458457., b2 = 3032295699695. / 12397907741132.,
                   b3 = 612618101729. / 6534652265123., b4 = 1155491934595. / 2954287928812.,
                   b5 = 707644755468. / 5028292464395.;
      double *xa = work, *z = work + N, *xb = vec_n;
      int rc;
      Epilogue e1{};  // stage 1: x2, z1 = u + dt a31 k1, y1 = u + dt b1 k1
      e1.b = vec_n; e1.o1 = xa; e1.a1 = dt * a21;
      e1.c = vec_n; e1.o2 = z; e1.a2 = dt * a31; e1.o3 = vec_np; e1.a3 = dt * b1;
      if ((rc = launch_epilogue(op, vec_n, e1, precision))) return rc;
      Epilogue e2{};  // stage 2: x3 = z1 + dt a32 k2, z2 = y1 + dt a42 k2, y2
      e2.b = z; e2.o1 = xb; e2.a1 = dt * a32;
      e2.c = vec_np; e2.o2 = z; e2.a2 = dt * a42; e2.o3 = vec_np; e2.a3 = dt * b2;
      if ((rc = launch_epilogue(op, xa, e2, precision))) return rc;
      Epilogue e3{};  // stage 3: x4 = z2 + dt a43 k3, z3 = y2 + dt a53 k3, y3
      e3.b = z; e3.o1 = xa; e3.a1 = dt * a43;
      e3.c = vec_np; e3.o2 = z; e3.a2 = dt * a53; e3.o3 = vec_np; e3.a3 = dt * b3;
      if ((rc = launch_epilogue(op, xb, e3, precision))) return rc;
      Epilogue e4{};  // stage 4: x5 = z3 + dt a54 k4, y4
      e4.b = z; e4.o1 = xb; e4.a1 = dt * a54;
      e4.c = vec_np; e4.o2 = vec_np; e4.a2 = dt * b4;
      if ((rc = launch_epilogue(op, xa, e4, precision))) return rc;
      Epilogue e5{};  // stage 5: vec_np = y4 + dt b5 k5
      e5.b = vec_np; e5.o1 = vec_np; e5.a1 = dt * b5;
      return launch_epilogue(op, xb, e5, precision);
    }
    case EXW_SSPRK:
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      return ssprk84(op, dt, vec_n, vec_np, work, precision);
    case EXW_ADER_ADCONFULL: {
      // WaveEquationOperationADERADCONFULL::apply_ader
      // (source/wave_equation_operations.cc:689-735) with ArbitraryHighOrderDG
      // (source/time_integrators.cc:37-50): vec_np = sum_{j=0}^{k+1} (dt A)^j/j! vec_n.
      // Sweep j reads w_{j-1} (w_0 = vec_n) and writes w_j = (dt/j) A w_{j-1} and
      // vec_np += w_j in its epilogue; w alternates between the two work vectors.
      if (!work) return fail(EXW_EINVAL, "exw_time_step: work is NULL");
      const int J = op->degree + 1;
      const double *w = vec_n;
      for (int j = 1; j <= J; ++j) {
        double *wn = work + ((j - 1) & 1) * N;
        Epilogue e{};
        if (j < J) {
          e.o1 = wn;
          e.a1 = dt / j;
          e.c = j == 1 ? vec_n : vec_np;
          e.o2 = vec_np;
          e.a2 = dt / j;
        } else {
          e.b = J == 1 ? vec_n : vec_np;
          e.o1 = vec_np;
          e.a1 = dt / j;
        }
        if (int rc = launch_epilogue(op, w, e, precision)) return rc;
        w = wn;
      }
      return EXW_OK;
    }
    default:
      return fail(EXW_EINVAL, "unknown time integrator " + std::to_string(integrator) +
                                  " requested");
  }
}

static DevModes to_dev(const exw_modes *m) {
  DevModes d{};
  d.n_modes = m->n_modes;
  for (int k = 0; k < 8; ++k) {
    d.amp[k] = m->amp[k];
    for (int j = 0; j < 3; ++j) d.m[k][j] = m->m[k][j];
  }
  d.rho = m->rho;
  d.c = m->c;
  return d;
}

int exw_project(exw_op op, const exw_modes *modes, double t, double *dst) {
  if (!op || !modes || !dst) return fail(EXW_EINVAL, "exw_project: null argument");
  if (modes->n_modes < 0 || modes->n_modes > 8) return fail(EXW_EINVAL, "n_modes must be 0..8");
  EXW_HIP(kernels(op->dim).project[op->degree](op->view(), op->d_tab, to_dev(modes), t, dst,
                                                op->ctx->stream));
  return EXW_OK;
}

int exw_l2_errors(exw_op op, const exw_modes *modes, double t, const double *u,
                  double *out_host) {
  if (!op || !modes || !u || !out_host) return fail(EXW_EINVAL, "exw_l2_errors: null argument");
  if (modes->n_modes < 0 || modes->n_modes > 8) return fail(EXW_EINVAL, "n_modes must be 0..8");
  const int nb = op->n_partial;
  EXW_HIP(kernels(op->dim).l2err[op->degree](op->view(), op->d_errtab, to_dev(modes), t, u,
                                              op->d_partial, nb, op->ctx->stream));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, op->ctx->stream, op->d_partial,
                     nb, 3, op->d_sum);
  EXW_HIP(hipGetLastError());
  double s[3];
  EXW_HIP(hipMemcpyAsync(s, op->d_sum, sizeof(s), hipMemcpyDeviceToHost, op->ctx->stream));
  EXW_HIP(hipStreamSynchronize(op->ctx->stream));
  for (int k = 0; k < 3; ++k) out_host[k] = std::sqrt(s[k]);
  return EXW_OK;
}

int exw_l2_errors_noalloc(exw_op op, const exw_modes *modes, double t, const double *u,
                          double *out_dev) {
  if (!op || !modes || !u || !out_dev)
    return fail(EXW_EINVAL, "exw_l2_errors_noalloc: null argument");
  if (modes->n_modes < 0 || modes->n_modes > 8) return fail(EXW_EINVAL, "n_modes must be 0..8");
  const int nb = op->n_partial;  // the block count of exw_l2_errors: same sums
  if (nb < 1 || nb > EXW_NOALLOC_MAX_PARTIALS)
    return fail(EXW_EINVAL, "exw_l2_errors_noalloc: " + std::to_string(nb) +
                                " partial blocks exceed the scratch array (" +
                                std::to_string(EXW_NOALLOC_MAX_PARTIALS) + ")");
  double *scr = op->ctx->scratch;
  if (!scr) return fail(EXW_EDEVICE, "exw_l2_errors_noalloc: no scratch on this context");
  // every one of the 3*nb partial slots is written by its block, then summed
  // in the fixed order of exw_l2_errors into scr[3*nb .. 3*nb+2]
  EXW_HIP(kernels(op->dim).l2err[op->degree](op->view(), op->d_errtab, to_dev(modes), t, u, scr,
                                              nb, op->ctx->stream));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, op->ctx->stream, scr, nb, 3,
                     scr + 3 * nb);
  EXW_HIP(hipGetLastError());
  hipLaunchKernelGGL(sqrt3_kernel, dim3(1), dim3(64), 0, op->ctx->stream, scr + 3 * nb, out_dev);
  EXW_HIP(hipGetLastError());
  return EXW_OK;
}

}  // extern "C"

namespace {
// host multilinear map of cell corners V at xi: point x and Jacobian inverse
// (Ji[e*dim+i] = dxi_e/dx_i); returns det J
double host_map(int dim, const double *V, const double *xi, double *x, double *Ji) {
  const int nv = 1 << dim;
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < dim; ++i) x[i] = 0.0;
  for (int v = 0; v < nv; ++v) {
    double w = 1.0;
    for (int d = 0; d < dim; ++d) w *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
    for (int i = 0; i < dim; ++i) x[i] += w * V[v * dim + i];
    for (int e = 0; e < dim; ++e) {
      double g = ((v >> e) & 1) ? 1.0 : -1.0;
      for (int d = 0; d < dim; ++d)
        if (d != e) g *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
      for (int i = 0; i < dim; ++i) J[i * dim + e] += g * V[v * dim + i];
    }
  }
  double det;
  if (dim == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / det;
    Ji[1] = -J[1] / det;
    Ji[2] = -J[2] / det;
    Ji[3] = J[0] / det;
  } else {
    const double A = J[4] * J[8] - J[5] * J[7], B = J[5] * J[6] - J[3] * J[8],
                 C = J[3] * J[7] - J[4] * J[6];
    det = J[0] * A + J[1] * B + J[2] * C;
    Ji[0] = A / det;
    Ji[3] = B / det;
    Ji[6] = C / det;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) / det;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) / det;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) / det;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) / det;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) / det;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) / det;
  }
  return det;
}

// reference coordinates of x0 in the cell by Newton's method; true if they lie
// in the closed unit cell (1e-10)
bool locate(int dim, const double *V, const double *x0, double *xi) {
  const int nv = 1 << dim;
  double h = 0.0, x[3], Ji[9];
  for (int d = 0; d < dim; ++d) {
    xi[d] = 0.5;
    h = std::max(h, std::fabs(V[(nv - 1) * dim + d] - V[d]));
  }
  for (int it = 0; it < 60; ++it) {
    host_map(dim, V, xi, x, Ji);
    double rn = 0.0, r[3];
    for (int i = 0; i < dim; ++i) {
      r[i] = x[i] - x0[i];
      rn = std::max(rn, std::fabs(r[i]));
    }
    for (int e = 0; e < dim; ++e) {
      double dx = 0.0;
      for (int i = 0; i < dim; ++i) dx += Ji[e * dim + i] * r[i];
      xi[e] -= dx;
      if (!(std::fabs(xi[e]) < 10.0)) return false;
    }
    if (rn <= 1e-14 * h) break;
  }
  host_map(dim, V, xi, x, Ji);
  for (int i = 0; i < dim; ++i)
    if (std::fabs(x[i] - x0[i]) > 1e-12 * h) return false;
  for (int d = 0; d < dim; ++d)
    if (xi[d] < -1e-10 || xi[d] > 1.0 + 1e-10) return false;
  return true;
}

double ricker_integral(const exw_source &s, double t1, double t2) {
  const double a = (M_PI * s.f0) * (M_PI * s.f0);
  const double u1 = t1 - s.t0, u2 = t2 - s.t0;
  return s.amplitude * (u2 * std::exp(-a * u2 * u2) - u1 * std::exp(-a * u1 * u1));
}
}  // namespace

namespace exw {
// u += (int_{t1}^{t2} r) b on the operator's source cells selected by sel
int source_add_selected(exw_op op, double t1, double t2, unsigned sel, double *u) {
  if (op->n_src == 0 || sel == 0) return EXW_OK;
  int nd = 1;
  for (int d = 0; d < op->dim; ++d) nd *= op->degree + 1;
  const double I = ricker_integral(op->src, t1, t2);
  const int total = op->n_src * nd;
  hipLaunchKernelGGL(source_add_kernel, dim3((total + 255) / 256), dim3(256), 0,
                     op->ctx->stream, u, op->d_src_cells, op->d_src_b, op->n_src, nd, op->dim,
                     sel, I);
  EXW_HIP(hipGetLastError());
  return EXW_OK;
}
}  // namespace exw

extern "C" {

int exw_op_set_source(exw_op op, const exw_source *src) {
  if (!op) return fail(EXW_EINVAL, "exw_op_set_source: null operator");
  hipFree(op->d_src_cells);
  hipFree(op->d_src_b);
  op->d_src_cells = nullptr;
  op->d_src_b = nullptr;
  op->n_src = 0;
  op->h_src_cells.clear();
  if (!src) return EXW_OK;
  if (!(src->f0 > 0)) return fail(EXW_EINVAL, "exw_op_set_source: f0 must be positive");
  const int dim = op->dim, nv = 1 << dim, N = op->degree + 1;
  int nd = 1;
  for (int d = 0; d < dim; ++d) nd *= N;
  std::vector<int> cells;
  std::vector<std::array<double, 3>> xis;
  for (int c = 0; c < op->n_cells && cells.size() < 8; ++c) {
    std::array<double, 3> xi{0, 0, 0};
    if (locate(dim, op->h_vertices.data() + (size_t)c * nv * dim, src->x, xi.data())) {
      cells.push_back(c);
      xis.push_back(xi);
    }
  }
  if (cells.empty()) return fail(EXW_EINVAL, "exw_op_set_source: no cell holds the point");
  const int ns = (int)cells.size();
  double X[16], W[16], l[3][16];
  cx::gauss01(N, X, W);
  std::vector<double> b((size_t)ns * nd);
  for (int k = 0; k < ns; ++k) {
    const double *V = op->h_vertices.data() + (size_t)cells[k] * nv * dim;
    for (int d = 0; d < dim; ++d) cx::lagrange(N, X, xis[k][d], l[d], nullptr);
    for (int q = 0; q < nd; ++q) {
      double xi[3], x[3], Ji[9], w = 1.0, phi = 1.0;
      int t = q;
      for (int d = 0; d < dim; ++d) {
        xi[d] = X[t % N];
        w *= W[t % N];
        phi *= l[d][t % N];
        t /= N;
      }
      const double det = host_map(dim, V, xi, x, Ji);
      b[(size_t)k * nd + q] = phi / ns / (w * det);
    }
  }
  int rc;
  if ((rc = upload(op, &op->d_src_cells, cells.data(), cells.size())) ||
      (rc = upload(op, &op->d_src_b, b.data(), b.size())))
    return rc;
  op->n_src = ns;
  op->src = *src;
  op->h_src_cells = cells;
  return EXW_OK;
}

int exw_source_add(exw_op op, double t1, double t2, double *u) {
  if (!op || !u) return fail(EXW_EINVAL, "exw_source_add: null argument");
  if (op->n_src == 0) return fail(EXW_EINVAL, "exw_source_add: the operator has no source");
  return source_add_selected(op, t1, t2, (1u << op->n_src) - 1u, u);
}

int exw_estimate_error(exw_op op, const double *solution, double *tmp, double *error_per_cell) {
  if (!op || !solution || !tmp || !error_per_cell)
    return fail(EXW_EINVAL, "exw_estimate_error: null argument");
  if (tmp == solution) return fail(EXW_EINVAL, "exw_estimate_error: tmp aliases solution");
  // apply(solution, tmp_vector) first (source/wave_equation_operations.cc:876)
  EstLaunch le = kernels(op->dim).estimate[op->degree];
  if (!le) return fail(EXW_EUNSUP, "error estimator not compiled for this degree");
  if (int rc = exw_op_apply(op, solution, tmp, EXW_FP64)) return rc;
  EXW_HIP(le(op->view(), op->d_tab, solution, tmp, error_per_cell, op->ctx->stream));
  return EXW_OK;
}

int64_t exw_post_n_values(exw_op op) { return op ? op->n_post() : -1; }

int exw_post_pressure(exw_op op, const double *solution, double *tmp, double *post,
                      int *iterations) {
  if (!op || !solution || !tmp || !post)
    return fail(EXW_EINVAL, "exw_post_pressure: null argument");
  if (tmp == solution) return fail(EXW_EINVAL, "exw_post_pressure: tmp aliases solution");
  PostLaunch l = kernels(op->dim).post[op->degree];
  if (!l) return fail(EXW_EUNSUP, "post-processing kernel not compiled");
  // compute_post_pressure calls apply(solution, tmp_vector) first (:795)
  if (int rc = exw_op_apply(op, solution, tmp, EXW_FP64)) return rc;
  int np = 1;
  for (int d = 0; d < op->dim; ++d) np *= op->degree + 2;
  const PostSolver cfg{EXW_POST_ABS_TOL, EXW_POST_REL_TOL, EXW_POST_MAX_ITER_FACTOR * np};
  EXW_HIP(l(op->view(), op->d_posttab, solution, tmp, post, iterations, cfg, op->ctx->stream));
  return EXW_OK;
}

int exw_post_error(exw_op op, const exw_modes *modes, double t, const double *post,
                   double *out_host) {
  if (!op || !modes || !post || !out_host) return fail(EXW_EINVAL, "exw_post_error: null argument");
  if (modes->n_modes < 0 || modes->n_modes > 8) return fail(EXW_EINVAL, "n_modes must be 0..8");
  PostErrLaunch l = kernels(op->dim).post_error[op->degree];
  if (!l) return fail(EXW_EUNSUP, "post-processing error kernel not compiled");
  const int nb = op->n_partial;
  EXW_HIP(l(op->view(), op->d_posttab, to_dev(modes), t, post, op->d_partial, nb,
            op->ctx->stream));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, op->ctx->stream, op->d_partial,
                     nb, 1, op->d_sum);
  EXW_HIP(hipGetLastError());
  double s = 0.0;
  EXW_HIP(hipMemcpyAsync(&s, op->d_sum, sizeof(double), hipMemcpyDeviceToHost, op->ctx->stream));
  EXW_HIP(hipStreamSynchronize(op->ctx->stream));
  *out_host = std::sqrt(s);
  return EXW_OK;
}

int exw_energy(exw_op op, const double *u, double *out_host) {
  if (!op || !u || !out_host) return fail(EXW_EINVAL, "exw_energy: null argument");
  const int nb = op->n_partial;
  EXW_HIP(kernels(op->dim).energy[op->degree](op->view(), op->d_tab, u, op->d_partial, nb,
                                               op->ctx->stream));
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, op->ctx->stream, op->d_partial,
                     nb, 1, op->d_sum);
  EXW_HIP(hipGetLastError());
  EXW_HIP(hipMemcpyAsync(out_host, op->d_sum, sizeof(double), hipMemcpyDeviceToHost,
                         op->ctx->stream));
  EXW_HIP(hipStreamSynchronize(op->ctx->stream));
  return EXW_OK;
}

}  // extern "C"
