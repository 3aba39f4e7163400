// exwave_lts.hip -- ADER with local time stepping on the device
// (WaveEquationOperationADERLTS, source/wave_equation_operations.cc:
// 1385-1692; ClusterManager, include/cluster_manager.templates.h:24-1049).
//
// Host side: the cluster categorisation and the update schedule are computed
// once; every cell set a cycle touches depends only on the cluster being
// updated, so the cell lists (with per-cell face masks and the target of each
// face sum) are built once per cluster and uploaded.  A time step is then a
// fixed sequence of list-driven launches on the context's stream:
//   per update cycle (cluster ac):
//     for the faster and the slower neighbour cluster, when their overlap
//     interval is not empty: CK time integrals of the cells on both sides of
//     the cluster boundary, then the face sums (ac's side -> face sum, the
//     other side -> its flux memory);
//     CK over ac's own interval of ac's cells, the faces inside ac and on the
//     boundary;
//     CK over [0, dt] of ac's cells and the second sweep (state update);
//   with the reconstruction (use_ader_post): the state is copied, the cells of
//   the neighbouring clusters next to ac are brought to ac's new time level
//   in the copy (no flux memory), and M^-1 A of the copy is evaluated on ac's
//   cells into the reconstruction vector.
// The reference also updates further cells of the neighbouring cluster in the
// copy and subtracts stray face terms from cells two layers away from ac
// (include/cluster_manager.templates.h:631-647, 1046); none of those values is
// read afterwards, so these launches leave them out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace exw;

namespace {

int lfail(int code, const std::string &msg) {
  set_last_error(msg);
  return code;
}
#define LHIP(call)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return lfail(EXW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

const DimKernels &kernels(int dim) { return dim == 2 ? kernels_2d() : kernels_3d(); }

double cell_min_vertex_distance(int dim, const double *V) {
  const int nv = 1 << dim;
  double mind = 1e300;
  for (int a = 0; a < nv; ++a)
    for (int b = a + 1; b < nv; ++b) {
      double s = 0.0;
      for (int d = 0; d < dim; ++d) {
        const double df = V[a * dim + d] - V[b * dim + d];
        s += df * df;
      }
      mind = std::min(mind, s);
    }
  return std::sqrt(mind);
}

struct List {
  int off = 0, n = 0;
};
// the hanging-face entries of a face pass, by target (0: face sum, 1: flux
// memory / the reconstruction output)
struct Pass {
  List ck, face, nc[2];
};
// launches of one update_elements call for one cluster and one update set
struct UpdatePlan {
  Pass side[2];  // 0: toward the faster cluster (ac-1), 1: toward the slower (ac+1)
  Pass same;
  List upd;
  unsigned src_sel = 0;  // source cells (bit s) among the update cells
  // Box meshes: the update cells whose neighbours all belong to the cluster
  // (no hanging face) take the plain operator sweep, state += M^-1 A(CK), as
  // global ADER does; same.face and upd restricted to the others
  List plain, rest_face, rest_upd;
};
struct ClusterPlan {
  UpdatePlan main;
  UpdatePlan recon[2];  // neighbour cluster ac-1 / ac+1 brought to ac's time level
  List apply, apply_nc;  // ac's cells (reconstruction)
};

// One neighbour relation of a cell: face f, subface slot j (the fine cell of
// subface j seen from the coarse side of a 2:1 face; 0 otherwise), the cell
// across (-1 at a boundary).  Mask bit 4f+j addresses it.
struct Slot {
  int f, j, nb;
};

// per-cell neighbour slots in face order, CSR
void build_slots(int dim, int n, const int *nbr, const int *nc_kind, const int *nc_other,
                 std::vector<int> &off, std::vector<Slot> &slots) {
  const int nface = 2 * dim, nsub = 1 << (dim - 1);
  off.assign(n + 1, 0);
  slots.clear();
  for (int c = 0; c < n; ++c) {
    off[c] = (int)slots.size();
    for (int f = 0; f < nface; ++f) {
      const int nb = nbr[(size_t)c * nface + f];
      if (nb >= 0) {
        slots.push_back({f, 0, nb});
      } else if (nb == EXW_NC_FACE && nc_kind) {
        const int kd = nc_kind[(size_t)c * nface + f];
        const int *o = nc_other + ((size_t)c * nface + f) * 4;
        if (kd == 1)
          slots.push_back({f, 0, o[0]});
        else
          for (int b = 0; b < nsub; ++b) slots.push_back({f, b, o[b]});
      } else {
        slots.push_back({f, 0, -1});
      }
    }
  }
  off[n] = (int)slots.size();
}

// host staging of the concatenated lists
struct ListBuilder {
  std::vector<int> cells;
  std::vector<unsigned> mask;
  std::vector<unsigned char> target;
  List add(const std::vector<int> &c, const std::vector<unsigned> &m,
           const std::vector<unsigned char> &t) {
    List l;
    l.off = (int)cells.size();
    l.n = (int)c.size();
    cells.insert(cells.end(), c.begin(), c.end());
    mask.insert(mask.end(), m.begin(), m.end());
    target.insert(target.end(), t.begin(), t.end());
    return l;
  }
};
// hanging-face entries: cell, its face kinds and partners (NcArgs rows), mask
struct NcBuilder {
  std::vector<int> cells, kind, other;
  std::vector<unsigned> mask;
  List add(exw_op op, const std::vector<int> &c, const std::vector<unsigned> &m) {
    const int nface = 2 * op->dim;
    List l;
    l.off = (int)cells.size();
    l.n = (int)c.size();
    for (size_t i = 0; i < c.size(); ++i) {
      cells.push_back(c[i]);
      kind.insert(kind.end(), op->h_nc_kind.begin() + (size_t)c[i] * nface,
                  op->h_nc_kind.begin() + (size_t)(c[i] + 1) * nface);
      other.insert(other.end(), op->h_nc_other.begin() + (size_t)c[i] * nface * 4,
                   op->h_nc_other.begin() + (size_t)(c[i] + 1) * nface * 4);
      mask.push_back(m[i]);
    }
    return l;
  }
};

// iterate_cluster_categorization (include/cluster_manager.templates.h:24-262)
// over neighbour slots: cells visited in order, ids lowered in place;
// speed == NULL is the reference's step cfl * min vertex distance, otherwise
// it is divided by the cell's wave speed
int categorize(int dim, int n, const double *vertices, const double *speed,
               const std::vector<int> &off, const std::vector<Slot> &slots, double cfl,
               int max_clusters, int max_diff, int *cluster, int *flags, int *n_clusters,
               int *cluster_diff, int *iterations) {
  const int nv = 1 << dim;
  std::vector<double> ts(n);
  double minimal = 1e300;
  for (int c = 0; c < n; ++c) {
    ts[c] = cfl * cell_min_vertex_distance(dim, vertices + (size_t)c * nv * dim);
    if (speed) ts[c] /= speed[c];
    minimal = std::min(minimal, ts[c]);
  }
  const double eps = minimal / 10.0;
  int nc = 0;
  for (int c = 0; c < n; ++c) {
    cluster[c] = (int)((ts[c] - eps) / minimal);
    nc = std::max(nc, cluster[c]);
  }
  ++nc;
  int diff = 1;
  if (nc > max_clusters) {
    diff = std::min(nc / max_clusters + 1, max_diff);
    for (int c = 0; c < n; ++c) cluster[c] /= diff;
  }
  // neighbouring cells must differ by at most one cluster and no cell may have
  // both a faster and a slower neighbour: lower the slower side in passes over
  // the clusters, in cell order and in place (:100-242)
  bool again = true;
  int count = 0;
  while (again && count < 100) {
    ++count;
    for (int x = 0; x < nc; ++x)
      for (int c = 0; c < n; ++c) {
        if (cluster[c] != x) continue;
        bool faster = false;
        for (int i = off[c]; i < off[c + 1]; ++i)
          if (slots[i].nb >= 0 && cluster[slots[i].nb] < x) faster = true;
        for (int i = off[c]; i < off[c + 1]; ++i) {
          const int nb = slots[i].nb;
          if (nb >= 0 && cluster[nb] > x) cluster[nb] = faster ? x : x + 1;
        }
      }
    again = false;
    for (int c = 0; c < n; ++c) {
      const int cur = cluster[c];
      int fl = 0;
      for (int i = off[c]; i < off[c + 1]; ++i) {
        const int nb = slots[i].nb;
        if (nb < 0) continue;
        if (cluster[nb] < cur)
          fl |= 2;
        else if (cluster[nb] > cur)
          fl |= 1;
        if (cluster[nb] < cur - 1 || cluster[nb] > cur + 1) again = true;
      }
      if (fl == 3) again = true;
      flags[c] = fl;
    }
  }
  if (again) return lfail(EXW_EINVAL, "could not find suited categories in 100 iterations");
  nc = 0;
  for (int c = 0; c < n; ++c) nc = std::max(nc, cluster[c]);
  if (n_clusters) *n_clusters = nc + 1;
  if (cluster_diff) *cluster_diff = diff;
  if (iterations) *iterations = count;
  return EXW_OK;
}

}  // namespace

struct exw_lts_s {
  exw_op op;
  int recon;
  int n_clusters = 0, cluster_diff = 1, iterations = 0;
  double fastest = 0.0, global_dt = 0.0, rel_tol = 1e-4;
  std::vector<int> cluster, flags;  // flags: 1 slower, 2 faster neighbour
  std::vector<int> mult;
  std::vector<double> cl_time;
  std::vector<int> order;
  std::vector<double> times;  // [n_updates][2]
  std::vector<ClusterPlan> plan;
  std::vector<int> soff;
  std::vector<Slot> slots;
  int *d_cluster = nullptr, *d_cells = nullptr;
  unsigned *d_mask = nullptr;
  unsigned char *d_target = nullptr;
  int *d_nc_cells = nullptr, *d_nc_kind = nullptr, *d_nc_other = nullptr;
  unsigned *d_nc_mask = nullptr;
  double *d_buf = nullptr;  // tempsrc, V, face sum, flux memory, reconstruction, copy
  double *tempsrc = nullptr, *V = nullptr, *fsum = nullptr, *fm = nullptr, *igd = nullptr,
         *rstate = nullptr;
};

namespace {

// the cell sets of ClusterManager::update_elements (include/cluster_manager.
// templates.h:807-1049) for cluster X and update set U, as launch lists; a
// face is taken per neighbour slot, so the subfaces of a 2:1 face are faces
// of their own (the reference's MatrixFree faces)
UpdatePlan plan_update(exw_lts L, ListBuilder &B, NcBuilder &NB, const std::vector<char> &U, int X,
                       bool write_fm) {
  exw_op op = L->op;
  const int n = op->n_cells, nface = 2 * op->dim, nc = L->n_clusters;
  const int *nbr = op->h_neighbor.data();
  const int *cl = L->cluster.data();
  const std::vector<int> &off = L->soff;
  const std::vector<Slot> &sl = L->slots;
  std::vector<char> nbrU(n, 0), E(n, 0), plain(n, 0);
  for (int c = 0; c < n; ++c) {
    if (U[c]) continue;
    for (int i = off[c]; i < off[c + 1]; ++i)
      if (sl[i].nb >= 0 && U[sl[i].nb]) nbrU[c] = 1;
  }
  auto is_nc = [&](int c, int f) { return nbr[(size_t)c * nface + f] == EXW_NC_FACE; };
  UpdatePlan P;
  for (int side = 0; side < 2; ++side) {
    const int other = side == 0 ? X - 1 : X + 1;
    if (other < 0 || other >= nc) continue;
    const int fX = side == 0 ? 2 : 1, fO = side == 0 ? 1 : 2;  // flag bits
    std::vector<int> ck, fc, ncc[2];
    std::vector<unsigned> fmask, ncm[2];
    std::vector<unsigned char> none, ftarget;
    for (int c = 0; c < n; ++c) {
      E[c] = (cl[c] == X && (L->flags[c] & fX) && U[c]) ||
             (cl[c] == other && (L->flags[c] & fO) && nbrU[c]);
      if (E[c]) ck.push_back(c);
    }
    for (int c = 0; c < n; ++c) {
      if (!E[c]) continue;
      unsigned mc = 0, mn = 0;
      int tgt = 0;
      for (int i = off[c]; i < off[c + 1]; ++i) {
        const int nb = sl[i].nb;
        if (nb < 0 || !E[nb]) continue;
        const bool to_sum = cl[c] == X && cl[nb] != X && nbrU[nb];
        const bool to_fm = cl[c] != X && cl[nb] == X && nbrU[c] && write_fm;
        if (!to_sum && !to_fm) continue;
        const unsigned bit = 1u << (4 * sl[i].f + sl[i].j);
        (is_nc(c, sl[i].f) ? mn : mc) |= bit;
        if (to_fm) tgt = 1;
      }
      if (mc) {
        fc.push_back(c);
        fmask.push_back(mc);
        ftarget.push_back((unsigned char)tgt);
      }
      if (mn) {
        ncc[tgt].push_back(c);
        ncm[tgt].push_back(mn);
      }
    }
    std::vector<unsigned> zm(ck.size(), 0);
    none.assign(ck.size(), 0);
    P.side[side].ck = B.add(ck, zm, none);
    P.side[side].face = B.add(fc, fmask, ftarget);
    for (int t = 0; t < 2; ++t) P.side[side].nc[t] = NB.add(op, ncc[t], ncm[t]);
  }
  {
    std::vector<int> ck, fc, ncc;
    std::vector<unsigned> fmask, ncm;
    std::vector<unsigned char> none, ftarget;
    for (int c = 0; c < n; ++c) {
      E[c] = cl[c] == X && (U[c] || nbrU[c]);
      if (E[c]) ck.push_back(c);
    }
    std::vector<int> rfc;
    std::vector<unsigned> rfmask;
    for (int c = 0; c < n; ++c) {
      if (!U[c]) continue;
      unsigned mc = 0, mn = 0;
      bool inner = op->cart;  // every neighbour in the cluster, no hanging face
      for (int i = off[c]; i < off[c + 1]; ++i) {
        const int nb = sl[i].nb;
        if (is_nc(c, sl[i].f) || (nb >= 0 && cl[nb] != X)) inner = false;
        if (!(nb < 0 || U[nb] || (E[nb] && cl[nb] == X))) continue;
        const unsigned bit = 1u << (4 * sl[i].f + sl[i].j);
        (is_nc(c, sl[i].f) ? mn : mc) |= bit;
      }
      plain[c] = inner;
      if (mc) {
        fc.push_back(c);
        fmask.push_back(mc);
        ftarget.push_back(0);
        if (!inner) {
          rfc.push_back(c);
          rfmask.push_back(mc);
        }
      }
      if (mn) {
        ncc.push_back(c);
        ncm.push_back(mn);
      }
    }
    std::vector<unsigned> zm(ck.size(), 0);
    none.assign(ck.size(), 0);
    P.same.ck = B.add(ck, zm, none);
    P.same.face = B.add(fc, fmask, ftarget);
    P.same.nc[0] = NB.add(op, ncc, ncm);
    P.rest_face = B.add(rfc, rfmask, std::vector<unsigned char>(rfc.size(), 0));
  }
  std::vector<int> uc, pc, rc;
  for (int c = 0; c < n; ++c)
    if (U[c]) {
      uc.push_back(c);
      (plain[c] ? pc : rc).push_back(c);
    }
  std::vector<unsigned> zm(uc.size(), 0);
  std::vector<unsigned char> none(uc.size(), 0);
  P.upd = B.add(uc, zm, none);
  P.plain = B.add(pc, std::vector<unsigned>(pc.size(), 0),
                  std::vector<unsigned char>(pc.size(), 0));
  P.rest_upd = B.add(rc, std::vector<unsigned>(rc.size(), 0),
                     std::vector<unsigned char>(rc.size(), 0));
  for (int s = 0; s < op->n_src; ++s)
    if (U[op->h_src_cells[s]]) P.src_sel |= 1u << s;
  return P;
}

int launch_ck_list(exw_lts L, const List &l, const double *src, double t2, double t1,
                   bool per_cluster_te, double *out) {
  if (l.n == 0) return EXW_OK;
  exw_op op = L->op;
  CellSel sel{};
  sel.list = L->d_cells + l.off;
  sel.n = l.n;
  if (per_cluster_te) {
    sel.cluster = L->d_cluster;
    for (int c = 0; c < L->n_clusters; ++c) sel.te[c] = L->cl_time[c];
  }
  CkLaunch ck = kernels(op->dim).ck[op->degree][0][op->cart ? 1 : 0];
  LHIP(ck(op->view(), src, L->recon ? L->igd : nullptr, op->d_cktab, op->ck, t2, t1, 0.0, &sel,
          out, op->ctx->stream));
  return EXW_OK;
}

int launch_sweep(exw_lts L, const List &l, int mode, const double *src, double fac, double *out,
                 bool use_fm, bool with_mask) {
  if (l.n == 0) return EXW_OK;
  exw_op op = L->op;
  LtsArgs a{};
  a.cells = L->d_cells + l.off;
  a.mask = with_mask ? L->d_mask + l.off : nullptr;
  a.target = L->d_target + l.off;
  a.n = l.n;
  a.mode = mode;
  a.src = src;
  a.fac = fac;
  a.acc0 = L->fsum;
  a.acc1 = L->fm;
  a.out = out;
  a.use_fm = use_fm ? 1 : 0;
  LHIP(kernels(op->dim).lts[op->degree][op->cart ? 1 : 0](op->view(), op->d_tab, a,
                                                          op->ctx->stream));
  return EXW_OK;
}

// state += M^-1 A(src) on the listed cells: the Cartesian operator sweep
// with a cell list, the second sweep of global ADER (exw_ader_step)
int launch_plain(exw_lts L, const List &l, const double *src, double *state) {
  if (l.n == 0) return EXW_OK;
  exw_op op = L->op;
  OpLaunch fn = kernels(op->dim).op[op->degree][0][1];
  Epilogue ep{};
  ep.b = state;
  ep.o1 = state;
  ep.a1 = 1.0;
  ep.cell_list = L->d_cells + l.off;
  ep.n_list = l.n;
  LHIP(fn(op->view(), src, op->d_tab, op->host_tab.data(), ep, op->ctx->stream));
  return EXW_OK;
}

// the subface integrals of the listed hanging-face entries (masks per
// subface), from src on both sides, times fac, added into out
int launch_nc_list(exw_lts L, const List &l, const double *src, double fac, double *out) {
  if (l.n == 0) return EXW_OK;
  exw_op op = L->op;
  NcArgs a{l.n, L->d_nc_cells + l.off, L->d_nc_kind + (size_t)l.off * 2 * op->dim,
           L->d_nc_other + (size_t)l.off * 2 * op->dim * 4, L->d_nc_mask + l.off};
  Epilogue e{};
  e.o1 = out;
  e.a1 = fac;
  NcLaunch ln = kernels(op->dim).nc[op->degree];
  LHIP(ln(op->view(), op->d_nctab, a, src, e, op->ctx->stream));
  return EXW_OK;
}

// one update_elements call: t[side] = {t1, t2} of the faster/slower passes,
// ts = {t1, t2} of the own pass, dt of the update; the update cells start at
// time t_cells (the point source's interval is [t_cells, t_cells + dt])
int run_update(exw_lts L, const UpdatePlan &P, double *state, const double tside[2][2],
               const double ts[2], double dt, double t_cells, bool use_fm) {
  const double tol = L->rel_tol * L->fastest;
  int rc;
  for (int side = 0; side < 2; ++side) {
    const Pass &S = P.side[side];
    if (S.ck.n == 0 && S.face.n == 0 && S.nc[0].n == 0 && S.nc[1].n == 0) continue;
    const double t1 = tside[side][0], t2 = tside[side][1];
    if (!(std::fabs(t2 - t1) > tol)) continue;
    if ((rc = launch_ck_list(L, S.ck, state, t2, t1, true, L->tempsrc))) return rc;
    if ((rc = launch_sweep(L, S.face, LTS_ACC, L->tempsrc, -1.0, nullptr, false, true)))
      return rc;
    if ((rc = launch_nc_list(L, S.nc[0], L->tempsrc, -1.0, L->fsum))) return rc;
    if ((rc = launch_nc_list(L, S.nc[1], L->tempsrc, -1.0, L->fm))) return rc;
  }
  // the update's volume term needs the CK integral over [0, dt] about the
  // cells' own time: the same-cluster pass computes it over [ts0, ts1] about
  // ts0 for every update cell, the same integral when ts1 - ts0 = dt
  const bool reuse = std::fabs((ts[1] - ts[0]) - dt) <= 1e-14 * std::fabs(dt);
  // then the inner cells' face and volume terms all come from that one
  // integral: their update is the plain operator sweep (no flux memory or
  // face sum can be pending on them, all their neighbours being in the cluster)
  const bool fast = reuse && P.plain.n > 0;
  if ((rc = launch_ck_list(L, P.same.ck, state, ts[1], ts[0], true, L->tempsrc))) return rc;
  if ((rc = launch_sweep(L, fast ? P.rest_face : P.same.face, LTS_ACC, L->tempsrc, -1.0, nullptr,
                         false, true)))
    return rc;
  if ((rc = launch_nc_list(L, P.same.nc[0], L->tempsrc, -1.0, L->fsum))) return rc;
  const double *vsrc = L->tempsrc;
  if (!reuse) {
    if ((rc = launch_ck_list(L, P.upd, state, dt, 0.0, false, L->V))) return rc;
    vsrc = L->V;
  }
  if ((rc = launch_sweep(L, fast ? P.rest_upd : P.upd, LTS_UPDATE, vsrc, -1.0, state, use_fm,
                         false)))
    return rc;
  if (fast && (rc = launch_plain(L, P.plain, vsrc, state))) return rc;
  return source_add_selected(L->op, t_cells, t_cells + dt, P.src_sel, state);
}

}  // namespace

extern "C" {

int exw_lts_categorize(int dim, int n_cells, const double *vertices, const int *neighbor,
                       double cfl, int max_clusters, int max_diff, int *cluster, int *flags,
                       int *n_clusters, int *cluster_diff, int *iterations) {
  if ((dim != 2 && dim != 3) || n_cells < 1 || !vertices || !neighbor || !cluster || !flags ||
      !(cfl > 0) || max_clusters < 1 || max_diff < 1)
    return lfail(EXW_EINVAL, "exw_lts_categorize: invalid argument");
  const int nface = 2 * dim;
  for (int64_t i = 0; i < (int64_t)n_cells * nface; ++i)
    if (neighbor[i] == EXW_NC_FACE)
      return lfail(EXW_EUNSUP, "exw_lts_categorize: hanging faces need the operator's records "
                               "(exw_lts_create / exw_lts_create2)");
  // the reference's rule: cfl * minimum vertex distance, no wave speed
  // (include/cluster_manager.templates.h:57-66)
  std::vector<int> off;
  std::vector<Slot> slots;
  build_slots(dim, n_cells, neighbor, nullptr, nullptr, off, slots);
  return categorize(dim, n_cells, vertices, nullptr, off, slots, cfl, max_clusters, max_diff,
                    cluster, flags, n_clusters, cluster_diff, iterations);
}

int exw_lts_destroy(exw_lts L) {
  if (!L) return EXW_OK;
  hipFree(L->d_cluster);
  hipFree(L->d_cells);
  hipFree(L->d_mask);
  hipFree(L->d_target);
  hipFree(L->d_nc_cells);
  hipFree(L->d_nc_kind);
  hipFree(L->d_nc_other);
  hipFree(L->d_nc_mask);
  hipFree(L->d_buf);
  delete L;
  return EXW_OK;
}

int exw_lts_create(exw_op op, double cfl, double fastest_dt, int max_clusters, int max_diff,
                   double time0, int recon, exw_lts *out) {
  return exw_lts_create2(op, cfl, fastest_dt, max_clusters, max_diff, time0, recon, 0, out);
}

int exw_lts_create2(exw_op op, double cfl, double fastest_dt, int max_clusters, int max_diff,
                    double time0, int recon, int flags, exw_lts *out) {
  if (!op || !out || !(fastest_dt > 0) || !(cfl > 0) || max_clusters < 1 || max_diff < 1 ||
      (flags & ~EXW_LTS_SPEED_AWARE))
    return lfail(EXW_EINVAL, "exw_lts_create: bad argument");
  if (!kernels(op->dim).lts[op->degree][0] || !kernels(op->dim).ck[op->degree][0][0])
    return lfail(EXW_EUNSUP, "exw_lts_create: local time stepping not compiled for this degree");
  const int n = op->n_cells, dim = op->dim;
  exw_lts L = new exw_lts_s;
  L->op = op;
  L->recon = recon ? 1 : 0;
  L->cluster.resize(n);
  L->flags.resize(n);
  const bool nc_mesh = op->n_hanging > 0;
  build_slots(dim, n, op->h_neighbor.data(), nc_mesh ? op->h_nc_kind.data() : nullptr,
              nc_mesh ? op->h_nc_other.data() : nullptr, L->soff, L->slots);
  int rc = categorize(dim, n, op->h_vertices.data(),
                      (flags & EXW_LTS_SPEED_AWARE) ? op->h_speed.data() : nullptr, L->soff,
                      L->slots, cfl, max_clusters, max_diff, L->cluster.data(), L->flags.data(),
                      &L->n_clusters, &L->cluster_diff, &L->iterations);
  if (rc) {
    delete L;
    return rc;
  }
  const int nc = L->n_clusters, diff = L->cluster_diff;
  if (nc > MAX_CLUSTERS) {
    delete L;
    return lfail(EXW_EUNSUP, "exw_lts_create: more than 16 clusters");
  }
  L->fastest = fastest_dt;
  // update order and times (ClusterManager::setup, :402-475)
  L->mult.resize(nc);
  L->cl_time.assign(nc, time0);
  for (int c = 0; c < nc; ++c) L->mult[c] = diff * c + 1;
  const int max_level = diff * (nc - 1) + 1;
  std::vector<int> tl(nc, 0);
  auto capped = [&](int v) { return std::min(v, max_level); };
  while (*std::min_element(tl.begin(), tl.end()) < max_level) {
    for (int c = 0; c < nc; ++c) {
      if (tl[c] >= max_level) continue;
      const int mine = capped(tl[c] + L->mult[c]);
      bool ok = true;
      if (c > 0) ok = ok && mine <= capped(tl[c - 1] + L->mult[c - 1]);
      if (c < nc - 1) ok = ok && mine <= capped(tl[c + 1] + L->mult[c + 1]);
      if (!ok) continue;
      L->times.push_back(tl[c] * fastest_dt);
      tl[c] += L->mult[c];
      L->order.push_back(c);
      L->times.push_back(tl[c] * fastest_dt);
    }
  }
  L->global_dt = fastest_dt * (diff * (nc - 1) + 1);
  for (size_t u = 0; u < L->order.size(); ++u) {
    L->times[2 * u + 1] = std::min(L->times[2 * u + 1], max_level * fastest_dt);
    L->times[2 * u] += time0;
    L->times[2 * u + 1] += time0;
  }
  // launch lists per cluster
  ListBuilder B;
  NcBuilder NB;
  L->plan.resize(nc);
  for (int X = 0; X < nc; ++X) {
    std::vector<char> U(n);
    for (int c = 0; c < n; ++c) U[c] = L->cluster[c] == X;
    L->plan[X].main = plan_update(L, B, NB, U, X, true);
    for (int side = 0; side < 2; ++side) {
      const int other = side == 0 ? X - 1 : X + 1;
      if (other < 0 || other >= nc) continue;
      // cells of the neighbouring cluster that touch X
      std::vector<char> Ur(n, 0);
      for (int c = 0; c < n; ++c) {
        if (L->cluster[c] != other) continue;
        for (int i = L->soff[c]; i < L->soff[c + 1]; ++i)
          if (L->slots[i].nb >= 0 && L->cluster[L->slots[i].nb] == X) Ur[c] = 1;
      }
      L->plan[X].recon[side] = plan_update(L, B, NB, Ur, other, false);
    }
    std::vector<int> ac, acn;
    std::vector<unsigned> ncm;
    for (int c = 0; c < n; ++c) {
      if (L->cluster[c] != X) continue;
      ac.push_back(c);
      bool has_nc = false;
      for (int f = 0; f < 2 * dim && nc_mesh; ++f)
        has_nc |= op->h_nc_kind[(size_t)c * 2 * dim + f] != 0;
      if (has_nc) {
        acn.push_back(c);
        ncm.push_back(0xFFFFFFFFu);
      }
    }
    std::vector<unsigned> zm(ac.size(), 0);
    std::vector<unsigned char> none(ac.size(), 0);
    L->plan[X].apply = B.add(ac, zm, none);
    L->plan[X].apply_nc = NB.add(op, acn, ncm);
  }
  const int64_t N = op->n_dofs();
  const size_t nl = std::max<size_t>(B.cells.size(), 1), nn = std::max<size_t>(NB.cells.size(), 1);
  hipSetDevice(op->ctx->device);
  if (hipMalloc((void **)&L->d_cluster, sizeof(int) * n) != hipSuccess ||
      hipMalloc((void **)&L->d_cells, sizeof(int) * nl) != hipSuccess ||
      hipMalloc((void **)&L->d_mask, sizeof(unsigned) * nl) != hipSuccess ||
      hipMalloc((void **)&L->d_target, nl) != hipSuccess ||
      hipMalloc((void **)&L->d_nc_cells, sizeof(int) * nn) != hipSuccess ||
      hipMalloc((void **)&L->d_nc_kind, sizeof(int) * nn * 2 * dim) != hipSuccess ||
      hipMalloc((void **)&L->d_nc_other, sizeof(int) * nn * 2 * dim * 4) != hipSuccess ||
      hipMalloc((void **)&L->d_nc_mask, sizeof(unsigned) * nn) != hipSuccess ||
      hipMalloc((void **)&L->d_buf, sizeof(double) * 6 * (size_t)N) != hipSuccess) {
    exw_lts_destroy(L);
    return lfail(EXW_ENOMEM, "exw_lts_create: allocation failed");
  }
  L->tempsrc = L->d_buf;
  L->V = L->d_buf + N;
  L->fsum = L->d_buf + 2 * N;
  L->fm = L->d_buf + 3 * N;
  L->igd = L->d_buf + 4 * N;
  L->rstate = L->d_buf + 5 * N;
  hipStream_t st = op->ctx->stream;
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  };
  if (!h2d(L->d_cluster, L->cluster.data(), sizeof(int) * n) ||
      !h2d(L->d_cells, B.cells.data(), sizeof(int) * B.cells.size()) ||
      !h2d(L->d_mask, B.mask.data(), sizeof(unsigned) * B.mask.size()) ||
      !h2d(L->d_target, B.target.data(), B.target.size()) ||
      !h2d(L->d_nc_cells, NB.cells.data(), sizeof(int) * NB.cells.size()) ||
      !h2d(L->d_nc_kind, NB.kind.data(), sizeof(int) * NB.kind.size()) ||
      !h2d(L->d_nc_other, NB.other.data(), sizeof(int) * NB.other.size()) ||
      !h2d(L->d_nc_mask, NB.mask.data(), sizeof(unsigned) * NB.mask.size()) ||
      hipMemsetAsync(L->d_buf, 0, sizeof(double) * 6 * (size_t)N, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    exw_lts_destroy(L);
    return lfail(EXW_EDEVICE, "exw_lts_create: upload failed");
  }
  *out = L;
  return EXW_OK;
}

int exw_lts_info(exw_lts L, int *n_clusters, int *cluster_diff, int *n_updates,
                 double *global_dt) {
  if (!L) return lfail(EXW_EINVAL, "exw_lts_info: null handle");
  if (n_clusters) *n_clusters = L->n_clusters;
  if (cluster_diff) *cluster_diff = L->cluster_diff;
  if (n_updates) *n_updates = (int)L->order.size();
  if (global_dt) *global_dt = L->global_dt;
  return EXW_OK;
}

int exw_lts_cells(exw_lts L, int *cluster, int *flags) {
  if (!L || !cluster || !flags) return lfail(EXW_EINVAL, "exw_lts_cells: null argument");
  std::copy(L->cluster.begin(), L->cluster.end(), cluster);
  std::copy(L->flags.begin(), L->flags.end(), flags);
  return EXW_OK;
}

int exw_lts_step(exw_lts L, double time_now, const double *src, double *dst) {
  if (!L || !src || !dst) return lfail(EXW_EINVAL, "exw_lts_step: bad argument");
  exw_op op = L->op;
  const int64_t N = op->n_dofs();
  hipStream_t st = op->ctx->stream;
  const double D = L->fastest, tol = L->rel_tol * D;
  const int nc = L->n_clusters;
  int rc;
  // state = src, updated in place in dst (ClusterManager::perform_time_step,
  // :488-803); src is read by nothing else, so src == dst steps in place
  if (dst != src)
    LHIP(hipMemcpyAsync(dst, src, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  double *state = dst;
  if (L->recon && (rc = exw_op_apply(op, state, L->igd, EXW_FP64))) return rc;
  for (size_t cy = 0; cy < L->order.size(); ++cy) {
    const int ac = L->order[cy];
    const double act = L->cl_time[ac];
    double tside[2][2] = {{0, 0}, {0, 0}}, ts[2];
    if (ac > 0) {
      const double ft = L->cl_time[ac - 1];
      tside[0][0] = std::max(ft, act);
      tside[0][1] = std::min(std::min(ft + L->mult[ac - 1] * D, act + L->mult[ac] * D), time_now);
    }
    if (ac < nc - 1) {
      const double sl = L->cl_time[ac + 1];
      tside[1][0] = std::max(sl, act);
      tside[1][1] = std::min(std::min(sl + L->mult[ac + 1] * D, act + L->mult[ac] * D), time_now);
    }
    ts[0] = act;
    ts[1] = std::min(act + L->mult[ac] * D, time_now);
    const double dt = std::min(L->mult[ac] * D, time_now - L->cl_time[ac]);
    if ((rc = run_update(L, L->plan[ac].main, state, tside, ts, dt, act, true))) return rc;
    L->cl_time[ac] = L->times[2 * cy + 1];
    if (!L->recon) continue;
    // reconstruction at ac's new time level (:608-767)
    LHIP(hipMemcpyAsync(L->rstate, state, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    const double nt = L->cl_time[ac];
    for (int side = 0; side < 2; ++side) {
      const int other = side == 0 ? ac - 1 : ac + 1;
      if (other < 0 || other >= nc) continue;
      if (!(std::fabs(nt - L->cl_time[other]) > tol)) continue;
      const double t1 = L->cl_time[other], t2 = nt;
      const double tr[2][2] = {{t1, t2}, {t1, t2}}, tsr[2] = {t1, t2};
      if ((rc = run_update(L, L->plan[ac].recon[side], L->rstate, tr, tsr, t2 - t1, t1, false)))
        return rc;
    }
    if ((rc = launch_sweep(L, L->plan[ac].apply, LTS_APPLY, L->rstate, 1.0, L->igd, false, false)))
      return rc;
    if ((rc = launch_nc_list(L, L->plan[ac].apply_nc, L->rstate, 1.0, L->igd))) return rc;
  }
  for (auto &t : L->times) t += L->global_dt;
  return EXW_OK;
}

}  // extern "C"
