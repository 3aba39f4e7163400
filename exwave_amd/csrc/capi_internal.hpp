// capi_internal.hpp -- private definitions of the C-ABI handles shared by the
// library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "launchers.hpp"

struct exw_op_s;

namespace exw {
void set_last_error(const std::string &msg);
// u += (int_{t1}^{t2} r dt) b on the operator's source cells s with bit s of
// sel set (asynchronous on the operator's stream)
int source_add_selected(exw_op_s *op, double t1, double t2, unsigned sel, double *u);
}

struct exw_ctx_s {
  int device;
  hipStream_t stream;
  // this device's copy of the library's static scratch array (the device
  // address from hipGetSymbolAddress; exw_l2_errors_noalloc)
  double *scratch = nullptr;
};

struct exw_op_s {
  exw_ctx ctx;
  int dim, degree, n_cells;
  bool cart;
  double minvd;
  int *d_neighbor = nullptr;
  double *d_rho = nullptr, *d_speed = nullptr, *d_h = nullptr, *d_vertices = nullptr;
  double *d_coef = nullptr;  // Cartesian meshes: DevMesh::coef
  int *d_coef_class = nullptr;  // DevMesh::coef_class (NULL: one row per cell)
  int n_coef_rows = 0;
  double *d_tab = nullptr, *d_errtab = nullptr, *d_cktab = nullptr, *d_posttab = nullptr;
  double *d_partial = nullptr, *d_sum = nullptr;
  // non-conforming faces: the cells that touch them (NcArgs) and the table
  int n_nc = 0, n_hanging = 0;
  int *d_nc_cells = nullptr, *d_nc_kind = nullptr, *d_nc_other = nullptr;
  double *d_nctab = nullptr;
  int n_partial = 0;
  exw::CkTab ck{};
  std::vector<double> host_tab;
  // host copies of the connectivity, geometry and wave speeds (local-time-
  // stepping set-up); with hanging faces also the per-cell face kinds and
  // partners of NcArgs ([n_cells][2 dim] and [n_cells][2 dim][4])
  std::vector<double> h_vertices, h_speed;
  std::vector<int> h_neighbor, h_nc_kind, h_nc_other;
  // point source (exw_op_set_source): the cells holding the point and per
  // cell the pressure increment per unit time integral (M^-1 delta)
  int n_src = 0;
  exw_source src{};
  std::vector<int> h_src_cells;
  int *d_src_cells = nullptr;
  double *d_src_b = nullptr;
  exw::NcArgs nc_args() const { return exw::NcArgs{n_nc, d_nc_cells, d_nc_kind, d_nc_other}; }
  exw::DevMesh view() const {
    exw::DevMesh m;
    m.n_cells = n_cells;
    m.neighbor = d_neighbor;
    m.rho = d_rho;
    m.speed = d_speed;
    m.h = d_h;
    m.vertices = d_vertices;
    m.coef = d_coef;
    m.coef_class = d_coef_class;
    return m;
  }
  int64_t n_post() const {  // post-processed pressure: (k+2)^dim values per cell
    int64_t np = 1;
    for (int d = 0; d < dim; ++d) np *= degree + 2;
    return (int64_t)n_cells * np;
  }
  int64_t n_dofs() const {
    int64_t nd = 1;
    for (int d = 0; d < dim; ++d) nd *= degree + 1;
    return (int64_t)n_cells * (dim + 1) * nd;
  }
};

