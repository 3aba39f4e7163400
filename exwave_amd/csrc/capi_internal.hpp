// capi_internal.hpp -- private definitions of the C-ABI handles shared by the
// library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "launchers.hpp"

namespace exw {
void set_last_error(const std::string &msg);
}

struct exw_ctx_s {
  int device;
  hipStream_t stream;
};

struct exw_op_s {
  exw_ctx ctx;
  int dim, degree, n_cells;
  bool cart;
  double minvd;
  int *d_neighbor = nullptr;
  double *d_rho = nullptr, *d_speed = nullptr, *d_h = nullptr, *d_vertices = nullptr;
  double *d_coef = nullptr;  // Cartesian meshes: DevMesh::coef
  double *d_tab = nullptr, *d_errtab = nullptr, *d_cktab = nullptr, *d_posttab = nullptr;
  double *d_partial = nullptr, *d_sum = nullptr;
  // non-conforming faces: the cells that touch them (NcArgs) and the table
  int n_nc = 0, n_hanging = 0;
  int *d_nc_cells = nullptr, *d_nc_kind = nullptr, *d_nc_other = nullptr;
  double *d_nctab = nullptr;
  int n_partial = 0;
  exw::CkTab ck{};
  std::vector<double> host_tab;
  // host copies of the connectivity and geometry (local-time-stepping set-up)
  std::vector<double> h_vertices;
  std::vector<int> h_neighbor;
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

