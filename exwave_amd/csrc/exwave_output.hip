// exwave_output.hip -- VTU output of a solution (output_results,
// source/explicit_wave.cc:339-436), host side.
//
// Like DataOut::build_patches(mapping, fe_degree) every cell is subdivided k
// times per direction: the fields are evaluated at the (k+1)^dim equispaced
// points of the cell (the DG polynomials, so points are not shared between
// cells) and written as k^dim linear VTK cells.  Field names are the
// reference's.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/exwave_amd.h"
#include "gauss_cx.hpp"

namespace {

// values at the equispaced points (xi_0 fastest) of a field given by its
// nodal values on n^dim points, E[i*n + j] = l_j(i/k)
void tensor_eval(int dim, int n, int m, const double *E, const double *in, double *out,
                 std::vector<double> &w1, std::vector<double> &w2) {
  int sz[3] = {n, n, n};
  const double *src = in;
  for (int d = 0; d < dim; ++d) {
    std::vector<double> &dstv = (d & 1) ? w2 : w1;
    double *dst = d == dim - 1 ? out : dstv.data();
    const int lo = d == 0 ? 1 : (d == 1 ? sz[0] : sz[0] * sz[1]);
    const int hi = d == dim - 1 ? 1 : (d == 0 ? (dim == 3 ? sz[1] * sz[2] : sz[1]) : sz[2]);
    const int nin = sz[d];
    for (int h = 0; h < hi; ++h)
      for (int i = 0; i < m; ++i)
        for (int l = 0; l < lo; ++l) {
          double s = 0.0;
          for (int j = 0; j < nin; ++j) s += E[i * nin + j] * src[(h * nin + j) * lo + l];
          dst[(h * m + i) * lo + l] = s;
        }
    sz[d] = m;
    src = dst;
  }
}

void write_array(FILE *f, const char *name, int ncomp, const std::vector<double> &v) {
  std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" "
                  "format=\"ascii\">\n", name, ncomp);
  for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%.10g\n", v[i]);
  std::fprintf(f, "</DataArray>\n");
}

}  // namespace

extern "C" {

int exw_write_vtu(int dim, int degree, int n_cells, const double *vertices,
                  const double *solution, const double *error, const double *error_estimate,
                  const double *post_pressure, const double *cluster_id, const double *time_step,
                  const char *filename) {
  if ((dim != 2 && dim != 3) || degree < 1 || degree > 12 || n_cells < 1 || !vertices ||
      !solution || !filename)
    return EXW_EINVAL;
  const int n = degree + 1, m = degree + 1, nf = dim + 1, nv = 1 << dim;
  const int nd = dim == 2 ? n * n : n * n * n, np = nd;  // points per cell: m^dim = n^dim
  const int P = degree + 2, npost = dim == 2 ? P * P : P * P * P;
  const int nsub = dim == 2 ? degree * degree : degree * degree * degree;
  FILE *f = std::fopen(filename, "w");
  if (!f) return EXW_EINVAL;
  // 1D evaluation matrices at the equispaced points i/k
  double xg[16], wg[16], xl[16];
  exw::cx::gauss01(n, xg, wg);
  exw::cx::gauss_lobatto01(P, xl);
  std::vector<double> E(m * n), Ep(m * P);
  for (int i = 0; i < m; ++i) {
    exw::cx::lagrange(n, xg, double(i) / degree, &E[i * n], nullptr);
    exw::cx::lagrange(P, xl, double(i) / degree, &Ep[i * P], nullptr);
  }
  const long npts = (long)n_cells * np, ncel = (long)n_cells * nsub;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\"%ld\" "
                  "NumberOfCells=\"%ld\">\n<Points>\n",
               npts, ncel);
  std::vector<double> pts(npts * 3, 0.0);
  for (int c = 0; c < n_cells; ++c) {
    const double *V = vertices + (size_t)c * nv * dim;
    for (int q = 0; q < np; ++q) {
      double xi[3] = {double(q % m) / degree, double((q / m) % m) / degree,
                      dim == 3 ? double(q / (m * m)) / degree : 0.0};
      for (int u = 0; u < nv; ++u) {
        double s = 1.0;
        for (int d = 0; d < dim; ++d) s *= ((u >> d) & 1) ? xi[d] : 1.0 - xi[d];
        for (int i = 0; i < dim; ++i) pts[((size_t)c * np + q) * 3 + i] += s * V[u * dim + i];
      }
    }
  }
  write_array(f, "Points", 3, pts);
  std::fprintf(f, "</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" "
                  "format=\"ascii\">\n");
  // VTK corner order: quad (0,0),(1,0),(1,1),(0,1); hexahedron the same at z = 0, then z = 1
  const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int c = 0; c < n_cells; ++c)
    for (int s = 0; s < nsub; ++s) {
      const int s0 = s % degree, s1 = (s / degree) % degree, s2 = dim == 3 ? s / (degree * degree) : 0;
      for (int v = 0; v < nv; ++v) {
        const int i0 = s0 + corner[v][0], i1 = s1 + corner[v][1], i2 = s2 + corner[v][2];
        std::fprintf(f, "%ld ", (long)c * np + i0 + m * i1 + (dim == 3 ? m * m * i2 : 0));
      }
      std::fprintf(f, "\n");
    }
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (long s = 1; s <= ncel; ++s) std::fprintf(f, "%ld\n", s * nv);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (long s = 0; s < ncel; ++s) std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);
  std::fprintf(f, "</DataArray>\n</Cells>\n<PointData>\n");
  std::vector<double> w1(np * n), w2(np * n), val(np);
  auto vec_and_scalar = [&](const double *u, const char *vname, const char *sname) {
    std::vector<double> vv(npts * 3, 0.0), pv(npts);
    for (int c = 0; c < n_cells; ++c) {
      for (int k = 0; k < nf; ++k) {
        tensor_eval(dim, n, m, E.data(), u + ((size_t)c * nf + k) * nd, val.data(), w1, w2);
        for (int q = 0; q < np; ++q) {
          if (k < dim)
            vv[((size_t)c * np + q) * 3 + k] = val[q];
          else
            pv[(size_t)c * np + q] = val[q];
        }
      }
    }
    write_array(f, vname, 3, vv);
    write_array(f, sname, 1, pv);
  };
  vec_and_scalar(solution, "solution_velocity", "solution_pressure");
  if (error) vec_and_scalar(error, "error_velocity", "error_pressure");
  if (post_pressure) {
    std::vector<double> pp(npts), pw1(np * P), pw2(np * P);
    for (int c = 0; c < n_cells; ++c) {
      tensor_eval(dim, P, m, Ep.data(), post_pressure + (size_t)c * npost, val.data(), pw1, pw2);
      for (int q = 0; q < np; ++q) pp[(size_t)c * np + q] = val[q];
    }
    write_array(f, "post_pressure", 1, pp);
  }
  std::fprintf(f, "</PointData>\n<CellData>\n");
  auto cell_field = [&](const double *a, const char *name) {
    if (!a) return;
    std::vector<double> cv(ncel);
    for (long s = 0; s < ncel; ++s) cv[s] = a[s / nsub];
    write_array(f, name, 1, cv);
  };
  cell_field(error_estimate, "Error_estimate");
  cell_field(cluster_id, "cluster_id");
  cell_field(time_step, "time_step");
  std::fprintf(f, "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? EXW_OK : EXW_EDEVICE;
}

}  // extern "C"
