/*
 * exwave_oracle.c -- CPU FP64 restatement of ExWave's hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (tests/, smoke(),
 * bench.py's cpu_baseline).  See exwave_oracle.h for the layout and for the
 * list of reference functions restated here; each function below cites the
 * reference lines it follows.
 */
#include "exwave_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* 1D building blocks                                                        */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre rule with n points on [0,1], ascending (deal.II QGauss<1>). */
void ox_gauss(int n, double *x, double *w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) {
        /* one more evaluation for the derivative at the converged point */
        p1 = 1.0;
        p2 = 0.0;
        for (int j = 1; j <= n; ++j) {
          double p3 = p2;
          p2 = p1;
          p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        break;
      }
    }
    /* z is the i-th largest root on [-1,1] */
    double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    x[i] = 0.5 * (1.0 - z);
    w[i] = w[n - 1 - i] = 0.5 * wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

void ox_lagrange(int n, const double *nodes, double x, double *val, double *der) {
  for (int i = 0; i < n; ++i) {
    double v = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != i) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    if (val) val[i] = v;
    if (der) {
      double d = 0.0;
      for (int m = 0; m < n; ++m) {
        if (m == i) continue;
        double t = 1.0 / (nodes[i] - nodes[m]);
        for (int j = 0; j < n; ++j)
          if (j != i && j != m) t *= (x - nodes[j]) / (nodes[i] - nodes[j]);
        d += t;
      }
      der[i] = d;
    }
  }
}

void ox_basis_init(ox_basis *b, int degree) {
  int n = degree + 1;
  b->n = n;
  ox_gauss(n, b->x, b->w);
  for (int q = 0; q < n; ++q) ox_lagrange(n, b->x, b->x[q], NULL, b->D + q * n);
  ox_lagrange(n, b->x, 0.0, b->face[0], NULL);
  ox_lagrange(n, b->x, 1.0, b->face[1], NULL);
}

static int ipow(int a, int e) {
  int r = 1;
  while (e-- > 0) r *= a;
  return r;
}

/* Apply the 1D matrix A (n_out x n_in, row-major) along direction `dir` of the
 * tensor `in` (sizes sz[0..dim-1], xi_0 fastest).  transpose: use A^T (then A is
 * n_in x n_out).  add: accumulate into out. */
static void tensor_apply(int dim, const int *sz, int dir, const double *A, int n_out,
                         const double *in, double *out, int transpose, int add) {
  int n_in = sz[dir];
  int stride = 1;
  for (int d = 0; d < dir; ++d) stride *= sz[d];
  int outer = 1;
  for (int d = dir + 1; d < dim; ++d) outer *= sz[d];
  for (int o = 0; o < outer; ++o)
    for (int s = 0; s < stride; ++s) {
      const double *ip = in + (long)o * stride * n_in + s;
      double *op = out + (long)o * stride * n_out + s;
      for (int r = 0; r < n_out; ++r) {
        double acc = 0.0;
        for (int c = 0; c < n_in; ++c)
          acc += (transpose ? A[c * n_out + r] : A[r * n_in + c]) * ip[c * stride];
        if (add)
          op[r * stride] += acc;
        else
          op[r * stride] = acc;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* geometry: multilinear cells (the reference's cells are multilinear images  */
/* of their 2^dim vertices, include/input_parameters.h:66-87 with a flat      */
/* manifold and MappingQGeneric(k) reproducing the multilinear map)           */
/* ------------------------------------------------------------------------ */

static void multilinear_point(int dim, const double *V, const double *xi, double *x) {
  int nv = 1 << dim;
  for (int i = 0; i < dim; ++i) x[i] = 0.0;
  for (int v = 0; v < nv; ++v) {
    double s = 1.0;
    for (int d = 0; d < dim; ++d) s *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
    for (int i = 0; i < dim; ++i) x[i] += s * V[v * dim + i];
  }
}

/* J[i*dim+e] = dx_i/dxi_e, Jinv[e*dim+i] = dxi_e/dx_i; returns det J */
static double multilinear_jacobian(int dim, const double *V, const double *xi, double *J,
                                   double *Jinv) {
  int nv = 1 << dim;
  for (int k = 0; k < dim * dim; ++k) J[k] = 0.0;
  for (int v = 0; v < nv; ++v)
    for (int e = 0; e < dim; ++e) {
      double s = ((v >> e) & 1) ? 1.0 : -1.0;
      for (int d = 0; d < dim; ++d)
        if (d != e) s *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
      for (int i = 0; i < dim; ++i) J[i * dim + e] += s * V[v * dim + i];
    }
  double det;
  if (dim == 2) {
    det = J[0] * J[3] - J[1] * J[2];
    Jinv[0] = J[3] / det;
    Jinv[1] = -J[1] / det;
    Jinv[2] = -J[2] / det;
    Jinv[3] = J[0] / det;
  } else {
    double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7],
           k = J[8];
    double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    det = a * A + b * B + c * C;
    Jinv[0] = A / det;
    Jinv[3] = B / det;
    Jinv[6] = C / det;
    Jinv[1] = -(b * k - c * h) / det;
    Jinv[4] = (a * k - c * g) / det;
    Jinv[7] = -(a * h - b * g) / det;
    Jinv[2] = (b * f - c * e) / det;
    Jinv[5] = -(a * f - c * d) / det;
    Jinv[8] = (a * e - b * d) / det;
  }
  return det;
}

/* ------------------------------------------------------------------------ */
/* mesh: GridGenerator::subdivided_hyper_cube(n,0,1) + transform of the       */
/* coarse vertices + refine_global (include/input_parameters.h:59-88)         */
/* ------------------------------------------------------------------------ */

static void transform_vertex(int dim, int type, double f, double *x) {
  double in[3] = {x[0], x[1], dim > 2 ? x[2] : 0.0};
  if (type == OX_TRANSFORM_REFERENCE_SINE) {
    /* include/input_parameters.h:41-50 */
    double s = f;
    for (int d = 0; d < dim; ++d) s *= sin(M_PI * in[d]);
    x[0] = in[0] + s;
  } else if (type == OX_TRANSFORM_AXIS_2PI) {
    for (int i = 0; i < dim; ++i) {
      double s = f;
      for (int e = 0; e < dim; ++e)
        if (e != i) s *= sin(2.0 * M_PI * in[e]);
      x[i] = in[i] + s;
    }
  } else if (type == OX_TRANSFORM_PRODUCT_2PI) {
    double s = f;
    for (int e = 0; e < dim; ++e) s *= sin(2.0 * M_PI * in[e]);
    for (int i = 0; i < dim; ++i) x[i] = in[i] + s;
  }
}

ox_mesh *ox_mesh_hypercube(int dim, int n_intervals, int n_refinements, int transform,
                           double factor, int boundary_id) {
  ox_mesh *m = (ox_mesh *)calloc(1, sizeof(ox_mesh));
  int N = n_intervals, R = 1 << n_refinements, Nf = N * R;
  int nv = 1 << dim;
  m->dim = dim;
  m->n_cells = ipow(Nf, dim);
  m->vertices = (double *)malloc(sizeof(double) * m->n_cells * nv * dim);
  m->neighbor = (int *)malloc(sizeof(int) * m->n_cells * 2 * dim);
  m->rho = (double *)malloc(sizeof(double) * m->n_cells);
  m->speed = (double *)malloc(sizeof(double) * m->n_cells);
  double mind = 1e300;
  for (int c = 0; c < m->n_cells; ++c) {
    int idx[3] = {0, 0, 0};
    int t = c;
    for (int d = 0; d < dim; ++d) {
      idx[d] = t % Nf;
      t /= Nf;
    }
    /* coarse parent vertices (transformed) */
    double PV[8 * 3];
    for (int v = 0; v < nv; ++v) {
      double x[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < dim; ++d) x[d] = (double)(idx[d] / R + ((v >> d) & 1)) / N;
      transform_vertex(dim, transform, factor, x);
      for (int d = 0; d < dim; ++d) PV[v * dim + d] = x[d];
    }
    /* child vertices: multilinear map of the parent at dyadic points */
    for (int v = 0; v < nv; ++v) {
      double xi[3];
      for (int d = 0; d < dim; ++d) xi[d] = (double)(idx[d] % R + ((v >> d) & 1)) / R;
      multilinear_point(dim, PV, xi, m->vertices + ((long)c * nv + v) * dim);
    }
    /* neighbours */
    for (int d = 0; d < dim; ++d)
      for (int s = 0; s < 2; ++s) {
        int nidx = idx[d] + (s ? 1 : -1);
        int f = 2 * d + s;
        if (nidx < 0 || nidx >= Nf)
          m->neighbor[c * 2 * dim + f] = -boundary_id;
        else {
          int stride = ipow(Nf, d);
          m->neighbor[c * 2 * dim + f] = c + (s ? stride : -stride);
        }
      }
    m->rho[c] = 1.0;
    m->speed[c] = 1.0;
    /* TriaAccessor::minimum_vertex_distance: min over all vertex pairs */
    const double *V = m->vertices + (long)c * nv * dim;
    for (int a = 0; a < nv; ++a)
      for (int b = a + 1; b < nv; ++b) {
        double s = 0.0;
        for (int d = 0; d < dim; ++d) s += (V[a * dim + d] - V[b * dim + d]) * (V[a * dim + d] - V[b * dim + d]);
        if (s < mind) mind = s;
      }
  }
  m->min_vertex_distance_over_c = sqrt(mind);
  return m;
}

void ox_mesh_set_material(ox_mesh *m, const double *rho, const double *speed) {
  int nv = 1 << m->dim;
  double mind = 1e300;
  for (int c = 0; c < m->n_cells; ++c) {
    m->rho[c] = rho[c];
    m->speed[c] = speed[c];
    const double *V = m->vertices + (long)c * nv * m->dim;
    for (int a = 0; a < nv; ++a)
      for (int b = a + 1; b < nv; ++b) {
        double s = 0.0;
        for (int d = 0; d < m->dim; ++d)
          s += (V[a * m->dim + d] - V[b * m->dim + d]) * (V[a * m->dim + d] - V[b * m->dim + d]);
        s = sqrt(s) / speed[c]; /* source/explicit_wave.cc:537-540 */
        if (s < mind) mind = s;
      }
  }
  m->min_vertex_distance_over_c = mind;
}

static double min_vertex_distance(int dim, const double *V) {
  const int nv = 1 << dim;
  double mind = 1e300;
  for (int a = 0; a < nv; ++a)
    for (int b = a + 1; b < nv; ++b) {
      double s = 0.0;
      for (int d = 0; d < dim; ++d)
        s += (V[a * dim + d] - V[b * dim + d]) * (V[a * dim + d] - V[b * dim + d]);
      if (s < mind) mind = s;
    }
  return sqrt(mind);
}

ox_mesh *ox_mesh_box_refined(int dim, int n_intervals, int transform, double factor,
                             int boundary_id, const int *refine) {
  const int N = n_intervals, nv = 1 << dim, nface = 2 * dim, nbase = ipow(N, dim);
  const int nsub = 1 << (dim - 1), rec = 2 + nsub;
  int *first = (int *)malloc(sizeof(int) * (nbase + 1));
  first[0] = 0;
  for (int B = 0; B < nbase; ++B) first[B + 1] = first[B] + (refine[B] ? nv : 1);
  ox_mesh *m = (ox_mesh *)calloc(1, sizeof(ox_mesh));
  m->dim = dim;
  m->n_cells = first[nbase];
  m->vertices = (double *)malloc(sizeof(double) * m->n_cells * nv * dim);
  m->neighbor = (int *)malloc(sizeof(int) * m->n_cells * nface);
  m->rho = (double *)malloc(sizeof(double) * m->n_cells);
  m->speed = (double *)malloc(sizeof(double) * m->n_cells);
  m->nc_rec = (int *)malloc(sizeof(int) * m->n_cells * nface);
  m->nc_pos = (int *)malloc(sizeof(int) * m->n_cells * nface);
  m->hanging = (int *)malloc(sizeof(int) * (nbase * nface + 1) * rec);
  m->n_hanging = 0;
  double mind = 1e300;
  for (int B = 0; B < nbase; ++B) {
    int idx[3] = {0, 0, 0}, t = B;
    for (int d = 0; d < dim; ++d) {
      idx[d] = t % N;
      t /= N;
    }
    double PV[8 * 3];
    for (int v = 0; v < nv; ++v) {
      double x[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < dim; ++d) x[d] = (double)(idx[d] + ((v >> d) & 1)) / N;
      transform_vertex(dim, transform, factor, x);
      for (int d = 0; d < dim; ++d) PV[v * dim + d] = x[d];
    }
    const int nch = refine[B] ? nv : 1;
    for (int ch = 0; ch < nch; ++ch) {
      const int c = first[B] + ch;
      for (int v = 0; v < nv; ++v) {
        double xi[3];
        for (int d = 0; d < dim; ++d)
          xi[d] = refine[B] ? 0.5 * (((ch >> d) & 1) + ((v >> d) & 1)) : (double)((v >> d) & 1);
        multilinear_point(dim, PV, xi, m->vertices + ((long)c * nv + v) * dim);
      }
      for (int d = 0; d < dim; ++d)
        for (int s = 0; s < 2; ++s) {
          const int f = 2 * d + s, cf = c * nface + f;
          m->nc_rec[cf] = m->nc_pos[cf] = -1;
          const int cd = (ch >> d) & 1;
          if (refine[B] && cd != s) {  /* sibling inside the parent */
            m->neighbor[cf] = first[B] + (ch ^ (1 << d));
            continue;
          }
          const int ni = idx[d] + (s ? 1 : -1);
          if (ni < 0 || ni >= N) {
            m->neighbor[cf] = -boundary_id;
            continue;
          }
          const int B2 = B + (s ? 1 : -1) * ipow(N, d);
          if (!refine[B2]) {
            m->neighbor[cf] = refine[B] ? OX_NC_FACE : first[B2];  /* fine side filled below */
          } else if (refine[B]) {
            m->neighbor[cf] = first[B2] + (ch ^ (1 << d));
          } else {
            /* coarse side of a hanging face: the children of B2 touching it */
            int *r = m->hanging + (long)m->n_hanging * rec;
            r[0] = c;
            r[1] = f;
            for (int b = 0; b < nsub; ++b) {
              int chn = (s ? 0 : 1) << d, j = 0;
              for (int e = 0; e < dim; ++e)
                if (e != d) chn |= ((b >> j++) & 1) << e;
              r[2 + b] = first[B2] + chn;
            }
            m->neighbor[cf] = OX_NC_FACE;
            m->nc_rec[cf] = m->n_hanging++;
          }
        }
      m->rho[c] = 1.0;
      m->speed[c] = 1.0;
      const double md = min_vertex_distance(dim, m->vertices + (long)c * nv * dim);
      if (md < mind) mind = md;
    }
  }
  /* fine sides */
  for (int h = 0; h < m->n_hanging; ++h) {
    const int *r = m->hanging + (long)h * rec;
    for (int b = 0; b < nsub; ++b) {
      const int cf = r[2 + b] * nface + (r[1] ^ 1);
      m->nc_rec[cf] = h;
      m->nc_pos[cf] = b;
    }
  }
  m->min_vertex_distance_over_c = mind;
  free(first);
  return m;
}

void ox_mesh_free(ox_mesh *m) {
  if (!m) return;
  free(m->hanging);
  free(m->nc_rec);
  free(m->nc_pos);
  free(m->vertices);
  free(m->neighbor);
  free(m->rho);
  free(m->speed);
  free(m);
}

long ox_n_dofs(const ox_mesh *m, int degree) {
  return (long)m->n_cells * (m->dim + 1) * ipow(degree + 1, m->dim);
}

/* ------------------------------------------------------------------------ */
/* DG operator                                                               */
/* ------------------------------------------------------------------------ */

/* numerical flux on one face quadrature point.
 * source/wave_equation_operations.cc:480-551 (inner) and 569-629 (boundary).
 * up/um: (v_0..v_{dim-1}, p) on this (plus) side and on the neighbour;
 * nb_bid < 0 selects the boundary branch with id -nb_bid.  out: the test
 * values submitted on the plus side, multiplied by boundary_fac. */
static void face_flux(int dim, const double *nrm, const double *up, const double *um,
                      double rho_p, double c_p, double rho_m, double c_m, int boundary_id,
                      double fac, double *out) {
  double nvp = 0.0;
  for (int d = 0; d < dim; ++d) nvp += up[d] * nrm[d];
  double pp = up[dim];
  double tau_p = 1.0 / c_p / rho_p;
  double lambda;
  if (boundary_id == 0) {
    double nvm = 0.0;
    for (int d = 0; d < dim; ++d) nvm -= um[d] * nrm[d];
    double tau_m = 1.0 / c_m / rho_m;
    double tau_inv = 1.0 / (tau_p + tau_m);
    lambda = tau_inv * (nvp + nvm + tau_p * pp + tau_m * um[dim]);
  } else if (boundary_id == 1) {
    lambda = 1.0 / tau_p * nvp + pp;
  } else if (boundary_id == 2) {
    lambda = 0.0;
  } else {
    double t2 = 1.0 / c_p / rho_p;
    lambda = tau_p / (tau_p + t2) * pp + 1.0 / (tau_p + t2) * nvp;
  }
  for (int d = 0; d < dim; ++d) out[d] = fac * (pp - lambda) / rho_p * nrm[d];
  out[dim] = fac * c_p * c_p * rho_p * (-nvp + tau_p * (lambda - pp));
}

/* value of field uk (n^dim nodal values) at the reference point xi */
static double eval_at(const ox_basis *b, int dim, const double *uk, const double *xi) {
  const int n = b->n;
  double l[3][OX_MAX_N];
  for (int d = 0; d < dim; ++d) ox_lagrange(n, b->x, xi[d], l[d], NULL);
  double s = 0.0;
  const int nd = ipow(n, dim);
  for (int q = 0; q < nd; ++q) {
    double w = 1.0;
    int t = q;
    for (int d = 0; d < dim; ++d) {
      w *= l[d][t % n];
      t /= n;
    }
    s += w * uk[q];
  }
  return s;
}

/* Face f of cell c on a non-conforming (2:1) face.  deal.II integrates such a
 * face on the fine cell's face with its QGauss(k+1) points and evaluates the
 * coarse cell there (FEFaceEvaluation on subfaces):
 *   fine side: own trace and geometry as usual, the coarse neighbour's
 *     polynomial at the matching points of its face quarter;
 *   coarse side: the sum over the 2^(dim-1) subfaces of the same integrals,
 *     with this cell's trace, geometry and test functions at the mapped
 *     points (xi_t + b_t)/2 and the weights w_t/2 per tangential direction.
 * Same flux (source/wave_equation_operations.cc:480-551) on both sides,
 * times boundary_fac `fac`.  sb_only >= 0 restricts the coarse side to that
 * subface (one MatrixFree face of the pair; local time stepping). */
static void cell_face_nc(const ox_mesh *m, const ox_basis *b, int c, int f, const double *src,
                         int sb_only, double fac, double *r) {
  const int dim = m->dim, n = b->n, nf = dim + 1, nface = 2 * dim;
  const int nd = ipow(n, dim), nv = 1 << dim, nt = ipow(n, dim - 1), nsub = 1 << (dim - 1);
  const int rec = 2 + nsub;
  const int d = f >> 1, s = f & 1;
  const int h = m->nc_rec[c * nface + f], pos = m->nc_pos[c * nface + f];
  const int *R = m->hanging + (long)h * rec;
  const double *V = m->vertices + (long)c * nv * dim;
  const double *u = src + (long)c * nf * nd;
  const double rho = m->rho[c], cc = m->speed[c];
  const int stride_d = ipow(n, d);
  const int nsb = pos >= 0 ? 1 : nsub;
  for (int sb = 0; sb < nsb; ++sb) {
    if (pos < 0 && sb_only >= 0 && sb != sb_only) continue;
    const int bpos = pos >= 0 ? pos : sb;
    const int other = pos >= 0 ? R[0] : R[2 + sb];
    const double *uo = src + (long)other * nf * nd;
    for (int t = 0; t < nt; ++t) {
      int base = 0, tt = t, j = 0;
      double xs[3], xc[3], wt = 1.0; /* fine-face point, coarse-face point */
      xs[d] = xc[d] = (double)s;
      for (int e = 0; e < dim; ++e) {
        if (e == d) continue;
        const int te = tt % n;
        tt /= n;
        base += te * ipow(n, e);
        xs[e] = b->x[te];
        xc[e] = 0.5 * (b->x[te] + ((bpos >> j) & 1));
        wt *= b->w[te];
        ++j;
      }
      /* own side: fine -> the usual face point; coarse -> the mapped point */
      const double *xo = pos >= 0 ? xs : xc;
      if (pos < 0) wt /= nsub;
      double J[9], Ji[9];
      const double det = multilinear_jacobian(dim, V, xo, J, Ji);
      double nrm[3], na = 0.0;
      for (int i = 0; i < dim; ++i) {
        nrm[i] = Ji[d * dim + i];
        na += nrm[i] * nrm[i];
      }
      na = sqrt(na);
      for (int i = 0; i < dim; ++i) nrm[i] = (s ? 1.0 : -1.0) * nrm[i] / na;
      const double jxw_f = fabs(det) * na * wt;
      double up[4], um[4], val[4];
      double xn[3]; /* the other side's reference point */
      for (int e = 0; e < dim; ++e) xn[e] = pos >= 0 ? xc[e] : xs[e];
      xn[d] = (double)(1 - s);
      for (int k = 0; k < nf; ++k) {
        if (pos >= 0) {
          double a = 0.0;
          for (int i = 0; i < n; ++i) a += b->face[s][i] * u[k * nd + base + i * stride_d];
          up[k] = a;
        } else {
          up[k] = eval_at(b, dim, u + k * nd, xo);
        }
        um[k] = eval_at(b, dim, uo + k * nd, xn);
      }
      face_flux(dim, nrm, up, um, rho, cc, m->rho[other], m->speed[other], 0, fac, val);
      if (pos >= 0) {
        for (int k = 0; k < nf; ++k)
          for (int i = 0; i < n; ++i)
            r[k * nd + base + i * stride_d] += b->face[s][i] * val[k] * jxw_f;
      } else {
        double l[3][OX_MAX_N];
        for (int e = 0; e < dim; ++e) ox_lagrange(n, b->x, xo[e], l[e], NULL);
        for (int q = 0; q < nd; ++q) {
          double phi = 1.0;
          int tq = q;
          for (int e = 0; e < dim; ++e) {
            phi *= l[e][tq % n];
            tq /= n;
          }
          for (int k = 0; k < nf; ++k) r[k * nd + q] += phi * val[k] * jxw_f;
        }
      }
    }
  }
}

typedef struct work {
  double *r, *g, *tmp, *jxw;
} work;

/* evaluate_cell (source/wave_equation_operations.cc:424-456) of cell c with
 * values u: the cell integral is added into r (unscaled by the inverse mass);
 * jxw receives det(J) w_q at the Gauss points. */
static void cell_volume(const ox_mesh *m, const ox_basis *b, int c, const double *u, double *r,
                        double *jxw, work *W) {
  const int dim = m->dim, n = b->n;
  const int nd = ipow(n, dim), nv = 1 << dim;
  const int sz[3] = {n, n, n};
  const double *V = m->vertices + (long)c * nv * dim;
  const double rho = m->rho[c], cc = m->speed[c];
  double *g = W->g, *tmp = W->tmp;
  const double *p = u + dim * nd;
  for (int e = 0; e < dim; ++e) tensor_apply(dim, sz, e, b->D, n, p, tmp + e * nd, 0, 0);
  for (int q = 0; q < nd; ++q) {
    double xi[3], J[9], Ji[9];
    int t = q;
    double wq = 1.0;
    for (int d = 0; d < dim; ++d) {
      xi[d] = b->x[t % n];
      wq *= b->w[t % n];
      t /= n;
    }
    double det = multilinear_jacobian(dim, V, xi, J, Ji);
    double dw = det * wq;
    jxw[q] = dw;
    for (int i = 0; i < dim; ++i) {
      double gp = 0.0;
      for (int e = 0; e < dim; ++e) gp += Ji[e * dim + i] * tmp[e * nd + q];
      r[i * nd + q] += -gp / rho * dw;
    }
    for (int e = 0; e < dim; ++e) {
      double s = 0.0;
      for (int i = 0; i < dim; ++i) s += Ji[e * dim + i] * u[i * nd + q];
      g[e * nd + q] = rho * cc * cc * dw * s;
    }
  }
  for (int e = 0; e < dim; ++e) tensor_apply(dim, sz, e, b->D, n, g + e * nd, r + dim * nd, 1, 1);
}

/* Face f of cell c evaluated from this side (evaluate_inner_face :480-551,
 * evaluate_boundary_face :569-629): u = this cell's values, un = the
 * neighbour's values (ignored at a boundary).  The test-function terms of this
 * side, times boundary_fac `fac`, are added into r. */
static void cell_face(const ox_mesh *m, const ox_basis *b, int c, int f, const double *u,
                      const double *un, double fac, double *r) {
  const int dim = m->dim, n = b->n, nf = dim + 1;
  const int nd = ipow(n, dim), nv = 1 << dim, nt = ipow(n, dim - 1);
  const double *V = m->vertices + (long)c * nv * dim;
  const double rho = m->rho[c], cc = m->speed[c];
  const int d = f >> 1, s = f & 1;
  const int nb = m->neighbor[c * 2 * dim + f];
  const int stride_d = ipow(n, d);
  if (nb < 0) un = NULL;
  for (int t = 0; t < nt; ++t) {
    int base = 0, tt = t;
    double xi[3], wt = 1.0;
    xi[d] = (double)s;
    for (int e = 0; e < dim; ++e) {
      if (e == d) continue;
      int te = tt % n;
      tt /= n;
      base += te * ipow(n, e);
      xi[e] = b->x[te];
      wt *= b->w[te];
    }
    double J[9], Ji[9];
    double det = multilinear_jacobian(dim, V, xi, J, Ji);
    double nrm[3], na = 0.0;
    for (int i = 0; i < dim; ++i) {
      nrm[i] = Ji[d * dim + i];
      na += nrm[i] * nrm[i];
    }
    na = sqrt(na);
    for (int i = 0; i < dim; ++i) nrm[i] = (s ? 1.0 : -1.0) * nrm[i] / na;
    double jxw_f = fabs(det) * na * wt;
    double up[4], um[4] = {0, 0, 0, 0}, val[4];
    for (int k = 0; k < nf; ++k) {
      double a = 0.0, bb = 0.0;
      for (int i = 0; i < n; ++i) {
        a += b->face[s][i] * u[k * nd + base + i * stride_d];
        if (un) bb += b->face[1 - s][i] * un[k * nd + base + i * stride_d];
      }
      up[k] = a;
      um[k] = bb;
    }
    face_flux(dim, nrm, up, um, rho, cc, nb >= 0 ? m->rho[nb] : rho,
              nb >= 0 ? m->speed[nb] : cc, nb >= 0 ? 0 : -nb, fac, val);
    for (int k = 0; k < nf; ++k)
      for (int i = 0; i < n; ++i)
        r[k * nd + base + i * stride_d] += b->face[s][i] * val[k] * jxw_f;
  }
}

/* One cell of M^-1 A(src): evaluate_cell + the faces of this cell +
 * inverse mass (GAUSS_POINTS_VECTOR_OPERATION form,
 * source/wave_equation_operations.cc:640-657). */
static void apply_cell(const ox_mesh *m, const ox_basis *b, int c, const double *src,
                       double *dst, double sign, work *W) {
  const int dim = m->dim, n = b->n, nf = dim + 1;
  const int nd = ipow(n, dim);
  const double *u = src + (long)c * nf * nd;
  double *r = W->r, *jxw = W->jxw;
  memset(r, 0, sizeof(double) * nf * nd);
  cell_volume(m, b, c, u, r, jxw, W);
  for (int f = 0; f < 2 * dim; ++f) {
    const int nb = m->neighbor[c * 2 * dim + f];
    if (nb == OX_NC_FACE)
      cell_face_nc(m, b, c, f, src, -1, 1.0, r);
    else
      cell_face(m, b, c, f, u, nb >= 0 ? src + (long)nb * nf * nd : NULL, 1.0, r);
  }
  /* inverse mass: collocation at Gauss points -> divide by JxW */
  double *o = dst + (long)c * nf * nd;
  for (int k = 0; k < nf; ++k)
    for (int q = 0; q < nd; ++q) o[k * nd + q] = sign * r[k * nd + q] / jxw[q];
}

static work work_alloc(int dim, int n) {
  int nd = ipow(n, dim);
  work W;
  W.r = (double *)malloc(sizeof(double) * 4 * nd);
  W.g = (double *)malloc(sizeof(double) * 4 * nd);
  W.tmp = (double *)malloc(sizeof(double) * 16 * nd);
  W.jxw = (double *)malloc(sizeof(double) * nd);
  return W;
}
static void work_free(work *W) {
  free(W->r);
  free(W->g);
  free(W->tmp);
  free(W->jxw);
}

/* WaveEquationOperation::apply (source/wave_equation_operations.cc:665-685)
 * multiplied by `sign` (-1 reproduces local_apply_secondader_domain +
 * the ADER face sweeps with boundary_fac = -1, :1265-1342). */
void ox_apply(const ox_mesh *m, int degree, const double *src, double *dst, double sign) {
  ox_basis b;
  ox_basis_init(&b, degree);
#pragma omp parallel
  {
    work W = work_alloc(m->dim, b.n);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) apply_cell(m, &b, c, src, dst, sign, &W);
    work_free(&W);
  }
}

/* RKVectorUpdater::apply_to_subrange (source/time_integrators.cc:128-155) */
void ox_rk_update(long n, double f1, double f2, int is_last, const double *v1, double *v2,
                  double *v3) {
  if (is_last) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) v2[i] = v3[i] + f1 * v1[i];
  } else {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      double up = v1[i];
      v2[i] += f1 * up;
      v3[i] = v2[i] + f2 * up;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* ADER Cauchy-Kowalevski (source/wave_equation_operations.cc:961-1262)      */
/* ------------------------------------------------------------------------ */

typedef struct ck_ctx {
  int dim, degree;
  const double *V;
  double rho, c2rho, rho_inv;
  double tstart, tend;
} ck_ctx;

/* integrate_taylor_cauchykovalewski_step<step_no, add_into_contrib>
 * (:1063-1262) with spectral_evaluation = true.  s holds (dim+1) fields at
 * the (my_degree+1)^dim Gauss points of this step's degree. */
static void ck_step(const ck_ctx *X, int step_no, int add, double *s, double *contrib) {
  const int dim = X->dim, k = X->degree;
  double fac = -0.5;
  for (int kk = 2; kk <= step_no + 2; ++kk) fac /= -(double)(kk + 1);
  const double fac_t = fac * (pow(X->tend, step_no + 3) - pow(X->tstart, step_no + 3));
  const int reduce_step = step_no / 2;
  const int at_reduce_step = (step_no - 2 * reduce_step) == 1;
  const int my_degree = reduce_step > 0 ? k - 2 * reduce_step : k;
  const int n = my_degree + 1, nd = ipow(n, dim);
  const int sz[3] = {n, n, n};
  double x[OX_MAX_N], w[OX_MAX_N], D[OX_MAX_N * OX_MAX_N];
  ox_gauss(n, x, w);
  for (int q = 0; q < n; ++q) ox_lagrange(n, x, x[q], NULL, D + q * n);

  /* reference-coordinate gradients of every field */
  double *gr = (double *)malloc(sizeof(double) * (dim + 1) * dim * nd);
  for (int f = 0; f <= dim; ++f)
    for (int e = 0; e < dim; ++e)
      tensor_apply(dim, sz, e, D, n, s + f * nd, gr + (f * dim + e) * nd, 0, 0);
  for (int q = 0; q < nd; ++q) {
    double xi[3], J[9], Ji[9];
    int t = q;
    for (int d = 0; d < dim; ++d) {
      xi[d] = x[t % n];
      t /= n;
    }
    multilinear_jacobian(dim, X->V, xi, J, Ji);
    double div = 0.0, pg[3];
    for (int d = 0; d < dim; ++d)
      for (int e = 0; e < dim; ++e) div += gr[(d * dim + e) * nd + q] * Ji[e * dim + d];
    for (int d = 0; d < dim; ++d) {
      pg[d] = 0.0;
      for (int e = 0; e < dim; ++e) pg[d] += gr[(dim * dim + e) * nd + q] * Ji[e * dim + d];
    }
    for (int d = 0; d < dim; ++d) {
      double val = X->rho_inv * pg[d];
      contrib[d * nd + q] = (add ? contrib[d * nd + q] : 0.0) + fac_t * val;
      s[d * nd + q] = val;
    }
    double val = X->c2rho * div;
    contrib[dim * nd + q] = (add ? contrib[dim * nd + q] : 0.0) + fac_t * val;
    s[dim * nd + q] = val;
  }
  free(gr);

  if (step_no < k - 2) {
    if (at_reduce_step) {
      /* weighted L2 projection degree my_degree -> next_degree on the unit
       * cell (setup :287-313, do_backward :1244) */
      const int next_degree = my_degree >= 2 ? my_degree - 2 : 1;
      const int nl = next_degree + 1, ndl = ipow(nl, dim);
      double xl[OX_MAX_N], wl[OX_MAX_N];
      ox_gauss(nl, xl, wl);
      double Vlh[OX_MAX_N * OX_MAX_N]; /* Vlh[q*nl+i] = l^low_i(x^high_q) */
      double P[OX_MAX_N * OX_MAX_N];   /* P[i*n+q] = Vlh[q][i] w_q / wl_i */
      for (int q = 0; q < n; ++q) ox_lagrange(nl, xl, x[q], Vlh + q * nl, NULL);
      for (int i = 0; i < nl; ++i)
        for (int q = 0; q < n; ++q) P[i * n + q] = Vlh[q * nl + i] * w[q] / wl[i];
      double *a = (double *)malloc(sizeof(double) * nd);
      double *bb = (double *)malloc(sizeof(double) * nd);
      double *slow = (double *)malloc(sizeof(double) * (dim + 1) * ndl);
      double *clow = (double *)malloc(sizeof(double) * (dim + 1) * ndl);
      for (int f = 0; f <= dim; ++f) {
        /* apply P along each direction, shrinking one direction at a time */
        int szc[3] = {n, n, n};
        const double *in = s + f * nd;
        double *bufs[2] = {a, bb};
        for (int e = 0; e < dim; ++e) {
          double *out = (e == dim - 1) ? slow + f * ndl : bufs[e & 1];
          tensor_apply(dim, szc, e, P, nl, in, out, 0, 0);
          szc[e] = nl;
          in = out;
        }
      }
      ck_step(X, step_no + 1, 0, slow, clow);
      /* interpolate back (do_forward :1252) and add */
      for (int f = 0; f <= dim; ++f) {
        int szc[3] = {nl, nl, nl};
        const double *in = clow + f * ndl;
        double *bufs[2] = {a, bb};
        for (int e = 0; e < dim; ++e) {
          double *out = bufs[e & 1];
          tensor_apply(dim, szc, e, Vlh, n, in, out, 0, 0);
          szc[e] = n;
          in = out;
        }
        for (int q = 0; q < nd; ++q) contrib[f * nd + q] += in[q];
      }
      free(a);
      free(bb);
      free(slow);
      free(clow);
    } else {
      ck_step(X, step_no + 1, 1, s, contrib);
    }
  }
}

/* integrate_taylor_cauchykovalewski (:961-1059) for all cells */
void ox_ader_ck(const ox_mesh *m, int degree, const double *src, const double *recon, double t2,
                double t1, double te, double *tempsrc) {
  ox_basis b;
  ox_basis_init(&b, degree);
  const int dim = m->dim, n = b.n, nf = dim + 1, nd = ipow(n, dim), nv = 1 << dim;
  const int sz[3] = {n, n, n};
#pragma omp parallel
  {
    double *s = (double *)malloc(sizeof(double) * nf * nd);
    double *contrib = (double *)malloc(sizeof(double) * nf * nd);
    double *gr = (double *)malloc(sizeof(double) * nf * dim * nd);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) {
      const double *u = src + (long)c * nf * nd;
      ck_ctx X;
      X.dim = dim;
      X.degree = degree;
      X.V = m->vertices + (long)c * nv * dim;
      X.rho = m->rho[c];
      X.rho_inv = 1.0 / m->rho[c];
      X.c2rho = m->speed[c] * m->speed[c] * m->rho[c];
      X.tstart = t1 - te;
      X.tend = t2 - te;
      const double c1 = 0.5 * ((t2 - te) * (t2 - te) - (t1 - te) * (t1 - te));
      for (int i = 0; i < nf * nd; ++i) contrib[i] = (t2 - t1) * u[i];
      if (recon) {
        const double *r = recon + (long)c * nf * nd;
        for (int i = 0; i < nf * nd; ++i) {
          contrib[i] += c1 * r[i];
          s[i] = -r[i];
        }
      } else {
        for (int f = 0; f < nf; ++f)
          for (int e = 0; e < dim; ++e)
            tensor_apply(dim, sz, e, b.D, n, u + f * nd, gr + (f * dim + e) * nd, 0, 0);
        for (int q = 0; q < nd; ++q) {
          double xi[3], J[9], Ji[9];
          int t = q;
          for (int d = 0; d < dim; ++d) {
            xi[d] = b.x[t % n];
            t /= n;
          }
          multilinear_jacobian(dim, X.V, xi, J, Ji);
          double div = 0.0;
          for (int d = 0; d < dim; ++d) {
            double pg = 0.0;
            for (int e = 0; e < dim; ++e) {
              pg += gr[(dim * dim + e) * nd + q] * Ji[e * dim + d];
              div += gr[(d * dim + e) * nd + q] * Ji[e * dim + d];
            }
            double nx = -X.rho_inv * pg;
            contrib[d * nd + q] += c1 * nx;
            s[d * nd + q] = -nx;
          }
          double np = -X.c2rho * div;
          contrib[dim * nd + q] += c1 * np;
          s[dim * nd + q] = -np;
        }
      }
      /* step 0 runs for every degree (the recursion guard is inside) */
      ck_step(&X, 0, 1, s, contrib);
      memcpy(tempsrc + (long)c * nf * nd, contrib, sizeof(double) * nf * nd);
    }
    free(s);
    free(contrib);
    free(gr);
  }
}

/* WaveEquationOperationADER::apply_ader (:1346-1381); scratch: 2 vectors */
void ox_apply_ader(const ox_mesh *m, int degree, const double *src, double *dst, double dt,
                   int recon, double *scratch) {
  long N = ox_n_dofs(m, degree);
  double *tempsrc = scratch, *rec = scratch + N;
  if (recon) ox_apply(m, degree, src, rec, 1.0);
  ox_ader_ck(m, degree, src, recon ? rec : NULL, dt, 0.0, 0.0, tempsrc);
  ox_apply(m, degree, tempsrc, dst, -1.0);
}

/* ------------------------------------------------------------------------ */
/* time integrators (source/time_integrators.cc)                             */
/* ------------------------------------------------------------------------ */

/* Shu-Osher coefficients of the optimal SSPRK with 8 stages and order 4
 * (S. J. Ruuth, Math. Comp. 75 (2006) 183-207), as tabulated in
 * SSPRK::initialize_coeffs (source/time_integrators.cc:726-858); row s-1
 * forms stage s from stages 0..s-1 */
static const double ox_ssprk84_a[8][8] = {
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
static const double ox_ssprk84_b[8][8] = {
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

static void lsrk_2reg(const ox_mesh *m, int degree, int S, const double *a, const double *bcoef,
                      double dt, double *vec_n, double *vec_np, double *tmp1) {
  long N = ox_n_dofs(m, degree);
  /* stage 1: apply(vec_n) ; update(v2 = vec_n, v3 = vec_np) then the
   * registers swap roles every stage (:260-280) */
  double *src = vec_n, *v2 = vec_n, *v3 = vec_np;
  for (int s = 0; s < S; ++s) {
    ox_apply(m, degree, src, tmp1, 1.0);
    if (s < S - 1) {
      ox_rk_update(N, a[s] * dt, (bcoef[s] - a[s]) * dt, 0, tmp1, v2, v3);
      /* next stage input is v2, accumulator v3 */
      double *nv2 = v3, *nv3 = v2;
      src = v2;
      v2 = nv2;
      v3 = nv3;
    } else {
      /* last stage: the input register receives accumulator + f1*k
       * (e.g. stage 5: update(vec_tmp1, vec_np, vec_n), :279-280) */
      ox_rk_update(N, bcoef[s] * dt, 0.0, 1, tmp1, src, v2);
    }
  }
  if (src != vec_np) memcpy(vec_np, src, sizeof(double) * N);
}

void ox_time_step(const ox_mesh *m, int degree, int integrator, double dt, double *vec_n,
                  double *vec_np, double *tmp1, double *tmp2, int ader_recon) {
  long N = ox_n_dofs(m, degree);
  if (integrator == OX_EXPL_EULER) {
    /* :24-34 */
    ox_apply(m, degree, vec_n, vec_np, 1.0);
    for (long i = 0; i < N; ++i) vec_np[i] = dt * vec_np[i] + vec_n[i];
  } else if (integrator == OX_CLRK4) {
    /* :66-102 */
    const double w[4] = {dt / 6., dt / 3., dt / 3., dt / 6.};
    const double cfac[3] = {0.5 * dt, 0.5 * dt, dt};
    ox_apply(m, degree, vec_n, tmp1, 1.0);
    for (int st = 0; st < 4; ++st) {
      if (st > 0) ox_apply(m, degree, tmp2, tmp1, 1.0);
      if (st == 0)
        for (long i = 0; i < N; ++i) vec_np[i] = vec_n[i] + w[0] * tmp1[i];
      else
        for (long i = 0; i < N; ++i) vec_np[i] += w[st] * tmp1[i];
      if (st < 3)
        for (long i = 0; i < N; ++i) tmp2[i] = vec_n[i] + cfac[st] * tmp1[i];
    }
  } else if (integrator == OX_LSRK45R2) {
    /* :235-283 */
    const double a[5] = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
                         2251764453980. / 15575788980749., 26877169314380. / 34165994151039.,
                         0.0};
    const double b[5] = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
                         -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
                         5198255086312. / 14908931495163.};
    lsrk_2reg(m, degree, 5, a, b, dt, vec_n, vec_np, tmp1);
  } else if (integrator == OX_LSRK33R2) {
    /* :200-231 */
    const double a[3] = {0.755726351946097, 0.386954477304099, 0.0};
    const double b[3] = {0.245170287303492, 0.184896052186740, 0.569933660509768};
    lsrk_2reg(m, degree, 3, a, b, dt, vec_n, vec_np, tmp1);
  } else if (integrator == OX_LSRK59R2) {
    /* :286-363 */
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
    lsrk_2reg(m, degree, 9, a, b, dt, vec_n, vec_np, tmp1);
  } else if (integrator == OX_LSRK45R3) {
    /* LowStorageRK45Reg3::perform_time_step (:366-425), its vector operations
     * in order; vec_n is overwritten (stage 3 writes into it) */
    const double a21 = 2365592473904. / 8146167614645., a32 = 4278267785271. / 6823155464066.,
                 a43 = 2789585899612. / 8986505720531., a54 = 15310836689591. / 24358012670437.,
                 a31 = -722262345248. / 10870640012513., a42 = 1365858020701. / 8494387045469.,
                 a53 = 3819021186. / 2763618202291.;
    const double b1 = 846876320697. / 6523801458457., b2 = 3032295699695. / 12397907741132.,
                 b3 = 612618101729. / 6534652265123., b4 = 1155491934595. / 2954287928812.,
                 b5 = 707644755468. / 5028292464395.;
    const double h = dt;
    ox_apply(m, degree, vec_n, vec_np, 1.0);
    for (long i = 0; i < N; ++i) {
      vec_n[i] += a21 * h * vec_np[i];
      tmp1[i] = vec_n[i] + (b1 - a21) * h * vec_np[i];
    }
    ox_apply(m, degree, vec_n, tmp2, 1.0);
    for (long i = 0; i < N; ++i) {
      tmp1[i] += a32 * h * tmp2[i] + (a31 - b1) * h * vec_np[i];
      vec_np[i] = (b1 - a31) * h * vec_np[i] + tmp1[i];
      vec_np[i] += (b2 - a32) * h * tmp2[i];
    }
    ox_apply(m, degree, tmp1, vec_n, 1.0);
    for (long i = 0; i < N; ++i) {
      vec_np[i] += a43 * h * vec_n[i] + (a42 - b2) * h * tmp2[i];
      tmp2[i] = (b2 - a42) * h * tmp2[i] + vec_np[i];
      tmp2[i] += (b3 - a43) * h * vec_n[i];
    }
    ox_apply(m, degree, vec_np, tmp1, 1.0);
    for (long i = 0; i < N; ++i) {
      tmp2[i] += a54 * h * tmp1[i] + (a53 - b3) * h * vec_n[i];
      vec_np[i] = tmp2[i];
      vec_np[i] += (b3 - a53) * h * vec_n[i] + (b4 - a54) * h * tmp1[i];
    }
    ox_apply(m, degree, tmp2, tmp1, 1.0);
    for (long i = 0; i < N; ++i) vec_np[i] += b5 * h * tmp1[i];
  } else if (integrator == OX_SSPRK) {
    /* SSPRK::perform_time_step (:440-479) with order 4, 8 stages: u_0 = vec_n,
     * u_s = sum_{l<s} A[s-1][l] u_l + dt B[s-1][l] F(u_l), vec_np = u_8 */
    const int S = 8;
    double *u = (double *)malloc(sizeof(double) * N * (2 * S + 1)), *F = u + N * (S + 1);
    memcpy(u, vec_n, sizeof(double) * N);
    for (int st = 1; st <= S; ++st) {
      ox_apply(m, degree, u + (long)(st - 1) * N, F + (long)(st - 1) * N, 1.0);
      double *us = u + (long)st * N;
      for (long i = 0; i < N; ++i) us[i] = 0.0;
      for (int l = 0; l < st; ++l) {
        const double al = ox_ssprk84_a[st - 1][l], bl = ox_ssprk84_b[st - 1][l] * dt;
        const double *ul = u + (long)l * N, *Fl = F + (long)l * N;
        for (long i = 0; i < N; ++i) us[i] += al * ul[i] + bl * Fl[i];
      }
    }
    memcpy(vec_np, u + (long)S * N, sizeof(double) * N);
    free(u);
  } else if (integrator == OX_ADER_ADCONFULL) {
    /* WaveEquationOperationADERADCONFULL::apply_ader (wave_equation_operations.cc:
     * 689-735): the Taylor series of the full operator; then ArbitraryHighOrderDG
     * (time_integrators.cc:37-50) vec_np = vec_n - apply_ader(vec_n) */
    double *tempsrc = tmp1, *tempvals = tmp2;
    ox_apply(m, degree, vec_n, vec_np, 1.0);
    for (long i = 0; i < N; ++i) {
      tempsrc[i] = -dt * vec_np[i];
      tempvals[i] = -vec_np[i];
    }
    double fac = -0.5 * dt * dt;
    for (int k = 1; k <= degree; ++k) {
      ox_apply(m, degree, tempvals, vec_np, 1.0);
      for (long i = 0; i < N; ++i) {
        tempsrc[i] += -fac * vec_np[i];
        tempvals[i] = -vec_np[i];
      }
      fac *= -dt / (k + 2);
    }
    for (long i = 0; i < N; ++i) vec_np[i] = -tempsrc[i] + vec_n[i];
  } else if (integrator == OX_ADER) {
    /* ArbitraryHighOrderDG (:37-50): vec_np = vec_n - apply_ader(vec_n);
     * tmp1 and tmp2 must be adjacent (2N scratch) -> use both explicitly */
    double *scr = (double *)malloc(sizeof(double) * 2 * N);
    ox_apply_ader(m, degree, vec_n, vec_np, dt, ader_recon, scr);
    free(scr);
    for (long i = 0; i < N; ++i) vec_np[i] = -vec_np[i] + vec_n[i];
    (void)tmp1;
    (void)tmp2;
  }
}

/* ------------------------------------------------------------------------ */
/* exact solution, projection, norms                                          */
/* ------------------------------------------------------------------------ */

/* Superposition of membrane modes; for one mode with m_d = M and rho=c=1 it is
 * input_exact_solution (include/input_parameters.h:94-128). */
double ox_exact_value(const ox_exact *e, const double *x, double t, int component) {
  const int dim = e->dim;
  double val = 0.0;
  for (int k = 0; k < e->n_modes; ++k) {
    double mm = 0.0;
    for (int d = 0; d < dim; ++d) mm += (double)e->m[k][d] * e->m[k][d];
    mm = sqrt(mm);
    double omega = e->c * M_PI * mm;
    double v;
    if (component == dim)
      v = e->amp[k] * cos(omega * t);
    else
      v = -e->amp[k] * e->m[k][component] / (e->rho * e->c * mm) * sin(omega * t);
    for (int d = 0; d < dim; ++d)
      v *= (d != component) ? sin(e->m[k][d] * M_PI * x[d]) : cos(e->m[k][d] * M_PI * x[d]);
    val += v;
  }
  return val;
}

/* project_initial_field (:902-935): with the Gauss-collocated basis the
 * QGauss(k+1) L2 projection is the interpolation at the mapped Gauss points. */
void ox_project(const ox_mesh *m, int degree, const ox_exact *e, double t, double *dst) {
  ox_basis b;
  ox_basis_init(&b, degree);
  const int dim = m->dim, n = b.n, nf = dim + 1, nd = ipow(n, dim), nv = 1 << dim;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < m->n_cells; ++c) {
    const double *V = m->vertices + (long)c * nv * dim;
    for (int q = 0; q < nd; ++q) {
      double xi[3], x[3];
      int t2 = q;
      for (int d = 0; d < dim; ++d) {
        xi[d] = b.x[t2 % n];
        t2 /= n;
      }
      multilinear_point(dim, V, xi, x);
      for (int k = 0; k < nf; ++k) dst[((long)c * nf + k) * nd + q] = ox_exact_value(e, x, t, k);
    }
  }
}

/* VectorTools::integrate_difference with QGauss(k+2) (source/explicit_wave.cc:452-496) */
void ox_l2_errors(const ox_mesh *m, int degree, const ox_exact *e, double t, const double *u,
                  double *out) {
  ox_basis b;
  ox_basis_init(&b, degree);
  const int dim = m->dim, n = b.n, ne = n + 1, nf = dim + 1, nd = ipow(n, dim),
            nde = ipow(ne, dim), nv = 1 << dim;
  double xe[OX_MAX_N], we[OX_MAX_N], I[OX_MAX_N * OX_MAX_N];
  ox_gauss(ne, xe, we);
  for (int q = 0; q < ne; ++q) ox_lagrange(n, b.x, xe[q], I + q * n, NULL);
  double ep = 0.0, ev = 0.0, mp = 0.0;
#pragma omp parallel reduction(+ : ep, ev, mp)
  {
    double *a = (double *)malloc(sizeof(double) * nde);
    double *bb = (double *)malloc(sizeof(double) * nde);
    double *ue = (double *)malloc(sizeof(double) * nf * nde);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) {
      const double *V = m->vertices + (long)c * nv * dim;
      for (int k = 0; k < nf; ++k) {
        int szc[3] = {n, n, n};
        const double *in = u + ((long)c * nf + k) * nd;
        double *bufs[2] = {a, bb};
        for (int d = 0; d < dim; ++d) {
          double *o = (d == dim - 1) ? ue + k * nde : bufs[d & 1];
          tensor_apply(dim, szc, d, I, ne, in, o, 0, 0);
          szc[d] = ne;
          in = o;
        }
      }
      for (int q = 0; q < nde; ++q) {
        double xi[3], x[3], J[9], Ji[9], w = 1.0;
        int t2 = q;
        for (int d = 0; d < dim; ++d) {
          xi[d] = xe[t2 % ne];
          w *= we[t2 % ne];
          t2 /= ne;
        }
        multilinear_point(dim, V, xi, x);
        double jxw = fabs(multilinear_jacobian(dim, V, xi, J, Ji)) * w;
        double pex = ox_exact_value(e, x, t, dim);
        double dp = ue[dim * nde + q] - pex;
        ep += dp * dp * jxw;
        mp += pex * pex * jxw;
        for (int d = 0; d < dim; ++d) {
          double dv = ue[d * nde + q] - ox_exact_value(e, x, t, d);
          ev += dv * dv * jxw;
        }
      }
    }
    free(a);
    free(bb);
    free(ue);
  }
  out[0] = sqrt(ep);
  out[1] = sqrt(ev);
  out[2] = sqrt(mp);
}

double ox_energy(const ox_mesh *m, int degree, const double *u) {
  ox_basis b;
  ox_basis_init(&b, degree);
  const int dim = m->dim, n = b.n, nf = dim + 1, nd = ipow(n, dim), nv = 1 << dim;
  double E = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : E)
  for (int c = 0; c < m->n_cells; ++c) {
    const double *V = m->vertices + (long)c * nv * dim;
    const double rho = m->rho[c], c2 = m->speed[c] * m->speed[c];
    for (int q = 0; q < nd; ++q) {
      double xi[3], J[9], Ji[9], w = 1.0;
      int t2 = q;
      for (int d = 0; d < dim; ++d) {
        xi[d] = b.x[t2 % n];
        w *= b.w[t2 % n];
        t2 /= n;
      }
      double jxw = fabs(multilinear_jacobian(dim, V, xi, J, Ji)) * w;
      const double *uc = u + (long)c * nf * nd;
      double v2 = 0.0;
      for (int d = 0; d < dim; ++d) v2 += uc[d * nd + q] * uc[d * nd + q];
      double p = uc[dim * nd + q];
      E += 0.5 * jxw * (rho * v2 + p * p / (rho * c2));
    }
  }
  return E;
}

/* WaveEquationOperation::estimate_error (source/wave_equation_operations.cc:
 * 871-898): tmp = apply(solution); per cell
 * err_K = sqrt( int_K |grad p_h - v_t|^2 dx ) at the QGauss(k+1) points (the
 * nodes), v_t the velocity part of tmp, restated literally. */
void ox_estimate_error(const ox_mesh *m, int degree, const double *solution, double *err) {
  const int dim = m->dim, n = degree + 1, nf = dim + 1, nd = ipow(n, dim), nv = 1 << dim;
  const long N = ox_n_dofs(m, degree);
  double *tmp = (double *)malloc(sizeof(double) * N);
  ox_apply(m, degree, solution, tmp, 1.0);
  ox_basis b;
  ox_basis_init(&b, degree);
#pragma omp parallel
  {
    double *g = (double *)malloc(sizeof(double) * dim * nd);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) {
      const double *V = m->vertices + (long)c * nv * dim;
      const double *p = solution + ((long)c * nf + dim) * nd;
      const double *vt = tmp + (long)c * nf * nd;
      const int sz[3] = {n, n, n};
      for (int e = 0; e < dim; ++e) tensor_apply(dim, sz, e, b.D, n, p, g + (long)e * nd, 0, 0);
      double s = 0.0;
      for (int q = 0; q < nd; ++q) {
        double xi[3], J[9], Ji[9], w = 1.0;
        int t = q;
        for (int d = 0; d < dim; ++d) {
          xi[d] = b.x[t % n];
          w *= b.w[t % n];
          t /= n;
        }
        const double det = multilinear_jacobian(dim, V, xi, J, Ji);
        for (int i = 0; i < dim; ++i) {
          double gp = 0.0;
          for (int e = 0; e < dim; ++e) gp += Ji[e * dim + i] * g[(long)e * nd + q];
          const double dlt = gp - vt[(long)i * nd + q];
          s += dlt * dlt * fabs(det) * w;
        }
      }
      err[c] = sqrt(s);
    }
    free(g);
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* adaptive refinement, one level (source/explicit_wave.cc:241-330)          */
/* ------------------------------------------------------------------------ */

static int cmp_desc(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return x < y ? 1 : (x > y ? -1 : 0);
}

void ox_refinement_flags(int n, const double *err, const int *level, int min_level, int max_level,
                         double max_error_init, int *flags) {
  /* refine_and_coarsen_fixed_number(tria, err, 0.1, 0.6): the int(0.1 n)-th
   * largest value is the refinement threshold (cells >= it), the int(0.6 n)-th
   * smallest the coarsening threshold (cells <= it, not flagged for
   * refinement); a full sort gives the same order statistics */
  double *s = (double *)malloc(sizeof(double) * n);
  memcpy(s, err, sizeof(double) * n);
  qsort(s, n, sizeof(double), cmp_desc);
  const int nr = (int)(0.1 * n), ncs = (int)(0.6 * n);
  for (int c = 0; c < n; ++c) flags[c] = 0;
  if (nr)
    for (int c = 0; c < n; ++c)
      if (fabs(err[c]) >= s[nr - 1]) flags[c] = 1;
  if (ncs)
    for (int c = 0; c < n; ++c)
      if (fabs(err[c]) <= s[n - ncs] && flags[c] != 1) flags[c] = -1;
  free(s);
  /* adapt_mesh (:281-296) */
  for (int c = 0; c < n; ++c) {
    if (flags[c] == 1 && level[c] == max_level)
      flags[c] = 0;
    else if (flags[c] == -1 && level[c] == min_level)
      flags[c] = 0;
    if (flags[c] == 1 && err[c] < 0.1 * max_error_init) flags[c] = 0;
    if (err[c] < 0.05 * max_error_init) flags[c] = -1;
  }
}

void ox_adapt_refine(int dim, int n_intervals, const int *refine_old, const int *flags,
                     int *refine_new) {
  const int nb = ipow(n_intervals, dim), nch = 1 << dim;
  int c = 0;
  for (int B = 0; B < nb; ++B) {
    if (!refine_old[B]) {
      refine_new[B] = flags[c++] == 1;
    } else {
      int all = 1;
      for (int ch = 0; ch < nch; ++ch) all &= flags[c + ch] == -1;
      refine_new[B] = !all;
      c += nch;
    }
  }
}

/* a child ch of a cell: the parent's polynomial at the child's nodes (all
 * fields; up and out hold (dim+1) (k+1)^dim values) */
static void transfer_child(const ox_basis *b, int dim, int ch, const double *up, double *out) {
  const int n = b->n, nd = ipow(n, dim), nf = dim + 1;
  for (int f = 0; f < nf; ++f)
    for (int q = 0; q < nd; ++q) {
      double xi[3];
      int t = q;
      for (int d = 0; d < dim; ++d) {
        xi[d] = 0.5 * (b->x[t % n] + ((ch >> d) & 1));
        t /= n;
      }
      out[f * nd + q] = eval_at(b, dim, up + f * nd, xi);
    }
}

/* a parent from its 2^dim children (consecutive in uc): parent_i =
 * (1/w_i) int_ref l_i u with QGauss(k+2) on each child */
static void transfer_parent(const ox_basis *b, int dim, const double *uc, double *out) {
  const int n = b->n, nd = ipow(n, dim), nf = dim + 1, nch = 1 << dim;
  const int nq = n + 1, nqd = ipow(nq, dim);
  double xq[OX_MAX_N], wq[OX_MAX_N];
  ox_gauss(nq, xq, wq);
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < nd; ++i) {
      double acc = 0.0, wi = 1.0, li[3][OX_MAX_N];
      int ti = i;
      int id[3];
      for (int d = 0; d < dim; ++d) {
        id[d] = ti % n;
        ti /= n;
        wi *= b->w[id[d]];
      }
      for (int ch = 0; ch < nch; ++ch)
        for (int q = 0; q < nqd; ++q) {
          double xc[3], xp[3], w = 1.0 / nch;
          int t = q;
          for (int d = 0; d < dim; ++d) {
            xc[d] = xq[t % nq];
            w *= wq[t % nq];
            t /= nq;
            xp[d] = 0.5 * (xc[d] + ((ch >> d) & 1));
          }
          double phi = 1.0;
          for (int d = 0; d < dim; ++d) {
            ox_lagrange(n, b->x, xp[d], li[d], NULL);
            phi *= li[d][id[d]];
          }
          acc += w * phi * eval_at(b, dim, uc + ((long)ch * nf + f) * nd, xc);
        }
      out[f * nd + i] = acc / wi;
    }
}

void ox_adapt_transfer(int dim, int degree, int n_intervals, const int *refine_old,
                       const int *refine_new, const double *old_u, double *new_u) {
  ox_basis b;
  ox_basis_init(&b, degree);
  const int nd = ipow(b.n, dim), nf = dim + 1, nch = 1 << dim, nb = ipow(n_intervals, dim);
  const long sz = (long)nf * nd;
  long co = 0, cn = 0;
  for (int B = 0; B < nb; ++B) {
    const int ro = refine_old[B] != 0, rn = refine_new[B] != 0;
    if (ro == rn) {
      const int m = ro ? nch : 1;
      memcpy(new_u + cn * sz, old_u + co * sz, sizeof(double) * m * sz);
      cn += m;
    } else if (rn) {
      for (int ch = 0; ch < nch; ++ch, ++cn) transfer_child(&b, dim, ch, old_u + co * sz, new_u + cn * sz);
    } else {
      transfer_parent(&b, dim, old_u + co * sz, new_u + cn * sz);
      ++cn;
    }
    co += ro ? nch : 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Forests of refinement trees: several levels of local refinement          */
/* (adapt_mesh, source/explicit_wave.cc:265-330, with min_level =           */
/* n_refinements and max_level = min_level + n_adaptive_refinements,        */
/* :272-273).  tree[] lists the base cells of the n_intervals^dim box       */
/* lexicographically, each depth first: 1 = refined (its 2^dim children     */
/* follow, lexicographic), 0 = an active cell.  The oracle works on the     */
/* dense picture: lev[g] over the finest grid (n_intervals 2^L cells per    */
/* direction) is the level of the active cell that covers fine cell g.      */
/* ------------------------------------------------------------------------ */

typedef struct fcells {
  int n, max_level, n_tree;
  int *level;
  long *coord; /* n * 3: integer coordinates at the cell's own level */
} fcells;

static void fcells_push(fcells *C, int *cap, int l, const long *I) {
  if (C->n == *cap) {
    *cap = *cap ? 2 * *cap : 64;
    C->level = (int *)realloc(C->level, sizeof(int) * *cap);
    C->coord = (long *)realloc(C->coord, sizeof(long) * 3 * *cap);
  }
  C->level[C->n] = l;
  for (int d = 0; d < 3; ++d) C->coord[3 * C->n + d] = I[d];
  if (l > C->max_level) C->max_level = l;
  ++C->n;
}

static int fparse(int dim, const int *tree, int n_tree, int *pos, int l, const long *I,
                  fcells *C, int *cap) {
  if (*pos >= n_tree || l > 20) return -1;
  const int t = tree[(*pos)++];
  if (t == 0) {
    fcells_push(C, cap, l, I);
    return 0;
  }
  if (t != 1) return -1;
  for (int ch = 0; ch < (1 << dim); ++ch) {
    long J[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) J[d] = 2 * I[d] + ((ch >> d) & 1);
    if (fparse(dim, tree, n_tree, pos, l + 1, J, C, cap)) return -1;
  }
  return 0;
}

/* the active cells of a forest in its order; n_tree < 0: read to the end of
 * the last base cell */
static int forest_cells(int dim, int N, const int *tree, int n_tree, fcells *C) {
  memset(C, 0, sizeof(*C));
  int cap = 0, pos = 0;
  const int nb = ipow(N, dim);
  for (int B = 0; B < nb; ++B) {
    long I[3] = {0, 0, 0};
    int t = B;
    for (int d = 0; d < dim; ++d) {
      I[d] = t % N;
      t /= N;
    }
    if (fparse(dim, tree, n_tree < 0 ? 1 << 30 : n_tree, &pos, 0, I, C, &cap)) return -1;
  }
  C->n_tree = pos;
  return 0;
}

static void fcells_free(fcells *C) {
  free(C->level);
  free(C->coord);
}

typedef struct fgrid {
  int dim, L;
  long G, ng;
} fgrid;

static long fg_at(const fgrid *F, const long *g) {
  long i = 0;
  for (int d = F->dim - 1; d >= 0; --d) i = i * F->G + g[d];
  return i;
}
static void fg_coord(const fgrid *F, long i, long *g) {
  g[0] = g[1] = g[2] = 0;
  for (int d = 0; d < F->dim; ++d) {
    g[d] = i % F->G;
    i /= F->G;
  }
}
/* calls the block of fine cells of the cell (l, I) */
static void fg_fill(const fgrid *F, int *arr, int l, const long *I, int val) {
  const long s = 1L << (F->L - l);
  long lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  for (int d = 0; d < F->dim; ++d) {
    lo[d] = I[d] * s;
    hi[d] = lo[d] + s;
  }
  long g[3];
  for (g[2] = lo[2]; g[2] < hi[2]; ++g[2])
    for (g[1] = lo[1]; g[1] < hi[1]; ++g[1])
      for (g[0] = lo[0]; g[0] < hi[0]; ++g[0]) arr[fg_at(F, g)] = val;
}

int ox_forest_levels(int dim, int n_intervals, const int *tree, int n_tree, int *level) {
  fcells C;
  if (forest_cells(dim, n_intervals, tree, n_tree, &C)) return -1;
  if (level) memcpy(level, C.level, sizeof(int) * C.n);
  const int n = C.n;
  fcells_free(&C);
  return n;
}

ox_mesh *ox_mesh_forest(int dim, int n_intervals, int transform, double factor, int boundary_id,
                        const int *tree, int n_tree) {
  fcells C;
  if (forest_cells(dim, n_intervals, tree, n_tree, &C)) return NULL;
  const int nv = 1 << dim, nface = 2 * dim, nsub = 1 << (dim - 1), rec = 2 + nsub;
  fgrid F = {dim, C.max_level, (long)n_intervals << C.max_level, 0};
  F.ng = 1;
  for (int d = 0; d < dim; ++d) F.ng *= F.G;
  int *owner = (int *)malloc(sizeof(int) * F.ng);
  for (int c = 0; c < C.n; ++c) fg_fill(&F, owner, C.level[c], C.coord + 3 * c, c);
  ox_mesh *m = (ox_mesh *)calloc(1, sizeof(ox_mesh));
  m->dim = dim;
  m->n_cells = C.n;
  m->vertices = (double *)malloc(sizeof(double) * C.n * nv * dim);
  m->neighbor = (int *)malloc(sizeof(int) * C.n * nface);
  m->rho = (double *)malloc(sizeof(double) * C.n);
  m->speed = (double *)malloc(sizeof(double) * C.n);
  m->nc_rec = (int *)malloc(sizeof(int) * C.n * nface);
  m->nc_pos = (int *)malloc(sizeof(int) * C.n * nface);
  m->hanging = (int *)malloc(sizeof(int) * ((long)C.n * nface + 1) * rec);
  m->n_hanging = 0;
  double mind = 1e300;
  int bad = 0;
  for (int c = 0; c < C.n && !bad; ++c) {
    const int l = C.level[c];
    const long *I = C.coord + 3 * c, s = 1L << (F.L - l);
    /* vertices: the base cell's (transformed) multilinear map at dyadic points */
    long B[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) B[d] = I[d] >> l;
    double PV[8 * 3];
    for (int v = 0; v < nv; ++v) {
      double x[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < dim; ++d) x[d] = (double)(B[d] + ((v >> d) & 1)) / n_intervals;
      transform_vertex(dim, transform, factor, x);
      for (int d = 0; d < dim; ++d) PV[v * dim + d] = x[d];
    }
    for (int v = 0; v < nv; ++v) {
      double xi[3];
      for (int d = 0; d < dim; ++d)
        xi[d] = (double)(I[d] - (B[d] << l) + ((v >> d) & 1)) / (double)(1L << l);
      multilinear_point(dim, PV, xi, m->vertices + ((long)c * nv + v) * dim);
    }
    m->rho[c] = m->speed[c] = 1.0;
    const double md = min_vertex_distance(dim, m->vertices + (long)c * nv * dim);
    if (md < mind) mind = md;
    for (int f = 0; f < nface; ++f) {
      const int d = f >> 1, sd = f & 1, cf = c * nface + f;
      m->nc_rec[cf] = m->nc_pos[cf] = -1;
      long g[3] = {0, 0, 0};
      for (int e = 0; e < dim; ++e) g[e] = I[e] * s;
      g[d] = sd ? (I[d] + 1) * s : I[d] * s - 1;
      if (g[d] < 0 || g[d] >= F.G) {
        m->neighbor[cf] = -boundary_id;
        continue;
      }
      const int y = owner[fg_at(&F, g)], ly = C.level[y];
      if (ly == l) {
        m->neighbor[cf] = y;
      } else if (ly == l - 1) {
        m->neighbor[cf] = OX_NC_FACE; /* fine side: filled from the coarse side's record */
      } else if (ly == l + 1) {
        int *r = m->hanging + (long)m->n_hanging * rec;
        r[0] = c;
        r[1] = f;
        for (int b = 0; b < nsub; ++b) {
          long gb[3] = {g[0], g[1], g[2]};
          int j = 0;
          for (int e = 0; e < dim; ++e)
            if (e != d) gb[e] += ((b >> j++) & 1) * (s / 2);
          r[2 + b] = owner[fg_at(&F, gb)];
          if (C.level[r[2 + b]] != l + 1) bad = 1;
        }
        m->neighbor[cf] = OX_NC_FACE;
        m->nc_rec[cf] = m->n_hanging++;
      } else {
        bad = 1; /* more than one level across a face */
      }
    }
  }
  for (int h = 0; h < m->n_hanging && !bad; ++h) {
    const int *r = m->hanging + (long)h * rec;
    for (int b = 0; b < nsub; ++b) {
      const int cf = r[2 + b] * nface + (r[1] ^ 1);
      m->nc_rec[cf] = h;
      m->nc_pos[cf] = b;
    }
  }
  m->min_vertex_distance_over_c = mind;
  free(owner);
  fcells_free(&C);
  if (bad) {
    ox_mesh_free(m);
    return NULL;
  }
  return m;
}

/* a fine cell and a vertex-adjacent one (faces, edges, corners) */
static int fg_neighbor(const fgrid *F, const long *g, int o, long *h) {
  for (int d = 0; d < F->dim; ++d) {
    const int od = (o % 3) - 1;
    o /= 3;
    h[d] = g[d] + od;
    if (h[d] < 0 || h[d] >= F->G) return 0;
  }
  if (F->dim == 2) h[2] = 0;
  return 1;
}

static int femit(const fgrid *F, const int *lev, int l, const long *I, int *out, int pos,
                 int cap) {
  const long s = 1L << (F->L - l);
  long lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  for (int d = 0; d < F->dim; ++d) {
    lo[d] = I[d] * s;
    hi[d] = lo[d] + s;
  }
  int finer = 0;
  long g[3];
  for (g[2] = lo[2]; g[2] < hi[2]; ++g[2])
    for (g[1] = lo[1]; g[1] < hi[1]; ++g[1])
      for (g[0] = lo[0]; g[0] < hi[0]; ++g[0]) {
        const int v = lev[fg_at(F, g)];
        if (v < l) return -1;
        if (v > l) finer = 1;
      }
  if (pos < 0 || (out && pos >= cap)) return -1;
  if (out) out[pos] = finer;
  ++pos;
  if (!finer) return pos;
  for (int ch = 0; ch < (1 << F->dim); ++ch) {
    long J[3] = {0, 0, 0};
    for (int d = 0; d < F->dim; ++d) J[d] = 2 * I[d] + ((ch >> d) & 1);
    pos = femit(F, lev, l + 1, J, out, pos, cap);
    if (pos < 0) return -1;
  }
  return pos;
}

int ox_adapt_forest(int dim, int n_intervals, const int *tree, const int *flags, int *tree_new,
                    int cap) {
  fcells C;
  if (forest_cells(dim, n_intervals, tree, -1, &C)) return -1;
  fgrid F = {dim, C.max_level + 1, (long)n_intervals << (C.max_level + 1), 0};
  F.ng = 1;
  for (int d = 0; d < dim; ++d) F.ng *= F.G;
  const int nn = ipow(3, dim), nch = 1 << dim;
  int *lev = (int *)malloc(sizeof(int) * F.ng), *owner = (int *)malloc(sizeof(int) * F.ng);
  int *tmp = (int *)malloc(sizeof(int) * F.ng);
  for (int c = 0; c < C.n; ++c) {
    fg_fill(&F, owner, C.level[c], C.coord + 3 * c, c);
    fg_fill(&F, lev, C.level[c], C.coord + 3 * c, C.level[c]);
  }
  /* the cells flagged for refinement */
  for (int c = 0; c < C.n; ++c)
    if (flags[c] == 1) fg_fill(&F, lev, C.level[c], C.coord + 3 * c, C.level[c] + 1);
  /* 2:1 balance across faces, edges and corners by refinement of the coarser side */
  for (int changed = 1; changed;) {
    changed = 0;
    for (long i = 0; i < F.ng; ++i) {
      long g[3], h[3];
      fg_coord(&F, i, g);
      for (int o = 0; o < nn; ++o)
        if (fg_neighbor(&F, g, o, h) && lev[fg_at(&F, h)] >= lev[i] + 2) {
          const int l = lev[i];
          long J[3] = {0, 0, 0};
          for (int d = 0; d < dim; ++d) J[d] = g[d] >> (F.L - l);
          fg_fill(&F, lev, l, J, l + 1);
          changed = 1;
          break;
        }
    }
  }
  /* coarsening candidates: families of active cells, all flagged -1 and none
   * refined above; the parent (level l-1) is given by its first child */
  int *cand = (int *)malloc(sizeof(int) * (C.n + 1)), nc = 0;
  for (int c = 0; c < C.n; ++c) {
    const int l = C.level[c];
    const long *I = C.coord + 3 * c;
    int first = l > 0 && flags[c] == -1;
    for (int d = 0; d < dim && first; ++d) first = (I[d] & 1) == 0;
    if (!first) continue;
    int ok = 1;
    for (int ch = 0; ch < nch && ok; ++ch) {
      long J[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) J[d] = (I[d] + ((ch >> d) & 1)) << (F.L - l);
      const int y = owner[fg_at(&F, J)];
      ok = C.level[y] == l && flags[y] == -1;
      if (ok) { /* still one active cell of level l after the refinement */
        long K[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) K[d] = I[d] + ((ch >> d) & 1);
        const long s = 1L << (F.L - l);
        long g[3];
        long lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) {
          lo[d] = K[d] * s;
          hi[d] = lo[d] + s;
        }
        for (g[2] = lo[2]; g[2] < hi[2] && ok; ++g[2])
          for (g[1] = lo[1]; g[1] < hi[1] && ok; ++g[1])
            for (g[0] = lo[0]; g[0] < hi[0] && ok; ++g[0]) ok = lev[fg_at(&F, g)] == l;
      }
    }
    if (ok) cand[nc++] = c;
  }
  /* drop the coarsenings that would break the balance, until none does */
  for (int dropped = 1; dropped;) {
    dropped = 0;
    memcpy(tmp, lev, sizeof(int) * F.ng);
    for (int i = 0; i < nc; ++i) {
      const int c = cand[i], l = C.level[c];
      long P[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) P[d] = C.coord[3 * c + d] >> 1;
      fg_fill(&F, tmp, l - 1, P, l - 1);
    }
    for (int i = 0; i < nc; ++i) {
      const int c = cand[i], l = C.level[c];
      const long s = 2L << (F.L - l);
      long lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      for (int d = 0; d < dim; ++d) {
        lo[d] = (C.coord[3 * c + d] >> 1) * s;
        hi[d] = lo[d] + s;
      }
      int viol = 0;
      long g[3], h[3];
      for (g[2] = lo[2]; g[2] < hi[2] && !viol; ++g[2])
        for (g[1] = lo[1]; g[1] < hi[1] && !viol; ++g[1])
          for (g[0] = lo[0]; g[0] < hi[0] && !viol; ++g[0])
            for (int o = 0; o < nn && !viol; ++o)
              if (fg_neighbor(&F, g, o, h) && tmp[fg_at(&F, h)] >= l + 1) viol = 1;
      if (viol) {
        cand[i--] = cand[--nc];
        dropped = 1;
      }
    }
  }
  for (int i = 0; i < nc; ++i) {
    const int c = cand[i], l = C.level[c];
    long P[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) P[d] = C.coord[3 * c + d] >> 1;
    fg_fill(&F, lev, l - 1, P, l - 1);
  }
  int pos = 0;
  const int nb = ipow(n_intervals, dim);
  for (int B = 0; B < nb && pos >= 0; ++B) {
    long I[3] = {0, 0, 0};
    int t = B;
    for (int d = 0; d < dim; ++d) {
      I[d] = t % n_intervals;
      t /= n_intervals;
    }
    pos = femit(&F, lev, 0, I, tree_new, pos, cap);
  }
  free(cand);
  free(lev);
  free(owner);
  free(tmp);
  fcells_free(&C);
  return pos;
}

/* SolutionTransfer between two forests that differ by at most one level per
 * cell: kept cells copied, refined ones interpolated, coarsened families
 * projected (the per-cell rules of ox_adapt_transfer) */
typedef struct ftr {
  int dim, nd, nf, nch;
  const ox_basis *b;
  const int *to, *tn;
  int po, pn;
  long co, cn;
  const double *old_u;
  double *new_u;
} ftr;

static int ftr_walk(ftr *T) {
  const int a = T->to[T->po++], b = T->tn[T->pn++];
  const long sz = (long)T->nf * T->nd;
  if (!a && !b) {
    memcpy(T->new_u + T->cn * sz, T->old_u + T->co * sz, sizeof(double) * sz);
    ++T->co;
    ++T->cn;
  } else if (!a) {
    for (int ch = 0; ch < T->nch; ++ch) {
      if (T->tn[T->pn++] != 0) return -1;
      transfer_child(T->b, T->dim, ch, T->old_u + T->co * sz, T->new_u + (T->cn + ch) * sz);
    }
    ++T->co;
    T->cn += T->nch;
  } else if (!b) {
    for (int ch = 0; ch < T->nch; ++ch)
      if (T->to[T->po++] != 0) return -1;
    transfer_parent(T->b, T->dim, T->old_u + T->co * sz, T->new_u + T->cn * sz);
    T->co += T->nch;
    ++T->cn;
  } else {
    for (int ch = 0; ch < T->nch; ++ch)
      if (ftr_walk(T)) return -1;
  }
  return 0;
}

int ox_adapt_forest_transfer(int dim, int degree, int n_intervals, const int *tree_old,
                             const int *tree_new, const double *old_u, double *new_u) {
  ox_basis b;
  ox_basis_init(&b, degree);
  ftr T = {dim, ipow(b.n, dim), dim + 1, 1 << dim, &b, tree_old, tree_new, 0, 0, 0, 0, old_u,
           new_u};
  const int nb = ipow(n_intervals, dim);
  for (int B = 0; B < nb; ++B)
    if (ftr_walk(&T)) return -1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* post-processed pressure (compute_post_pressure,                            */
/* source/wave_equation_operations.cc:786-848)                                */
/* ------------------------------------------------------------------------ */

/* n-point Gauss-Lobatto points on [0,1] (ends and the roots of P'_{n-1}) */
void ox_gauss_lobatto(int n, double *x) {
  const int m = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  for (int i = 1; i < n - 1; ++i) {
    double z = -cos(M_PI * i / m);
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= m; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      const double pm = m == 1 ? z : p1, pm1 = m == 1 ? 1.0 : p0;
      const double d1 = m * (z * pm - pm1) / (z * z - 1.0);          /* P'_m */
      const double d2 = (2.0 * z * d1 - m * (m + 1.0) * pm) / (1.0 - z * z); /* P''_m */
      const double dz = d1 / d2;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
}

/* tensor interpolation of one field: A (n_out x n_in) along every direction */
static void tensor_interp(int dim, int n_in, int n_out, const double *A, const double *in,
                          double *out, double *w1, double *w2) {
  int sz[3] = {n_in, n_in, n_in};
  const double *src = in;
  double *bufs[2] = {w1, w2};
  for (int d = 0; d < dim; ++d) {
    double *o = (d == dim - 1) ? out : bufs[d & 1];
    tensor_apply(dim, sz, d, A, n_out, src, o, 0, 0);
    sz[d] = n_out;
    src = o;
  }
}

/* reference gradient (dim components) at the nq^dim points of a field on the
 * np^dim nodes: along e the derivative matrix Pd, along the others Pv */
static void tensor_grad(int dim, int np, int nq, const double *Pv, const double *Pd,
                        const double *in, double *g, double *w1, double *w2) {
  for (int e = 0; e < dim; ++e) {
    int sz[3] = {np, np, np};
    const double *src = in;
    double *bufs[2] = {w1, w2};
    for (int d = 0; d < dim; ++d) {
      double *o = (d == dim - 1) ? g + (long)e * ipow(nq, dim) : bufs[d & 1];
      tensor_apply(dim, sz, d, d == e ? Pd : Pv, nq, src, o, 0, 0);
      sz[d] = nq;
      src = o;
    }
  }
}

/* the transpose of tensor_grad: sum_e grad_e^T f_e, added into out (np^dim) */
static void tensor_grad_t(int dim, int np, int nq, const double *Pv, const double *Pd,
                          const double *f, double *out, double *w1, double *w2) {
  const int npd = ipow(np, dim);
  for (int i = 0; i < npd; ++i) out[i] = 0.0;
  for (int e = 0; e < dim; ++e) {
    int sz[3] = {nq, nq, nq};
    const double *src = f + (long)e * ipow(nq, dim);
    double *bufs[2] = {w1, w2};
    for (int d = dim - 1; d >= 0; --d) {
      double *o = bufs[d & 1];
      tensor_apply(dim, sz, d, d == e ? Pd : Pv, np, src, o, 1, 0);
      sz[d] = np;
      src = o;
    }
    for (int i = 0; i < npd; ++i) out[i] += src[i];
  }
}

/* per-cell CG tolerances of compute_post_pressure (:798-799): absolute,
 * relative, and the iteration cap as a multiple of the unknowns */
static double ox_post_tol[2] = {1e-20, 1e-11};
static int ox_post_max_iter_factor = 6;
static int *ox_post_iterations = NULL;
static int ox_post_mode = 0;

void ox_post_set_solver(double abs_tol, double rel_tol, int max_iter_factor, int *iterations,
                        int mode) {
  ox_post_mode = mode;
  ox_post_tol[0] = abs_tol;
  ox_post_tol[1] = rel_tol;
  ox_post_max_iter_factor = max_iter_factor;
  ox_post_iterations = iterations;
}

void ox_post_pressure(const ox_mesh *m, int degree, const double *solution, double *post) {
  const int dim = m->dim, n = degree + 1, np = degree + 2, nq = degree + 2, nf = dim + 1;
  const int nd = ipow(n, dim), npd = ipow(np, dim), nqd = ipow(nq, dim), nv = 1 << dim;
  const long N = ox_n_dofs(m, degree);
  double *tmp = (double *)malloc(sizeof(double) * N);
  ox_apply(m, degree, solution, tmp, 1.0); /* :795 */
  ox_basis b;
  ox_basis_init(&b, degree);
  double xq[OX_MAX_N], wq[OX_MAX_N], xp[OX_MAX_N];
  ox_gauss(nq, xq, wq);
  ox_gauss_lobatto(np, xp);
  double Iq[OX_MAX_N * OX_MAX_N], Pv[OX_MAX_N * OX_MAX_N], Pd[OX_MAX_N * OX_MAX_N];
  for (int q = 0; q < nq; ++q) {
    ox_lagrange(n, b.x, xq[q], Iq + q * n, NULL);
    ox_lagrange(np, xp, xq[q], Pv + q * np, Pd + q * np);
  }
#pragma omp parallel
  {
    const int W = nqd > npd ? nqd : npd;
    double *w1 = (double *)malloc(sizeof(double) * W), *w2 = (double *)malloc(sizeof(double) * W);
    double *vt = (double *)malloc(sizeof(double) * dim * nqd);
    double *ph = (double *)malloc(sizeof(double) * nqd);
    double *G = (double *)malloc(sizeof(double) * dim * dim * nqd);
    double *jw = (double *)malloc(sizeof(double) * nqd);
    double *fr = (double *)malloc(sizeof(double) * dim * nqd);
    double *gr = (double *)malloc(sizeof(double) * dim * nqd);
    double *rhs = (double *)malloc(sizeof(double) * npd);
    double *u = (double *)malloc(sizeof(double) * npd);
    double *r = (double *)malloc(sizeof(double) * npd);
    double *pp = (double *)malloc(sizeof(double) * npd);
    double *Ap = (double *)malloc(sizeof(double) * npd);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) {
      const double *V = m->vertices + (long)c * nv * dim;
      const double rho = m->rho[c];
      for (int d = 0; d < dim; ++d)
        tensor_interp(dim, n, nq, Iq, tmp + ((long)c * nf + d) * nd, vt + (long)d * nqd, w1, w2);
      tensor_interp(dim, n, nq, Iq, solution + ((long)c * nf + dim) * nd, ph, w1, w2);
      /* geometry at QGauss(k+2); rhs integrand (grad phi, -rho v_t) in reference form */
      for (int q = 0; q < nqd; ++q) {
        double xi[3], J[9], Ji[9], w = 1.0;
        int t = q;
        for (int d = 0; d < dim; ++d) {
          xi[d] = xq[t % nq];
          w *= wq[t % nq];
          t /= nq;
        }
        const double det = multilinear_jacobian(dim, V, xi, J, Ji);
        jw[q] = det * w;
        for (int e = 0; e < dim; ++e) {
          double s = 0.0;
          for (int i = 0; i < dim; ++i) s += Ji[e * dim + i] * (-rho * vt[(long)i * nqd + q]);
          fr[(long)e * nqd + q] = s * det * w;
          for (int e2 = 0; e2 < dim; ++e2) {
            double g = 0.0;
            for (int i = 0; i < dim; ++i) g += Ji[e * dim + i] * Ji[e2 * dim + i];
            G[((long)e * dim + e2) * nqd + q] = g * det * w;
          }
        }
      }
      tensor_grad_t(dim, np, nq, Pv, Pd, fr, rhs, w1, w2);
      const double rhs0 = rhs[0];
      rhs[0] = 0.0; /* :808-811 */
      /* CG on the Laplacian with the first unknown fixed (the reference's
       * MatrixPostProcessing::vmult, :750-769) */
#define OX_POST_APPLY(in, out)                                                           \
  do {                                                                                   \
    const double keep = (in)[0];                                                         \
    (in)[0] = 0.0;                                                                       \
    tensor_grad(dim, np, nq, Pv, Pd, in, gr, w1, w2);                                    \
    for (int q = 0; q < nqd; ++q) {                                                      \
      double g[3];                                                                       \
      for (int e = 0; e < dim; ++e) {                                                    \
        g[e] = 0.0;                                                                      \
        for (int e2 = 0; e2 < dim; ++e2) g[e] += G[((long)e * dim + e2) * nqd + q] * gr[(long)e2 * nqd + q]; \
      }                                                                                  \
      for (int e = 0; e < dim; ++e) fr[(long)e * nqd + q] = g[e];                        \
    }                                                                                    \
    tensor_grad_t(dim, np, nq, Pv, Pd, fr, out, w1, w2);                                 \
    (out)[0] = keep;                                                                     \
    (in)[0] = keep;                                                                      \
  } while (0)
      /* SolverCGvect::solve (include/elementwise_cg.templates.h:66-145) for one
       * cell: zero start, identity preconditioner, stop on |r| < abs_tol,
       * |r|/|r0| < rel_tol or after max_iter iterations.  Mode 1 instead runs
       * CG on the Neumann problem itself, restricted to vectors with zero sum
       * (the matrix's null space is the constant vector): the same p* after
       * the average adjustment below, with a far smaller condition number. */
      const int mean_free = ox_post_mode == 1;
      if (mean_free) {
        rhs[0] = rhs0;
        double mean = 0.0;
        for (int i = 0; i < npd; ++i) mean += rhs[i];
        mean /= npd;
        for (int i = 0; i < npd; ++i) rhs[i] -= mean;
      }
#define OX_POST_NEUMANN(in, out)                                                         \
  do {                                                                                   \
    tensor_grad(dim, np, nq, Pv, Pd, in, gr, w1, w2);                                    \
    for (int q = 0; q < nqd; ++q) {                                                      \
      double g[3];                                                                       \
      for (int e = 0; e < dim; ++e) {                                                    \
        g[e] = 0.0;                                                                      \
        for (int e2 = 0; e2 < dim; ++e2) g[e] += G[((long)e * dim + e2) * nqd + q] * gr[(long)e2 * nqd + q]; \
      }                                                                                  \
      for (int e = 0; e < dim; ++e) fr[(long)e * nqd + q] = g[e];                        \
    }                                                                                    \
    tensor_grad_t(dim, np, nq, Pv, Pd, fr, out, w1, w2);                                 \
  } while (0)
      double rr = 0.0;
      for (int i = 0; i < npd; ++i) {
        u[i] = 0.0;
        r[i] = rhs[i];
        pp[i] = r[i];
        rr += r[i] * r[i];
      }
      const double nr0 = sqrt(rr);
      int it = 0;
      while (1) {
        if (mean_free)
          OX_POST_NEUMANN(pp, Ap);
        else
          OX_POST_APPLY(pp, Ap);
        double pAp = 0.0;
        for (int i = 0; i < npd; ++i) pAp += pp[i] * Ap[i];
        /* adjust_division_by_zero (include/elementwise_cg.templates.h:44-61) */
        if (pAp < fmax(1e-5 * ox_post_tol[0] * ox_post_tol[0], 1e4 * DBL_MIN)) pAp = 1.0;
        const double alpha = rr / pAp;
        double rn = 0.0, rs = 0.0;
        for (int i = 0; i < npd; ++i) {
          u[i] += alpha * pp[i];
          r[i] -= alpha * Ap[i];
          rs += r[i];
        }
        if (mean_free) /* keep the residual in the range of the matrix */
          for (int i = 0; i < npd; ++i) r[i] -= rs / npd;
        for (int i = 0; i < npd; ++i) rn += r[i] * r[i];
        ++it;
        if (sqrt(rn) < ox_post_tol[0] || sqrt(rn) < ox_post_tol[1] * nr0 ||
            it > ox_post_max_iter_factor * npd)
          break;
        const double beta = rn / rr;
        rr = rn;
        for (int i = 0; i < npd; ++i) pp[i] = r[i] + beta * pp[i];
      }
      if (ox_post_iterations) ox_post_iterations[c] = it;
#undef OX_POST_APPLY
#undef OX_POST_NEUMANN
      /* constant offset from the cell average of the pressure (:813-837) */
      double vol = 0.0, correct = 0.0, prelim = 0.0;
      tensor_interp(dim, np, nq, Pv, u, gr, w1, w2);
      for (int q = 0; q < nqd; ++q) {
        vol += jw[q];
        correct += jw[q] * ph[q];
        prelim += jw[q] * gr[q];
      }
      const double adjust = (correct - prelim) / vol;
      double *o = post + (long)c * npd;
      for (int i = 0; i < npd; ++i) o[i] = u[i] + adjust;
    }
    free(w1);
    free(w2);
    free(vt);
    free(ph);
    free(G);
    free(jw);
    free(fr);
    free(gr);
    free(rhs);
    free(u);
    free(r);
    free(pp);
    free(Ap);
  }
  free(tmp);
}

/* L2 error of the post-processed pressure with QGauss(k+3)
 * (source/explicit_wave.cc:477-484) */
double ox_post_error(const ox_mesh *m, int degree, const ox_exact *e, double t,
                     const double *post) {
  const int dim = m->dim, np = degree + 2, nq = degree + 3, npd = ipow(np, dim),
            nqd = ipow(nq, dim), nv = 1 << dim;
  double xq[OX_MAX_N + 2], wq[OX_MAX_N + 2], xp[OX_MAX_N], Pv[(OX_MAX_N + 2) * OX_MAX_N];
  ox_gauss(nq, xq, wq);
  ox_gauss_lobatto(np, xp);
  for (int q = 0; q < nq; ++q) ox_lagrange(np, xp, xq[q], Pv + q * np, NULL);
  double err = 0.0;
#pragma omp parallel reduction(+ : err)
  {
    double *w1 = (double *)malloc(sizeof(double) * nqd), *w2 = (double *)malloc(sizeof(double) * nqd);
    double *pq = (double *)malloc(sizeof(double) * nqd);
#pragma omp for schedule(static)
    for (int c = 0; c < m->n_cells; ++c) {
      const double *V = m->vertices + (long)c * nv * dim;
      tensor_interp(dim, np, nq, Pv, post + (long)c * npd, pq, w1, w2);
      for (int q = 0; q < nqd; ++q) {
        double xi[3], x[3], J[9], Ji[9], w = 1.0;
        int tq = q;
        for (int d = 0; d < dim; ++d) {
          xi[d] = xq[tq % nq];
          w *= wq[tq % nq];
          tq /= nq;
        }
        multilinear_point(dim, V, xi, x);
        const double jxw = fabs(multilinear_jacobian(dim, V, xi, J, Ji)) * w;
        const double dp = pq[q] - ox_exact_value(e, x, t, dim);
        err += dp * dp * jxw;
      }
    }
    free(w1);
    free(w2);
    free(pq);
  }
  return sqrt(err);
}

/* ------------------------------------------------------------------------ */
/* Point source (no counterpart in the reference; see exwave_oracle.h)        */
/* ------------------------------------------------------------------------ */

double ox_ricker(const ox_source *s, double t) {
  const double a = (M_PI * s->f0) * (M_PI * s->f0), tau = t - s->t0;
  return s->amplitude * (1.0 - 2.0 * a * tau * tau) * exp(-a * tau * tau);
}

/* d/dt [tau exp(-a tau^2)] = (1 - 2 a tau^2) exp(-a tau^2) */
double ox_ricker_integral(const ox_source *s, double t1, double t2) {
  const double a = (M_PI * s->f0) * (M_PI * s->f0);
  const double u1 = t1 - s->t0, u2 = t2 - s->t0;
  return s->amplitude * (u2 * exp(-a * u2 * u2) - u1 * exp(-a * u1 * u1));
}

/* reference coordinates of x in cell c by Newton's method on the multilinear
 * map; 1 if they lie in the closed unit cell (tolerance 1e-10) */
static int locate_in_cell(const ox_mesh *m, int c, const double *x, double *xi) {
  const int dim = m->dim, nv = 1 << dim;
  const double *V = m->vertices + (long)c * nv * dim;
  for (int d = 0; d < dim; ++d) xi[d] = 0.5;
  double h = 0.0;
  for (int d = 0; d < dim; ++d) h = fmax(h, fabs(V[(nv - 1) * dim + d] - V[d]));
  for (int it = 0; it < 60; ++it) {
    double y[3], J[9], Ji[9], r[3], rn = 0.0;
    multilinear_point(dim, V, xi, y);
    multilinear_jacobian(dim, V, xi, J, Ji);
    for (int i = 0; i < dim; ++i) {
      r[i] = y[i] - x[i];
      rn = fmax(rn, fabs(r[i]));
    }
    for (int e = 0; e < dim; ++e) {
      double dx = 0.0;
      for (int i = 0; i < dim; ++i) dx += Ji[e * dim + i] * r[i];
      xi[e] -= dx;
      if (!(fabs(xi[e]) < 10.0)) return 0; /* far outside: diverging */
    }
    if (rn <= 1e-14 * h) break;
  }
  double y[3], rn = 0.0;
  multilinear_point(dim, V, xi, y);
  for (int i = 0; i < dim; ++i) rn = fmax(rn, fabs(y[i] - x[i]));
  if (rn > 1e-12 * h) return 0;
  for (int d = 0; d < dim; ++d)
    if (xi[d] < -1e-10 || xi[d] > 1.0 + 1e-10) return 0;
  return 1;
}

int ox_source_cells(const ox_mesh *m, int degree, const ox_source *s, int *cells, double *b) {
  const int dim = m->dim, nv = 1 << dim;
  ox_basis B;
  ox_basis_init(&B, degree);
  const int n = B.n, nd = ipow(n, dim);
  double xis[8][3];
  int ns = 0;
  for (int c = 0; c < m->n_cells && ns < 8; ++c)
    if (locate_in_cell(m, c, s->x, xis[ns])) cells[ns++] = c;
  for (int k = 0; k < ns; ++k) {
    const double *V = m->vertices + (long)cells[k] * nv * dim;
    double l[3][OX_MAX_N];
    for (int d = 0; d < dim; ++d) ox_lagrange(n, B.x, xis[k][d], l[d], NULL);
    for (int q = 0; q < nd; ++q) {
      double xi[3], J[9], Ji[9], w = 1.0, phi = 1.0;
      int t = q;
      for (int d = 0; d < dim; ++d) {
        xi[d] = B.x[t % n];
        w *= B.w[t % n];
        phi *= l[d][t % n];
        t /= n;
      }
      const double det = multilinear_jacobian(dim, V, xi, J, Ji);
      b[(long)k * nd + q] = phi / ns / (w * det);
    }
  }
  return ns;
}

void ox_source_add(const ox_mesh *m, int degree, const ox_source *s, double t1, double t2,
                   double *u) {
  const int nd = ipow(degree + 1, m->dim);
  int cells[8];
  double *b = (double *)malloc(sizeof(double) * 8 * nd);
  const int ns = ox_source_cells(m, degree, s, cells, b);
  const double I = ox_ricker_integral(s, t1, t2);
  for (int k = 0; k < ns; ++k) {
    double *p = u + ((long)cells[k] * (m->dim + 1) + m->dim) * nd;
    for (int q = 0; q < nd; ++q) p[q] += I * b[(long)k * nd + q];
  }
  free(b);
}

/* ------------------------------------------------------------------------ */
/* ADER with local time stepping (WaveEquationOperationADERLTS,               */
/* source/wave_equation_operations.cc:1385-1692, and ClusterManager,          */
/* include/cluster_manager.templates.h:24-1049), restated per cell.           */
/*                                                                            */
/* The reference keeps its flags per cell batch of VectorizedArray width and  */
/* its face masks per lane; batches are homogeneous in (cluster, has faster   */
/* neighbour, has slower neighbour) because the categories of                 */
/* iterate_cluster_categorization are handed to MatrixFree.  Every quantity   */
/* that reaches a used result is a per-cell one, so the restatement works on  */
/* cells (batch width 1).                                                     */
/* ------------------------------------------------------------------------ */

struct ox_lts {
  const ox_mesh *m;
  int degree, recon;
  ox_basis b;
  long N;
  int nd, nf;
  int n_clusters, cluster_diff, max_level;
  double fastest_dt, global_dt, rel_tol;
  int *cluster;          /* per cell */
  char *faster, *slower; /* per cell: has a faster / slower face neighbour */
  int *mult;             /* per cluster: time step multiple */
  double *cl_time;       /* per cluster: time level */
  int n_updates;
  int *order;            /* cluster_update_order */
  double *times;         /* cluster_update_times [n_updates][2] */
  int iterations;
  /* vectors */
  double *state, *tempsrc, *dst, *fm, *igd, *rstate;
  /* per-cell flags of the current pass */
  char *U, *E, *nbrU;
  double *cell_time;
  /* times of the current cycle */
  double t1fa, t2fa, t1sl, t2sl, t1sa, t2sa, dt;
  int fm_considered;
  /* point source (ox_lts_set_source): cells and their M^-1 delta vectors */
  int n_src;
  int src_cells[8];
  double *src_b;
  ox_source source;
};

static double cell_min_vertex_distance(const ox_mesh *m, int c) {
  const int dim = m->dim, nv = 1 << dim;
  const double *V = m->vertices + (long)c * nv * dim;
  double mind = 1e300;
  for (int a = 0; a < nv; ++a)
    for (int b = a + 1; b < nv; ++b) {
      double s = 0.0;
      for (int d = 0; d < dim; ++d) s += (V[a * dim + d] - V[b * dim + d]) * (V[a * dim + d] - V[b * dim + d]);
      if (s < mind) mind = s;
    }
  return sqrt(mind);
}

/* The cells across face f of cell c, as the MatrixFree faces the reference
 * loops over see them: none at a boundary, one across a conforming face or
 * from the fine side of a 2:1 face (the coarse cell), the 2^(dim-1) fine cells
 * by subface position from the coarse side (cell->neighbor_child_on_subface,
 * include/cluster_manager.templates.h:118-125).  Returns the count. */
static int face_cells(const ox_mesh *m, int c, int f, int *nbs) {
  const int nface = 2 * m->dim, nsub = 1 << (m->dim - 1);
  const int nb = m->neighbor[c * nface + f];
  if (nb >= 0) {
    nbs[0] = nb;
    return 1;
  }
  if (nb != OX_NC_FACE) return 0;
  const int *R = m->hanging + (long)m->nc_rec[c * nface + f] * (2 + nsub);
  if (m->nc_pos[c * nface + f] >= 0) {
    nbs[0] = R[0];
    return 1;
  }
  for (int b = 0; b < nsub; ++b) nbs[b] = R[2 + b];
  return nsub;
}

/* iterate_cluster_categorization (include/cluster_manager.templates.h:24-262)
 * on one process: cells visited in active-cell order, ids changed in place,
 * neighbours across 2:1 faces through their subfaces (:118-125).  The
 * reference's per-cell step is cfl * minimum_vertex_distance (:57-66, no wave
 * speed); with speed_aware the step is divided by the cell's wave speed, the
 * CFL rule of the global step (source/explicit_wave.cc:529-545). */
static void lts_categorize(ox_lts *L, double cfl, int max_clusters, int max_diff,
                           int speed_aware) {
  const ox_mesh *m = L->m;
  const int n = m->n_cells, nface = 2 * m->dim;
  int nbs[4];
  double *ts = (double *)malloc(sizeof(double) * n);
  double minimal = 1e300;
  for (int c = 0; c < n; ++c) {
    ts[c] = cfl * cell_min_vertex_distance(m, c);
    if (speed_aware) ts[c] /= m->speed[c];
    if (ts[c] < minimal) minimal = ts[c];
  }
  const double eps = minimal / 10.0;
  int *id = L->cluster;
  int nc = 0;
  for (int c = 0; c < n; ++c) {
    id[c] = (int)((ts[c] - eps) / minimal);
    if (id[c] > nc) nc = id[c];
  }
  nc++;
  int diff = 1;
  if (nc > max_clusters) {
    diff = nc / max_clusters + 1;
    if (diff > max_diff) diff = max_diff;
    for (int c = 0; c < n; ++c) id[c] /= diff;
  }
  int fs = 1, count = 0;
  while (fs && count < 100) {
    count++;
    for (int cl = 0; cl < nc; ++cl)
      for (int c = 0; c < n; ++c) {
        if (id[c] != cl) continue;
        int anyf = 0;
        for (int f = 0; f < nface; ++f) {
          const int k = face_cells(m, c, f, nbs);
          for (int j = 0; j < k; ++j)
            if (id[nbs[j]] < cl) anyf = 1;
        }
        for (int f = 0; f < nface; ++f) {
          const int k = face_cells(m, c, f, nbs);
          for (int j = 0; j < k; ++j)
            if (id[nbs[j]] > cl) id[nbs[j]] = anyf ? cl : cl + 1;
        }
      }
    fs = 0;
    for (int c = 0; c < n; ++c) {
      const int cur = id[c];
      int fa = 0, sl = 0;
      for (int f = 0; f < nface; ++f) {
        const int k = face_cells(m, c, f, nbs);
        for (int j = 0; j < k; ++j) {
          const int nb = nbs[j];
          if (id[nb] < cur)
            fa = 1;
          else if (id[nb] > cur)
            sl = 1;
          if (id[nb] < cur - 1 || id[nb] > cur + 1) fs = 1;
        }
      }
      L->faster[c] = (char)fa;
      L->slower[c] = (char)sl;
      if (fa && sl) fs = 1;
    }
  }
  L->iterations = count;
  nc = 0;
  for (int c = 0; c < n; ++c)
    if (id[c] > nc) nc = id[c];
  L->n_clusters = nc + 1;
  L->cluster_diff = diff;
  free(ts);
}

/* ClusterManager::setup (include/cluster_manager.templates.h:330-484) */
static void lts_schedule(ox_lts *L, double time0) {
  const int nc = L->n_clusters, diff = L->cluster_diff;
  const double D = L->fastest_dt;
  L->mult = (int *)malloc(sizeof(int) * nc);
  L->cl_time = (double *)malloc(sizeof(double) * nc);
  for (int c = 0; c < nc; ++c) {
    L->mult[c] = diff * c + 1;
    L->cl_time[c] = time0;
  }
  const int max_level = diff * (nc - 1) + 1;
  L->max_level = max_level;
  int *tl = (int *)calloc(nc, sizeof(int));
  int cap = 64;
  L->order = (int *)malloc(sizeof(int) * cap);
  L->times = (double *)malloc(sizeof(double) * 2 * cap);
  int nu = 0;
#define MINL(a) ((a) < max_level ? (a) : max_level)
  for (;;) {
    int mn = tl[0];
    for (int c = 1; c < nc; ++c)
      if (tl[c] < mn) mn = tl[c];
    if (mn >= max_level) break;
    for (int c = 0; c < nc; ++c) {
      if (tl[c] >= max_level) continue;
      const int mine = MINL(tl[c] + L->mult[c]);
      int ok;
      if (nc == 1)
        ok = 1;
      else if (c == 0)
        ok = mine <= MINL(tl[c + 1] + L->mult[c + 1]);
      else if (c == nc - 1)
        ok = mine <= MINL(tl[c - 1] + L->mult[c - 1]);
      else
        ok = mine <= MINL(tl[c + 1] + L->mult[c + 1]) && mine <= MINL(tl[c - 1] + L->mult[c - 1]);
      if (!ok) continue;
      if (nu == cap) {
        cap *= 2;
        L->order = (int *)realloc(L->order, sizeof(int) * cap);
        L->times = (double *)realloc(L->times, sizeof(double) * 2 * cap);
      }
      L->times[2 * nu] = tl[c] * D;
      tl[c] += L->mult[c];
      L->order[nu] = c;
      L->times[2 * nu + 1] = tl[c] * D;
      nu++;
    }
  }
#undef MINL
  free(tl);
  L->n_updates = nu;
  L->global_dt = D * (diff * (nc - 1) + 1);
  for (int u = 0; u < nu; ++u) {
    if (L->times[2 * u + 1] > max_level * D) L->times[2 * u + 1] = max_level * D;
    L->times[2 * u] += time0;
    L->times[2 * u + 1] += time0;
  }
}

ox_lts *ox_lts_create(const ox_mesh *m, int degree, double cfl, double fastest_dt,
                      int max_clusters, int max_diff, double time0, int recon) {
  return ox_lts_create2(m, degree, cfl, fastest_dt, max_clusters, max_diff, time0, recon, 0);
}

ox_lts *ox_lts_create2(const ox_mesh *m, int degree, double cfl, double fastest_dt,
                       int max_clusters, int max_diff, double time0, int recon, int flags) {
  ox_lts *L = (ox_lts *)calloc(1, sizeof(ox_lts));
  L->m = m;
  L->degree = degree;
  L->recon = recon;
  ox_basis_init(&L->b, degree);
  L->N = ox_n_dofs(m, degree);
  L->nf = m->dim + 1;
  L->nd = ipow(degree + 1, m->dim);
  L->rel_tol = 1e-4; /* ClusterManager(rel_tol = 1e-4), include/cluster_manager.h:34 */
  const int n = m->n_cells;
  L->cluster = (int *)calloc(n, sizeof(int));
  L->faster = (char *)calloc(n, 1);
  L->slower = (char *)calloc(n, 1);
  L->U = (char *)calloc(n, 1);
  L->E = (char *)calloc(n, 1);
  L->nbrU = (char *)calloc(n, 1);
  L->cell_time = (double *)calloc(n, sizeof(double));
  lts_categorize(L, cfl, max_clusters, max_diff, (flags & OX_LTS_SPEED_AWARE) != 0);
  L->fastest_dt = fastest_dt;
  lts_schedule(L, time0);
  const long N = L->N;
  L->state = (double *)calloc(N, sizeof(double));
  L->tempsrc = (double *)calloc(N, sizeof(double));
  L->dst = (double *)calloc(N, sizeof(double));
  L->fm = (double *)calloc(N, sizeof(double));
  L->igd = (double *)calloc(N, sizeof(double));
  L->rstate = (double *)calloc(N, sizeof(double));
  return L;
}

void ox_lts_free(ox_lts *L) {
  if (!L) return;
  free(L->cluster);
  free(L->faster);
  free(L->slower);
  free(L->U);
  free(L->E);
  free(L->nbrU);
  free(L->cell_time);
  free(L->mult);
  free(L->cl_time);
  free(L->order);
  free(L->times);
  free(L->state);
  free(L->tempsrc);
  free(L->dst);
  free(L->fm);
  free(L->igd);
  free(L->rstate);
  free(L->src_b);
  free(L);
}

void ox_lts_set_source(ox_lts *L, const ox_source *src) {
  free(L->src_b);
  L->src_b = NULL;
  L->n_src = 0;
  if (!src) return;
  L->source = *src;
  L->src_b = (double *)malloc(sizeof(double) * 8 * L->nd);
  L->n_src = ox_source_cells(L->m, L->degree, src, L->src_cells, L->src_b);
}

int ox_lts_n_clusters(const ox_lts *L) { return L->n_clusters; }
int ox_lts_cluster_diff(const ox_lts *L) { return L->cluster_diff; }
int ox_lts_iterations(const ox_lts *L) { return L->iterations; }
double ox_lts_global_dt(const ox_lts *L) { return L->global_dt; }
int ox_lts_n_updates(const ox_lts *L) { return L->n_updates; }
void ox_lts_schedule(const ox_lts *L, int *order, double *times) {
  for (int u = 0; u < L->n_updates; ++u) {
    order[u] = L->order[u];
    times[2 * u] = L->times[2 * u];
    times[2 * u + 1] = L->times[2 * u + 1];
  }
}
void ox_lts_cells(const ox_lts *L, int *cluster, int *flags) {
  for (int c = 0; c < L->m->n_cells; ++c) {
    cluster[c] = L->cluster[c];
    flags[c] = L->slower[c] + 2 * L->faster[c]; /* element_categories % 3 (:252) */
  }
}

/* CK time integral over [t1,t2] about each flagged cell's own time level
 * (local_apply_firstader_domain :1385-1408), into tempsrc */
static void lts_ck_cells(ox_lts *L, const double *src, const char *flag, double t1, double t2) {
  const ox_mesh *m = L->m;
  const long cs = (long)L->nf * L->nd;
  for (int c = 0; c < m->n_cells; ++c) {
    if (!flag[c]) continue;
    ox_mesh one = *m;
    one.n_cells = 1;
    one.vertices = m->vertices + (long)c * (1 << m->dim) * m->dim;
    one.rho = m->rho + c;
    one.speed = m->speed + c;
    one.neighbor = m->neighbor + (long)c * 2 * m->dim;
    ox_ader_ck(&one, L->degree, src + c * cs, L->recon ? L->igd + c * cs : NULL, t2, t1,
               L->cell_time[c], L->tempsrc + c * cs);
  }
}

/* contribution of face f of cell c, across slot j of face_cells (the subface
 * of a 2:1 face seen from its coarse side), evaluated from tempsrc on both
 * sides with boundary_fac -1 as in local_apply_ader_face :1514-1537, added
 * into out */
static void lts_face_into(ox_lts *L, int c, int f, int j, double *out, work *W) {
  const ox_mesh *m = L->m;
  const long cs = (long)L->nf * L->nd;
  const int nb = m->neighbor[c * 2 * m->dim + f];
  memset(W->r, 0, sizeof(double) * cs);
  if (nb == OX_NC_FACE)
    cell_face_nc(m, &L->b, c, f, L->tempsrc, j, -1.0, W->r);
  else
    cell_face(m, &L->b, c, f, L->tempsrc + c * cs, nb >= 0 ? L->tempsrc + (long)nb * cs : NULL,
              -1.0, W->r);
  for (long i = 0; i < cs; ++i) out[c * cs + i] += W->r[i];
}

/* ClusterManager::update_elements (include/cluster_manager.templates.h:807-1049)
 * for cluster `ac` and the update set L->U */
static void lts_update_elements(ox_lts *L, double *lstate, int ac, int write_fm) {
  const ox_mesh *m = L->m;
  const int n = m->n_cells, nface = 2 * m->dim, nc = L->n_clusters;
  const long cs = (long)L->nf * L->nd;
  const double tol = L->rel_tol * L->fastest_dt;
  const int *cl = L->cluster;
  char *U = L->U, *E = L->E, *nbrU = L->nbrU;
  int nbs[4];
  work W = work_alloc(m->dim, L->b.n);
  for (int c = 0; c < n; ++c) {
    nbrU[c] = 0;
    if (U[c]) continue;
    for (int f = 0; f < nface; ++f) {
      const int k = face_cells(m, c, f, nbs);
      for (int j = 0; j < k; ++j)
        if (U[nbs[j]]) nbrU[c] = 1;
    }
  }
  /* contributions across the faces to the faster (ac-1) and slower (ac+1)
   * cluster: the actual cluster's side goes to dst, the other side to the
   * flux memory (:862-993) */
  for (int side = 0; side < 2; ++side) {
    const int other = side == 0 ? ac - 1 : ac + 1;
    if (other < 0 || other >= nc) continue;
    const double t1 = side == 0 ? L->t1fa : L->t1sl, t2 = side == 0 ? L->t2fa : L->t2sl;
    if (!(fabs(t2 - t1) > tol)) continue;
    for (int c = 0; c < n; ++c)
      E[c] = (cl[c] == ac && (side == 0 ? L->faster[c] : L->slower[c]) && U[c]) ||
             (cl[c] == other && (side == 0 ? L->slower[c] : L->faster[c]) && nbrU[c]);
    lts_ck_cells(L, lstate, E, t1, t2);
    for (int c = 0; c < n; ++c) {
      if (!E[c]) continue;
      for (int f = 0; f < nface; ++f) {
        const int k = face_cells(m, c, f, nbs);
        for (int j = 0; j < k; ++j) {
          const int nb = nbs[j];
          if (!E[nb]) continue;
          if (cl[c] == ac && cl[nb] != ac && nbrU[nb])
            lts_face_into(L, c, f, j, L->dst, &W);
          else if (cl[c] != ac && cl[nb] == ac && nbrU[c] && write_fm)
            lts_face_into(L, c, f, j, L->fm, &W);
        }
      }
    }
  }
  /* contribution within the actual cluster and its boundary (:996-1049) */
  for (int c = 0; c < n; ++c) E[c] = cl[c] == ac && (U[c] || nbrU[c]);
  lts_ck_cells(L, lstate, E, L->t1sa, L->t2sa);
  for (int c = 0; c < n; ++c) {
    if (!E[c]) continue;
    for (int f = 0; f < nface; ++f) {
      const int k = face_cells(m, c, f, nbs);
      if (k == 0) {
        lts_face_into(L, c, f, 0, L->dst, &W); /* boundary face */
        continue;
      }
      for (int j = 0; j < k; ++j) {
        const int nb = nbs[j];
        if ((U[c] && U[nb]) || (U[c] && E[nb] && cl[nb] == ac) || (E[c] && U[nb] && cl[c] == ac))
          lts_face_into(L, c, f, j, L->dst, &W);
      }
    }
  }
  /* second ADER sweep on the update cells (local_apply_secondader_domain
   * :1413-1497): -volume(CK over [0,dt]) + flux memory + faces, inverse mass */
  double *V = (double *)malloc(sizeof(double) * cs);
  for (int c = 0; c < n; ++c) {
    if (!U[c]) continue;
    ox_mesh one = *m;
    one.n_cells = 1;
    one.vertices = m->vertices + (long)c * (1 << m->dim) * m->dim;
    one.rho = m->rho + c;
    one.speed = m->speed + c;
    one.neighbor = m->neighbor + (long)c * nface;
    ox_ader_ck(&one, L->degree, lstate + c * cs, L->recon ? L->igd + c * cs : NULL, L->dt, 0.0,
               0.0, V);
    double *r = W.r;
    memset(r, 0, sizeof(double) * cs);
    cell_volume(m, &L->b, c, V, r, W.jxw, &W);
    for (long i = 0; i < cs; ++i) r[i] = -r[i];
    if (L->fm_considered)
      for (long i = 0; i < cs; ++i) {
        r[i] += L->fm[c * cs + i];
        L->fm[c * cs + i] = 0.0;
      }
    for (long i = 0; i < cs; ++i) r[i] += L->dst[c * cs + i];
    for (int k = 0; k < L->nf; ++k)
      for (int q = 0; q < L->nd; ++q) L->dst[c * cs + k * L->nd + q] = r[k * L->nd + q] / W.jxw[q];
  }
  free(V);
  /* local_state -= dst; dst = 0 (:1046-1047).  Cells outside U that received
   * face terms are at distance >= 2 from every cell the caller reads (see
   * ox_lts_step), and get the same subtraction as in the reference. */
  for (long i = 0; i < L->N; ++i) {
    lstate[i] -= L->dst[i];
    L->dst[i] = 0.0;
  }
  /* the point source over each updated source cell's own interval
   * [t_K, t_K + dt] (no counterpart in the reference) */
  for (int s = 0; s < L->n_src; ++s) {
    const int c = L->src_cells[s];
    if (!U[c]) continue;
    const double I = ox_ricker_integral(&L->source, L->cell_time[c], L->cell_time[c] + L->dt);
    double *pc = lstate + c * cs + (long)m->dim * L->nd;
    for (int q = 0; q < L->nd; ++q) pc[q] += I * L->src_b[(long)s * L->nd + q];
  }
  work_free(&W);
}

/* M^-1 A(src) on the flagged cells only (reconstruct_div_grad :1593-1680) */
static void lts_reconstruct(ox_lts *L, const double *src, const char *flag) {
  const ox_mesh *m = L->m;
  ox_basis *b = &L->b;
  work W = work_alloc(m->dim, b->n);
  for (int c = 0; c < m->n_cells; ++c)
    if (flag[c]) apply_cell(m, b, c, src, L->igd, 1.0, &W);
  work_free(&W);
}

/* ClusterManager::perform_time_step (include/cluster_manager.templates.h:
 * 488-803); time_now is the time control's time after advance_time_step. */
void ox_lts_step(ox_lts *L, double time_now, const double *src, double *dst) {
  const ox_mesh *m = L->m;
  const int n = m->n_cells, nc = L->n_clusters, nface = 2 * m->dim;
  const double D = L->fastest_dt, tol = L->rel_tol * D;
  const int *cl = L->cluster;
  memcpy(L->state, src, sizeof(double) * L->N);
  memset(L->dst, 0, sizeof(double) * L->N);
  if (L->recon) ox_apply(m, L->degree, L->state, L->igd, 1.0);
  for (int c = 0; c < n; ++c) L->cell_time[c] = L->cl_time[cl[c]];
  char *tf = (char *)malloc(n), *Ucopy = (char *)malloc(n);
  for (int cy = 0; cy < L->n_updates; ++cy) {
    const int ac = L->order[cy];
    double act = L->cl_time[ac];
    L->fm_considered = 1;
    for (int c = 0; c < n; ++c) L->U[c] = cl[c] == ac;
    if (ac > 0) {
      const double ft = L->cl_time[ac - 1];
      L->t1fa = fmax(ft, act);
      L->t2fa = fmin(ft + L->mult[ac - 1] * D, act + L->mult[ac] * D);
      L->t2fa = fmin(L->t2fa, time_now);
    } else
      L->t1fa = L->t2fa = 0.0;
    if (ac < nc - 1) {
      const double st = L->cl_time[ac + 1];
      L->t1sl = fmax(st, act);
      L->t2sl = fmin(st + L->mult[ac + 1] * D, act + L->mult[ac] * D);
      L->t2sl = fmin(L->t2sl, time_now);
    } else
      L->t1sl = L->t2sl = 0.0;
    L->t1sa = act;
    L->t2sa = fmin(act + L->mult[ac] * D, time_now);
    L->dt = fmin(L->mult[ac] * D, time_now - L->cl_time[ac]);
    lts_update_elements(L, L->state, ac, 1);
    L->cl_time[ac] = L->times[2 * cy + 1];
    for (int c = 0; c < n; ++c) L->cell_time[c] = L->cl_time[cl[c]];
    if (!L->recon) continue;
    /* reconstruction with the neighbouring clusters brought to the new time
     * level of `ac` (:608-767); only cells next to `ac` are read by it */
    L->fm_considered = 0;
    memcpy(L->rstate, L->state, sizeof(double) * L->N);
    act = L->cl_time[ac];
    for (int side = 0; side < 2; ++side) {
      const int other = side == 0 ? ac - 1 : ac + 1;
      if (other < 0 || other >= nc) continue;
      if (!(fabs(act - L->cl_time[other]) > tol)) continue;
      for (int c = 0; c < n; ++c) tf[c] = cl[c] == other && (side == 0 ? L->slower[c] : L->faster[c]);
      /* the reference widens the set by further cells of `other` (:631-647,
       * :682-694); their updates are never read, the same cells are updated
       * here for the same arithmetic */
      for (int c = 0; c < n; ++c) {
        Ucopy[c] = 0;
        if (cl[c] != other) continue;
        int nbs[4];
        if (side == 0) {
          /* the last face with a neighbour decides (:635-645) */
          for (int f = 0; f < nface; ++f) {
            const int k = face_cells(m, c, f, nbs);
            for (int j = 0; j < k; ++j) Ucopy[c] = tf[nbs[j]];
          }
        } else {
          for (int f = 0; f < nface; ++f) {
            const int k = face_cells(m, c, f, nbs);
            for (int j = 0; j < k; ++j)
              if (tf[nbs[j]]) Ucopy[c] = 1;
          }
        }
      }
      for (int c = 0; c < n; ++c) L->U[c] = Ucopy[c] || tf[c];
      L->t1fa = L->cl_time[other];
      L->t2fa = act;
      L->t1sl = L->t1sa = L->t1fa;
      L->t2sl = L->t2sa = L->t2fa;
      L->dt = L->t2fa - L->t1fa;
      lts_update_elements(L, L->rstate, other, 0);
    }
    for (int c = 0; c < n; ++c) tf[c] = cl[c] == ac;
    lts_reconstruct(L, L->rstate, tf);
  }
  for (int u = 0; u < L->n_updates; ++u) {
    L->times[2 * u] += L->global_dt;
    L->times[2 * u + 1] += L->global_dt;
  }
  free(tf);
  free(Ucopy);
  memcpy(dst, L->state, sizeof(double) * L->N);
}

/* ------------------------------------------------------------------------ */
/* driver: WaveEquationProblem::run with TimeControl                          */
/* ------------------------------------------------------------------------ */

/* the depth-first entries of a base cell refined uniformly `depth` times */
static int uniform_tree(int dim, int depth, int *out, int pos) {
  out[pos++] = depth > 0;
  if (depth > 0)
    for (int ch = 0; ch < (1 << dim); ++ch) pos = uniform_tree(dim, depth - 1, out, pos);
  return pos;
}

int ox_run(const ox_problem *P, char *log, int log_len, double *final_energy) {
  const int adaptive = P->n_adaptive_refinements > 0;
  /* adaptive runs: the forest of the globally refined box, every base cell
   * refined n_refinements times (levels min_level .. max_level,
   * source/explicit_wave.cc:272-273) */
  const int nbase = ipow(P->n_intervals, P->dim), R = P->n_refinements;
  const int min_level = R, max_level = R + P->n_adaptive_refinements;
  int tcap = 0, ntree = 0;
  int *tree = NULL, *tnew = NULL;
  if (adaptive) {
    long per = 0, pw = 1; /* entries of a uniform tree of depth R */
    for (int l = 0; l <= R; ++l, pw *= 1 << P->dim) per += pw;
    tcap = (int)(per * nbase);
    tree = (int *)malloc(sizeof(int) * tcap);
    for (int B = 0; B < nbase; ++B) ntree = uniform_tree(P->dim, R, tree, ntree);
  }
  ox_mesh *m = adaptive ? ox_mesh_forest(P->dim, P->n_intervals, P->transform,
                                         P->transform_factor, P->boundary_id, tree, ntree)
                        : ox_mesh_hypercube(P->dim, P->n_intervals, P->n_refinements,
                                            P->transform, P->transform_factor,
                                            P->boundary_id);
  const int k = P->degree;
  long N = ox_n_dofs(m, k);
  /* source/parameters.cc:195 and source/explicit_wave.cc:529-545 */
  const double cfl = P->cfl_number / pow((double)k, 1.5);
  const double dt_cfl = cfl * m->min_vertex_distance_over_c;
  /* TimeControl::setup (source/utilities.cc:31-46) */
  const double T = P->final_time, tick = P->output_every_time;
  double time = 0.0, dt = dt_cfl;
  unsigned step = 0, max_step;
  dt = (T - time) / fmax(round((T - time) / dt), 1.0);
  max_step = P->max_time_steps == 0 ? (unsigned)round((T - time) / dt) : (unsigned)P->max_time_steps;
  /* make_dofs overwrites with the unrounded CFL step (source/explicit_wave.cc:222) */
  dt = dt_cfl;
  ox_lts *lts = NULL;
  if (P->integrator == OX_ADER_LTS) {
    /* WaveEquationOperationADERLTS::setup: clusters from the unrounded CFL
     * step, then the time control advances by the slowest cluster's step
     * (include/cluster_manager.templates.h:463) */
    lts = ox_lts_create(m, k, cfl, dt_cfl, P->lts_max_clusters, P->lts_max_diff, 0.0,
                        P->ader_recon);
    dt = ox_lts_global_dt(lts);
  }

  double *sol = (double *)malloc(sizeof(double) * N);
  double *tmp = (double *)malloc(sizeof(double) * N);
  double *t1 = (double *)malloc(sizeof(double) * N);
  double *t2 = (double *)malloc(sizeof(double) * N);
  double *post = P->post_pressure
                     ? (double *)malloc(sizeof(double) * m->n_cells * ipow(k + 2, P->dim))
                     : NULL;
  ox_project(m, k, &P->exact, 0.0, sol);
  /* adapt_mesh (source/explicit_wave.cc:265-330) on the box mesh: marking,
   * new mesh, solution transfer, new CFL step */
  double max_err_init = -1.0;
#define OX_ADAPT()                                                                         \
  do {                                                                                     \
    double *err = (double *)malloc(sizeof(double) * m->n_cells);                           \
    int *lev = (int *)malloc(sizeof(int) * m->n_cells), *fl = (int *)malloc(sizeof(int) * m->n_cells); \
    ox_estimate_error(m, k, sol, err);                                                     \
    ox_forest_levels(P->dim, P->n_intervals, tree, ntree, lev);                            \
    ox_refinement_flags(m->n_cells, err, lev, min_level, max_level, max_err_init, fl);     \
    const int nt2 = ox_adapt_forest(P->dim, P->n_intervals, tree, fl, NULL, 0);            \
    tnew = (int *)realloc(tnew, sizeof(int) * nt2);                                        \
    ox_adapt_forest(P->dim, P->n_intervals, tree, fl, tnew, nt2);                          \
    ox_mesh *m2 = ox_mesh_forest(P->dim, P->n_intervals, P->transform,                     \
                                 P->transform_factor, P->boundary_id, tnew, nt2);          \
    const long N2 = ox_n_dofs(m2, k);                                                      \
    double *s2 = (double *)malloc(sizeof(double) * N2);                                    \
    ox_adapt_forest_transfer(P->dim, k, P->n_intervals, tree, tnew, sol, s2);              \
    free(sol);                                                                             \
    sol = s2;                                                                              \
    tmp = (double *)realloc(tmp, sizeof(double) * N2);                                     \
    t1 = (double *)realloc(t1, sizeof(double) * N2);                                       \
    t2 = (double *)realloc(t2, sizeof(double) * N2);                                       \
    if (post) post = (double *)realloc(post, sizeof(double) * m2->n_cells * ipow(k + 2, P->dim)); \
    ox_mesh_free(m);                                                                       \
    m = m2;                                                                                \
    N = N2;                                                                                \
    int *sw_ = tree;                                                                       \
    tree = tnew;                                                                           \
    tnew = sw_;                                                                            \
    ntree = nt2;                                                                           \
    /* make_dofs: compute_time_step_size (explicit_wave.cc:222, 529-545); with   \
     * ADER-LTS the operator's setup then builds new clusters from it, starting   \
     * at the current time, and sets the slowest cluster's step (:1353-1379,     \
     * include/cluster_manager.templates.h:320-484); adapt_mesh keeps that step \
     * (explicit_wave.cc:324-327) */                                             \
    dt = cfl * m->min_vertex_distance_over_c;                                              \
    if (lts) {                                                                             \
      ox_lts_free(lts);                                                                    \
      lts = ox_lts_create(m, k, cfl, dt, P->lts_max_clusters, P->lts_max_diff, time,       \
                          P->ader_recon);                                                  \
      dt = ox_lts_global_dt(lts);                                                          \
    }                                                                                      \
    free(err);                                                                             \
    free(lev);                                                                             \
    free(fl);                                                                              \
  } while (0)
  if (adaptive) {
    /* the initial adaptation (explicit_wave.cc:704-716) */
    for (int left = P->n_adaptive_refinements; left > 0; --left) {
      OX_ADAPT();
      ox_project(m, k, &P->exact, 0.0, sol);
    }
    double *err = (double *)malloc(sizeof(double) * m->n_cells);
    ox_estimate_error(m, k, sol, err);
    max_err_init = 0.0;
    for (int c = 0; c < m->n_cells; ++c) max_err_init = fmax(max_err_init, err[c]);
    free(err);
  }
  int pos = 0;
  if (log && log_len > 0) log[0] = 0;
#define OX_LOG()                                                                           \
  do {                                                                                     \
    double err[3], epp = -1.0;                                                             \
    ox_l2_errors(m, k, &P->exact, time, sol, err);                                         \
    if (P->post_pressure) {                                                                \
      ox_post_pressure(m, k, sol, post);                                                   \
      epp = ox_post_error(m, k, &P->exact, time, post);                                    \
    }                                                                                      \
    if (log && pos < log_len)                                                              \
      pos += snprintf(log + pos, log_len - pos, "%u %.17g %.10e %.10e %.10e %.10e\n", step,  \
                      time, err[0], err[1], err[2], epp);                                  \
  } while (0)
  OX_LOG();
  /* while (!done()) : time >= final_time || step > max_step (utilities.cc:84-87) */
  while (!(time >= T || step > max_step)) {
    ++step;
    time += dt;
    double *sw = tmp;
    tmp = sol;
    sol = sw;
    if (lts)
      ox_lts_step(lts, time, tmp, sol); /* ArbitraryHighOrderDGLTS, time_integrators.cc:54-63 */
    else
      ox_time_step(m, k, P->integrator, dt, tmp, sol, t1, t2, P->ader_recon);
    if (adaptive && step % P->adaptive_refinement_interval == 0) OX_ADAPT();
    /* TimeControl::at_tick (utilities.cc:89-95), with the step size as the
     * adaptation left it */
    if ((int)(time / tick) != (int)((time - dt) / tick)) OX_LOG();
  }
#undef OX_LOG
#undef OX_ADAPT
  free(tree);
  free(tnew);
  if (final_energy) *final_energy = ox_energy(m, k, sol);
  ox_lts_free(lts);
  free(post);
  free(sol);
  free(tmp);
  free(t1);
  free(t2);
  ox_mesh_free(m);
  return (int)step;
}
