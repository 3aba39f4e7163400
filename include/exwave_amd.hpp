// exwave_amd.hpp -- C++ host mirror of ExWave's hot-path interface over the
// C-ABI of exwave_amd.h.  Same class names, argument meaning and error
// behaviour (exceptions, caught by the caller's main) as the reference:
//
//   WaveEquationOperationBase<dim>   include/wave_equation_operations.h:56-98
//   WaveEquationOperation<dim,k>     include/wave_equation_operations.h:104-224
//   WaveEquationOperationADER<dim,k> include/wave_equation_operations.h:226-293
//   ExplicitIntegrator<V,Op>         include/time_integrators.h:25-39
//   ExplicitEuler, ClassRK4, LowStorageRK33Reg2, LowStorageRK45Reg2,
//   LowStorageRK59Reg2, ArbitraryHighOrderDG   source/time_integrators.cc:24-363
//   RKVectorUpdatesRange                       source/time_integrators.cc:165-195
//   TimeControl                                include/utilities.h:37-73
//
// The vector type is DeviceVector (FP64 device memory in the cell-major
// layout of exwave_amd.h) instead of LinearAlgebra::distributed::Vector.
// Everything is asynchronous on the context's stream; DeviceVector reads
// synchronise.
#pragma once
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "exwave_amd.h"

namespace exwave_amd {

inline void check(int rc) {
  if (rc < 0) throw std::runtime_error(std::string("exwave_amd: ") + exw_last_error());
}

class Context {
 public:
  explicit Context(int device = 0, void *stream = nullptr) { check(exw_ctx_create(device, stream, &h_)); }
  ~Context() { exw_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  exw_ctx handle() const { return h_; }
  void sync() const { check(exw_sync(h_)); }

 private:
  exw_ctx h_ = nullptr;
};

// FP64 device vector (the role of LinearAlgebra::distributed::Vector<double>).
class DeviceVector {
 public:
  typedef double value_type;
  DeviceVector() = default;
  DeviceVector(const Context &ctx, int64_t n) { reinit(ctx, n); }
  ~DeviceVector() { release(); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept { *this = std::move(o); }
  DeviceVector &operator=(DeviceVector &&o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(data_, o.data_);
    std::swap(n_, o.n_);
    return *this;
  }
  void reinit(const Context &ctx, int64_t n) {
    release();
    ctx_ = ctx.handle();
    n_ = n;
    check(exw_alloc(ctx_, sizeof(double) * (size_t)n, (void **)&data_));
  }
  // same meaning as deal.II's Vector::swap: exchanges the storage
  void swap(DeviceVector &o) {
    std::swap(ctx_, o.ctx_);
    std::swap(data_, o.data_);
    std::swap(n_, o.n_);
  }
  int64_t size() const { return n_; }
  double *data() { return data_; }
  const double *data() const { return data_; }
  void upload(const std::vector<double> &h) {
    if ((int64_t)h.size() != n_) throw std::runtime_error("DeviceVector::upload: size mismatch");
    check(exw_memcpy_h2d(ctx_, data_, h.data(), sizeof(double) * (size_t)n_));
  }
  std::vector<double> download() const {
    std::vector<double> h((size_t)n_);
    check(exw_memcpy_d2h(ctx_, h.data(), data_, sizeof(double) * (size_t)n_));
    return h;
  }

 private:
  void release() {
    if (data_) exw_free(ctx_, data_);
    data_ = nullptr;
    n_ = 0;
  }
  exw_ctx ctx_ = nullptr;
  double *data_ = nullptr;
  int64_t n_ = 0;
};

// TimeControl (source/utilities.cc:31-95)
class TimeControl {
 public:
  void setup(double final_time_in, double tick_time, double time_step_in, int max_time_step_in = 0) {
    final_time = final_time_in;
    time = 0.0;
    tick_size = tick_time;
    time_step_number = 0;
    time_step = (final_time - time) / std::max(std::round((final_time - time) / time_step_in), 1.0);
    max_time_step = max_time_step_in == 0 ? (unsigned)std::round((final_time - time) / time_step)
                                          : (unsigned)max_time_step_in;
  }
  void advance_time_step() {
    ++time_step_number;
    time += time_step;
  }
  void set_time_step(double dt) { time_step = dt; }
  void set_time(double t) { time = t; }
  double get_time() const { return time; }
  double get_time_step() const { return time_step; }
  unsigned get_step_number() const { return time_step_number; }
  unsigned get_output_step_number() const { return (unsigned)(time / tick_size); }
  bool done() const { return time >= final_time || time_step_number > max_time_step; }
  bool at_tick() const { return int(time / tick_size) != int((time - time_step) / tick_size); }

 private:
  double final_time = 1.0, time_step = 0.0, time = 0.0, tick_size = 0.1;
  unsigned time_step_number = 0, max_time_step = 0;
};

// Material (include/utilities.h:28-34)
struct Material {
  double density = 1.0;
  double speed = 1.0;
};

// The geometry a WaveEquationOperation is set up on: multilinear cells with
// the connectivity conventions of exw_op_create.
struct Mesh {
  int dim = 0;
  int n_cells = 0;
  std::vector<double> vertices;  // [n_cells][2^dim][dim]
  std::vector<int> neighbor;     // [n_cells][2*dim]
  std::vector<int> material_id;  // [n_cells]

  // input_geometry_description (include/input_parameters.h:59-88)
  static Mesh hypercube(int dim, int n_intervals, int n_refinements, int transform, double factor,
                        int boundary_id) {
    Mesh m;
    m.dim = dim;
    check(exw_mesh_hypercube(dim, n_intervals, n_refinements, transform, factor, boundary_id,
                             &m.n_cells, nullptr, nullptr));
    m.vertices.resize((size_t)m.n_cells * (1 << dim) * dim);
    m.neighbor.resize((size_t)m.n_cells * 2 * dim);
    m.material_id.assign(m.n_cells, 0);
    check(exw_mesh_hypercube(dim, n_intervals, n_refinements, transform, factor, boundary_id,
                             &m.n_cells, m.vertices.data(), m.neighbor.data()));
    return m;
  }
};

template <int dim>
class WaveEquationOperationBase {
 public:
  typedef double value_type;
  virtual ~WaveEquationOperationBase() = default;
  virtual void setup(const Mesh &mesh, const std::vector<Material> &mats) = 0;
  virtual std::string Name() = 0;
  virtual TimeControl &get_time_control() const = 0;
  virtual void apply(const DeviceVector &src, DeviceVector &dst) const = 0;
  virtual void apply_ader(const DeviceVector &src, DeviceVector &dst) const = 0;
  virtual void project_initial_field(DeviceVector &solution, const exw_modes &function) const = 0;
  // the RK stage with RKVectorUpdater fused (no counterpart in the reference:
  // the reference calls apply and then RKVectorUpdatesRange)
  virtual void apply_rk_stage(const DeviceVector &src, DeviceVector &b, DeviceVector &y_out,
                              DeviceVector *b_out, double f1, double f2, bool is_last) const = 0;
  virtual exw_op handle() const = 0;
  virtual const Context &context() const = 0;
  virtual int precision() const = 0;
  // compute_post_pressure (source/wave_equation_operations.cc:786-848):
  // tmp_vector <- apply(solution), post_pressure <- p* ((k+2)^dim Gauss-Lobatto
  // nodal values per cell, exw_post_n_values)
  virtual void compute_post_pressure(const DeviceVector &solution, DeviceVector &tmp_vector,
                                     DeviceVector &post_pressure) const {
    check(exw_post_pressure(handle(), solution.data(), tmp_vector.data(), post_pressure.data(),
                            nullptr));
  }
  virtual unsigned int cluster_id(unsigned int) const { return (unsigned)-1; }
  virtual value_type time_step(unsigned int) const { return get_time_control().get_time_step(); }
};

template <int dim, int fe_degree>
class WaveEquationOperation : public WaveEquationOperationBase<dim> {
 public:
  typedef double value_type;
  static const int dimension = dim;
  WaveEquationOperation(const Context &ctx, TimeControl &time_control_in, int precision = EXW_FP64)
      : ctx_(ctx), time_control(time_control_in), precision_(precision) {}
  ~WaveEquationOperation() override {
    if (op_) exw_op_destroy(op_);
  }
  // WaveEquationOperation::setup + reset_data_vectors
  // (source/wave_equation_operations.cc:200-250, 383-400)
  void setup(const Mesh &mesh, const std::vector<Material> &mats) override {
    if (mesh.dim != dim) throw std::runtime_error("WaveEquationOperation::setup: dim mismatch");
    std::vector<double> rho(mesh.n_cells), speed(mesh.n_cells);
    for (int c = 0; c < mesh.n_cells; ++c) {
      rho[c] = mats.at(mesh.material_id[c]).density;
      speed[c] = mats.at(mesh.material_id[c]).speed;
    }
    if (op_) exw_op_destroy(op_);
    op_ = nullptr;
    check(exw_op_create(ctx_.handle(), dim, fe_degree, mesh.n_cells, mesh.vertices.data(),
                        mesh.neighbor.data(), rho.data(), speed.data(), &op_));
  }
  std::string Name() override { return "Runge-Kutta"; }
  TimeControl &get_time_control() const override { return time_control; }
  int64_t n_dofs() const { return exw_op_n_dofs(op_); }
  double min_vertex_distance_over_c() const { return exw_op_min_vertex_distance_over_c(op_); }
  void apply(const DeviceVector &src, DeviceVector &dst) const override {
    check(exw_op_apply(op_, src.data(), dst.data(), precision_));
  }
  void apply_ader(const DeviceVector &, DeviceVector &) const override {
    throw std::runtime_error("apply_ader is not implemented for the Runge-Kutta operator");
  }
  void project_initial_field(DeviceVector &solution, const exw_modes &function) const override {
    check(exw_project(op_, &function, time_control.get_time(), solution.data()));
  }
  void apply_rk_stage(const DeviceVector &src, DeviceVector &b, DeviceVector &y_out,
                      DeviceVector *b_out, double f1, double f2, bool is_last) const override {
    check(exw_rk_stage(op_, src.data(), b.data(), y_out.data(), b_out ? b_out->data() : nullptr,
                       f1, f2, is_last ? 1 : 0, precision_));
  }
  exw_op handle() const override { return op_; }
  const Context &context() const override { return ctx_; }
  int precision() const override { return precision_; }

 protected:
  const Context &ctx_;
  TimeControl &time_control;
  int precision_;
  exw_op op_ = nullptr;
};

template <int dim, int fe_degree>
class WaveEquationOperationADER : public WaveEquationOperation<dim, fe_degree> {
 public:
  WaveEquationOperationADER(const Context &ctx, TimeControl &tc, bool use_ader_post,
                            int precision = EXW_FP64)
      : WaveEquationOperation<dim, fe_degree>(ctx, tc, precision), recon_(use_ader_post) {}
  std::string Name() override { return "ADER"; }
  void setup(const Mesh &mesh, const std::vector<Material> &mats) override {
    WaveEquationOperation<dim, fe_degree>::setup(mesh, mats);
    work_.reinit(this->ctx_, 2 * this->n_dofs());
  }
  // dst = -M^-1 A(int u dt) like WaveEquationOperationADER::apply_ader
  // (source/wave_equation_operations.cc:1346-1381)
  void apply_ader(const DeviceVector &src, DeviceVector &dst) const override {
    const int64_t N = this->n_dofs();
    double *tempsrc = const_cast<double *>(work_.data()), *rec = tempsrc + N;
    if (recon_) check(exw_op_apply(this->op_, src.data(), rec, this->precision_));
    check(exw_ader_ck(this->op_, src.data(), recon_ ? rec : nullptr,
                      this->time_control.get_time_step(), 0.0, 0.0, tempsrc, this->precision_));
    exw_epilogue e{nullptr, dst.data(), -1.0, nullptr, nullptr, 0.0, 0, nullptr};
    check(exw_op_apply_fused(this->op_, tempsrc, &e, this->precision_));
  }
  bool use_ader_post() const { return recon_; }
  double *work() const { return const_cast<double *>(work_.data()); }

 private:
  bool recon_;
  DeviceVector work_;
};

template <typename VectorType, typename Operator>
class ExplicitIntegrator {
 public:
  virtual ~ExplicitIntegrator() = default;
  // vec_n is clobbered, the result is left in vec_np (include/time_integrators.h:25-39)
  virtual void perform_time_step(VectorType &vec_n, VectorType &vec_np, const double time_step,
                                 Operator &op) = 0;
};

namespace internal {
template <typename VectorType, typename Operator>
class FusedIntegrator : public ExplicitIntegrator<VectorType, Operator> {
 public:
  explicit FusedIntegrator(int id) : id_(id) {}
  void perform_time_step(VectorType &vec_n, VectorType &vec_np, const double time_step,
                         Operator &op) override {
    const int nw = exw_time_step_work_vectors(id_);
    const int64_t need = std::max<int64_t>(nw, 1) * vec_n.size();
    if (work_.size() != need) work_.reinit(op.context(), need);
    check(exw_time_step(op.handle(), id_, time_step, vec_n.data(), vec_np.data(), work_.data(),
                        recon_, op.precision()));
  }
  void set_options(int recon) { recon_ = recon; }

 private:
  int id_;
  int recon_ = 0;
  DeviceVector work_;
};
}  // namespace internal

// source/time_integrators.cc:24-34
template <typename V, typename Op>
struct ExplicitEuler : internal::FusedIntegrator<V, Op> {
  ExplicitEuler() : internal::FusedIntegrator<V, Op>(EXW_EXPL_EULER) {}
};
// source/time_integrators.cc:66-102
template <typename V, typename Op>
struct ClassRK4 : internal::FusedIntegrator<V, Op> {
  ClassRK4() : internal::FusedIntegrator<V, Op>(EXW_CLRK4) {}
};
// source/time_integrators.cc:200-231
template <typename V, typename Op>
struct LowStorageRK33Reg2 : internal::FusedIntegrator<V, Op> {
  LowStorageRK33Reg2() : internal::FusedIntegrator<V, Op>(EXW_LSRK33R2) {}
};
// source/time_integrators.cc:235-283
template <typename V, typename Op>
struct LowStorageRK45Reg2 : internal::FusedIntegrator<V, Op> {
  LowStorageRK45Reg2() : internal::FusedIntegrator<V, Op>(EXW_LSRK45R2) {}
};
// source/time_integrators.cc:286-363
template <typename V, typename Op>
struct LowStorageRK59Reg2 : internal::FusedIntegrator<V, Op> {
  LowStorageRK59Reg2() : internal::FusedIntegrator<V, Op>(EXW_LSRK59R2) {}
};
// source/time_integrators.cc:37-50
template <typename V, typename Op>
struct ArbitraryHighOrderDG : internal::FusedIntegrator<V, Op> {
  explicit ArbitraryHighOrderDG(bool use_ader_post = false)
      : internal::FusedIntegrator<V, Op>(EXW_ADER) {
    this->set_options(use_ader_post ? 1 : 0);
  }
};

// RKVectorUpdatesRange (source/time_integrators.cc:165-195): the unfused update
inline void RKVectorUpdatesRange(const Context &ctx, double factor1, double factor2, bool is_last,
                                 const DeviceVector &vector1, DeviceVector &vector2,
                                 DeviceVector &vector3) {
  if (vector2.size() != vector1.size() || vector3.size() != vector1.size())
    throw std::runtime_error("RKVectorUpdatesRange: dimension mismatch");
  check(exw_rk_update(ctx.handle(), vector1.size(), factor1, factor2, is_last ? 1 : 0,
                      vector1.data(), vector2.data(), vector3.data()));
}

}  // namespace exwave_amd
