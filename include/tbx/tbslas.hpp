// tbx/tbslas.hpp: header-only C++ shim that restores arashb/tbslas's functor
// and step API over the C ABI of include/tbx.h, so callers written against
// the reference's templates run on one MI355X ("one MPI rank's whole tree on
// one GPU").
//
//   reference                                   here
//   tbslas::NodeFieldFunctor<real_t, Tree>      tbx::NodeFieldFunctor<double>
//     (tree_functor.h:793-815)                    over tbx_eval
//   tbslas::ComputeTrajRK2 (traj.inc:48-68)     tbx::ComputeTrajRK2
//   tbslas::SolveSemilagRK2 (semilag.inc:26-45) tbx::SolveSemilagRK2
//   tbslas::SolveSemilagInSitu                  tbx::SolveSemilagInSitu
//     (tree_semilag.h:92-181)
//   tbslas::FieldExtrapFunctor                  tbx::FieldExtrapFunctor
//     (tree_extrap_functor.h:27-90)               over tbx_velocity_eval
//   tbslas::FieldSetFunctor                     tbx::FieldSetFunctor
//     (tree_set_functor.h:27-95)
//   the extrapolating overloads of              the same overloads
//     IntegrateRK2 / ComputeTrajRK2
//     (traj.inc:70-115), SolveSemilagRK2
//     (semilag.inc:47-69)
//
// The functor keeps the reference's calling convention: host pointers,
// AoS points, point-major values, and the in-place periodic wrap of the
// caller's points.  Calls with host arrays copy them over PCIe; when both
// functors are device trees, ComputeTrajRK2 / SolveSemilagRK2 /
// SolveSemilagInSitu run as single fused device calls instead of the
// reference's host loops.  Errors throw tbx::Error (the C ABI itself never
// throws).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../tbx.h"

namespace tbx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != TBX_OK) throw Error(rc, tbx_last_error());
}

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(TBX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer owned by the shim.
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { if (p_) (void)hipFree(p_); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void resize(size_t n) {
    if (n <= cap_) { n_ = n; return; }
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    hip_check(hipMalloc(&p_, sizeof(T) * (n ? n : 1)), "hipMalloc");
    cap_ = n;
    n_ = n;
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T *p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// One rank's octree on the GPU: Morton-ordered leaves (anchor in units of
// 2^-depth, depth), degree q, dof components, coefficients [L][dof][Nc] in
// the reference's packed-triangle layout.
class Tree {
 public:
  Tree(const std::vector<uint32_t> &anchor, const std::vector<uint8_t> &depth, int q, int dof,
       const std::vector<double> &coeff, int bc = TBX_BC_FREESPACE)
      : bc_(bc), q_(q), dof_(dof), L_((int64_t)depth.size()) {
    if (anchor.size() != 3 * depth.size()) throw Error(TBX_EINVAL, "anchor/depth size mismatch");
    tbx_tree_desc d{L_, anchor.data(), depth.data(), q, dof, coeff.empty() ? nullptr : coeff.data()};
    check(tbx_tree_create(&d, &t_));
  }
  ~Tree() { if (t_) tbx_tree_destroy(t_); }
  Tree(const Tree &) = delete;
  Tree &operator=(const Tree &) = delete;

  tbx_tree *handle() const { return t_; }
  int bc() const { return bc_; }  // the sim_config->bc the functors use
  int q() const { return q_; }
  int dof() const { return dof_; }
  int64_t num_leaves() const { return L_; }
  int64_t grid_points_per_leaf() const { return (int64_t)(q_ + 1) * (q_ + 1) * (q_ + 1); }

  std::vector<double> coeffs() const {
    std::vector<double> c((size_t)L_ * dof_ * (q_ + 1) * (q_ + 2) * (q_ + 3) / 6);
    check(tbx_tree_get_coeffs(t_, c.data(), 0, nullptr));
    return c;
  }

 private:
  tbx_tree *t_ = nullptr;
  int bc_, q_, dof_;
  int64_t L_;
};

// Arithmetic of the contraction used by the functors below.
enum class Arith : uint32_t {
  kFp64 = TBX_FP64,
  kExact = TBX_EXACT,
  kFp32 = TBX_FP32,
  kFp32Head64 = TBX_FP32_HEAD64
};

// The reference's functor concept over a device tree.
template <typename real_t = double>
class NodeFieldFunctor {
  static_assert(std::is_same<real_t, double>::value, "the engine's coordinates are FP64");

 public:
  // Copies share their staging buffers: the reference passes functors by
  // value (semilag.h:22) and calls them from one host thread.
  explicit NodeFieldFunctor(Tree *tree, Arith a = Arith::kFp64)
      : tree_(tree), arith_(a), buf_(std::make_shared<Staging>()) {}
  virtual ~NodeFieldFunctor() {}

  // EvalTree(tree, const_cast(points), n, out, bc) (tree_functor.h:800-806):
  // host arrays, points wrapped in place when periodic.
  void operator()(const real_t *points_pos, int num_points, real_t *out) {
    if (num_points <= 0) return;
    const size_t n = (size_t)num_points;
    DeviceBuffer<double> &pts = buf_->pts, &vals = buf_->vals;
    pts.resize(3 * n);
    vals.resize(n * tree_->dof());
    hip_check(hipMemcpy(pts.get(), points_pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice),
              "hipMemcpy");
    check(tbx_eval(tree_->handle(), pts.get(), (int64_t)n, tree_->bc(),
                   (uint32_t)arith_ | TBX_WRITE_BACK_WRAP, vals.get(), nullptr));
    hip_check(hipMemcpy(out, vals.get(), sizeof(double) * n * tree_->dof(), hipMemcpyDeviceToHost),
              "hipMemcpy");
    if (tree_->bc() == TBX_BC_PERIODIC)
      hip_check(hipMemcpy(const_cast<real_t *>(points_pos), pts.get(), sizeof(double) * 3 * n,
                          hipMemcpyDeviceToHost),
                "hipMemcpy");
  }

  // The time argument is ignored, as in the reference (tree_functor.h:808-811).
  void operator()(const real_t *points_pos, int num_points, real_t /*time*/, real_t *out) {
    (*this)(points_pos, num_points, out);
  }

  Tree *tree() const { return tree_; }
  Arith arith() const { return arith_; }

 private:
  struct Staging {
    DeviceBuffer<double> pts, vals;
  };
  Tree *tree_;
  Arith arith_;
  std::shared_ptr<Staging> buf_;
};

// A steady velocity tree scaled by g(t): v(x, t) = g(t) V(x).
template <typename real_t = double>
class ScaledFieldFunctor {
 public:
  ScaledFieldFunctor(Tree *tree, tbx_tfac_fn g, void *ctx, Arith a = Arith::kFp64)
      : base_(tree, a), g_(g), ctx_(ctx) {}
  void operator()(const real_t *pts, int n, real_t time, real_t *out) {
    base_(pts, n, out);
    const double f = g_(time, ctx_);
    for (size_t i = 0; i < (size_t)n * base_.tree()->dof(); ++i) out[i] = out[i] * f;
  }
  Tree *tree() const { return base_.tree(); }
  Arith arith() const { return base_.arith(); }
  tbx_tfac_fn g() const { return g_; }
  void *ctx() const { return ctx_; }

 private:
  NodeFieldFunctor<real_t> base_;
  tbx_tfac_fn g_;
  void *ctx_;
};

// Evaluation of a tbx_velocity on host arrays through device staging, the
// points wrapped in place when periodic (each NodeFieldFunctor the
// reference's composite functors build does so).
class VelocityStaging {
 public:
  void eval(const tbx_velocity &v, int bc, Arith a, const double *pts, int num_points,
            double t, double *out) {
    if (num_points <= 0) return;
    const size_t n = (size_t)num_points;
    pts_.resize(3 * n);
    vals_.resize(3 * n);
    hip_check(hipMemcpy(pts_.get(), pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice),
              "hipMemcpy");
    check(tbx_velocity_eval(&v, pts_.get(), (int64_t)n, t, bc,
                            (uint32_t)a | TBX_WRITE_BACK_WRAP, vals_.get(), nullptr));
    hip_check(hipMemcpy(out, vals_.get(), sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
              "hipMemcpy");
    if (bc == TBX_BC_PERIODIC)
      hip_check(hipMemcpy(const_cast<double *>(pts), pts_.get(), sizeof(double) * 3 * n,
                          hipMemcpyDeviceToHost),
                "hipMemcpy");
  }

 private:
  DeviceBuffer<double> pts_, vals_;
};

// FieldExtrapFunctor (tree_extrap_functor.h:27-90): 1.5 v(t^n) - 0.5
// v(t^n-1) from the current and previous dof-3 velocity trees.
template <typename real_t = double>
class FieldExtrapFunctor {
  static_assert(std::is_same<real_t, double>::value, "the engine's coordinates are FP64");

 public:
  FieldExtrapFunctor(Tree *tp, Tree *tc, Arith a = Arith::kFp64)
      : tp_(tp), tc_(tc), arith_(a), buf_(std::make_shared<VelocityStaging>()) {}
  virtual ~FieldExtrapFunctor() {}

  void operator()(const real_t *points_pos, int num_points, real_t *out) {
    buf_->eval(velocity(), tc_->bc(), arith_, points_pos, num_points, 0.0, out);
  }

  // Shifts the pair: previous <- current <- new_tree.  The reference deletes
  // the dropped tree; here trees belong to the caller, so it is returned.
  Tree *update(Tree *new_tree, real_t /*time*/) {
    Tree *dropped = tp_;
    tp_ = tc_;
    tc_ = new_tree;
    return dropped;
  }

  // The device functor pairing NodeFieldFunctor(current) for the first RK
  // stage with this extrapolation for the second (traj.inc:80, 86).
  tbx_velocity velocity() const {
    tbx_velocity v{};
    v.kind = TBX_VEL_EXTRAP;
    v.trees[0] = tc_->handle();
    v.trees[1] = tp_->handle();
    return v;
  }
  Tree *current() const { return tc_; }
  Tree *previous() const { return tp_; }
  Arith arith() const { return arith_; }

 private:
  Tree *tp_, *tc_;
  Arith arith_;
  std::shared_ptr<VelocityStaging> buf_;
};

// FieldSetFunctor (tree_set_functor.h:27-95): four dof-3 trees at four
// times, cubic Hermite interpolation in time (utils/cubic.h:42-56).
template <typename real_t = double>
class FieldSetFunctor {
  static_assert(std::is_same<real_t, double>::value, "the engine's coordinates are FP64");

 public:
  FieldSetFunctor(std::vector<Tree *> field_set_elems, std::vector<real_t> field_set_times,
                  Arith a = Arith::kFp64)
      : elems_(std::move(field_set_elems)), times_(std::move(field_set_times)), arith_(a),
        buf_(std::make_shared<VelocityStaging>()) {
    if (elems_.size() != 4 || times_.size() != 4)
      throw Error(TBX_EINVAL, "FieldSetFunctor takes four trees and four times");
  }
  virtual ~FieldSetFunctor() {}

  void operator()(const real_t *points_pos, int num_points, real_t query_time, real_t *out) {
    buf_->eval(velocity(), elems_[0]->bc(), arith_, points_pos, num_points, query_time, out);
  }

  // Pops the oldest tree and time, appends the new ones; the dropped tree
  // is returned to its owner (the reference deletes it).
  Tree *update(Tree *new_tree, real_t time) {
    Tree *dropped = elems_.front();
    elems_.erase(elems_.begin());
    times_.erase(times_.begin());
    elems_.push_back(new_tree);
    times_.push_back(time);
    return dropped;
  }

  tbx_velocity velocity() const {
    tbx_velocity v{};
    v.kind = TBX_VEL_SET;
    for (int k = 0; k < 4; ++k) {
      v.trees[k] = elems_[k]->handle();
      v.times[k] = times_[k];
    }
    return v;
  }
  Tree *tree() const { return elems_[0]; }
  Arith arith() const { return arith_; }

 private:
  std::vector<Tree *> elems_;
  std::vector<real_t> times_;
  Arith arith_;
  std::shared_ptr<VelocityStaging> buf_;
};

namespace detail {
// Functors that are device velocity sources, as a tbx_velocity plus the
// boundary condition and arithmetic they evaluate with.
template <typename F>
struct device_field : std::false_type {};
template <typename R>
struct device_field<NodeFieldFunctor<R>> : std::true_type {
  static tbx_velocity velocity(const NodeFieldFunctor<R> &f) {
    tbx_velocity v{};
    v.kind = TBX_VEL_TREE;
    v.trees[0] = f.tree()->handle();
    return v;
  }
  static int bc(const NodeFieldFunctor<R> &f) { return f.tree()->bc(); }
};
template <typename R>
struct device_field<ScaledFieldFunctor<R>> : std::true_type {
  static tbx_velocity velocity(const ScaledFieldFunctor<R> &f) {
    tbx_velocity v{};
    v.kind = TBX_VEL_TREE;
    v.trees[0] = f.tree()->handle();
    v.g = f.g();
    v.gctx = f.ctx();
    return v;
  }
  static int bc(const ScaledFieldFunctor<R> &f) { return f.tree()->bc(); }
};
template <typename R>
struct device_field<FieldSetFunctor<R>> : std::true_type {
  static tbx_velocity velocity(const FieldSetFunctor<R> &f) { return f.velocity(); }
  static int bc(const FieldSetFunctor<R> &f) { return f.tree()->bc(); }
};

// (field_fn, extrap_fn) pairs that one device velocity covers:
// NodeFieldFunctor(current) with FieldExtrapFunctor(previous, current).
template <typename F, typename E>
struct device_extrap_pair : std::false_type {};
template <typename R>
struct device_extrap_pair<NodeFieldFunctor<R>, FieldExtrapFunctor<R>> : std::true_type {
  static bool covers(const NodeFieldFunctor<R> &f, const FieldExtrapFunctor<R> &e) {
    return f.tree() == e.current() && f.arith() == e.arith();
  }
};
}  // namespace detail

// IntegrateRK2 (traj.inc:20-46) for any functor of the concept: the
// reference's host algorithm, used when a functor is not a device tree.
template <typename real_t, typename FieldFunctor>
void IntegrateRK2(FieldFunctor field_fn, const std::vector<real_t> &xinit, const real_t tinit,
                  const real_t dt, std::vector<real_t> &xsol) {
  const int num_points = (int)(xinit.size() / 3);
  std::vector<real_t> xtmp(xinit.size()), vtmp(xinit.size());
  field_fn(xinit.data(), num_points, tinit, vtmp.data());
  for (size_t i = 0; i < xinit.size(); ++i) xtmp[i] = xinit[i] + 0.5 * dt * vtmp[i];
  field_fn(xtmp.data(), num_points, tinit + 0.5 * dt, vtmp.data());
  for (size_t i = 0; i < xinit.size(); ++i) xsol[i] = xinit[i] + dt * vtmp[i];
}

// ComputeTrajRK2 (traj.inc:48-68).  A device-tree functor runs the whole
// trajectory in one device call.
template <typename real_t, typename Functor>
void ComputeTrajRK2(Functor field_fn, const std::vector<real_t> &xinit, const real_t tinit,
                    const real_t tfinal, const int num_rk_step, std::vector<real_t> &xsol) {
  if (xsol.size() != xinit.size()) xsol.resize(xinit.size());
  if constexpr (detail::device_field<Functor>::value) {
    const size_t n = xinit.size() / 3;
    if (n == 0) return;
    DeviceBuffer<double> x(3 * n);
    hip_check(hipMemcpy(x.get(), xinit.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice),
              "hipMemcpy");
    const tbx_velocity v = detail::device_field<Functor>::velocity(field_fn);
    check(tbx_traj_rk2_v(&v, x.get(), (int64_t)n, tinit, tfinal, num_rk_step,
                         detail::device_field<Functor>::bc(field_fn), (uint32_t)field_fn.arith(),
                         x.get(), nullptr));
    hip_check(hipMemcpy(xsol.data(), x.get(), sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
              "hipMemcpy");
  } else {
    const real_t tau = (tfinal - tinit) / num_rk_step;
    for (size_t i = 0; i < xinit.size(); ++i) xsol[i] = xinit[i];
    real_t t = tinit;
    for (int s = 0; s < num_rk_step; ++s) {
      IntegrateRK2(field_fn, xsol, t, tau, xsol);
      t = t + tau;
    }
  }
}

// IntegrateRK2 with a separate functor for the second stage (traj.inc:70-95):
// field_fn at xinit, extrap_fn at the midpoint, both without a time.
template <typename real_t, typename FieldFunctor, typename ExtrapFunctor>
void IntegrateRK2(FieldFunctor field_fn, ExtrapFunctor extrap_fn,
                  const std::vector<real_t> &xinit, const real_t /*tinit*/, const real_t dt,
                  std::vector<real_t> &xsol) {
  const int num_points = (int)(xinit.size() / 3);
  std::vector<real_t> xtmp(xinit.size()), vtmp(xinit.size());
  field_fn(xinit.data(), num_points, vtmp.data());
  for (size_t i = 0; i < xinit.size(); ++i) xtmp[i] = xinit[i] + 0.5 * dt * vtmp[i];
  extrap_fn(xtmp.data(), num_points, vtmp.data());
  for (size_t i = 0; i < xinit.size(); ++i) xsol[i] = xinit[i] + dt * vtmp[i];
}

// ComputeTrajRK2 with the extrapolating second stage (traj.inc:97-115).
// NodeFieldFunctor(current) + FieldExtrapFunctor(previous, current), the
// pairing the reference's drivers use (advtvextrap.cpp:294-300), runs as
// one device call; other pairs run the host loop over their functors.
template <typename real_t, typename FieldFunctor, typename ExtrapFunctor>
void ComputeTrajRK2(FieldFunctor field_fn, ExtrapFunctor extrap_fn,
                    const std::vector<real_t> &xinit, const real_t tinit, const real_t tfinal,
                    const int num_rk_step, std::vector<real_t> &xsol) {
  if (xsol.size() != xinit.size()) xsol.resize(xinit.size());
  if constexpr (detail::device_extrap_pair<FieldFunctor, ExtrapFunctor>::value) {
    if (detail::device_extrap_pair<FieldFunctor, ExtrapFunctor>::covers(field_fn, extrap_fn)) {
      const size_t n = xinit.size() / 3;
      if (n == 0) return;
      DeviceBuffer<double> x(3 * n);
      hip_check(hipMemcpy(x.get(), xinit.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice),
                "hipMemcpy");
      const tbx_velocity v = extrap_fn.velocity();
      check(tbx_traj_rk2_v(&v, x.get(), (int64_t)n, tinit, tfinal, num_rk_step,
                           field_fn.tree()->bc(), (uint32_t)field_fn.arith(), x.get(), nullptr));
      hip_check(hipMemcpy(xsol.data(), x.get(), sizeof(double) * 3 * n, hipMemcpyDeviceToHost),
                "hipMemcpy");
      return;
    }
  }
  const real_t tau = (tfinal - tinit) / num_rk_step;
  for (size_t i = 0; i < xinit.size(); ++i) xsol[i] = xinit[i];
  real_t t = tinit;
  for (int s = 0; s < num_rk_step; ++s) {
    IntegrateRK2(field_fn, extrap_fn, xsol, t, tau, xsol);
    t = t + tau;
  }
}

// SolveSemilagRK2 (semilag.inc:26-45).
template <typename real_t, typename VFieldFunctor, typename CFieldFunctor>
void SolveSemilagRK2(VFieldFunctor vel_evaluator, CFieldFunctor con_evaluator,
                     const std::vector<real_t> &points_pos, const int sdim, const int timestep,
                     const real_t dt, const int num_rk_step, std::vector<real_t> &points_vals) {
  const size_t num_points = points_pos.size() / sdim;
  const real_t tinit = timestep * dt;
  const real_t tfinal = tinit - dt;
  std::vector<real_t> departure(points_pos.size());
  ComputeTrajRK2(vel_evaluator, points_pos, tinit, tfinal, num_rk_step, departure);
  con_evaluator(departure.data(), (int)num_points, points_vals.data());
}

// SolveSemilagRK2 with the extrapolating second stage (semilag.inc:47-69).
template <typename real_t, typename VFieldFunctor, typename VExtrapFunctor,
          typename CFieldFunctor>
void SolveSemilagRK2(VFieldFunctor vel_evaluator, VExtrapFunctor vel_extrap_evaluator,
                     CFieldFunctor con_evaluator, const std::vector<real_t> &points_pos,
                     const int sdim, const int timestep, const real_t dt, const int num_rk_step,
                     std::vector<real_t> &points_vals) {
  const size_t num_points = points_pos.size() / sdim;
  const real_t tinit = timestep * dt;
  const real_t tfinal = tinit - dt;
  std::vector<real_t> departure(points_pos.size());
  ComputeTrajRK2(vel_evaluator, vel_extrap_evaluator, points_pos, tinit, tfinal, num_rk_step,
                 departure);
  con_evaluator(departure.data(), (int)num_points, points_vals.data());
}

// SolveSemilagInSitu (tree_semilag.h:92-135) on a device concentration
// tree: one fused device step (arrival points, RK2, evaluation, refit).
template <typename VelFunctor>
void SolveSemilagInSitu(VelFunctor &tvel_func, Tree &tree_curr, const int timestep,
                        const double dt, int num_rk_step = 1, bool /*adaptive*/ = true,
                        Arith a = Arith::kFp64) {
  static_assert(detail::device_field<VelFunctor>::value,
                "SolveSemilagInSitu needs a device velocity functor");
  const tbx_velocity v = detail::device_field<VelFunctor>::velocity(tvel_func);
  check(tbx_semilag_insitu_v(&v, tree_curr.handle(), timestep, dt, num_rk_step, tree_curr.bc(),
                             (uint32_t)a, nullptr, nullptr));
}

// SolveSemilagInSitu with the extrapolating second stage
// (tree_semilag.h:137-181): NodeFieldFunctor(current) and
// FieldExtrapFunctor(previous, current).
template <typename R>
void SolveSemilagInSitu(NodeFieldFunctor<R> &tvel_func, FieldExtrapFunctor<R> &tvel_extrap,
                        Tree &tree_curr, const int timestep, const double dt,
                        int num_rk_step = 1, bool /*adaptive*/ = true,
                        Arith a = Arith::kFp64) {
  if (!detail::device_extrap_pair<NodeFieldFunctor<R>, FieldExtrapFunctor<R>>::covers(
          tvel_func, tvel_extrap))
    throw Error(TBX_EUNSUPPORTED,
                "SolveSemilagInSitu: the field functor must evaluate the extrapolation's "
                "current tree");
  const tbx_velocity v = tvel_extrap.velocity();
  check(tbx_semilag_insitu_v(&v, tree_curr.handle(), timestep, dt, num_rk_step, tree_curr.bc(),
                             (uint32_t)a, nullptr, nullptr));
}

}  // namespace tbx
