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
//     (tree_semilag.h:92-135)
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
enum class Arith : uint32_t { kFp64 = TBX_FP64, kExact = TBX_EXACT, kFp32 = TBX_FP32 };

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

namespace detail {
template <typename F>
struct device_field : std::false_type {};
template <typename R>
struct device_field<NodeFieldFunctor<R>> : std::true_type {
  static tbx_tfac_fn g(const NodeFieldFunctor<R> &) { return nullptr; }
  static void *ctx(const NodeFieldFunctor<R> &) { return nullptr; }
};
template <typename R>
struct device_field<ScaledFieldFunctor<R>> : std::true_type {
  static tbx_tfac_fn g(const ScaledFieldFunctor<R> &f) { return f.g(); }
  static void *ctx(const ScaledFieldFunctor<R> &f) { return f.ctx(); }
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
    check(tbx_traj_rk2(field_fn.tree()->handle(), detail::device_field<Functor>::g(field_fn),
                       detail::device_field<Functor>::ctx(field_fn), x.get(), (int64_t)n, tinit,
                       tfinal, num_rk_step, field_fn.tree()->bc(), (uint32_t)field_fn.arith(),
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

// SolveSemilagInSitu (tree_semilag.h:92-135) on a device concentration
// tree: one fused device step (arrival points, RK2, evaluation, refit).
template <typename VelFunctor>
void SolveSemilagInSitu(VelFunctor &tvel_func, Tree &tree_curr, const int timestep,
                        const double dt, int num_rk_step = 1, bool /*adaptive*/ = true,
                        Arith a = Arith::kFp64) {
  static_assert(detail::device_field<VelFunctor>::value,
                "SolveSemilagInSitu needs a device velocity functor");
  check(tbx_semilag_insitu(tvel_func.tree()->handle(), detail::device_field<VelFunctor>::g(tvel_func),
                           detail::device_field<VelFunctor>::ctx(tvel_func), tree_curr.handle(),
                           timestep, dt, num_rk_step, tree_curr.bc(), (uint32_t)a, nullptr,
                           nullptr));
}

}  // namespace tbx
