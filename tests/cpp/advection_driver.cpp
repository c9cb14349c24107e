// Drives the engine through the reference's functor/step API
// (include/tbx/tbslas.hpp), as a tbslas caller would: the rotating-Gaussian
// case of advection.cpp (test 1) at small size.
//
// Prints one line "traj_diff=<a> rk2_diff=<b> linf=<c> sum_drift=<d>
// extrap_diff=<e> set_diff=<f> insitu_extrap_diff=<g>":
//   a: fused device ComputeTrajRK2 vs the reference's host loop over the
//      same device functor (both TBX_EXACT: expected 0)
//   b: SolveSemilagRK2 with device functors vs an explicit traj + eval
//      (expected 0)
//   c: relative L-inf of the concentration after 20 SolveSemilagInSitu
//      steps against the analytically rotated blob
//   d: relative change of the plain sum of the grid-point values (moves with
//      the blob across the grid; semi-Lagrangian steps do not conserve it)
//   e: ComputeTrajRK2(NodeFieldFunctor, FieldExtrapFunctor) fused vs the
//      reference's host loop over the same functors (expected 0)
//   f: ComputeTrajRK2(FieldSetFunctor) fused vs the host loop (expected 0)
//   g: SolveSemilagInSitu with FieldExtrapFunctor(v, v) vs the plain step
//      (1.5 v - 0.5 v: rounding only)
#include <cmath>
#include <cstdio>
#include <vector>

#include "tbx/tbslas.hpp"

namespace {

const double kPi = 3.14159265358979323846;
const double kOmega = 2 * kPi;

void rotation(const double *p, size_t n, double *v) {
  for (size_t i = 0; i < n; ++i) {
    v[3 * i + 0] = kOmega * (0.5 - p[3 * i + 1]);
    v[3 * i + 1] = kOmega * (p[3 * i + 0] - 0.5);
    v[3 * i + 2] = 0.0;
  }
}

void blob(const double *p, size_t n, double t, double *c) {
  const double th = kOmega * t, r = 0.2;
  const double xc = 0.5 + r * std::cos(th), yc = 0.5 + r * std::sin(th), zc = 0.5;
  const double s2 = 2 * 0.06 * 0.06;
  for (size_t i = 0; i < n; ++i) {
    const double dx = p[3 * i] - xc, dy = p[3 * i + 1] - yc, dz = p[3 * i + 2] - zc;
    c[i] = std::exp(-(dx * dx + dy * dy + dz * dz) / s2);
  }
}

// The reference's host-loop RK2 over a functor (any functor that is not a
// device tree takes ComputeTrajRK2's generic path).
struct HostLoop {
  tbx::NodeFieldFunctor<double> *f;
  void operator()(const double *p, int n, double t, double *o) { (*f)(p, n, t, o); }
  void operator()(const double *p, int n, double *o) { (*f)(p, n, o); }
};

template <class F>
struct HostLoopOf {
  F *f;
  void operator()(const double *p, int n, double t, double *o) { (*f)(p, n, t, o); }
  void operator()(const double *p, int n, double *o) { (*f)(p, n, o); }
};

double max_diff(const std::vector<double> &a, const std::vector<double> &b) {
  double d = 0;
  for (size_t i = 0; i < a.size(); ++i) d = std::fmax(d, std::fabs(a[i] - b[i]));
  return d;
}

}  // namespace

int main() {
  const int depth = 3, q = 8;
  const int64_t L = (int64_t)1 << (3 * depth);
  std::vector<uint32_t> anchor(3 * L);
  std::vector<uint8_t> dep(L);
  tbx::check(tbx_uniform_leaves(depth, anchor.data(), dep.data()));
  const int64_t P = (int64_t)(q + 1) * (q + 1) * (q + 1);
  std::vector<double> fk(3 * L * P);
  tbx::check(tbx_first_kind_points(L, anchor.data(), dep.data(), q, fk.data()));
  std::vector<double> vv(3 * L * P), cv(L * P);
  rotation(fk.data(), L * P, vv.data());
  blob(fk.data(), L * P, 0.0, cv.data());
  const int nc = (q + 1) * (q + 2) * (q + 3) / 6;
  std::vector<double> vcf(3 * L * nc), ccf(L * nc);
  tbx::check(tbx_fit_leaves(L, q, 3, vv.data(), vcf.data()));
  tbx::check(tbx_fit_leaves(L, q, 1, cv.data(), ccf.data()));

  tbx::Tree vel(anchor, dep, q, 3, vcf), conc(anchor, dep, q, 1, ccf);
  tbx::NodeFieldFunctor<double> vf(&vel, tbx::Arith::kExact);
  tbx::NodeFieldFunctor<double> cf(&conc, tbx::Arith::kExact);

  // trajectories: fused device call vs the reference's host loop
  std::vector<double> x0(3 * 4096);
  for (size_t i = 0; i < x0.size(); ++i) x0[i] = 0.05 + 0.9 * std::fmod(0.6180339887 * (i + 1), 1.0);
  std::vector<double> xa, xb;
  tbx::ComputeTrajRK2(vf, x0, 0.3, 0.25, 3, xa);
  tbx::ComputeTrajRK2(HostLoop{&vf}, x0, 0.3, 0.25, 3, xb);
  double traj_diff = 0;
  for (size_t i = 0; i < xa.size(); ++i) traj_diff = std::fmax(traj_diff, std::fabs(xa[i] - xb[i]));

  // SolveSemilagRK2 vs traj + eval
  const double dt = 2 * kPi / 200 / kOmega;  // 1/200 of a revolution
  std::vector<double> va(4096), vb(4096), dpt;
  tbx::SolveSemilagRK2(vf, cf, x0, 3, 1, dt, 1, va);
  tbx::ComputeTrajRK2(vf, x0, dt, 0.0, 1, dpt);
  cf(dpt.data(), 4096, vb.data());
  double rk2_diff = 0;
  for (size_t i = 0; i < va.size(); ++i) rk2_diff = std::fmax(rk2_diff, std::fabs(va[i] - vb[i]));

  // 20 in-situ steps, compared with the analytically rotated blob
  std::vector<double> g(3 * L * P);
  tbx::check(tbx_first_kind_points(L, anchor.data(), dep.data(), q, g.data()));
  std::vector<double> before(L * P), after(L * P), exact(L * P);
  cf(g.data(), (int)(L * P), before.data());
  const int steps = 20;
  for (int s = 1; s <= steps; ++s) tbx::SolveSemilagInSitu(vf, conc, s, dt, 1, true);
  cf(g.data(), (int)(L * P), after.data());
  blob(g.data(), L * P, steps * dt, exact.data());
  double err = 0, mx = 0, m0 = 0, m1 = 0;
  for (int64_t i = 0; i < L * P; ++i) {
    err = std::fmax(err, std::fabs(after[i] - exact[i]));
    mx = std::fmax(mx, std::fabs(exact[i]));
    m0 += before[i];
    m1 += after[i];
  }

  // time-dependent velocity functors: a previous field at 0.8 of the speed
  std::vector<double> vcf_prev(vcf.size());
  for (size_t i = 0; i < vcf.size(); ++i) vcf_prev[i] = 0.8 * vcf[i];
  tbx::Tree velp(anchor, dep, q, 3, vcf_prev);
  tbx::FieldExtrapFunctor<double> ef(&velp, &vel, tbx::Arith::kExact);
  tbx::ComputeTrajRK2(vf, ef, x0, 0.3, 0.25, 3, xa);
  tbx::ComputeTrajRK2(HostLoop{&vf}, HostLoopOf<tbx::FieldExtrapFunctor<double>>{&ef}, x0, 0.3,
                      0.25, 3, xb);
  const double extrap_diff = max_diff(xa, xb);
  tbx::FieldSetFunctor<double> sf({&velp, &vel, &velp, &vel}, {0.2, 0.24, 0.28, 0.32},
                                  tbx::Arith::kExact);
  tbx::ComputeTrajRK2(sf, x0, 0.3, 0.25, 3, xa);
  tbx::ComputeTrajRK2(HostLoopOf<tbx::FieldSetFunctor<double>>{&sf}, x0, 0.3, 0.25, 3, xb);
  const double set_diff = max_diff(xa, xb);
  tbx::Tree c1(anchor, dep, q, 1, ccf), c2(anchor, dep, q, 1, ccf);
  tbx::FieldExtrapFunctor<double> same(&vel, &vel, tbx::Arith::kExact);
  for (int s = 1; s <= 2; ++s) {
    tbx::SolveSemilagInSitu(vf, c1, s, dt, 1, true, tbx::Arith::kExact);
    tbx::SolveSemilagInSitu(vf, same, c2, s, dt, 1, true, tbx::Arith::kExact);
  }
  const double insitu_extrap_diff = max_diff(c1.coeffs(), c2.coeffs());

  std::printf("traj_diff=%.3e rk2_diff=%.3e linf=%.3e sum_drift=%.3e extrap_diff=%.3e "
              "set_diff=%.3e insitu_extrap_diff=%.3e\n",
              traj_diff, rk2_diff, err / mx, std::fabs(m1 - m0) / std::fabs(m0), extrap_diff,
              set_diff, insitu_extrap_diff);
  return 0;
}
