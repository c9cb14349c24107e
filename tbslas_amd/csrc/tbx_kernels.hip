// Point kernels of libtbx for one Chebyshev degree.  This file is compiled
// once per degree with -DTBX_Q=<q>; each build defines the launchers
// launch_{eval,traj,step}_q<q> declared in tbx_internal.h.
#include <hip/hip_runtime.h>

#include "tbx_device.h"

#ifndef TBX_Q
#error "compile with -DTBX_Q=<degree>"
#endif

namespace tbx {
namespace {

constexpr int D = TBX_Q + 1;

// EvalTree at caller points (tree_functor.h:397-791): wrap, locate,
// contract.  out[i][l] point-major.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_eval(TreeView t,
                                                 double *__restrict__ pts,
                                                 int64_t n, int bc,
                                                 int write_back,
                                                 double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x = pts[3 * i + 0], y = pts[3 * i + 1], z = pts[3 * i + 2];
  if (bc == TBX_BC_PERIODIC) {
    x = wrap1(x);
    y = wrap1(y);
    z = wrap1(z);
    if (write_back) {
      pts[3 * i + 0] = x;
      pts[3 * i + 1] = y;
      pts[3 * i + 2] = z;
    }
  }
  const int dof = t.dof;
  double *o = out + i * dof;
  eval_dof<D, MODE>(t, x, y, z, [&](int l, double v) { o[l] = v; });
}

// A velocity functor (FieldExtrapFunctor / FieldSetFunctor / a scaled tree)
// at caller points at time t; the extrapolating kind evaluates its
// extrapolation (the second-stage functor).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_veval(VelSource vel,
                                                  double *__restrict__ pts,
                                                  int64_t n, int bc,
                                                  int write_back, double t,
                                                  int scaled, double g,
                                                  double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x = pts[3 * i + 0], y = pts[3 * i + 1], z = pts[3 * i + 2];
  if (bc == TBX_BC_PERIODIC) {
    x = wrap1(x);
    y = wrap1(y);
    z = wrap1(z);
    if (write_back) {
      pts[3 * i + 0] = x;
      pts[3 * i + 1] = y;
      pts[3 * i + 2] = z;
    }
  }
  double v0, v1, v2;
  eval_velocity<D, MODE>(vel, 2, t, scaled, g, x, y, z, v0, v1, v2);
  out[3 * i + 0] = v0;
  out[3 * i + 1] = v1;
  out[3 * i + 2] = v2;
}

// One IntegrateRK2 substep for caller points, in place.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_traj(VelSource vel,
                                                 double *__restrict__ xs,
                                                 int64_t n, SubArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x = xs[3 * i + 0], y = xs[3 * i + 1], z = xs[3 * i + 2];
  rk2_substep<D, MODE>(vel, a.bc, a.htau, a.tau, a.scaled, a.g1, a.g2, a.ts1,
                       a.ts2, x, y, z);
  xs[3 * i + 0] = x;
  xs[3 * i + 1] = y;
  xs[3 * i + 2] = z;
}

// SolveSemilagInSitu fused (tree_semilag.h:92-135): arrival point of
// (leaf, node) -> num_rk_step RK2 substeps through vel -> conc at the
// departure point -> vals[leaf][l][node] for the refit.  Workgroups are
// remapped so that each XCD runs a contiguous range of arrival points and a
// leaf's coefficients are fetched into one XCD's L2.  BLOCK and WPE (the
// minimum waves per SIMD the register budget must allow) select the launch
// shape: wide workgroups keep the waves resident on a CU inside one or two
// leaves, so their scalar coefficient loads hit the same cache lines.
template <int MODE, int BLOCK, int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_step(
    VelSource vel, TreeView conc, StepArgs a, double *__restrict__ vals,
    double *__restrict__ dep) {
  const uint32_t b = blockIdx.x;
  const int64_t lb = (int64_t)(b & 7u) * a.per_xcd + (int64_t)(b >> 3);
  if (lb >= a.nblocks) return;
  const int64_t i = lb * BLOCK + threadIdx.x;
  if (i >= a.npts) return;
  const int64_t leaf = i / a.P;
  const int m = (int)(i - leaf * a.P);
  // arrival point: corner + 2^-depth * node (tree_utils.h:480-491)
  const uint4 inf = conc.info[leaf];
  const double len = 1.0 / (double)(1u << inf.w);
  double x = (double)inf.x * len + len * a.node[m % D];
  double y = (double)inf.y * len + len * a.node[(m / D) % D];
  double z = (double)inf.z * len + len * a.node[m / (D * D)];
  // ComputeTrajRK2 (traj.inc:48-68), tinit -> tfinal = tinit - dt
  for (int s = 0; s < a.nsub; ++s)
    rk2_substep<D, MODE>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[2 * s],
                         a.g[2 * s + 1], a.ts[2 * s], a.ts[2 * s + 1], x, y, z);
  // con_evaluator(departure points) wraps them in place (semilag.inc:43)
  if (a.bc == TBX_BC_PERIODIC) {
    x = wrap1(x);
    y = wrap1(y);
    z = wrap1(z);
  }
  if (dep) {
    dep[3 * i + 0] = x;
    dep[3 * i + 1] = y;
    dep[3 * i + 2] = z;
  }
  const int dof = conc.dof;
  double *o = vals + leaf * dof * a.P + m;
  eval_dof<D, MODE>(conc, x, y, z,
                    [&](int l, double v) { o[(int64_t)l * a.P] = v; });
}

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

}  // namespace

#define TBX_CAT2(a, b) a##b
#define TBX_CAT(a, b) TBX_CAT2(a, b)

hipError_t TBX_CAT(launch_eval_q, TBX_Q)(int mode, const TreeView &t,
                                         double *pts, int64_t n, int bc,
                                         int write_back, double *out,
                                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(grid_for(n)), blk(kBlock);
  switch (mode) {
    case kFma64:
      hipLaunchKernelGGL(k_eval<kFma64>, g, blk, 0, s, t, pts, n, bc,
                         write_back, out);
      break;
    case kExact64:
      hipLaunchKernelGGL(k_eval<kExact64>, g, blk, 0, s, t, pts, n, bc,
                         write_back, out);
      break;
    case kFp32:
      hipLaunchKernelGGL(k_eval<kFp32>, g, blk, 0, s, t, pts, n, bc,
                         write_back, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t TBX_CAT(launch_veval_q, TBX_Q)(int mode, const VelSource &vel,
                                          double *pts, int64_t n, int bc,
                                          int write_back, double t, int scaled,
                                          double g, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 gr(grid_for(n)), blk(kBlock);
  switch (mode) {
    case kFma64:
      hipLaunchKernelGGL(k_veval<kFma64>, gr, blk, 0, s, vel, pts, n, bc, write_back, t,
                         scaled, g, out);
      break;
    case kExact64:
      hipLaunchKernelGGL(k_veval<kExact64>, gr, blk, 0, s, vel, pts, n, bc, write_back, t,
                         scaled, g, out);
      break;
    case kFp32:
      hipLaunchKernelGGL(k_veval<kFp32>, gr, blk, 0, s, vel, pts, n, bc, write_back, t,
                         scaled, g, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t TBX_CAT(launch_traj_q, TBX_Q)(int mode, const VelSource &vel,
                                         double *x, int64_t n,
                                         const SubArgs &a, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(grid_for(n)), blk(kBlock);
  switch (mode) {
    case kFma64:
      hipLaunchKernelGGL(k_traj<kFma64>, g, blk, 0, s, vel, x, n, a);
      break;
    case kExact64:
      hipLaunchKernelGGL(k_traj<kExact64>, g, blk, 0, s, vel, x, n, a);
      break;
    case kFp32:
      hipLaunchKernelGGL(k_traj<kFp32>, g, blk, 0, s, vel, x, n, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

namespace {

template <int MODE, int BLOCK, int WPE>
hipError_t launch_step_shape(const VelSource &vel, const TreeView &conc, const StepArgs &a,
                             double *vals, double *dep, hipStream_t s) {
  hipLaunchKernelGGL((k_step<MODE, BLOCK, WPE>), dim3((unsigned)(8 * a.per_xcd)), dim3(BLOCK), 0,
                     s, vel, conc, a, vals, dep);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_step_mode(const VelSource &vel, const TreeView &conc, const StepArgs &a,
                            double *vals, double *dep, hipStream_t s) {
  // launch shapes of kStepShapes (tbx_internal.h); the extra shapes are
  // built for the degrees of the benchmark workloads only
  switch (a.shape) {
#if TBX_Q == 8 || TBX_Q == 14
    case 1:
      return launch_step_shape<MODE, 256, 4>(vel, conc, a, vals, dep, s);
    case 2:
      return launch_step_shape<MODE, 1024, 4>(vel, conc, a, vals, dep, s);
    case 3:
      return launch_step_shape<MODE, 512, 4>(vel, conc, a, vals, dep, s);
#endif
    default:
      return launch_step_shape<MODE, 256, 1>(vel, conc, a, vals, dep, s);
  }
}

}  // namespace

hipError_t TBX_CAT(launch_step_q, TBX_Q)(int mode, const VelSource &vel,
                                         const TreeView &conc,
                                         const StepArgs &a, double *vals,
                                         double *dep, hipStream_t s) {
  if (a.npts <= 0) return hipSuccess;
  switch (mode) {
    case kFma64:
      return launch_step_mode<kFma64>(vel, conc, a, vals, dep, s);
    case kExact64:
      return launch_step_mode<kExact64>(vel, conc, a, vals, dep, s);
    case kFp32:
      return launch_step_mode<kFp32>(vel, conc, a, vals, dep, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace tbx
