// Point kernels of libtbx for one Chebyshev degree.  This file is compiled
// once per degree with -DTBX_Q=<q>; each build defines the launchers
// launch_{eval,traj,step}_q<q> declared in tbx_internal.h.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tbx_device.h"
#include "tbx_mfma.h"

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
template <int MODE, int KIND>
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
  }
  double v0, v1, v2;
  eval_velocity<D, MODE, KIND>(vel, 2, t, scaled, g, bc, x, y, z, v0, v1, v2);
  if (bc == TBX_BC_PERIODIC && write_back) {  // wrapped once per tree
    pts[3 * i + 0] = x;
    pts[3 * i + 1] = y;
    pts[3 * i + 2] = z;
  }
  out[3 * i + 0] = v0;
  out[3 * i + 1] = v1;
  out[3 * i + 2] = v2;
}

// One IntegrateRK2 substep for caller points, in place.
template <int MODE, int KIND>
__global__ __launch_bounds__(kBlock) void k_traj(VelSource vel,
                                                 double *__restrict__ xs,
                                                 int64_t n, SubArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x = xs[3 * i + 0], y = xs[3 * i + 1], z = xs[3 * i + 2];
  rk2_substep<D, MODE, KIND>(vel, a.bc, a.htau, a.tau, a.scaled, a.g1, a.g2, a.ts1,
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
template <int MODE, int KIND, int BLOCK, int WPE>
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
  int s = 0;
  if constexpr (KIND == kVelTree || KIND == kVelExtrap) {
    if (a.vnode) {  // the first stage at the arrival node from the nodal values
      if (a.bc == TBX_BC_PERIODIC) {
        x = wrap1(x);
        y = wrap1(y);
        z = wrap1(z);
      }
      const int lf = locate(vel.t[0], x, y, z);
      double v0, v1, v2;
      if (!nodal_velocity<D>((const NodalT<MODE> *)a.vnode, vel.t[0], inf, m, lf, a.P, a.scaled, a.g[0], v0, v1, v2))
        eval_velocity<D, MODE, KIND>(vel, 1, a.ts[0], a.scaled, a.g[0], a.bc, x, y, z, v0, v1,
                                     v2, lf);
      rk2_finish<D, MODE, KIND>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[1], a.ts[1], x, y, z,
                                v0, v1, v2);
      s = 1;
    }
  }
  for (; s < a.nsub; ++s)
    rk2_substep<D, MODE, KIND>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[2 * s],
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

// The start point of this thread's RK2 substep, read back from its row of
// the regroup buffer after a velocity evaluation instead of being held in
// registers through it (the row is this thread's until the next regroup,
// whose first barrier every thread reaches after its evaluation).  The
// empty asm with a memory clobber makes the compiler load it again rather
// than keep the value it loaded at the regroup.
template <int BLOCK, int NF>
__device__ __forceinline__ void reload_start(RegroupLds<BLOCK, NF> &sh, double (&st)[6]) {
  asm volatile("" ::: "memory");
  st[0] = sh.f[0][threadIdx.x];
  st[1] = sh.f[1][threadIdx.x];
  st[2] = sh.f[2][threadIdx.x];
}

// Leaf and node of arrival point i0 + t of a workgroup starting at point i0
// (t < the workgroup size), points past the last taken as the last: the
// workgroup-uniform i0 / P in scalar registers, the rest in 32 bits.
__device__ __forceinline__ void arrival_node(const StepArgs &a, int64_t i0, int t,
                                             int64_t &leaf, int &m) {
  if (i0 + t >= a.npts) {
    leaf = a.npts / a.P - 1;
    m = a.P - 1;
    return;
  }
  const int64_t l0 = i0 / a.P;
  const uint32_t r = (uint32_t)(i0 - l0 * a.P) + (uint32_t)t;
  const uint32_t q = r / (uint32_t)a.P;
  leaf = l0 + q;
  m = (int)(r - q * (uint32_t)a.P);
}

// The same step with the workgroup's points regrouped by leaf before each
// of the 2 num_rk_step + 1 evaluations (regroup_pos in tbx_device.h; the
// concentration's only in some modes, see below).  The
// thread that holds a point after a regroup carries it on: the RK2 state
// (start point, wrapped midpoint) travels with the point through LDS, and
// the leaf found for the regroup is the one the evaluation uses.  Lanes
// past the last arrival point carry a copy of it and store nothing.  MF:
// the concentration's components go through eval_dof_mfma.
template <int MODE, int KIND, int BLOCK, bool MF>
__device__ __forceinline__ void k_step_rg_body(const VelSource &vel, const TreeView &conc,
                                               const StepArgs &a, double *__restrict__ vals,
                                               double *__restrict__ dep) {
  __shared__ RegroupLds<BLOCK, 6> sh;
  const uint32_t b = blockIdx.x;
  const int64_t lb = (int64_t)(b & 7u) * a.per_xcd + (int64_t)(b >> 3);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  const int64_t i0 = lb * BLOCK;
  int orig = (int)threadIdx.x;
  {
    const int64_t i = i0 + orig < a.npts ? i0 + orig : a.npts - 1;
    const int64_t leaf = i / a.P;
    const int m = (int)(i - leaf * a.P);
    const uint4 inf = conc.info[leaf];
    const double len = 1.0 / (double)(1u << inf.w);
    double st[6];
    st[0] = (double)inf.x * len + len * a.node[m % D];
    st[1] = (double)inf.y * len + len * a.node[(m / D) % D];
    st[2] = (double)inf.z * len + len * a.node[m / (D * D)];
    int lf = 0;
    for (int s = 0; s < a.nsub; ++s) {
      // stage 1 at the (wrapped in place) start point (traj.inc:33)
      if (a.bc == TBX_BC_PERIODIC) {
        st[0] = wrap1(st[0]);
        st[1] = wrap1(st[1]);
        st[2] = wrap1(st[2]);
      }
      lf = locate(vel.t[0], st[0], st[1], st[2]);
      double v0, v1, v2;
      bool nodal = false;
      if constexpr (KIND == kVelTree || KIND == kVelExtrap) {
        // the first substep starts at the arrival nodes: their velocities
        // from the nodal values, and the regroup and contraction only in
        // workgroups holding a node of another depth's leaf (whose points
        // then take the nodal value again where they have one, so that a
        // point's result does not depend on the workgroup it falls in)
        if (s == 0 && a.vnode) {
          // (the arrival node recomputed from the thread index: nothing of it
          // stays live through the substeps)
          int64_t jl;
          int jm;
          arrival_node(a, i0, (int)threadIdx.x, jl, jm);
          const bool got = nodal_velocity<D>((const NodalT<MODE> *)a.vnode, vel.t[0], conc.info[jl], jm, lf, a.P,
                                             a.scaled, a.g[0], v0, v1, v2);
          nodal = !__syncthreads_or(!got);
        }
      }
      if (!nodal) {
        {
          double p3[3] = {st[0], st[1], st[2]};
          regroup(sh, orig, lf, p3);
          st[0] = p3[0];
          st[1] = p3[1];
          st[2] = p3[2];
        }
        eval_velocity<D, MODE, KIND>(vel, 1, a.ts[2 * s], a.scaled, a.g[2 * s], a.bc, st[0],
                                     st[1], st[2], v0, v1, v2, lf);
        reload_start(sh, st);
        if constexpr (KIND == kVelTree || KIND == kVelExtrap) {
          if (s == 0 && a.vnode) {  // the point this thread now holds
            int64_t jl;
            int jm;
            arrival_node(a, i0, orig, jl, jm);
            nodal_velocity<D>((const NodalT<MODE> *)a.vnode, vel.t[0], conc.info[jl], jm, lf, a.P, a.scaled, a.g[0],
                              v0, v1, v2);
          }
        }
      }
      // midpoint xtmp = xinit + (0.5 * dt) * vtmp (traj.inc:36), wrapped
      // for its evaluation only
      st[3] = st[0] + a.htau * v0;
      st[4] = st[1] + a.htau * v1;
      st[5] = st[2] + a.htau * v2;
      if (a.bc == TBX_BC_PERIODIC) {
        st[3] = wrap1(st[3]);
        st[4] = wrap1(st[4]);
        st[5] = wrap1(st[5]);
      }
      lf = locate(vel.t[0], st[3], st[4], st[5]);
      regroup(sh, orig, lf, st);
      eval_velocity<D, MODE, KIND>(vel, 2, a.ts[2 * s + 1], a.scaled, a.g[2 * s + 1], a.bc, st[3],
                             st[4], st[5], v0, v1, v2, lf);
      reload_start(sh, st);
      // xsol = xinit + dt * vtmp (traj.inc:41-42)
      st[0] = st[0] + a.tau * v0;
      st[1] = st[1] + a.tau * v1;
      st[2] = st[2] + a.tau * v2;
    }
    // con_evaluator(departure points) wraps them in place (semilag.inc:43)
    if (a.bc == TBX_BC_PERIODIC) {
      st[0] = wrap1(st[0]);
      st[1] = wrap1(st[1]);
      st[2] = wrap1(st[2]);
    }
    lf = locate(conc, st[0], st[1], st[2]);
    double p3[3] = {st[0], st[1], st[2]};
    // A departure point lies mostly in its midpoint's leaf, by which the
    // points were just regrouped.  Regrouping again before a one-component
    // contraction costs the FP32 + FP64 head and FP64 kernels more than the
    // leaf passes it saves (headline kernel 25.3 against 26.1 ms at q = 14);
    // the FP32 kernel at its 64-register budget, and many components on the
    // matrix cores, keep it.
    constexpr bool kConcRegroup = MF || !(MODE == kFp32h || MODE == kFma64);
    if constexpr (kConcRegroup) regroup(sh, orig, lf, p3);
    const int64_t gi = i0 + orig;
    const bool live = gi < a.npts;
    if (live && dep) {
      dep[3 * gi + 0] = p3[0];
      dep[3 * gi + 1] = p3[1];
      dep[3 * gi + 2] = p3[2];
    }
    const int64_t al = gi / a.P;
    const int am = (int)(gi - al * a.P);
    if constexpr (MF) {
      using T = typename ModeType<MODE>::T;
      eval_dof_mfma<D, T>(conc, lf, p3[0], p3[1], p3[2], vals, al * conc.dof * a.P + am, a.P,
                          live);
    } else {
      double *o = vals + al * conc.dof * a.P + am;
      eval_dof<D, MODE>(
          conc, p3[0], p3[1], p3[2],
          [&](int l, double v) {
            if (live) o[(int64_t)l * a.P] = v;
          },
          lf);
    }
  }
}

template <int MODE, int KIND, int BLOCK, int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_step_rg(
    VelSource vel, TreeView conc, StepArgs a, double *__restrict__ vals,
    double *__restrict__ dep) {
  k_step_rg_body<MODE, KIND, BLOCK, false>(vel, conc, a, vals, dep);
}

// The same step with the concentration's components contracted on the
// matrix cores (conc.dof a multiple of 16; FP64 or FP32 arithmetic).
template <int MODE, int KIND, int BLOCK, int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void k_step_rg_mfma(
    VelSource vel, TreeView conc, StepArgs a, double *__restrict__ vals,
    double *__restrict__ dep) {
  k_step_rg_body<MODE, KIND, BLOCK, true>(vel, conc, a, vals, dep);
}

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

// Calls f with the arithmetic mode as a compile-time constant.
template <typename F>
hipError_t with_mode(int mode, F f) {
  switch (mode) {
    case kFma64:
      return f(std::integral_constant<int, kFma64>{});
    case kExact64:
      return f(std::integral_constant<int, kExact64>{});
    case kFp32:
      return f(std::integral_constant<int, kFp32>{});
    case kFp32h:
      return f(std::integral_constant<int, kFp32h>{});
    default:
      return hipErrorInvalidValue;
  }
}

// The leaf contraction alone, without the step around it: point i sits in
// leaf (w / waves_per_leaf) mod L, w = i / 64 its wavefront, at a position in
// the leaf's box from a hash of i; every component of the leaf is
// contracted there one after another (as in a velocity evaluation) and
// out[i] = ((0 + u_0) + u_1) + ... .  pts (optional) receives the points.
// It measures the contraction's own rate, the ceiling of the step's.
__device__ __forceinline__ double unit_hash(uint64_t v) {
  v += 0x9E3779B97F4A7C15ull;  // SplitMix64 finaliser
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
  v ^= v >> 31;
  return (double)(v >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_contract_only(TreeView t, int64_t n,
                                                          int waves_per_leaf,
                                                          double *__restrict__ out,
                                                          double *__restrict__ pts) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t w = __builtin_amdgcn_readfirstlane((int)((i >> 6) / waves_per_leaf % t.L));
  const int lf = (int)w;
  const uint4 inf = t.info[lf];
  const double len = 1.0 / (double)(1u << inf.w);
  const double x = (double)inf.x * len + len * unit_hash(3 * (uint64_t)i);
  const double y = (double)inf.y * len + len * unit_hash(3 * (uint64_t)i + 1);
  const double z = (double)inf.z * len + len * unit_hash(3 * (uint64_t)i + 2);
  PointBasis<D, MODE> b;
  b.build(t, x, y, z, lf);
  double acc = 0.0;
#pragma unroll 1
  for (int c = 0; c < t.dof; ++c) acc += leaf_value<D, MODE>(t, lf, c, b);
  if (i < n) {
    out[i] = acc;
    if (pts) {
      pts[3 * i + 0] = x;
      pts[3 * i + 1] = y;
      pts[3 * i + 2] = z;
    }
  }
}

// Calls f with the velocity kind as a compile-time constant.
template <typename F>
hipError_t with_kind(int kind, F f) {
  switch (kind) {
    case kVelTree:
      return f(std::integral_constant<int, kVelTree>{});
    case kVelExtrap:
      return f(std::integral_constant<int, kVelExtrap>{});
    case kVelSet:
      return f(std::integral_constant<int, kVelSet>{});
    default:
      return hipErrorInvalidValue;
  }
}

template <int MODE, int KIND, int BLOCK, int WPE>
hipError_t launch_step_shape(const VelSource &vel, const TreeView &conc, const StepArgs &a,
                             double *vals, double *dep, hipStream_t s) {
  hipLaunchKernelGGL((k_step<MODE, KIND, BLOCK, WPE>), dim3((unsigned)(8 * a.per_xcd)),
                     dim3(BLOCK), 0, s, vel, conc, a, vals, dep);
  return hipGetLastError();
}

template <int MODE, int KIND, int BLOCK, int WPE>
hipError_t launch_step_rg_shape(const VelSource &vel, const TreeView &conc, const StepArgs &a,
                                double *vals, double *dep, hipStream_t s) {
  hipLaunchKernelGGL((k_step_rg<MODE, KIND, BLOCK, WPE>), dim3((unsigned)(8 * a.per_xcd)),
                     dim3(BLOCK), 0, s, vel, conc, a, vals, dep);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_step_mode(const VelSource &vel, const TreeView &conc, const StepArgs &a,
                            double *vals, double *dep, hipStream_t s) {
  // many components on the matrix cores: workgroups of mfma_step_block
  // threads (the tile operands and accumulators need the registers)
  if (a.mfma) {
    if constexpr (MODE == kFma64 || MODE == kFp32) {
      if (vel.kind != kVelTree || a.shape != kMfmaStepShape) return hipErrorInvalidValue;
      constexpr int B = mfma_step_block(MODE), W = mfma_step_waves(MODE);
      hipLaunchKernelGGL((k_step_rg_mfma<MODE, kVelTree, B, W>), dim3((unsigned)(8 * a.per_xcd)),
                         dim3(B), 0, s, vel, conc, a, vals, dep);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  // the composite velocity kinds run the default shape only (the host
  // sizes the grid for the shape it asks for: kDefaultStepShape then)
  if (vel.kind == kVelExtrap)
    return launch_step_rg_shape<MODE, kVelExtrap, 1024, 4>(vel, conc, a, vals, dep, s);
  if (vel.kind == kVelSet)
    return launch_step_rg_shape<MODE, kVelSet, 1024, 4>(vel, conc, a, vals, dep, s);
  // launch shapes of kStepShapes (tbx_internal.h); the extra shapes are
  // built for the degrees of the benchmark workloads only
  constexpr int T = kVelTree;
  switch (a.shape) {
    case 0:
      return launch_step_shape<MODE, T, 256, 1>(vel, conc, a, vals, dep, s);
#if TBX_Q == 8 || TBX_Q == 14
    case 4:
      return launch_step_rg_shape<MODE, T, 256, 2>(vel, conc, a, vals, dep, s);
    case 1:
      return launch_step_shape<MODE, T, 256, 4>(vel, conc, a, vals, dep, s);
    case 2:
      return launch_step_shape<MODE, T, 1024, 4>(vel, conc, a, vals, dep, s);
    case 3:
      return launch_step_shape<MODE, T, 512, 4>(vel, conc, a, vals, dep, s);
    case 5:
      return launch_step_rg_shape<MODE, T, 512, 2>(vel, conc, a, vals, dep, s);
    case 7:
      return launch_step_rg_shape<MODE, T, 512, 4>(vel, conc, a, vals, dep, s);
    case 8:
      return launch_step_rg_shape<MODE, T, 1024, 8>(vel, conc, a, vals, dep, s);
    case 9:
      return launch_step_rg_shape<MODE, T, 512, 6>(vel, conc, a, vals, dep, s);
    case 10:
      return launch_step_rg_shape<MODE, T, 768, 6>(vel, conc, a, vals, dep, s);
#endif
    default:  // shape 6, and the shapes not built for this degree
      return launch_step_rg_shape<MODE, T, 1024, 4>(vel, conc, a, vals, dep, s);
  }
}

}  // namespace

#define TBX_CAT2(a, b) a##b
#define TBX_CAT(a, b) TBX_CAT2(a, b)

hipError_t TBX_CAT(launch_eval_q, TBX_Q)(int mode, const TreeView &t,
                                         double *pts, int64_t n, int bc,
                                         int write_back, double *out,
                                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return with_mode(mode, [&](auto m) {
    hipLaunchKernelGGL(k_eval<decltype(m)::value>, dim3(grid_for(n)), dim3(kBlock), 0, s, t, pts,
                       n, bc, write_back, out);
    return hipGetLastError();
  });
}

hipError_t TBX_CAT(launch_contract_q, TBX_Q)(int mode, const TreeView &t, int64_t n,
                                             int waves_per_leaf, double *out, double *pts,
                                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return with_mode(mode, [&](auto m) {
    hipLaunchKernelGGL(k_contract_only<decltype(m)::value>, dim3(grid_for(n)), dim3(kBlock), 0, s,
                       t, n, waves_per_leaf, out, pts);
    return hipGetLastError();
  });
}

hipError_t TBX_CAT(launch_veval_q, TBX_Q)(int mode, const VelSource &vel,
                                          double *pts, int64_t n, int bc,
                                          int write_back, double t, int scaled,
                                          double g, double *out,
                                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return with_mode(mode, [&](auto m) {
    return with_kind(vel.kind, [&](auto k) {
      hipLaunchKernelGGL((k_veval<decltype(m)::value, decltype(k)::value>), dim3(grid_for(n)),
                         dim3(kBlock), 0, s, vel, pts, n, bc, write_back, t, scaled, g, out);
      return hipGetLastError();
    });
  });
}

hipError_t TBX_CAT(launch_traj_q, TBX_Q)(int mode, const VelSource &vel,
                                         double *x, int64_t n,
                                         const SubArgs &a, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return with_mode(mode, [&](auto m) {
    return with_kind(vel.kind, [&](auto k) {
      hipLaunchKernelGGL((k_traj<decltype(m)::value, decltype(k)::value>), dim3(grid_for(n)),
                         dim3(kBlock), 0, s, vel, x, n, a);
      return hipGetLastError();
    });
  });
}

hipError_t TBX_CAT(launch_step_q, TBX_Q)(int mode, const VelSource &vel,
                                         const TreeView &conc,
                                         const StepArgs &a, double *vals,
                                         double *dep, hipStream_t s) {
  if (a.npts <= 0) return hipSuccess;
  return with_mode(mode, [&](auto m) {
    return launch_step_mode<decltype(m)::value>(vel, conc, a, vals, dep, s);
  });
}

}  // namespace tbx
