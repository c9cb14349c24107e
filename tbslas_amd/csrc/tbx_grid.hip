// Leaves holding regular N^3 sample grids: tricubic evaluation and the
// fused in-situ step (the large cubic-interpolation variant of the path).
//
// A leaf's sample (a, b, c) sits at corner + 2^-depth (a, b, c) / (N - 1),
// x fastest, samples [L][dof][N^3].  A point is located like any other
// (Morton key, directory, binary search) and interpolated by the
// reference's fast_interp (tree_functor.h:89-153) in the leaf's unit
// coordinates u = (x - corner) 2^depth: a 4 x 4 x 4 Lagrange stencil whose
// first index is (int)(u (N - 1)) - 1 clamped to [0, N - 4], zero outside
// [0, 1]^3.  kExact64 keeps fast_interp's operation order (bit-identical
// with the oracle); kFma64 sums the stencil axis by axis with fused
// multiply-adds; kFp32 reads FP32 samples and sums in FP32 (coordinates,
// stencil position and RK2 stay FP64).  The step has no refit: the new
// samples are the values at the departure points of the sample points.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tbx_device.h"

namespace tbx {
namespace {

struct GridData {
  const double *s;    // [L][dof][N^3]
  const float *s32;   // FP32 mirror (kFp32)
  int n;
  int dof;
};

// lagr_denom of fast_interp (tree_functor.h:93-99), same divisions.
__device__ __forceinline__ double lagr_denom(int i) {
  double d = 1;
  for (int j = 0; j < 4; j++)
    if (i != j) d /= (double)(i - j);
  return d;
}

// All dof components of a leaf at (x, y, z) through store(c, value): inf is
// the leaf's anchor and depth, data its samples [dof][N^3] (FP32 samples for
// kFp32, FP64 otherwise), in global memory or in LDS.
// U2: unroll of the outer stencil loop (4: every gather issued up front,
// for global memory; 1: a row of 16 at a time, for LDS).
template <int MODE, int U2 = 4, typename T, typename Store>
__device__ __forceinline__ void grid_interp_at(const uint4 inf, const T *data, int N, int dof,
                                               double x, double y, double z, Store store) {
  const double s = (double)(1u << inf.w), len = 1.0 / s;
  double u[3] = {(x - (double)inf.x * len) * s, (y - (double)inf.y * len) * s,
                 (z - (double)inf.z * len) * s};
  const int N3 = N * N * N;
  if (!(u[0] >= 0 && u[0] <= 1.0 && u[1] >= 0 && u[1] <= 1.0 && u[2] >= 0 && u[2] <= 1.0)) {
    for (int c = 0; c < dof; ++c) store(c, 0.0);  // tree_functor.h:106-114
    return;
  }
  int gi[3];
  double M[3][4];
#pragma unroll
  for (int j = 0; j < 3; ++j) {  // tree_functor.h:118-135
    double p = u[j] * (N - 1);
    gi[j] = ((int)p) - 1;
    if (gi[j] < 0) gi[j] = 0;
    if (gi[j] > N - 4) gi[j] = N - 4;
    p -= gi[j];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double m = lagr_denom(k);
#pragma unroll
      for (int l = 0; l < 4; ++l)
        if (k != l) m *= (p - l);
      M[j][k] = m;
    }
  }
  const int base = gi[0];
  for (int c = 0; c < dof; ++c) {
    const T *cb = data + c * N3 + base;
    if constexpr (MODE == kExact64) {
      double val = 0;  // tree_functor.h:137-151
#pragma unroll U2
      for (int j2 = 0; j2 < 4; ++j2)
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const double M1M2 = M[1][j1] * M[2][j2];
          const T *row = cb + N * ((gi[1] + j1) + N * (gi[2] + j2));
#pragma unroll
          for (int j0 = 0; j0 < 4; ++j0) val += M[0][j0] * M1M2 * row[j0];
        }
      store(c, val);
    } else if constexpr (MODE == kFp32) {
      float mf[3][4];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) mf[j][k] = (float)M[j][k];
      float val = 0.f;
#pragma unroll U2
      for (int j2 = 0; j2 < 4; ++j2) {
        float vz = 0.f;
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const T *row = cb + N * ((gi[1] + j1) + N * (gi[2] + j2));
          float r = 0.f;
#pragma unroll
          for (int j0 = 0; j0 < 4; ++j0) r = __builtin_fmaf(mf[0][j0], row[j0], r);
          vz = __builtin_fmaf(mf[1][j1], r, vz);
        }
        val = __builtin_fmaf(mf[2][j2], vz, val);
      }
      store(c, (double)val);
    } else {
      double val = 0.0;
#pragma unroll U2
      for (int j2 = 0; j2 < 4; ++j2) {
        double vz = 0.0;
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const T *row = cb + N * ((gi[1] + j1) + N * (gi[2] + j2));
          double r = 0.0;
#pragma unroll
          for (int j0 = 0; j0 < 4; ++j0) r = __builtin_fma(M[0][j0], row[j0], r);
          vz = __builtin_fma(M[1][j1], r, vz);
        }
        val = __builtin_fma(M[2][j2], vz, val);
      }
      store(c, val);
    }
  }
}

template <int MODE, int U2 = 4, typename Store>
__device__ __forceinline__ void grid_interp(const TreeView &t, const GridData &g, int leaf,
                                            double x, double y, double z, Store store) {
  const int64_t off = (int64_t)leaf * g.dof * g.n * g.n * g.n;
  if constexpr (MODE == kFp32)
    grid_interp_at<MODE, U2>(t.info[leaf], g.s32 + off, g.n, g.dof, x, y, z, store);
  else
    grid_interp_at<MODE, U2>(t.info[leaf], g.s + off, g.n, g.dof, x, y, z, store);
}

template <int MODE>
__device__ __forceinline__ void grid_vel(const TreeView &t, const GridData &g, double x,
                                         double y, double z, int scaled, double f, double &v0,
                                         double &v1, double &v2) {
  const int leaf = locate(t, x, y, z);
  grid_interp<MODE>(t, g, leaf, x, y, z, [&](int c, double v) {
    if (c == 0) v0 = v;
    else if (c == 1) v1 = v;
    else v2 = v;
  });
  if (scaled) {
    v0 = v0 * f;
    v1 = v1 * f;
    v2 = v2 * f;
  }
}

// The first RK2 stage at sample point m of leaf `own`: where the velocity
// leaf lf the point is located in has own's depth, the point is one of
// lf's sample points (node_of) and its velocity is that sample (the
// interpolation reproduces it up to rounding).  lf's samples start at
// s + slot * 3 N^3 (slot = lf for the grid's own array, the leaf's place
// for a staged copy); T: the sample precision the mode reads.  False when
// the point is no sample of lf.
template <typename T>
__device__ __forceinline__ bool grid_node_velocity(const TreeView &tv, const T *__restrict__ s,
                                                   int64_t slot, uint4 own, int m, int N, int lf,
                                                   int scaled, double f, double &v0, double &v1,
                                                   double &v2) {
  const int mm = node_of(own, tv.info[lf], m, N);
  if (mm < 0) return false;
  const int64_t N3 = (int64_t)N * N * N;
  const T *p = s + slot * 3 * N3 + mm;
  v0 = (double)p[0];
  v1 = (double)p[N3];
  v2 = (double)p[2 * N3];
  if (scaled) {
    v0 = v0 * f;
    v1 = v1 * f;
    v2 = v2 * f;
  }
  return true;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_grid_eval(TreeView t, GridData g,
                                                      double *__restrict__ pts, int64_t n,
                                                      int bc, int write_back,
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
  const int leaf = locate(t, x, y, z);
  double *o = out + i * g.dof;
  grid_interp<MODE>(t, g, leaf, x, y, z, [&](int c, double v) { o[c] = v; });
}

// SolveSemilagInSitu for grid leaves, fused: sample point -> num_rk_step
// RK2 substeps through the velocity grid tree (traj.inc:20-68) -> conc at
// the departure point -> out[leaf][c][sample].
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_grid_step(TreeView tv, GridData vel, TreeView tc,
                                                      GridData conc, GridStepArgs a,
                                                      double *__restrict__ out,
                                                      double *__restrict__ dep) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.npts) return;
  const int N = conc.n;
  const int64_t N3 = (int64_t)N * N * N;
  const int64_t leaf = i / N3;
  const int64_t m = i - leaf * N3;
  const uint4 inf = tc.info[leaf];
  const double len = 1.0 / (double)(1u << inf.w);
  double x = (double)inf.x * len + len * a.node[m % N];
  double y = (double)inf.y * len + len * a.node[(m / N) % N];
  double z = (double)inf.z * len + len * a.node[m / ((int64_t)N * N)];
  for (int s = 0; s < a.nsub; ++s) {
    if (a.bc == TBX_BC_PERIODIC) {
      x = wrap1(x);
      y = wrap1(y);
      z = wrap1(z);
    }
    double v0, v1, v2;
    bool nodal = false;
    if (s == 0 && a.nodal) {
      const int lf = locate(tv, x, y, z);
      if constexpr (MODE == kFp32)
        nodal = grid_node_velocity(tv, vel.s32, lf, inf, (int)m, N, lf, a.scaled, a.g[0], v0, v1,
                                   v2);
      else
        nodal = grid_node_velocity(tv, vel.s, lf, inf, (int)m, N, lf, a.scaled, a.g[0], v0, v1,
                                   v2);
    }
    if (!nodal) grid_vel<MODE>(tv, vel, x, y, z, a.scaled, a.g[2 * s], v0, v1, v2);
    double hx = x + a.htau * v0, hy = y + a.htau * v1, hz = z + a.htau * v2;  // traj.inc:36
    if (a.bc == TBX_BC_PERIODIC) {
      hx = wrap1(hx);
      hy = wrap1(hy);
      hz = wrap1(hz);
    }
    grid_vel<MODE>(tv, vel, hx, hy, hz, a.scaled, a.g[2 * s + 1], v0, v1, v2);
    x = x + a.tau * v0;  // traj.inc:41-42
    y = y + a.tau * v1;
    z = z + a.tau * v2;
  }
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
  const int cl = locate(tc, x, y, z);
  double *o = out + leaf * conc.dof * N3 + m;
  grid_interp<MODE>(tc, conc, cl, x, y, z, [&](int c, double v) { o[c * N3] = v; });
}

// The same step for a uniform tree of 8^3-sample leaves whose velocity and
// concentration trees share their leaves (the large tricubic workload): a
// workgroup takes the 8 sibling leaves of one parent (4,096 sample points,
// 512 threads with 8 points each) and stages their velocity and
// concentration samples in LDS (FP64: 128 KiB; FP32: 64 KiB).  An
// evaluation whose point lies in one of the 8 leaves gathers its 64-sample
// stencils from LDS; one outside them (a block face's upper sample points
// belong to the next block under half-open boxes, and departure points
// cross faces) gathers from global memory as k_grid_step does.  Same
// arithmetic, same results as k_grid_step.
constexpr int kBlkLeaves = 8, kBlkN3 = 512, kBlkThreads = 512;
constexpr int kBlkPts = kBlkLeaves * kBlkN3 / kBlkThreads;

template <int MODE>
__global__ __launch_bounds__(kBlkThreads) void k_grid_step_blk(TreeView tv, GridData vel,
                                                               TreeView tc, GridData conc,
                                                               GridStepArgs a,
                                                               double *__restrict__ out,
                                                               double *__restrict__ dep) {
  typedef typename std::conditional<MODE == kFp32, float, double>::type T;
  __shared__ T sv[kBlkLeaves * 3 * kBlkN3];
  __shared__ T sc[kBlkLeaves * kBlkN3];
  const int tid = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * kBlkLeaves;
  const T *gv, *gc;
  if constexpr (MODE == kFp32) {
    gv = vel.s32;
    gc = conc.s32;
  } else {
    gv = vel.s;
    gc = conc.s;
  }
  const T *g0 = gv;  // the velocity samples of every leaf
  gv += L0 * 3 * kBlkN3;
  gc += L0 * kBlkN3;
  for (int i = tid; i < kBlkLeaves * 3 * kBlkN3; i += kBlkThreads) sv[i] = gv[i];
  for (int i = tid; i < kBlkLeaves * kBlkN3; i += kBlkThreads) sc[i] = gc[i];
  __syncthreads();
  constexpr int N = 8;
  auto velocity = [&](double x, double y, double z, double f, double &v0, double &v1,
                      double &v2) {
    const int leaf = locate(tv, x, y, z);
    const int64_t lb = (int64_t)leaf - L0;
    auto st = [&](int c, double v) {
      if (c == 0) v0 = v;
      else if (c == 1) v1 = v;
      else v2 = v;
    };
    if (lb >= 0 && lb < kBlkLeaves)
      grid_interp_at<MODE, 1>(tv.info[leaf], sv + lb * 3 * kBlkN3, N, 3, x, y, z, st);
    else
      grid_interp<MODE, 1>(tv, vel, leaf, x, y, z, st);
    if (a.scaled) {
      v0 = v0 * f;
      v1 = v1 * f;
      v2 = v2 * f;
    }
  };
#pragma unroll 1
  for (int k = 0; k < kBlkPts; ++k) {
    const int li = k * kBlkThreads + tid;
    const int64_t leaf = L0 + li / kBlkN3;
    const int m = li % kBlkN3;
    const int64_t i = leaf * kBlkN3 + m;
    const uint4 inf = tc.info[leaf];
    const double len = 1.0 / (double)(1u << inf.w);
    double x = (double)inf.x * len + len * a.node[m % N];
    double y = (double)inf.y * len + len * a.node[(m / N) % N];
    double z = (double)inf.z * len + len * a.node[m / (N * N)];
    for (int s = 0; s < a.nsub; ++s) {
      if (a.bc == TBX_BC_PERIODIC) {
        x = wrap1(x);
        y = wrap1(y);
        z = wrap1(z);
      }
      double v0, v1, v2;
      bool nodal = false;
      if (s == 0 && a.nodal) {  // the sample point's own velocity sample
        const int lf = locate(tv, x, y, z);
        const int64_t lb = (int64_t)lf - L0;
        if (lb >= 0 && lb < kBlkLeaves)
          nodal = grid_node_velocity(tv, (const T *)sv, lb, inf, m, N, lf, a.scaled, a.g[0], v0,
                                     v1, v2);
        else
          nodal = grid_node_velocity(tv, g0, lf, inf, m, N, lf, a.scaled, a.g[0], v0, v1, v2);
      }
      if (!nodal) velocity(x, y, z, a.g[2 * s], v0, v1, v2);
      double hx = x + a.htau * v0, hy = y + a.htau * v1, hz = z + a.htau * v2;  // traj.inc:36
      if (a.bc == TBX_BC_PERIODIC) {
        hx = wrap1(hx);
        hy = wrap1(hy);
        hz = wrap1(hz);
      }
      velocity(hx, hy, hz, a.g[2 * s + 1], v0, v1, v2);
      x = x + a.tau * v0;  // traj.inc:41-42
      y = y + a.tau * v1;
      z = z + a.tau * v2;
    }
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
    const int cl = locate(tc, x, y, z);
    const int64_t lb = (int64_t)cl - L0;
    double *o = out + leaf * kBlkN3 + m;
    auto st = [&](int, double v) { *o = v; };
    if (lb >= 0 && lb < kBlkLeaves)
      grid_interp_at<MODE, 1>(tc.info[cl], sc + lb * kBlkN3, N, 1, x, y, z, st);
    else
      grid_interp<MODE, 1>(tc, conc, cl, x, y, z, st);
  }
}

template <typename F>
hipError_t with_grid_mode(int mode, F f) {
  switch (mode) {
    case kFma64:
      return f(std::integral_constant<int, kFma64>{});
    case kExact64:
      return f(std::integral_constant<int, kExact64>{});
    case kFp32:
      return f(std::integral_constant<int, kFp32>{});
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_grid_eval(int mode, const TreeView &t, const double *s, const float *s32,
                            int n, int dof, double *pts, int64_t npts, int bc, int write_back,
                            double *out, hipStream_t st) {
  if (npts <= 0) return hipSuccess;
  const GridData g{s, s32, n, dof};
  return with_grid_mode(mode, [&](auto md) {
    hipLaunchKernelGGL(k_grid_eval<decltype(md)::value>,
                       dim3((unsigned)((npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, t, g,
                       pts, npts, bc, write_back, out);
    return hipGetLastError();
  });
}

hipError_t launch_grid_step(int mode, const TreeView &tv, const double *vs, const float *vs32,
                            int vn, const TreeView &tc, const double *cs, const float *cs32,
                            int cn, int cdof, const GridStepArgs &a, double *out, double *dep,
                            hipStream_t st) {
  if (a.npts <= 0) return hipSuccess;
  const GridData vel{vs, vs32, vn, 3}, conc{cs, cs32, cn, cdof};
  if (a.blk) {
    if (vn != 8 || cn != 8 || cdof != 1 || a.npts % (kBlkLeaves * kBlkN3))
      return hipErrorInvalidValue;
    const int64_t nb = a.npts / (kBlkLeaves * kBlkN3);
    return with_grid_mode(mode, [&](auto md) {
      hipLaunchKernelGGL(k_grid_step_blk<decltype(md)::value>, dim3((unsigned)nb),
                         dim3(kBlkThreads), 0, st, tv, vel, tc, conc, a, out, dep);
      return hipGetLastError();
    });
  }
  return with_grid_mode(mode, [&](auto md) {
    hipLaunchKernelGGL(k_grid_step<decltype(md)::value>,
                       dim3((unsigned)((a.npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, tv,
                       vel, tc, conc, a, out, dep);
    return hipGetLastError();
  });
}

}  // namespace tbx
