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

// All dof components of leaf `leaf` at (x, y, z) through store(c, value).
template <int MODE, typename Store>
__device__ __forceinline__ void grid_interp(const TreeView &t, const GridData &g, int leaf,
                                            double x, double y, double z, Store store) {
  const uint4 inf = t.info[leaf];
  const double s = (double)(1u << inf.w), len = 1.0 / s;
  double u[3] = {(x - (double)inf.x * len) * s, (y - (double)inf.y * len) * s,
                 (z - (double)inf.z * len) * s};
  const int N = g.n;
  const int64_t N3 = (int64_t)N * N * N;
  if (!(u[0] >= 0 && u[0] <= 1.0 && u[1] >= 0 && u[1] <= 1.0 && u[2] >= 0 && u[2] <= 1.0)) {
    for (int c = 0; c < g.dof; ++c) store(c, 0.0);  // tree_functor.h:106-114
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
  const int64_t base = (int64_t)leaf * g.dof * N3 + gi[0];
  for (int c = 0; c < g.dof; ++c) {
    const int64_t cb = base + c * N3;
    if constexpr (MODE == kExact64) {
      double val = 0;  // tree_functor.h:137-151
#pragma unroll
      for (int j2 = 0; j2 < 4; ++j2)
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const double M1M2 = M[1][j1] * M[2][j2];
          const double *row = g.s + cb + N * ((gi[1] + j1) + (int64_t)N * (gi[2] + j2));
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
#pragma unroll
      for (int j2 = 0; j2 < 4; ++j2) {
        float vz = 0.f;
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const float *row = g.s32 + cb + N * ((gi[1] + j1) + (int64_t)N * (gi[2] + j2));
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
#pragma unroll
      for (int j2 = 0; j2 < 4; ++j2) {
        double vz = 0.0;
#pragma unroll
        for (int j1 = 0; j1 < 4; ++j1) {
          const double *row = g.s + cb + N * ((gi[1] + j1) + (int64_t)N * (gi[2] + j2));
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
    grid_vel<MODE>(tv, vel, x, y, z, a.scaled, a.g[2 * s], v0, v1, v2);
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
  return with_grid_mode(mode, [&](auto md) {
    hipLaunchKernelGGL(k_grid_step<decltype(md)::value>,
                       dim3((unsigned)((a.npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, tv,
                       vel, tc, conc, a, out, dep);
    return hipGetLastError();
  });
}

}  // namespace tbx
