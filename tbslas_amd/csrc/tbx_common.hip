// Degree-independent kernels of libtbx: point location, grid points and the
// FP32 coefficient mirror.
#include <hip/hip_runtime.h>

#include "tbx_device.h"

namespace tbx {
namespace {

// EvalTree's wrap + keys + leaf assignment (tree_functor.h:442-489,
// :190-198) without evaluation; pts are left untouched.
__global__ __launch_bounds__(kBlock) void k_locate(TreeView t,
                                                   const double *__restrict__ pts,
                                                   int64_t n, int bc,
                                                   int32_t *__restrict__ leaf) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double x = pts[3 * i + 0], y = pts[3 * i + 1], z = pts[3 * i + 2];
  if (bc == TBX_BC_PERIODIC) {
    x = wrap1(x);
    y = wrap1(y);
    z = wrap1(z);
  }
  leaf[i] = locate(t, x, y, z);
}

struct NodeTable {
  double node[TBX_MAX_Q + 1];
  int d;
};

// CollectChebTreeGridPoints (tree_utils.h:478-491): corner + len * node,
// node index x fastest (cheb.h:60-67).
__global__ __launch_bounds__(kBlock) void k_grid_points(TreeView t,
                                                        NodeTable nt,
                                                        double *__restrict__ out) {
  const int d = nt.d, P = d * d * d;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= t.L * P) return;
  const int64_t leaf = i / P;
  const int m = (int)(i - leaf * P);
  const uint4 inf = t.info[leaf];
  const double len = 1.0 / (double)(1u << inf.w);
  out[3 * i + 0] = (double)inf.x * len + len * nt.node[m % d];
  out[3 * i + 1] = (double)inf.y * len + len * nt.node[(m / d) % d];
  out[3 * i + 2] = (double)inf.z * len + len * nt.node[m / (d * d)];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(const double *__restrict__ in,
                                                   float *__restrict__ out,
                                                   int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

inline unsigned grid_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

}  // namespace

hipError_t launch_locate(const TreeView &t, const double *pts, int64_t n,
                         int bc, int32_t *leaf, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_locate, dim3(grid_for(n)), dim3(kBlock), 0, s, t, pts,
                     n, bc, leaf);
  return hipGetLastError();
}

hipError_t launch_grid_points(const TreeView &t, const double *node, int d,
                              double *out, hipStream_t s) {
  const int64_t n = t.L * (int64_t)d * d * d;
  if (n <= 0) return hipSuccess;
  NodeTable nt{};
  for (int i = 0; i < d; ++i) nt.node[i] = node[i];
  nt.d = d;
  hipLaunchKernelGGL(k_grid_points, dim3(grid_for(n)), dim3(kBlock), 0, s, t,
                     nt, out);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double *in, float *out, int64_t n,
                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out,
                     n);
  return hipGetLastError();
}

}  // namespace tbx
