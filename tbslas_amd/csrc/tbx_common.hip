// Degree-independent kernels of libtbx: point location, grid points and the
// FP32 coefficient mirror.
#include <hip/hip_runtime.h>

#include "tbx_device.h"

namespace tbx {
namespace {

// EvalTree's wrap + keys + leaf assignment (tree_functor.h:442-489,
// :190-198) without evaluation; pts are left untouched.  A block takes
// kLocatePts x kBlock consecutive points: their coordinates (24 bytes a
// point) come in through LDS with 16-byte loads, consecutive per lane, and
// each thread then locates kLocatePts points a block apart, all their
// directory loads issued before the first search.  Coordinates and leaf ids
// are streamed (non-temporal), which leaves the caches to the directory and
// the keys.
constexpr int kLocatePts = 4;
typedef double stage2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(kBlock) void k_locate(TreeView t,
                                                   const double *__restrict__ pts,
                                                   int64_t n, int bc,
                                                   int32_t *__restrict__ leaf) {
  constexpr int NP = kBlock * kLocatePts;  // points of the block
  __shared__ stage2 stage[NP * 3 / 2];     // NP x 3 doubles
  const int64_t p0 = (int64_t)blockIdx.x * NP;
  const int64_t np = n - p0 < NP ? n - p0 : NP;
  const int tid = threadIdx.x;
  if (np == NP && (reinterpret_cast<uintptr_t>(pts) & 15u) == 0) {
    // the block's 3 NP doubles as NP * 3 / 2 16-byte pieces
    const stage2 *src = reinterpret_cast<const stage2 *>(pts + 3 * p0);
#pragma unroll
    for (int k = 0; k < NP * 3 / 2 / kBlock; ++k)
      stage[k * kBlock + tid] = __builtin_nontemporal_load(src + k * kBlock + tid);
  } else {
    double *dst = reinterpret_cast<double *>(stage);
    for (int k = tid; k < 3 * np; k += kBlock) dst[k] = pts[3 * p0 + k];
  }
  __syncthreads();
  const double *c = reinterpret_cast<const double *>(stage);
  uint64_t key[kLocatePts];
  int lo[kLocatePts], hi[kLocatePts];
#pragma unroll
  for (int k = 0; k < kLocatePts; ++k) {
    // lanes past the end locate the block's first point and store nothing
    const int j = k * kBlock + tid < np ? k * kBlock + tid : 0;
    double a = c[3 * j + 0], b = c[3 * j + 1], z = c[3 * j + 2];
    if (bc == TBX_BC_PERIODIC) {
      a = wrap1(a);
      b = wrap1(b);
      z = wrap1(z);
    }
    key[k] = point_key(a, b, z);
    lo[k] = dir_first(t, key[k]);
    hi[k] = dir_last(t, key[k]);
  }
#pragma unroll
  for (int k = 0; k < kLocatePts; ++k) {
    const int j = k * kBlock + tid;
    const int r = dir_search(t, key[k], lo[k], hi[k]);
    if (j < np) __builtin_nontemporal_store(r, leaf + p0 + j);
  }
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
  const int64_t per = (int64_t)kBlock * kLocatePts;
  hipLaunchKernelGGL(k_locate, dim3((unsigned)((n + per - 1) / per)), dim3(kBlock), 0, s, t,
                     pts, n, bc, leaf);
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
