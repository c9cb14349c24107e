// Kernels of the on-device adaptive refinement (the RefineTree + Balance21
// roles, advection.cpp:305-314, tree_utils.h:117-123):
//
//   k_adapt_tail      the refinement criterion per leaf: the largest over
//                     components of sum |c_ijk| over the top shell
//                     i + j + k = q (the same sums, in the same order, as
//                     tbx_tail_norm);
//   k_leaf_resample   a new leaf's values at its stretched grid nodes from
//                     the polynomial of one given old leaf (prolongation:
//                     the parent of a child) or of the old child holding
//                     each node (restriction: the 8 children of a new
//                     parent); the refit of these values gives the new
//                     coefficients (exact for prolongation: a polynomial of
//                     total degree <= q restricted to a sub-box is one);
//   k_balance_probe   2:1 balance: every leaf looks at the centres of its 26
//                     same-size neighbour cells and marks the leaf holding
//                     one if that leaf is more than one level coarser.
#include <hip/hip_runtime.h>

#include "tbx_device.h"

namespace tbx {
namespace {

constexpr int kAdaptBlock = 256;

__device__ __forceinline__ int tri_base(int D, int i) {
  const int n = D - i;
  return (D * (D + 1) * (D + 2) - n * (n + 1) * (n + 2)) / 6;
}

__global__ __launch_bounds__(kAdaptBlock) void k_adapt_tail(const double *__restrict__ coeff,
                                                            int64_t L, int dof, int q,
                                                            double *__restrict__ tail) {
  const int D = q + 1, nc = D * (D + 1) * (D + 2) / 6;
  for (int64_t j = (int64_t)blockIdx.x * kAdaptBlock + threadIdx.x; j < L;
       j += (int64_t)gridDim.x * kAdaptBlock) {
  double best = 0.0;
  for (int l = 0; l < dof; ++l) {
    const double *c = coeff + (j * dof + l) * nc;
    double s = 0.0;
    for (int i = 0; i <= q; ++i) {
      const int bi = tri_base(D, i), n = D - i;
      for (int jj = 0; i + jj <= q; ++jj)  // k = q - i - jj, the last of row jj
        s += fabs(c[bi + jj * n - jj * (jj - 1) / 2 + (q - i - jj)]);
    }
    best = fmax(best, s);
  }
  tail[j] = best;
  }
}

// One thread per (new leaf n, node p): the node of leaf n (anchor and depth
// in info_new[n]), the source leaf src[n] (restriction: src[n] + the
// child octant of the node, the upper child from the middle on), its
// coordinates in the source leaf clamped to [-1, 1] (a node on a shared
// face evaluates the source polynomial there, not the zero outside), and
// every component of the source's expansion at it.
struct Nodes {
  double v[TBX_MAX_Q + 1];
};

__global__ __launch_bounds__(kAdaptBlock) void k_leaf_resample(
    TreeView src, const int32_t *__restrict__ src_leaf, const uint4 *__restrict__ info_new,
    int restrict_children, int64_t nnew, int q, Nodes nd, double *__restrict__ vals) {
  const int D = q + 1, P = D * D * D, nc = D * (D + 1) * (D + 2) / 6;
  for (int64_t i = (int64_t)blockIdx.x * kAdaptBlock + threadIdx.x; i < nnew * P;
       i += (int64_t)gridDim.x * kAdaptBlock) {
  const int64_t n = i / P;
  const int p = (int)(i - n * P);
  const double *node = nd.v;
  const int a = p % D, b = (p / D) % D, c = p / (D * D);
  const uint4 nf = info_new[n];
  const double len = 1.0 / (double)(1u << nf.w);
  const double x = (double)nf.x * len + len * node[a];
  const double y = (double)nf.y * len + len * node[b];
  const double z = (double)nf.z * len + len * node[c];
  int s = src_leaf[n];
  if (restrict_children)
    s += (node[a] >= 0.5 ? 1 : 0) + (node[b] >= 0.5 ? 2 : 0) + (node[c] >= 0.5 ? 4 : 0);
  const uint4 sf = src.info[s];
  const double ss = (double)(1u << sf.w), sl = 1.0 / ss;
  double ex = ((x - (double)sf.x * sl) * 2.0) * ss - 1.0;
  double ey = ((y - (double)sf.y * sl) * 2.0) * ss - 1.0;
  double ez = ((z - (double)sf.z * sl) * 2.0) * ss - 1.0;
  ex = fmin(fmax(ex, -1.0), 1.0);
  ey = fmin(fmax(ey, -1.0), 1.0);
  ez = fmin(fmax(ez, -1.0), 1.0);
  double tx[TBX_MAX_Q + 1], ty[TBX_MAX_Q + 1], tz[TBX_MAX_Q + 1];
  cheb_poly<TBX_MAX_Q + 1>(ex, tx);
  cheb_poly<TBX_MAX_Q + 1>(ey, ty);
  cheb_poly<TBX_MAX_Q + 1>(ez, tz);
  for (int l = 0; l < src.dof; ++l) {
    const double *cf = src.coeff + ((int64_t)s * src.dof + l) * nc;
    double u = 0.0;
    int m = 0;
    for (int ii = 0; ii < D; ++ii) {
      double uj = 0.0;
      for (int jj = 0; ii + jj < D; ++jj) {
        double uk = 0.0;
        for (int kk = 0; ii + jj + kk < D; ++kk, ++m) uk = __builtin_fma(tx[kk], cf[m], uk);
        uj = __builtin_fma(ty[jj], uk, uj);
      }
      u = __builtin_fma(tz[ii], uj, u);
    }
    vals[(n * src.dof + l) * P + p] = u;
  }
  }
}

__global__ __launch_bounds__(kAdaptBlock) void k_balance_probe(TreeView t, int bc,
                                                               uint8_t *__restrict__ split) {
  for (int64_t m = (int64_t)blockIdx.x * kAdaptBlock + threadIdx.x; m < t.L;
       m += (int64_t)gridDim.x * kAdaptBlock) {
  const uint4 inf = t.info[m];
  const double len = 1.0 / (double)(1u << inf.w);
  const double cx = ((double)inf.x + 0.5) * len, cy = ((double)inf.y + 0.5) * len,
               cz = ((double)inf.z + 0.5) * len;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (!dx && !dy && !dz) continue;
        double px = cx + dx * len, py = cy + dy * len, pz = cz + dz * len;
        if (bc == TBX_BC_PERIODIC) {
          px = wrap1(px);
          py = wrap1(py);
          pz = wrap1(pz);
        } else if (px < 0.0 || px >= 1.0 || py < 0.0 || py >= 1.0 || pz < 0.0 || pz >= 1.0) {
          continue;
        }
        const int l = locate(t, px, py, pz);
        if ((int)t.info[l].w + 1 < (int)inf.w) split[l] = 1;
      }
  }
}

// dst row r = src row map[r] for map[r] >= 0 (rows of rowlen doubles).
__global__ __launch_bounds__(kAdaptBlock) void k_gather_rows(const double *__restrict__ src,
                                                             const int64_t *__restrict__ map,
                                                             int64_t nrows, int rowlen,
                                                             double *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kAdaptBlock + threadIdx.x; i < nrows * rowlen;
       i += (int64_t)gridDim.x * kAdaptBlock) {
    const int64_t r = i / rowlen;
    const int64_t m = map[r];
    if (m >= 0) dst[i] = src[m * rowlen + (i - r * rowlen)];
  }
}

// A dispatch holds at most 2^32 - 1 work-items: the element-wise kernels
// run at most 2^20 workgroups and stride over the rest.
inline unsigned blocks_for(int64_t n) {
  const int64_t b = (n + kAdaptBlock - 1) / kAdaptBlock;
  return (unsigned)(b < ((int64_t)1 << 20) ? b : ((int64_t)1 << 20));
}

}  // namespace

hipError_t launch_adapt_tail(const double *coeff, int64_t L, int dof, int q, double *tail,
                             hipStream_t s) {
  if (L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adapt_tail, dim3(blocks_for(L)), dim3(kAdaptBlock), 0, s, coeff, L, dof,
                     q, tail);
  return hipGetLastError();
}

hipError_t launch_leaf_resample(const TreeView &src, const int32_t *src_leaf,
                                const uint4 *info_new, int restrict_children, int64_t nnew,
                                const double *node, double *vals, hipStream_t s) {
  const int D = src.q + 1;
  const int64_t n = nnew * D * D * D;
  if (n <= 0) return hipSuccess;
  Nodes nd{};
  for (int a = 0; a < D; ++a) nd.v[a] = node[a];
  hipLaunchKernelGGL(k_leaf_resample, dim3(blocks_for(n)), dim3(kAdaptBlock), 0, s, src,
                     src_leaf, info_new, restrict_children, nnew, src.q, nd, vals);
  return hipGetLastError();
}

hipError_t launch_balance_probe(const TreeView &t, int bc, uint8_t *split, hipStream_t s) {
  if (t.L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_balance_probe, dim3(blocks_for(t.L)), dim3(kAdaptBlock), 0, s, t, bc,
                     split);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const double *src, const int64_t *map, int64_t nrows, int rowlen,
                              double *dst, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(nrows * rowlen)), dim3(kAdaptBlock), 0, s,
                     src, map, nrows, rowlen, dst);
  return hipGetLastError();
}

}  // namespace tbx
