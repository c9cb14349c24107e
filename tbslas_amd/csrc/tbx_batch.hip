// Many small problems in one persistent launch: every problem is one
// velocity tree and one single-component concentration tree (its own
// leaves, degree and time step), all advanced by one in-situ step in FP32
// in a single kernel.
//
// Work items are (problem, run of LPI leaves), numbered problem by problem;
// persistent workgroups claim them from a global counter until none are
// left.  An item computes the departure points of its leaves' grid points
// (one RK2 substep through the problem's velocity tree, FP32 contraction,
// points regrouped by leaf in LDS as in the fused step), the concentration
// there (FP32), keeps the values in LDS and refits them at once
// (c'[m] = sum_p M[p][m] v[p], the per-leaf product of SetTreeGridValues,
// tree_utils.h:534-546, summed in FP64), so no value buffer exists.  Coefficients
// are FP32 and ping-pong between two buffers; the step of parity k reads
// buffer k and writes buffer 1 - k.  With vnode (the stationary velocity's
// values at its own grid nodes, expanded once when the batch is made) the
// first RK2 stage at the arrival nodes reads them, as the single-tree step
// does.
#include <hip/hip_runtime.h>

#include "tbx_device.h"

namespace tbx {
namespace {

__device__ __forceinline__ int find_problem(const BatchProblem *pb, int nprob, int64_t item) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {  // last problem whose first item is <= item
    const int mid = (lo + hi + 1) >> 1;
    if (pb[mid].item0 <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One item of degree D - 1: leaves [leaf0, leaf0 + nleaf) of problem p.
template <int D, int BLOCK, int LPI>
__device__ __forceinline__ void batch_item(const BatchProblem &pb, int parity, int64_t leaf0,
                                           int nleaf, float *vals, RegroupLds<BLOCK, 6> &sh) {
  constexpr int P = D * D * D;
  constexpr int NC = D * (D + 1) * (D + 2) / 6;
  VelSource vs{};
  vs.t[0] = pb.vel;
  vs.kind = kVelTree;
  vs.ntree = 1;
  const TreeView &conc = pb.conc[parity];
  const int total = nleaf * P;
  for (int start = 0; start < total; start += BLOCK) {
    int orig = (int)threadIdx.x;
    const int idx = start + orig < total ? start + orig : total - 1;
    const int64_t leaf = leaf0 + idx / P;
    const int m = idx % P;
    const uint4 inf = conc.info[leaf];
    const double len = 1.0 / (double)(1u << inf.w);
    double st[6];
    st[0] = (double)inf.x * len + len * pb.node[m % D];
    st[1] = (double)inf.y * len + len * pb.node[(m / D) % D];
    st[2] = (double)inf.z * len + len * pb.node[m / (D * D)];
    if (pb.bc == TBX_BC_PERIODIC) {
      st[0] = wrap1(st[0]);
      st[1] = wrap1(st[1]);
      st[2] = wrap1(st[2]);
    }
    int lf = locate(pb.vel, st[0], st[1], st[2]);
    double v0, v1, v2;
    // stage 1 from the nodal values; the regroup and contraction only in
    // workgroups holding a node of another depth's leaf, whose points then
    // take the nodal value again where they have one (a point's result does
    // not depend on the workgroup it falls in)
    bool nodal = false;
    if (pb.vnode) {
      const bool got = nodal_velocity<D>(pb.vnode, pb.vel, inf, m, lf, P, 0, 1.0, v0, v1, v2);
      nodal = !__syncthreads_or(!got);
    }
    if (!nodal) {
      {
        double p3[3] = {st[0], st[1], st[2]};
        regroup(sh, orig, lf, p3);
        st[0] = p3[0];
        st[1] = p3[1];
        st[2] = p3[2];
      }
      eval_velocity<D, kFp32, kVelTree>(vs, 1, 0.0, 0, 1.0, pb.bc, st[0], st[1], st[2], v0, v1,
                                        v2, lf);
      if (pb.vnode) {  // the arrival node this thread now holds
        const int jdx = start + orig < total ? start + orig : total - 1;
        const int64_t jl = leaf0 + jdx / P;
        nodal_velocity<D>(pb.vnode, pb.vel, conc.info[jl], jdx % P, lf, P, 0, 1.0, v0, v1, v2);
      }
    }
    st[3] = st[0] + pb.htau * v0;  // traj.inc:36
    st[4] = st[1] + pb.htau * v1;
    st[5] = st[2] + pb.htau * v2;
    if (pb.bc == TBX_BC_PERIODIC) {
      st[3] = wrap1(st[3]);
      st[4] = wrap1(st[4]);
      st[5] = wrap1(st[5]);
    }
    lf = locate(pb.vel, st[3], st[4], st[5]);
    regroup(sh, orig, lf, st);
    eval_velocity<D, kFp32, kVelTree>(vs, 2, 0.0, 0, 1.0, pb.bc, st[3], st[4], st[5], v0, v1,
                                      v2, lf);
    st[0] = st[0] + pb.tau * v0;  // traj.inc:41-42
    st[1] = st[1] + pb.tau * v1;
    st[2] = st[2] + pb.tau * v2;
    if (pb.bc == TBX_BC_PERIODIC) {
      st[0] = wrap1(st[0]);
      st[1] = wrap1(st[1]);
      st[2] = wrap1(st[2]);
    }
    lf = locate(conc, st[0], st[1], st[2]);
    double p3[3] = {st[0], st[1], st[2]};
    regroup(sh, orig, lf, p3);
    const bool live = start + orig < total;
    eval_dof<D, kFp32>(
        conc, p3[0], p3[1], p3[2],
        [&](int, double v) {
          if (live) vals[start + orig] = (float)v;
        },
        lf);
  }
  __syncthreads();
  // refit of the item's leaves: c'[b][m] = sum_p M[p][m] v[b][p], summed in
  // FP64 (a sum of P terms; the contraction above is the FP32 part).  One
  // thread per m does every leaf of the item, so each entry of M (up to
  // 3 MB, streamed from L2) is loaded once per item, not once per leaf.
  // The loop has no branch, so the unrolled loads of M stay in flight
  // together; the sums of the item's missing leaves (b >= nleaf) run on
  // whatever the LDS holds there and are not stored.
  float *out = pb.out32[parity];
  const double *__restrict__ M = pb.M;
  for (int mm = threadIdx.x; mm < NC; mm += BLOCK) {
    double acc[LPI];
#pragma unroll
    for (int b = 0; b < LPI; ++b) acc[b] = 0.0;
#pragma unroll 16
    for (int p = 0; p < P; ++p) {
      const double m = M[p * NC + mm];
#pragma unroll
      for (int b = 0; b < LPI; ++b) acc[b] = __builtin_fma(m, (double)vals[b * P + p], acc[b]);
    }
#pragma unroll
    for (int b = 0; b < LPI; ++b)
      if (b < nleaf) out[(leaf0 + b) * NC + mm] = (float)acc[b];
  }
  __syncthreads();
}

template <int BLOCK, int LPI>
__global__ __launch_bounds__(BLOCK) void k_batch(const BatchProblem *__restrict__ pb, int nprob,
                                                 int parity, int64_t total_items,
                                                 unsigned long long *counter) {
  __shared__ RegroupLds<BLOCK, 6> sh;
  __shared__ float vals[LPI * 11 * 11 * 11];
  __shared__ long long s_item;
  while (true) {
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();
    if (item >= total_items) break;  // uniform over the workgroup
    const int p = find_problem(pb, nprob, item);
    const BatchProblem &b = pb[p];
    const int64_t leaf0 = (item - b.item0) * LPI;
    const int nleaf = (int)(b.L - leaf0 < LPI ? b.L - leaf0 : LPI);
    switch (b.q) {
      case 6:
        batch_item<7, BLOCK, LPI>(b, parity, leaf0, nleaf, vals, sh);
        break;
      case 8:
        batch_item<9, BLOCK, LPI>(b, parity, leaf0, nleaf, vals, sh);
        break;
      default:
        batch_item<11, BLOCK, LPI>(b, parity, leaf0, nleaf, vals, sh);
        break;
    }
  }
}

__global__ void k_batch_reset(unsigned long long *counter) { *counter = 0ull; }

template <int BLOCK, int LPI>
hipError_t launch_cfg(const BatchProblem *pb, int nprob, int parity, int64_t total_items,
                      unsigned long long *counter, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, counter);
  hipLaunchKernelGGL((k_batch<BLOCK, LPI>), dim3(grid), dim3(BLOCK), 0, s, pb, nprob, parity,
                     total_items, counter);
  return hipGetLastError();
}

template <int BLOCK, int LPI>
int occupancy() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch<BLOCK, LPI>, BLOCK, 0) !=
      hipSuccess)
    return 1;
  return n > 0 ? n : 1;
}

}  // namespace

int batch_blocks_per_cu(int cfg) {
  switch (cfg) {
    case 0: return occupancy<256, 1>();
    case 1: return occupancy<256, 2>();
    case 2: return occupancy<512, 1>();
    case 3: return occupancy<512, 2>();
    case 4: return occupancy<512, 4>();
    default: return occupancy<1024, 1>();
  }
}

hipError_t launch_batch(int cfg, const BatchProblem *pb, int nprob, int parity,
                        int64_t total_items, unsigned long long *counter, int grid,
                        hipStream_t s) {
  switch (cfg) {
    case 0: return launch_cfg<256, 1>(pb, nprob, parity, total_items, counter, grid, s);
    case 1: return launch_cfg<256, 2>(pb, nprob, parity, total_items, counter, grid, s);
    case 2: return launch_cfg<512, 1>(pb, nprob, parity, total_items, counter, grid, s);
    case 3: return launch_cfg<512, 2>(pb, nprob, parity, total_items, counter, grid, s);
    case 4: return launch_cfg<512, 4>(pb, nprob, parity, total_items, counter, grid, s);
    default: return launch_cfg<1024, 1>(pb, nprob, parity, total_items, counter, grid, s);
  }
}

}  // namespace tbx
