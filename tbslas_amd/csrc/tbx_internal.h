// Internal interface between the host side of libtbx (tbx_host.cpp) and the
// HIP kernel translation units.  Not installed; the public ABI is include/tbx.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbx.h"

namespace tbx {

// Morton keys resolve TBX_MAX_DEPTH bits per axis: 3*21 = 63-bit keys,
// z above y above x at every level (the child order PVFMM's MortonId uses).
constexpr int kKeyBits = TBX_MAX_DEPTH;
constexpr double kKeyScale = 2097152.0;  // 2^21

// kFp32h: FP32 tail of the contraction (total degree > kHeadDeg) plus the
// head (total degree <= kHeadDeg) in FP64.
enum Mode : int { kFma64 = 0, kExact64 = 1, kFp32 = 2, kFp32h = 3 };
constexpr int kHeadDeg = 1;

// Device view of one tree.  All pointers are device memory owned by the
// tree handle.
struct TreeView {
  const uint64_t *key;   // [L] Morton key of each leaf's anchor at 21 bits
  const uint4 *info;     // [L] anchor x, y, z (units 2^-depth), depth
  const int32_t *dir;    // [8^g + 1] leaf holding the first key of each
                         // level-g cell; dir[8^g] = L - 1
  const double *coeff;   // [L][dof][nc]
  const float *coeff32;  // [L][dof][nc] FP32 mirror (TBX_FP32 paths)
  int64_t L;
  int32_t dof;
  int32_t nc;
  int32_t q;
  int32_t dir_shift;     // 3 * (21 - g)
  int32_t dir_exact;     // no leaf deeper than g: dir[c] is the leaf
                         // that holds all of cell c
};

// A velocity functor of the kinds the reference's trajectory templates take.
//  kVelTree   NodeFieldFunctor(t[0]), optionally scaled by g(t);
//  kVelExtrap field_fn = NodeFieldFunctor(t[0]) for the first RK stage and
//             extrap_fn = FieldExtrapFunctor(t[1], t[0]) for the second
//             (traj.inc:70-95, tree_extrap_functor.h:47-78): t[0] current,
//             t[1] previous;
//  kVelSet    FieldSetFunctor(t[0..3], tt[0..3]) (tree_set_functor.h:49-79).
enum VelKind : int { kVelTree = 0, kVelExtrap = 1, kVelSet = 2 };

struct VelSource {
  TreeView t[4];
  double tt[4];
  int32_t kind;
  int32_t ntree;
};

constexpr int kMaxSub = 16;  // RK substeps fused into one step launch

// One IntegrateRK2 substep: stage factors of the time modulation.
struct SubArgs {
  double tau;   // (tfinal - tinit) / num_rk_step   (traj.inc:55)
  double htau;  // 0.5 * tau                         (traj.inc:36)
  double g1;    // g(t)            (1 when steady)
  double g2;    // g(t + 0.5 tau)
  double ts1;   // stage times t and t + 0.5 tau (traj.inc:33, 40)
  double ts2;
  int scaled;   // multiply by g1/g2
  int bc;
};

// Launch shapes of the fused step kernel (workgroup sizes below).  0-3 run
// the points in arrival order; 4-7 regroup each workgroup's points by leaf
// before every evaluation.  Shapes 0 and 6 (the default) are built for
// every degree, the others for the benchmark degrees 8 and 14 only; an
// unbuilt shape falls back to 6.
constexpr int kStepShapes = 11;
constexpr int kStepShapeBlock[kStepShapes] = {256, 256, 1024, 512, 256, 512, 1024, 512, 1024, 512, 768};
constexpr int kDefaultStepShape = 6;
// The step shape option's default, -1: the shape measured fastest for the
// step's arithmetic mode and degree.  At q = 14 (adaptive tree of 36,324
// leaves, 122,593,500 points, MI355X) the FP32 step kernel runs in 33.6 ms
// with shape 8 (1024 threads, 64 VGPRs: two workgroups per CU) against
// 42.6 ms with shape 6, and FP32 + FP64 head in 38.6 ms with shape 10 (768
// threads, 80 VGPRs, two per CU) against 41.0 ms with shape 9 (512
// threads, whose smaller regrouping leaves 1.49 leaf passes per wavefront)
// and 44.9 ms with shape 6; FP64 keeps shape 6 (73.3 ms against 92-102 ms,
// where the smaller budgets spill).
constexpr int kAutoStepShape = -1;
inline int auto_step_shape(int mode, int q) {
  if (q == 14 && mode == kFp32) return 8;
  if (q == 14 && mode == kFp32h) return 10;
  return kDefaultStepShape;
}
inline bool step_shape_built(int shape, int q) {
  return shape == 0 || shape == 6 || q == 8 || q == 14;
}
// Workgroup size of the matrix-core step (many-component concentrations,
// every degree): shape 5's 512 threads.
constexpr int kMfmaStepShape = 5;
// Workgroup of the matrix-core tracer step: FP64 at 512 threads (two waves
// per SIMD: its tiles need 256 registers), FP32 at 256 threads with a
// budget of three waves per SIMD (166 registers), three workgroups per CU.
constexpr int mfma_step_block(int mode) { return mode == kFma64 ? 512 : 256; }
constexpr int mfma_step_waves(int mode) { return mode == kFma64 ? 2 : 3; }

struct StepArgs {
  int32_t shape;      // index into kStepShapeBlock
  int32_t mfma;       // concentration components on the matrix cores
  int64_t npts;       // L * P arrival points
  int64_t nblocks;    // logical workgroups
  int64_t per_xcd;    // logical workgroups per XCD
  int32_t P;          // (q+1)^3
  int32_t nsub;       // RK substeps
  int32_t bc;
  int32_t scaled;
  double tau, htau;
  double g[2 * kMaxSub];
  double ts[2 * kMaxSub];  // stage times
  double node[TBX_MAX_Q + 1];  // stretched nodes (cheb.h:53-58)
  // the velocity tree's values at its own grid nodes, [L][3][P], FP32 in the
  // FP32 modes and FP64 otherwise (null: the first stage evaluates like the
  // second); single velocity trees and the extrapolating functor's current
  // tree (its first stage) only
  const void *vnode;
  // the velocity adjoint: per velocity grid node [L][3][P], the weights of
  // the first substep's arrival nodes that are grid nodes of their velocity
  // leaf (null: every point scatters)
  double *wnode;
};

constexpr int kBlock = 256;
constexpr int kAdjBlock = 256;  // workgroup of the adjoint concentration kernel

// Launchers instantiated once per degree Q (tbx_kernels.hip, -DTBX_Q=Q).
#define TBX_DECLARE_Q(Q)                                                      \
  hipError_t launch_eval_q##Q(int mode, const TreeView &t, double *pts,       \
                              int64_t n, int bc, int write_back, double *out, \
                              hipStream_t s);                                 \
  hipError_t launch_traj_q##Q(int mode, const VelSource &vel, double *x,      \
                              int64_t n, const SubArgs &a, hipStream_t s);    \
  hipError_t launch_step_q##Q(int mode, const VelSource &vel,                 \
                              const TreeView &conc, const StepArgs &a,        \
                              double *vals, double *dep, hipStream_t s);      \
  hipError_t launch_veval_q##Q(int mode, const VelSource &vel, double *pts,   \
                               int64_t n, int bc, int write_back,             \
                               double t, int scaled, double g, double *out,   \
                               hipStream_t s);                                \
  hipError_t launch_adj_conc_q##Q(int mode, const VelSource &vel,             \
                                  const TreeView &conc, const StepArgs &a,    \
                                  const double *dep, const double *ua,        \
                                  double *lam, double *A, hipStream_t s);     \
  hipError_t launch_adj_vel_q##Q(int mode, const VelSource &vel,              \
                                 const TreeView &conc, const StepArgs &a,     \
                                 const double *A, double *G, hipStream_t s);  \
  hipError_t launch_contract_q##Q(int mode, const TreeView &t, int64_t n,     \
                                  int waves_per_leaf, double *out,            \
                                  double *pts, hipStream_t s);

TBX_DECLARE_Q(1)
TBX_DECLARE_Q(2)
TBX_DECLARE_Q(3)
TBX_DECLARE_Q(4)
TBX_DECLARE_Q(5)
TBX_DECLARE_Q(6)
TBX_DECLARE_Q(7)
TBX_DECLARE_Q(8)
TBX_DECLARE_Q(9)
TBX_DECLARE_Q(10)
TBX_DECLARE_Q(11)
TBX_DECLARE_Q(12)
TBX_DECLARE_Q(13)
TBX_DECLARE_Q(14)
TBX_DECLARE_Q(15)
TBX_DECLARE_Q(16)

// Fused step of grid (tricubic) leaves.
struct GridStepArgs {
  int64_t npts;  // L * N^3 sample points
  int32_t nsub;
  int32_t bc;
  int32_t scaled;
  int32_t blk;   // uniform tree of 8^3-sample leaves shared by velocity and
                 // concentration: sibling blocks staged in LDS
  int32_t nodal;  // first RK2 stage at a sample point from the velocity's
                  // sample there (velocity and concentration of one N)
  double tau, htau;
  double g[2 * kMaxSub];
  double node[16];  // a / (N - 1)
};

// One problem of a persistent batch (tbx_batch.hip).
struct BatchProblem {
  TreeView vel;         // coeff32: the batch's FP32 velocity coefficients
  TreeView conc[2];     // coeff32: the FP32 state read by a step of parity k
  float *out32[2];      // written by a step of parity k (= conc[1 - k].coeff32)
  const double *M;      // [P][Nc] refit operator
  const float *vnode;   // the velocity at its grid nodes [Lv][3][P] (FP32), or null
  int64_t item0;        // first work item
  int64_t L;            // concentration leaves
  int32_t q;
  int32_t bc;
  double tau, htau;
  double node[17];
};
// Launch configurations the batch autotuner chooses from:
// {threads per workgroup, leaves per work item}.
constexpr int kBatchConfigs = 6;
constexpr int kBatchBlock[kBatchConfigs] = {256, 256, 512, 512, 512, 1024};
constexpr int kBatchLpi[kBatchConfigs] = {1, 2, 1, 2, 4, 1};
int batch_blocks_per_cu(int cfg);
hipError_t launch_batch(int cfg, const BatchProblem *pb, int nprob, int parity,
                        int64_t total_items, unsigned long long *counter, int grid,
                        hipStream_t s);

// Grid-leaf launchers (tbx_grid.hip).
hipError_t launch_grid_eval(int mode, const TreeView &t, const double *s, const float *s32,
                            int n, int dof, double *pts, int64_t npts, int bc, int write_back,
                            double *out, hipStream_t st);
hipError_t launch_grid_step(int mode, const TreeView &tv, const double *vs, const float *vs32,
                            int vn, const TreeView &tc, const double *cs, const float *cs32,
                            int cn, int cdof, const GridStepArgs &a, double *out, double *dep,
                            hipStream_t st);

// The refit (tbx_refit.hip): T[k][a] = T_k(node a mapped to [-1, 1]) for
// the d = q + 1 stretched nodes.
struct RefitTable {
  double T[TBX_MAX_Q + 1][TBX_MAX_Q + 1];
};
// C = A H on the matrix cores: A [rows][lda] (K columns used), H [KP][np]
// zero-padded, C [rows][ldc] (ncols columns written), C32 optional.
hipError_t launch_refit_gemm(const double *A, int64_t lda, int64_t rows, int K, int KP,
                             const double *H, int np, double *C, int64_t ldc, int ncols,
                             float *C32, hipStream_t s);
// W = S V per row (S = the table's one-axis matrix on each axis, restricted
// to the triangle): row r = leaf * dof + l reads V[leaf d^3 dof + l cs +
// p ps]; W [rows][ldw] (packed triangle, zero tail).
hipError_t launch_refit_project(int d, const double *V, int64_t rows, int dof, int64_t cs,
                                int64_t ps, const RefitTable &tab, double *W, int ldw,
                                hipStream_t s);
// The same with the result added to W's first Nc entries per row.
hipError_t launch_refit_project_add(int d, const double *V, int64_t rows, int dof, int64_t cs,
                                    int64_t ps, const RefitTable &tab, double *W, int ldw,
                                    hipStream_t s);
// V = B U per row: U [rows][KP] (packed triangle), V [rows][d^3].
hipError_t launch_refit_expand(int d, const double *U, int64_t rows, const RefitTable &tab,
                               int KP, double *V, hipStream_t s);
// The same with V rounded to FP32.
hipError_t launch_refit_expand32(int d, const double *U, int64_t rows, const RefitTable &tab,
                                 int KP, float *V, hipStream_t s);
// CheckChebOutput's block sums (metric_blocks(npts) blocks; pass 0: dof
// sums per block, pass 1: 4 values per block); dof <= 8.
int metric_blocks(int64_t npts);
hipError_t launch_metric(const uint4 *info, int n, const double *val, const double *ref,
                         int64_t npts, int dof, const double *avg, int pass, double *partial,
                         hipStream_t s);
// d = x - y and per-block partial sums of d^2 (diff_norm2_blocks(n) of them).
int diff_norm2_blocks(int64_t n);
hipError_t launch_diff_norm2(const double *x, const double *y, int64_t n, double *d,
                             double *partial, hipStream_t s);

// The allocation-free step's device scratch (tbx_noalloc.hip): 256 MiB per
// GPU; noalloc_scratch gives its address on the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
hipError_t noalloc_scratch(void **ptr, size_t *bytes);

// On-device adaptivity (tbx_adapt.hip).
hipError_t launch_adapt_tail(const double *coeff, int64_t L, int dof, int q, double *tail,
                             hipStream_t s);
// vals [nnew][dof][(q+1)^3]: each new leaf's stretched grid nodes (node[q+1])
// evaluated in the source leaf src_leaf[n] (restrict_children: in the child
// src_leaf[n] + octant of the node).
hipError_t launch_leaf_resample(const TreeView &src, const int32_t *src_leaf,
                                const uint4 *info_new, int restrict_children, int64_t nnew,
                                const double *node, double *vals, hipStream_t s);
// split[l] = 1 for every leaf l more than one level coarser than a leaf
// whose same-size neighbour cell it holds (split must start at 0).
hipError_t launch_balance_probe(const TreeView &t, int bc, uint8_t *split, hipStream_t s);
// dst row r = src row map[r] where map[r] >= 0.
hipError_t launch_gather_rows(const double *src, const int64_t *map, int64_t nrows, int rowlen,
                              double *dst, hipStream_t s);

// Degree-independent launchers (tbx_common.hip).
hipError_t launch_locate(const TreeView &t, const double *pts, int64_t n,
                         int bc, int32_t *leaf, hipStream_t s);
hipError_t launch_grid_points(const TreeView &t, const double *node, int d,
                              double *out, hipStream_t s);
hipError_t launch_to_f32(const double *in, float *out, int64_t n,
                         hipStream_t s);

}  // namespace tbx
