// Refit of leaf values at the stretched grid nodes to Chebyshev coefficients
// (SetTreeGridValues, tree_utils.h:500-552, with GetPt2CoeffMatrix,
// cheb.h:166-196) and its transpose, for every leaf and component at once.
//
// Per row (one leaf component) the reference applies the dense P x Nc
// matrix pinv(B)^T to the d^3 values, B[p][m] = T_k(x_a) T_j(y_b) T_i(z_c)
// the triangle basis at the nodes.  B has full column rank, so
// pinv(B) = H B^T with H = (B^T B)^-1 (Nc x Nc, symmetric), and B^T is a
// tensor product restricted to the triangle's columns.  The refit runs as
//   w = B^T v   three one-axis contractions per row (k_refit_project):
//               d^4 + about d^4 / 2 + Nc d multiply-adds;
//   c = w H     one GEMM over all rows on the matrix cores (k_refit_gemm):
//               Nc^2 multiply-adds per row;
// 548 k instead of 2.3 M multiply-adds per row at q = 14.  The transpose
// (the adjoint's pinv(B)^T lam = B (H lam)) is the same GEMM followed by
// the three contractions in reverse (k_refit_expand).
//
// The GEMM: 64 x 64 output tiles per 256-thread workgroup, each wave a
// 32 x 32 block of four v_mfma_f64_16x16x4 tiles; K in chunks of 16
// staged in LDS (double-buffered, the next chunk's global loads issued
// before the current chunk's MFMAs).  Operand layouts of
// v_mfma_f64_16x16x4 (measured on gfx950): A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i % 4), register i / 4.
#include <hip/hip_runtime.h>

#include "tbx_internal.h"

namespace tbx {
namespace {

constexpr int kGemmBM = 64, kGemmBN = 64, kGemmBK = 16;
constexpr int kGemmThreads = 256;
constexpr int kAStride = kGemmBK + 1;  // LDS row stride of the A chunk (doubles)
constexpr int kBStride = kGemmBN + 2;  // LDS row stride of the B chunk

typedef double acc4 __attribute__((ext_vector_type(4)));

// C[r][c] = sum_k A[r][k] H[k][c] for r < rows, c < ncols, k < K; A rows
// have stride lda, H is [KP][np] with zero padding past K and ncols
// (KP >= K a multiple of 16, np a multiple of 64).  C32 (may be null)
// receives C rounded to FP32 with the same layout.  Workgroups are remapped
// so each XCD runs a contiguous range of tiles (the column tiles of one row
// block share its A rows in that XCD's L2).
__global__ __launch_bounds__(kGemmThreads) void k_refit_gemm(
    const double *__restrict__ A, int64_t lda, int64_t rows, int K, int KP,
    const double *__restrict__ H, int np, double *__restrict__ C, int64_t ldc, int ncols,
    float *__restrict__ C32, int64_t ntiles, int64_t per_xcd) {
  __shared__ double As[2][kGemmBM * kAStride];
  __shared__ double Bs[2][kGemmBK * kBStride];
  const int64_t logical = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= ntiles) return;
  const int ncb = np / kGemmBN;
  const int64_t r0 = (logical / ncb) * kGemmBM;
  const int c0 = (int)(logical % ncb) * kGemmBN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lg = lane >> 4;

  // global -> register staging: A chunk 64 x 16 (4 per thread), B chunk
  // 16 x 64 (4 per thread)
  const int ar = tid >> 2, ak = (tid & 3) * 4;
  const int bk = tid >> 4, bc = (tid & 15) * 4;
  const int64_t arow = r0 + ar;
  const bool arow_ok = arow < rows;
  const double *ap = A + (arow_ok ? arow : 0) * lda;
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + ak + u;
      ra[u] = (arow_ok && k < K) ? ap[k] : 0.0;
    }
    const double *hp = H + (int64_t)(k0 + bk) * np + c0 + bc;
#pragma unroll
    for (int u = 0; u < 4; ++u) rb[u] = hp[u];
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u) As[buf][ar * kAStride + ak + u] = ra[u];
#pragma unroll
    for (int u = 0; u < 4; ++u) Bs[buf][bk * kBStride + bc + u] = rb[u];
  };

  acc4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc4{0.0, 0.0, 0.0, 0.0};

  const int nchunks = KP / kGemmBK;
  load(0);
  store(0);
  __syncthreads();
  for (int ch = 0; ch < nchunks; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nchunks) load((ch + 1) * kGemmBK);
    const double *as = As[buf];
    const double *bs = Bs[buf];
#pragma unroll
    for (int s = 0; s < kGemmBK / 4; ++s) {
      const int k = 4 * s + lg;
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = as[(wr + 16 * a + li) * kAStride + k];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = bs[k * kBStride + wc + 16 * b + li];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    if (ch + 1 < nchunks) {
      store(buf ^ 1);
      __syncthreads();
    }
  }
  // D[i][j] of tile (a, b): lane j + 16 (i % 4), register i / 4
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = c0 + wc + 16 * b + li;
      if (col >= ncols) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + wr + 16 * a + lg + 4 * r;
        if (row >= rows) continue;
        const double v = acc[a][b][r];
        C[row * ldc + col] = v;
        if (C32) C32[row * ldc + col] = (float)v;
      }
    }
}

// Packed index of term (i, j, k) of the triangle i -> j -> k, i+j+k < D.
template <int D>
__device__ __forceinline__ int tri_index(int i, int j, int k) {
  // terms before block i: sum_{a<i} (D-a)(D-a+1)/2; before row j of block
  // i: sum_{b<j} (D-i-b)
  const int n = D - i;
  const int before_i = (D * (D + 1) * (D + 2) - n * (n + 1) * (n + 2)) / 6;
  return before_i + j * n - j * (j - 1) / 2 + k;
}

constexpr int kRefitThreads = 256;

// Rows a workgroup works on at once: as many teams of D^2 threads (one per
// (c, b) pair of a row) as the workgroup holds, so that low degrees keep
// the workgroup busy (3 rows at q = 8, 1 from q = 11).
template <int D>
constexpr int refit_rows() {
  return D * D <= kRefitThreads ? kRefitThreads / (D * D) : 1;
}

// Whether a thread loads its first line of the next row ahead (when each
// thread has one line per stage; above q = 15 the registers are spent).
template <int D>
constexpr bool refit_prefetch() {
  return D * D <= kRefitThreads;
}

// w = S v per row with S = S_z (x) S_y (x) S_x restricted to the triangle:
// one team of threads per row (refit_rows rows per workgroup at a time).
//   a[c][b][k] = sum_a S[k][a] v[c][b][a]
//   b[c][j][k] = sum_b S[j][b] a[c][b][k]          j + k < D
//   w[i][j][k] = sum_c S[i][c] b[c][j][k]          i + j + k < D
// With S[k][a] = T_k(node a) this is B^T v (the refit); with the scaled
// discrete transform at first-kind nodes, the initial fit.  Row r = leaf *
// dof + l reads V[leaf P dof + l cs + p ps] at node p = a + D (b + D c)
// (cs = P, ps = 1 for [L][dof][P] values; cs = 1, ps = dof for point-major
// [L][P][dof]); W rows have stride ldw, entries Nc .. ldw-1 written as 0.
// ACC: w is added to W instead (entries past Nc untouched).
template <int D, bool ACC>
__global__ __launch_bounds__(kRefitThreads) void k_refit_project(const double *__restrict__ V,
                                                                 int64_t rows, int dof,
                                                                 int64_t cs, int64_t ps,
                                                                 RefitTable tab,
                                                                 double *__restrict__ W,
                                                                 int ldw) {
  constexpr int P = D * D * D, NC = D * (D + 1) * (D + 2) / 6;
  constexpr int R = refit_rows<D>();
  constexpr int TS = R > 1 ? D * D : kRefitThreads;  // threads of a row's team
  __shared__ double sa[R * P];
  __shared__ double sb[R * P];
  const int tid = threadIdx.x;
  const int rl = tid / TS < R ? tid / TS : R - 1;  // this thread's row of the group
  const int tl = tid - (tid / TS) * TS;
  const bool team = tid < R * TS;
  double *ra = sa + rl * P, *rb = sb + rl * P;
  // The one-axis table S is read from the kernel arguments: every index
  // into it is wave-uniform, so its entries arrive as scalar operands
  // instead of LDS broadcasts (k_refit_expand keeps its copy in LDS; read
  // this way there it ran slower).
  // The first (c, b) line of the thread's next row is loaded while the
  // current row goes through its LDS stages, so that each row does not
  // wait for its own loads (two workgroups per CU at q = 14).
  const int64_t stride = (int64_t)gridDim.x * R;
  double xn[D];
  auto fetch = [&](int64_t r) {
    if (refit_prefetch<D>() && team && tl < D * D && r < rows) {
      const double *v = V + (r / dof) * (int64_t)P * dof + (r % dof) * cs;
#pragma unroll
      for (int a = 0; a < D; ++a) xn[a] = v[(tl * D + a) * ps];
    }
  };
  fetch((int64_t)blockIdx.x * R + rl);
  __syncthreads();
  for (int64_t r0 = (int64_t)blockIdx.x * R; r0 < rows; r0 += stride) {
    const int64_t r = r0 + rl;
    const bool act = team && r < rows;
    if (act) {
      const double *v = V + (r / dof) * (int64_t)P * dof + (r % dof) * cs;
      for (int t = tl; t < D * D; t += TS) {  // t = c * D + b
        double x[D];
        if (refit_prefetch<D>() && t == tl) {
#pragma unroll
          for (int a = 0; a < D; ++a) x[a] = xn[a];
        } else {
#pragma unroll
          for (int a = 0; a < D; ++a) x[a] = v[(t * D + a) * ps];
        }
        auto out = [&](int k) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) s = __builtin_fma(tab.T[k][a], x[a], s);
          ra[t * D + k] = s;
        };
        // several rows at once: one output at a time keeps the registers
        // (and the waves per SIMD) that the rows' latency needs; one row:
        // all outputs unrolled for their independent chains
        if constexpr (R > 1) {
#pragma unroll 1
          for (int k = 0; k < D; ++k) out(k);
        } else {
#pragma unroll
          for (int k = 0; k < D; ++k) out(k);
        }
      }
    }
    fetch(r + stride);
    __syncthreads();
    if (act) {
      for (int t = tl; t < D * D; t += TS) {  // t = c * D + k
        const int c = t / D, k = t % D;
        double x[D];
#pragma unroll
        for (int b = 0; b < D; ++b) x[b] = ra[(c * D + b) * D + k];
        for (int j = 0; j + k < D; ++j) {
          double s = 0.0;
#pragma unroll
          for (int b = 0; b < D; ++b) s = __builtin_fma(tab.T[j][b], x[b], s);
          rb[(c * D + j) * D + k] = s;
        }
      }
    }
    __syncthreads();
    if (act) {
      double *w = W + r * ldw;
      for (int t = tl; t < D * D; t += TS) {  // t = j * D + k
        const int j = t / D, k = t % D;
        if (j + k >= D) continue;
        double x[D];
#pragma unroll
        for (int c = 0; c < D; ++c) x[c] = rb[(c * D + j) * D + k];
        for (int i = 0; i + j + k < D; ++i) {
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < D; ++c) s = __builtin_fma(tab.T[i][c], x[c], s);
          if constexpr (ACC)
            w[tri_index<D>(i, j, k)] += s;
          else
            w[tri_index<D>(i, j, k)] = s;
        }
      }
      if constexpr (!ACC)
        for (int m = NC + tl; m < ldw; m += TS) w[m] = 0.0;
    }
    __syncthreads();  // the next rows reuse sa and sb
  }
}

// The transpose: v = B u per row (u rows of stride KP in packed order), the
// row's polynomial at the D^3 grid nodes, v[c][b][a] (x fastest), by the
// three one-axis contractions z, y, x through LDS.  TO: the output precision
// (FP32 for the FP32 modes' nodal velocities).  (A variant contracting x
// first over the triangle, with y and z in registers, does a third fewer
// multiply-adds but ran 1.3x slower: its scalar-operand and register loads
// cost more than the LDS reads they replace.)
template <int D, typename TO>
__global__ __launch_bounds__(kRefitThreads) void k_refit_expand(const double *__restrict__ U,
                                                                int64_t rows, RefitTable tab,
                                                                int KP, TO *__restrict__ V) {
  constexpr int P = D * D * D;
  constexpr int R = refit_rows<D>();
  constexpr int TS = R > 1 ? D * D : kRefitThreads;  // threads of a row's team
  __shared__ double sa[R * P];
  __shared__ double sb[R * P];
  __shared__ double sT[D * D];
  const int tid = threadIdx.x;
  const int rl = tid / TS < R ? tid / TS : R - 1;  // this thread's row of the group
  const int tl = tid - (tid / TS) * TS;
  const bool team = tid < R * TS;
  double *ra = sa + rl * P, *rb = sb + rl * P;
  for (int t = tid; t < D * D; t += kRefitThreads) sT[t] = tab.T[t / D][t % D];
  // the first (j, k) line of the thread's next row loaded during the
  // current row's LDS stages, as in k_refit_project
  const int64_t stride = (int64_t)gridDim.x * R;
  double xn[D];
  auto fetch = [&](int64_t r) {
    const int j = tl / D, k = tl % D;
    if (refit_prefetch<D>() && team && tl < D * D && j + k < D && r < rows) {
      const double *u = U + r * KP;
      const int ni = D - j - k;
#pragma unroll
      for (int i = 0; i < D; ++i) xn[i] = i < ni ? u[tri_index<D>(i, j, k)] : 0.0;
    }
  };
  fetch((int64_t)blockIdx.x * R + rl);
  __syncthreads();
  for (int64_t r0 = (int64_t)blockIdx.x * R; r0 < rows; r0 += stride) {
    const int64_t r = r0 + rl;
    const bool act = team && r < rows;
    if (act) {
      const double *u = U + r * KP;
      for (int t = tl; t < D * D; t += TS) {  // t = j * D + k
        const int j = t / D, k = t % D;
        if (j + k >= D) continue;
        double x[D];
        const int ni = D - j - k;
        if (refit_prefetch<D>() && t == tl) {
#pragma unroll
          for (int i = 0; i < D; ++i) x[i] = xn[i];
        } else {
#pragma unroll
          for (int i = 0; i < D; ++i) x[i] = i < ni ? u[tri_index<D>(i, j, k)] : 0.0;
        }
        for (int c = 0; c < D; ++c) {
          double s = 0.0;
#pragma unroll
          for (int i = 0; i < D; ++i) s = __builtin_fma(sT[i * D + c], x[i], s);
          rb[(c * D + j) * D + k] = s;
        }
      }
    }
    fetch(r + stride);
    __syncthreads();
    if (act) {
      for (int t = tl; t < D * D; t += TS) {  // t = c * D + k
        const int c = t / D, k = t % D;
        double x[D];
#pragma unroll
        for (int j = 0; j < D; ++j) x[j] = j + k < D ? rb[(c * D + j) * D + k] : 0.0;
        for (int b = 0; b < D; ++b) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) s = __builtin_fma(sT[j * D + b], x[j], s);
          ra[(c * D + b) * D + k] = s;
        }
      }
    }
    __syncthreads();
    if (act) {
      TO *v = V + r * P;
      for (int t = tl; t < D * D; t += TS) {  // t = c * D + b
        double x[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = ra[t * D + k];
        auto out = [&](int a) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) s = __builtin_fma(sT[k * D + a], x[k], s);
          v[t * D + a] = (TO)s;
        };
        if constexpr (R > 1) {  // as in k_refit_project
#pragma unroll 1
          for (int a = 0; a < D; ++a) out(a);
        } else {
#pragma unroll
          for (int a = 0; a < D; ++a) out(a);
        }
      }
    }
    __syncthreads();  // the next rows reuse sa and sb
  }
}

__global__ __launch_bounds__(256) void k_axpy_dot(const double *__restrict__ x,
                                                  const double *__restrict__ y, int64_t n,
                                                  double *__restrict__ d,
                                                  double *__restrict__ partial) {
  // d = x - y; partial[block] = sum of d^2 over the block's elements
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const double v = x[i] - y[i];
    d[i] = v;
    s = __builtin_fma(v, v, s);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// CheckChebOutput's sums (utils.txx:158-293) over the metric's points: n^3
// first-kind nodes per leaf (point i in leaf i / n3), weight s^3 / n^3 with
// s the leaf edge.  Pass 0: per-block sums of w (val - ref) per component;
// pass 1: per-block sum w ref^2, sum w e^2, max |ref|, max |e| with
// e = val - ref - avg.  Block partials are summed on the host in block order.
constexpr int kMetricThreads = 256;
constexpr int kMetricMaxDof = 8;
struct MetricAvg {
  double avg[kMetricMaxDof];
};

__global__ __launch_bounds__(kMetricThreads) void k_metric(const uint4 *__restrict__ info, int n,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ ref,
                                                           int64_t npts, int dof, MetricAvg m,
                                                           int pass,
                                                           double *__restrict__ partial) {
  __shared__ double red[4][kMetricThreads];
  const int n3 = n * n * n;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  double sw[kMetricMaxDof];
  for (int l = 0; l < kMetricMaxDof; ++l) sw[l] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kMetricThreads + threadIdx.x; i < npts;
       i += (int64_t)gridDim.x * kMetricThreads) {
    const double s = 1.0 / (double)(1u << info[i / n3].w);
    const double w = s * s * s / (double)n3;
    for (int l = 0; l < dof; ++l) {
      const double r = ref[i * dof + l];
      const double e = val[i * dof + l] - r;
      if (pass == 0) {
        sw[l] += w * e;
      } else {
        const double ec = e - m.avg[l];
        acc[0] += w * r * r;
        acc[1] += w * ec * ec;
        acc[2] = fmax(acc[2], fabs(r));
        acc[3] = fmax(acc[3], fabs(ec));
      }
    }
  }
  const int tid = threadIdx.x;
  const int nq = pass == 0 ? dof : 4;
  for (int q0 = 0; q0 < nq; q0 += 4) {
    for (int k = 0; k < 4; ++k) {
      const int l = q0 + k;
      red[k][tid] = pass == 0 ? (l < dof ? sw[l] : 0.0) : acc[k];
    }
    __syncthreads();
    for (int w = kMetricThreads / 2; w > 0; w >>= 1) {
      if (tid < w)
        for (int k = 0; k < 4; ++k) {
          const bool is_max = pass == 1 && k >= 2;
          red[k][tid] = is_max ? fmax(red[k][tid], red[k][tid + w]) : red[k][tid] + red[k][tid + w];
        }
      __syncthreads();
    }
    if (tid < 4 && q0 + tid < nq) partial[(int64_t)blockIdx.x * nq + q0 + tid] = red[tid][0];
    __syncthreads();
  }
}

}  // namespace

int metric_blocks(int64_t n) {
  const int64_t b = (n + kMetricThreads - 1) / kMetricThreads;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

hipError_t launch_metric(const uint4 *info, int n, const double *val, const double *ref,
                         int64_t npts, int dof, const double *avg, int pass, double *partial,
                         hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  if (dof < 1 || dof > kMetricMaxDof) return hipErrorInvalidValue;
  MetricAvg m{};
  for (int l = 0; l < dof && avg; ++l) m.avg[l] = avg[l];
  hipLaunchKernelGGL(k_metric, dim3(metric_blocks(npts)), dim3(kMetricThreads), 0, s, info, n,
                     val, ref, npts, dof, m, pass, partial);
  return hipGetLastError();
}

hipError_t launch_refit_gemm(const double *A, int64_t lda, int64_t rows, int K, int KP,
                             const double *H, int np, double *C, int64_t ldc, int ncols,
                             float *C32, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  // at most 2^22 tiles (2^30 work-items) per dispatch
  const int64_t chunk = ((int64_t)1 << 22) / (np / kGemmBN) * kGemmBM;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int64_t nr = rows - r0 < chunk ? rows - r0 : chunk;
    const int64_t ntiles = ((nr + kGemmBM - 1) / kGemmBM) * (np / kGemmBN);
    const int64_t per_xcd = (ntiles + 7) / 8;
    hipLaunchKernelGGL(k_refit_gemm, dim3((unsigned)(8 * per_xcd)), dim3(kGemmThreads), 0, s,
                       A + r0 * lda, lda, nr, K, KP, H, np, C + r0 * ldc, ldc, ncols,
                       C32 ? C32 + r0 * ldc : nullptr, ntiles, per_xcd);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

#define TBX_REFIT_CASES(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)

// A dispatch holds at most 2^32 - 1 work-items: the row kernels run at most
// kRowBlocks workgroups, each taking every kRowBlocks-th row.
constexpr int64_t kRowBlocks = 1 << 20;

template <bool ACC>
hipError_t launch_project(int d, const double *V, int64_t rows, int dof, int64_t cs, int64_t ps,
                          const RefitTable &tab, double *W, int ldw, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  if (dof < 1) return hipErrorInvalidValue;
  switch (d) {
#define X(D)                                                                                 \
  case D: {                                                                                  \
    const int64_t groups = (rows + refit_rows<D>() - 1) / refit_rows<D>();                   \
    const unsigned nb = (unsigned)(groups < kRowBlocks ? groups : kRowBlocks);               \
    hipLaunchKernelGGL((k_refit_project<D, ACC>), dim3(nb), dim3(kRefitThreads), 0, s, V, rows, \
                       dof, cs, ps, tab, W, ldw);                                            \
    break;                                                                                   \
  }
    TBX_REFIT_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_refit_project(int d, const double *V, int64_t rows, int dof, int64_t cs,
                                int64_t ps, const RefitTable &tab, double *W, int ldw,
                                hipStream_t s) {
  return launch_project<false>(d, V, rows, dof, cs, ps, tab, W, ldw, s);
}

hipError_t launch_refit_project_add(int d, const double *V, int64_t rows, int dof, int64_t cs,
                                    int64_t ps, const RefitTable &tab, double *W, int ldw,
                                    hipStream_t s) {
  return launch_project<true>(d, V, rows, dof, cs, ps, tab, W, ldw, s);
}

template <typename TO>
hipError_t launch_expand(int d, const double *U, int64_t rows, const RefitTable &tab, int KP,
                         TO *V, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  // 8,192 row-strided workgroups: 9 % faster than one per row on 108,972
  // rows of degree 14 (the table set-up is paid once per workgroup)
  constexpr int64_t kExpandBlocks = 8192;
  switch (d) {
#define X(D)                                                                             \
  case D: {                                                                              \
    const int64_t groups = (rows + refit_rows<D>() - 1) / refit_rows<D>();               \
    const unsigned nb = (unsigned)(groups < kExpandBlocks ? groups : kExpandBlocks);     \
    hipLaunchKernelGGL((k_refit_expand<D, TO>), dim3(nb), dim3(kRefitThreads), 0, s, U, rows, \
                       tab, KP, V);                                                      \
    break;                                                                               \
  }
    TBX_REFIT_CASES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_refit_expand(int d, const double *U, int64_t rows, const RefitTable &tab,
                               int KP, double *V, hipStream_t s) {
  return launch_expand<double>(d, U, rows, tab, KP, V, s);
}

hipError_t launch_refit_expand32(int d, const double *U, int64_t rows, const RefitTable &tab,
                                 int KP, float *V, hipStream_t s) {
  return launch_expand<float>(d, U, rows, tab, KP, V, s);
}

int diff_norm2_blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

hipError_t launch_diff_norm2(const double *x, const double *y, int64_t n, double *d,
                             double *partial, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy_dot, dim3(diff_norm2_blocks(n)), dim3(256), 0, s, x, y, n, d,
                     partial);
  return hipGetLastError();
}

}  // namespace tbx
