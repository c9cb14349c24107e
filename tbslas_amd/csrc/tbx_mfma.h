// Many-component leaf evaluation on the matrix cores.
//
// For the points of one leaf, the values of all dof components are a
// small GEMM: vals[c][p] = sum_m coeff[c][m] B_m(p), with B_m the packed
// triangle basis T_i(z) T_j(y) T_k(x) (tree_functor.h:256-266).  With many
// components (tracers sharing one trajectory set) the VALU contraction
// streams dof * Nc coefficients through the scalar cache for every wave;
// here the product runs as 16 x 16 x 4 MFMA tiles instead:
//
//   M = 16 components, N = 16 points, K = 4 basis terms per instruction.
//
// K walks the triangle row by row: for each (i, j) the x-degrees k of the
// row go in chunks of four, the four lane groups g = lane / 16 taking
// k = 4c + g (rows padded to a multiple of 4: 244 of 165 terms at q = 8).
// A lane's B operand is then T_i(z) T_j(y) T_{4c+g}(x) of its tile point,
// and its A operand coeff[c0 + lane % 16][base(i, j) + 4c + g], so only the
// x-degree depends on the lane.
//
// Operand layouts of v_mfma_{f64,f32}_16x16x4 (both measured on gfx950):
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; the result D[i][j]
// is in lane j + 16 (i / 4), register i % 4 for f32, and in lane
// j + 16 (i % 4), register i / 4 for f64.
//
// The wave's 64 points (one per lane, all lanes active) go in two halves of
// 32 (two N tiles); the tile point of lane l in tile n of half h is
// 32 h + 16 n + l % 16, its coordinates and leaf fetched by cross-lane
// reads.  The leaves of a half are visited one after another (waterfall),
// the B operand zeroed for tile points outside the visited leaf, so every
// point accumulates its own leaf's terms only.  Components go 64 at a time
// (four M tiles).
#pragma once

#include "tbx_device.h"

namespace tbx {

template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result row held in register r of lane group g
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <>
struct MfmaAcc<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

template <typename T>
__device__ __forceinline__ T lane_read(T v, int src) {
  if constexpr (sizeof(T) == 8) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
  } else {
    return __builtin_amdgcn_ds_bpermute(src << 2, v);
  }
}

__device__ __forceinline__ double lane_read_f64(double v, int src) { return lane_read<double>(v, src); }
__device__ __forceinline__ int64_t lane_read_i64(int64_t v, int src) {
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)v);
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// The K chunks of the packed triangle in order: row (i, j) of length
// R = D - i - j, chunk c covers x-degrees 4c .. 4c + 3 of the row, packed
// index base + 4c + g for lane group g.
template <int D>
struct MfmaChunks {
  static constexpr int count() {
    int n = 0;
    for (int i = 0; i < D; ++i)
      for (int j = 0; i + j < D; ++j) n += (D - i - j + 3) / 4;
    return n;
  }
  static constexpr int N = count();
  int i[N], j[N], c[N], base[N], R[N];
  constexpr MfmaChunks() : i(), j(), c(), base(), R() {
    int n = 0, b = 0;
    for (int a = 0; a < D; ++a)
      for (int bb = 0; a + bb < D; ++bb) {
        const int r = D - a - bb;
        for (int cc = 0; 4 * cc < r; ++cc) {
          i[n] = a;
          j[n] = bb;
          c[n] = cc;
          base[n] = b;
          R[n] = r;
          ++n;
        }
        b += r;
      }
  }
};

// Raw buffer loads of one coefficient: the leaf's component block is the
// buffer (a wave-uniform descriptor, its extent in bytes), the lane's
// offset a 32-bit register and the chunk's a scalar.  Reads past the extent
// return 0.  Descriptor word 3 of the gfx9 family (raw, 32-bit data).
constexpr int kBufferWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ T buffer_coeff(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 8) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return __hiloint2double((int)v[1], (int)v[0]);
  } else {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  }
}

// One leaf's terms for NT M tiles (16 NT components) and the two N tiles:
// acc[mt][n] += coeff tile mt x basis tile n.  rsrc covers the leaf's
// components from the first of the tiles; voff = the lane's byte offset
// (component lane % 16, x-degree slot g); in[n]: the lane's tile point n
// lies in the leaf (its B operand is zero otherwise).  The chunk table is
// unrolled; the A operands are loaded two chunks ahead and a scheduling
// barrier closes every chunk, which keeps the loads in flight under the
// MFMAs without hoisting them all to the top.
template <int D, typename T, int NT>
__device__ __forceinline__ void mfma_leaf_terms(__amdgpu_buffer_rsrc_t rsrc, int voff, int g,
                                                const bool (&in)[2], const T (&tz)[2][D],
                                                const T (&ty)[2][D],
                                                const T (&txg)[2][(D + 3) / 4],
                                                typename MfmaAcc<T>::V (&acc)[4][2]) {
  using Acc = MfmaAcc<T>;
  constexpr int NC = D * (D + 1) * (D + 2) / 6;
  constexpr MfmaChunks<D> ch{};
  constexpr int NQ = MfmaChunks<D>::N;
  auto valid_at = [&](int q) { return 4 * ch.c[q] + 3 < ch.R[q] || 4 * ch.c[q] + g < ch.R[q]; };
  auto load = [&](int q, T (&a)[NT]) {
    const bool valid = valid_at(q);
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      const int soff = (int)sizeof(T) * (mt * 16 * NC + ch.base[q] + 4 * ch.c[q]);
      const T v = buffer_coeff<T>(rsrc, voff, soff);
      a[mt] = valid ? v : (T)0;
    }
  };
  T a0[NT], a1[NT];
  load(0, a0);
  if (NQ > 1) load(1, a1);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    T a2[NT];
    if (q + 2 < NQ) load(q + 2, a2);
    const int i = ch.i[q], j = ch.j[q], c = ch.c[q];
    const bool valid = valid_at(q);
    T bv[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      // the leaf mask first: products of the pass, not hoisted out of it
      const T zm = in[n] ? tz[n][i] : (T)0;
      bv[n] = valid ? zm * ty[n][j] * txg[n][c] : (T)0;
    }
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
      for (int n = 0; n < 2; ++n) acc[mt][n] = Acc::mma(a0[mt], bv[n], acc[mt][n]);
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      a0[mt] = a1[mt];
      if (q + 2 < NQ) a1[mt] = a2[mt];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Values of all t.dof components (dof a multiple of 16) at the wave's
// points: lane l's point (x, y, z) in leaf `leaf`; component c of the point
// is written to out[obase + c * cstride] when `live`.  T is the operand and
// accumulation type (double: FP64 coefficients; float: the FP32 mirror).
template <int D, typename T>
__device__ __forceinline__ void eval_dof_mfma(const TreeView &t, int leaf, double x, double y,
                                              double z, double *out, int64_t obase,
                                              int64_t cstride, bool live) {
  using Acc = MfmaAcc<T>;
  using V = typename Acc::V;
  constexpr int NCH = (D + 3) / 4;  // chunks of the longest row
  constexpr int NC = D * (D + 1) * (D + 2) / 6;
  const int lane = (int)(threadIdx.x & 63u);
  const int g = lane >> 4, col = lane & 15;
  const T *coeff;
  if constexpr (sizeof(T) == 8)
    coeff = reinterpret_cast<const T *>(t.coeff);
  else
    coeff = reinterpret_cast<const T *>(t.coeff32);
  double ex, ey, ez;
  leaf_xi(t, leaf, x, y, z, ex, ey, ez);
  const int dof = t.dof;
#pragma unroll 1
  for (int half = 0; half < 2; ++half) {
    // tile points of this lane: 32 half + 16 n + col
    T tz[2][D], ty[2][D], txg[2][NCH];
    int pleaf[2];
    int64_t pbase[2];
    bool plive[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const int src = 32 * half + 16 * n + col;
      const double px = lane_read_f64(ex, src), py = lane_read_f64(ey, src),
                   pz = lane_read_f64(ez, src);
      pleaf[n] = __builtin_amdgcn_ds_bpermute(src << 2, leaf);
      pbase[n] = lane_read_i64(obase, src);
      plive[n] = __builtin_amdgcn_ds_bpermute(src << 2, live ? 1 : 0) != 0;
      T tx[4 * NCH];
      cheb_poly<D>(px, tx);
#pragma unroll
      for (int k = D; k < 4 * NCH; ++k) tx[k] = (T)0;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const T a0 = tx[4 * c], a1 = tx[4 * c + 1], a2 = tx[4 * c + 2], a3 = tx[4 * c + 3];
        txg[n][c] = g == 0 ? a0 : g == 1 ? a1 : g == 2 ? a2 : a3;
      }
      cheb_poly<D>(py, ty[n]);
      cheb_poly<D>(pz, tz[n]);
    }
    const uint64_t halfmask = half ? 0xffffffff00000000ull : 0x00000000ffffffffull;
#pragma unroll 1
    for (int c0 = 0; c0 < dof; c0 += 64) {
      const int ntile = (dof - c0) / 16 < 4 ? (dof - c0) / 16 : 4;
      V acc[4][2];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int n = 0; n < 2; ++n) acc[mt][n] = V{0, 0, 0, 0};
      uint64_t pending = __builtin_amdgcn_ballot_w64(true) & halfmask;
      while (pending) {
        const int lf = __builtin_amdgcn_readlane(leaf, __builtin_ctzll(pending));
        pending &= ~__builtin_amdgcn_ballot_w64(leaf == lf);
        const bool in[2] = {pleaf[0] == lf, pleaf[1] == lf};
        // the buffer: components c0 .. c0 + 16 ntile - 1 of leaf lf
        const T *cb = coeff + ((int64_t)lf * dof + c0) * NC;
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<T *>(cb), 0, (int)sizeof(T) * 16 * ntile * NC, kBufferWord3);
        const int voff = (int)sizeof(T) * (col * NC + g);
        switch (ntile) {  // wave-uniform
          case 4:
            mfma_leaf_terms<D, T, 4>(rsrc, voff, g, in, tz, ty, txg, acc);
            break;
          case 3:
            mfma_leaf_terms<D, T, 3>(rsrc, voff, g, in, tz, ty, txg, acc);
            break;
          case 2:
            mfma_leaf_terms<D, T, 2>(rsrc, voff, g, in, tz, ty, txg, acc);
            break;
          default:
            mfma_leaf_terms<D, T, 1>(rsrc, voff, g, in, tz, ty, txg, acc);
            break;
        }
      }
      // lane l holds D[row(g, r)][col]: component c0 + 16 mt + row(g, r)
      // of tile point n
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        if (!plive[n]) continue;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          if (mt < ntile) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              out[pbase[n] + (int64_t)(c0 + 16 * mt + Acc::row(g, r)) * cstride] =
                  (double)acc[mt][n][r];
          }
        }
      }
    }
  }
}

}  // namespace tbx
