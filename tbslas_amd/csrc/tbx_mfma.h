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
// K walks the packed triangle four terms at a time, the four lane groups
// g = lane / 16 taking terms 4q + g (165 terms in 42 chunks at q = 8).  A
// lane's A operand is coeff[c0 + lane % 16][4q + g], its B operand
// T_i(z) T_j(y) T_k(x) of term 4q + g at its tile point: the chunk's four
// products are formed and the lane picks its own.
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

#include <utility>

#include "tbx_device.h"

namespace tbx {

// f(integral_constant<int, I>) for I = 0 .. N - 1, unrolled at compile time
// (the chunk loops are too long for the unroller's limits at high degree).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

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

__device__ __forceinline__ double fma_t(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fma_t(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

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

// The packed triangle in chunks of four consecutive terms: chunk q holds
// packed indices 4q .. 4q + 3 (lane group g takes 4q + g), with their
// (i, j, k); indices past Nc are padding.
template <int D>
struct MfmaChunks {
  static constexpr int NC = D * (D + 1) * (D + 2) / 6;
  static constexpr int N = (NC + 3) / 4;
  int i[4 * N], j[4 * N], k[4 * N];
  constexpr MfmaChunks() : i(), j(), k() {
    int m = 0;
    for (int a = 0; a < D; ++a)
      for (int b = 0; a + b < D; ++b)
        for (int c = 0; a + b + c < D; ++c) {
          i[m] = a;
          j[m] = b;
          k[m] = c;
          ++m;
        }
    for (; m < 4 * N; ++m) i[m] = j[m] = k[m] = 0;
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
// (component lane % 16, term slot g); in[n]: the lane's tile point n lies
// in the leaf (its B operand is zero otherwise).  K runs over the packed
// triangle four terms at a time; the lane's B operand is the product of
// its term's three basis values, picked among the chunk's four by 0/1
// weights (no divergent selection).  The
// chunks are unrolled; the A operands are loaded two chunks ahead and a
// scheduling barrier closes every chunk, which keeps the loads in flight
// under the MFMAs without hoisting them all to the top.
template <int D, typename T, int NT>
__device__ __forceinline__ void mfma_leaf_terms(__amdgpu_buffer_rsrc_t rsrc, int voff, int g,
                                                const bool (&in)[2], const T (&tz)[2][D],
                                                const T (&ty)[2][D], const T (&tx)[2][D],
                                                typename MfmaAcc<T>::V (&acc)[4][2]) {
  using Acc = MfmaAcc<T>;
  constexpr int NC = MfmaChunks<D>::NC;
  constexpr int NQ = MfmaChunks<D>::N;
  const bool last_valid = 4 * (NQ - 1) + g < NC;  // the final chunk's padding
  // the lane's pick among a chunk's four products as 0/1 weights: a
  // branch-free, exact selection (finite products times 0 and 1)
  T w[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) w[s] = g == s ? (T)1 : (T)0;
  auto load = [&](int q, T (&a)[NT]) {
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      const int soff = (int)sizeof(T) * (mt * 16 * NC + 4 * q);
      const T v = buffer_coeff<T>(rsrc, voff, soff);
      a[mt] = (q < NQ - 1 || last_valid) ? v : (T)0;
    }
  };
  T a0[NT], a1[NT];
  load(0, a0);
  if (NQ > 1) load(1, a1);
  static_for<NQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    T a2[NT];
    if constexpr (q + 2 < NQ) load(q + 2, a2);
    T bv[2];
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      // the leaf mask first: products of the pass, not hoisted out of it
      T p[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        constexpr MfmaChunks<D> tab{};
        const int m = 4 * q + s;
        const T zm = in[n] ? tz[n][tab.i[m]] : (T)0;
        p[s] = m < NC ? zm * ty[n][tab.j[m]] * tx[n][tab.k[m]] : (T)0;
      }
      T b = p[0] * w[0];
#pragma unroll
      for (int s = 1; s < 4; ++s) b = fma_t(p[s], w[s], b);
      bv[n] = b;
    }
#pragma unroll
    for (int mt = 0; mt < NT; ++mt)
#pragma unroll
      for (int n = 0; n < 2; ++n) acc[mt][n] = Acc::mma(a0[mt], bv[n], acc[mt][n]);
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      a0[mt] = a1[mt];
      if constexpr (q + 2 < NQ) a1[mt] = a2[mt];
    }
    __builtin_amdgcn_sched_barrier(0);
  });
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
    T tz[2][D], ty[2][D], tx[2][D];
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
      cheb_poly<D>(px, tx[n]);
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
            mfma_leaf_terms<D, T, 4>(rsrc, voff, g, in, tz, ty, tx, acc);
            break;
          case 3:
            mfma_leaf_terms<D, T, 3>(rsrc, voff, g, in, tz, ty, tx, acc);
            break;
          case 2:
            mfma_leaf_terms<D, T, 2>(rsrc, voff, g, in, tz, ty, tx, acc);
            break;
          default:
            mfma_leaf_terms<D, T, 1>(rsrc, voff, g, in, tz, ty, tx, acc);
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
