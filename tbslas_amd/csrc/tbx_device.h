// Device building blocks shared by the HIP kernels of libtbx (gfx950).
//
// Execution model of the point kernels: one departure point per lane.  A
// lane finds its leaf (Morton key -> coarse-cell directory -> bounded binary
// search), maps the point to the leaf's [-1,1]^3 and builds T_0..T_q per
// axis in registers.  The coefficient contraction then runs as a waterfall
// over the distinct leaves present in the wavefront: the first active lane's
// leaf is broadcast (readfirstlane), every lane in that leaf contracts with
// that leaf's coefficients, which are wave-uniform and therefore fetched by
// scalar loads into SGPRs (one fetch feeds 64 lanes), and the loop repeats
// for the remaining lanes.  Points generated leaf by leaf, or sorted by leaf,
// keep a wavefront inside one or two leaves, so the waterfall adds little.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "tbx_internal.h"

namespace tbx {

template <int MODE>
struct ModeType {
  using T = double;
};
template <>
struct ModeType<kFp32> {
  using T = float;
};
template <>
struct ModeType<kFp32h> {
  using T = float;
};

// acc + a*b.  kExact64 keeps the reference's separate multiply and add
// (vec_eval: add_intrin(w, mul_intrin(px, c)), tree_functor.h:63-76);
// the other modes fuse.
template <int MODE, typename T>
__device__ __forceinline__ T madd(T a, T b, T acc) {
  if constexpr (MODE == kExact64) {
    return acc + a * b;
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_fmaf(a, b, acc);
  } else {
    return __builtin_fma(a, b, acc);
  }
}

// Periodic wrap of one coordinate (tree_functor.h:442-449).
__device__ __forceinline__ double wrap1(double c) {
  if (c < 0.0) c = c + 1.0;
  if (c >= 1.0) c = c - 1.0;
  return c;
}

// Integer cell of a coordinate at 21 bits.  floor(c * 2^21) is exact for a
// double in [0,1).  c == 1.0 (and anything above) takes the last cell: this
// is the reference's "c - 2^-MAX_DEPTH" shift for c == 1.0 under a
// non-periodic bc (tree_functor.h:464-479), extended to a clamp outside the
// domain; negative and NaN coordinates take cell 0.
__device__ __forceinline__ uint32_t cell21(double c) {
  if (!(c >= 0.0)) return 0u;
  const double f = floor(c * kKeyScale);
  return f > kKeyScale - 1.0 ? (uint32_t)(kKeyScale - 1.0) : (uint32_t)f;
}

// Spread the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint64_t point_key(double x, double y, double z) {
  return spread3(cell21(x)) | (spread3(cell21(y)) << 1) |
         (spread3(cell21(z)) << 2);
}

// Index of the leaf whose key range holds key k: the last leaf with
// key <= k.  The directory narrows the search to the leaves overlapping the
// point's level-g cell, dir[c] .. dir[c + 1]; when no leaf is deeper than g
// (dir_exact) the leaf holding the cell's first key holds the whole cell and
// dir[c] is the answer.  dir_first / dir_search are the two halves, so that
// a thread locating several points can issue all their directory loads
// before it waits on any of them.
__device__ __forceinline__ int dir_first(const TreeView &t, uint64_t k) {
  return t.dir[(uint32_t)(k >> t.dir_shift)];
}

__device__ __forceinline__ int dir_last(const TreeView &t, uint64_t k) {
  return t.dir_exact ? 0 : t.dir[(uint32_t)(k >> t.dir_shift) + 1];
}

__device__ __forceinline__ int dir_search(const TreeView &t, uint64_t k, int lo,
                                          int hi) {
  if (t.dir_exact) return lo;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.key[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int locate_key(const TreeView &t, uint64_t k) {
  return dir_search(t, k, dir_first(t, k), dir_last(t, k));
}

__device__ __forceinline__ int locate(const TreeView &t, double x, double y,
                                      double z) {
  return locate_key(t, point_key(x, y, z));
}

// Map a point into leaf coordinates: xi = ((x - c) * 2.0) * s - 1.0 with
// c the leaf corner and s = 2^depth (tree_functor.h:283-295).
__device__ __forceinline__ void leaf_xi(const TreeView &t, int leaf, double x,
                                        double y, double z, double &ex,
                                        double &ey, double &ez) {
  const uint4 inf = t.info[leaf];
  const double s = (double)(1u << inf.w);
  const double len = 1.0 / s;
  const double cx = (double)inf.x * len;
  const double cy = (double)inf.y * len;
  const double cz = (double)inf.z * len;
  ex = (x - cx) * 2.0 * s - 1.0;
  ey = (y - cy) * 2.0 * s - 1.0;
  ez = (z - cz) * 2.0 * s - 1.0;
}

// T_0..T_{D-1}(x) by T_{k+1} = (2x) T_k - T_{k-1}, evaluated in FP64 and
// stored as T; |x| > 1 (or NaN) gives all zeros (pvfmm::cheb_poly, called at
// tree_functor.h:328-330).
template <int D, typename T>
__device__ __forceinline__ void cheb_poly(double x, T *p) {
  const bool in = fabs(x) <= 1.0;
  const double xx = in ? x : 0.0;
  double y0 = in ? 1.0 : 0.0;
  p[0] = (T)y0;
  if constexpr (D > 1) p[1] = (T)xx;
  double y1 = xx;
#pragma unroll
  for (int k = 2; k < D; ++k) {
    const double y2 = 2.0 * xx * y1 - y0;
    p[k] = (T)y2;
    y0 = y1;
    y1 = y2;
  }
}

// Triangular tensor contraction of one leaf and one component
// (vec_eval, tree_functor.h:38-77): u = sum_i pz_i (sum_j py_j
// (sum_k px_k c_ijk)), i+j+k < D, coefficients in packed order.
//
// The coefficients are wave-uniform and stream through SGPRs: the packed
// triangle is read as 64-byte vectors from the constant address space
// (s_load_dwordx16), two per block of 16 (FP64) or 32 (FP32) coefficients.
// Empty asm statements order the stream without touching it: the fence at
// the start of block b+1 names block b+1's registers (so the compiler's own
// wait for them lands there, after block b's FMAs), the running sums (so
// block b's FMAs cannot move past it) and the block pointer (so block b+2's
// loads, addressed from it, cannot be hoisted above it); a scheduling
// barrier after those loads keeps them ahead of block b+1's FMAs.  Left alone, the
// compiler hoists every load of the unrolled contraction to its top and
// spills the SGPRs into VGPR lanes.  A block may read up to 124 bytes past
// the coefficient array; the tree handle pads its buffers for that.
template <typename T>
struct Chunk;
template <>
struct Chunk<double> {
  typedef double V __attribute__((ext_vector_type(8), aligned(8)));
  static constexpr int N = 8;
};
template <>
struct Chunk<float> {
  typedef float V __attribute__((ext_vector_type(16), aligned(8)));
  static constexpr int N = 16;
};

// volatile: a plain constant-address-space load is "invariant" and the
// register allocator may rematerialise it at its use, which puts every
// block's load right before its FMAs and waits a full scalar-load round
// trip per block; a volatile load stays where the schedule put it, one
// block ahead.
template <typename T>
__device__ __forceinline__ typename Chunk<T>::V load_chunk(const T *p) {
  typedef typename Chunk<T>::V V;
  return *(const volatile __attribute__((address_space(4))) V *)(p);
}

template <typename V, typename P>
__device__ __forceinline__ void fence_block(V &a, V &b, P &ptr) {
  asm volatile("" : "+s"(a), "+s"(b), "+s"(ptr));
}

template <typename V, typename P, typename T>
__device__ __forceinline__ void fence_block(V &a, V &b, P &ptr, T &w, T &v, T &u) {
  asm volatile("" : "+s"(a), "+s"(b), "+s"(ptr), "+v"(w), "+v"(v), "+v"(u));
}

// T_i(z) for the outer index is generated inside the contraction by the
// recurrence of cheb_poly, with the same operations, from zx (the clamped
// coordinate, 0 outside [-1,1]) and z0 (T_0: 1, or 0 outside): the values
// are those cheb_poly would store, and the 3 x D basis array shrinks to 2 x D.
template <int D, int MODE, typename T>
__device__ __forceinline__ T contract(const T *c, const T *px, const T *py,
                                      double zx, double z0) {
  typedef typename Chunk<T>::V V;
  constexpr int CN = Chunk<T>::N;  // coefficients per chunk
  constexpr int BN = 2 * CN;       // coefficients per block
  constexpr int NC = D * (D + 1) * (D + 2) / 6;
  constexpr int NB = (NC + BN - 1) / BN;
  const T *bp = c;
  V cur0 = load_chunk(bp), cur1 = load_chunk(bp + CN);
  fence_block(cur0, cur1, bp);
  V nxt0 = cur0, nxt1 = cur1;
  if constexpr (NB > 1) {
    nxt0 = load_chunk(bp + BN);
    nxt1 = load_chunk(bp + BN + CN);
    __builtin_amdgcn_sched_barrier(0);  // issue them before block 0's FMAs
  }
  T u = T(0);
  int idx = 0;
  const double z2 = 2.0 * zx;
  double zm = z0, zc = zx;  // T_{i-1}, T_i for i >= 1
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double zi;
    if (i == 0) {
      zi = z0;
    } else if (i == 1) {
      zi = zx;
    } else {
      zi = z2 * zc - zm;  // (2.0 * xx) * y1 - y0, as cheb_poly
      zm = zc;
      zc = zi;
    }
    T v = T(0);
#pragma unroll
    for (int j = 0; i + j < D; ++j) {
      T w = T(0);
#pragma unroll
      for (int k = 0; i + j + k < D; ++k) {
        if (idx % BN == 0 && idx > 0) {
          // block idx/BN arrives here, after the previous block's FMAs
          fence_block(nxt0, nxt1, bp, w, v, u);
          cur0 = nxt0;
          cur1 = nxt1;
          bp = bp + BN;
          if (idx / BN + 1 < NB) {
            nxt0 = load_chunk(bp + BN);
            nxt1 = load_chunk(bp + BN + CN);
            __builtin_amdgcn_sched_barrier(0);  // before this block's FMAs
          }
        }
        const int r = idx % BN;
        const T ck = r < CN ? cur0[r] : cur1[r - CN];
        if (MODE != kFp32h || i + j + k > kHeadDeg)  // kFp32h: head in FP64
          w = madd<MODE>(px[k], ck, w);
        ++idx;
      }
      v = madd<MODE>(py[j], w, v);
    }
    u = madd<MODE>((T)zi, v, u);
  }
  return u;
}

template <int D>
constexpr int ncoef() {
  return D * (D + 1) * (D + 2) / 6;
}

template <int MODE>
__device__ __forceinline__ const typename ModeType<MODE>::T *coeff_base(
    const TreeView &t) {
  if constexpr (MODE == kFp32 || MODE == kFp32h)
    return t.coeff32;
  else
    return t.coeff;
}

// Inside a waterfall branch every active lane holds the same leaf.  Reading
// it again with readfirstlane there (instead of reusing the value compared
// against) keeps the address wave-uniform: the compiler would otherwise
// substitute the lane's own, divergent, copy of the equal value.
__device__ __forceinline__ int uniform_leaf(int leaf) {
  return __builtin_amdgcn_readfirstlane(leaf);
}

// Waterfall over the distinct leaves held by the active lanes: body(lf) runs
// once per distinct leaf, with exactly the lanes holding lf active and lf
// wave-uniform.  The loop is driven by a ballot mask of pending lanes, which
// shrinks every pass; a loop that only re-tests readfirstlane(leaf) == leaf
// looks, to the compiler, like one that either exits at once or never
// exits, and is folded to a single pass.
template <typename Body>
__device__ __forceinline__ void for_each_wave_leaf(int leaf, Body body) {
  uint64_t pending = __builtin_amdgcn_ballot_w64(true);
  while (pending) {
    const int src = __builtin_ctzll(pending);
    const int lf = __builtin_amdgcn_readlane(leaf, src);
    const bool mine = (leaf == lf);
    pending &= ~__builtin_amdgcn_ballot_w64(mine);
    if (mine) body(uniform_leaf(leaf));
  }
}

// Per-lane leaf location and basis values for one point.
template <int D, int MODE>
struct PointBasis {
  using T = typename ModeType<MODE>::T;
  int leaf;
  T px[D], py[D];
  double zx, z0;               // T_i(z) source for contract
  double hx[3], hy[3], hz[3];  // kFp32h: T_0..T_2 in FP64

  // hint: the point's leaf when the caller has located it already, else -1
  __device__ __forceinline__ void build(const TreeView &t, double x, double y,
                                        double z, int hint = -1) {
    leaf = hint >= 0 ? hint : locate(t, x, y, z);
    double ex, ey, ez;
    leaf_xi(t, leaf, x, y, z, ex, ey, ez);
    cheb_poly<D>(ex, px);
    cheb_poly<D>(ey, py);
    const bool zin = fabs(ez) <= 1.0;  // cheb_poly's range test
    zx = zin ? ez : 0.0;
    z0 = zin ? 1.0 : 0.0;
    if constexpr (MODE == kFp32h) {
      cheb_poly<3>(ex, hx);
      cheb_poly<3>(ey, hy);
      cheb_poly<3>(ez, hz);
    }
  }
};

// Packed index of coefficient (i, j, k) in a degree D - 1 triangle: the
// i(z) -> j(y) -> k(x) order of vec_eval (tree_functor.h:258-266).
__host__ __device__ constexpr int tri_index(int D, int i, int j, int k) {
  int idx = 0;
  for (int a = 0; a < i; ++a) idx += (D - a) * (D - a + 1) / 2;
  for (int b = 0; b < j; ++b) idx += D - i - b;
  return idx + k;
}

template <typename T>
__device__ __forceinline__ T load_uniform(const T *p) {
  return *(const __attribute__((address_space(4))) T *)(p);
}

// The head of one leaf's contraction in FP64: the terms with
// i + j + k <= kHeadDeg, nested as the full contraction (k, then j, then i),
// from the FP64 coefficients (wave-uniform: scalar loads).  load() reads
// them before the FP32 contraction starts: the loads are volatile, so they
// stay ahead of the contraction's own (volatile) coefficient stream and
// their latency hides behind it instead of following it.
template <int D>
struct Head {
  static constexpr int H = kHeadDeg < D - 1 ? kHeadDeg : D - 1;
  static constexpr int N = (H + 1) * (H + 2) * (H + 3) / 6;
  double c[N];

  __device__ __forceinline__ void load(const double *base) {
    int n = 0;
#pragma unroll
    for (int i = 0; i <= H; ++i)
#pragma unroll
      for (int j = 0; i + j <= H; ++j)
#pragma unroll
        for (int k = 0; i + j + k <= H; ++k)
          c[n++] = *(const volatile __attribute__((address_space(4))) double *)(
              base + tri_index(D, i, j, k));
  }

  __device__ __forceinline__ double eval(const double *hx, const double *hy,
                                         const double *hz) const {
    double u = 0.0;
    int n = 0;
#pragma unroll
    for (int i = 0; i <= H; ++i) {
      double v = 0.0;
#pragma unroll
      for (int j = 0; i + j <= H; ++j) {
        double w = 0.0;
#pragma unroll
        for (int k = 0; i + j + k <= H; ++k) w = __builtin_fma(hx[k], c[n++], w);
        v = __builtin_fma(hy[j], w, v);
      }
      u = __builtin_fma(hz[i], v, u);
    }
    return u;
  }
};

// One component of one leaf at the lane's point, in the mode's arithmetic.
template <int D, int MODE>
__device__ __forceinline__ double leaf_value(const TreeView &t, int lf, int comp,
                                             const PointBasis<D, MODE> &b) {
  constexpr int NC = D * (D + 1) * (D + 2) / 6;
  const size_t off = ((size_t)lf * t.dof + comp) * NC;
  if constexpr (MODE == kFp32h) {
    Head<D> h;
    h.load(t.coeff + off);
    const double u = (double)contract<D, MODE>(coeff_base<MODE>(t) + off, b.px, b.py, b.zx, b.z0);
    return u + h.eval(b.hx, b.hy, b.hz);
  } else {
    return (double)contract<D, MODE>(coeff_base<MODE>(t) + off, b.px, b.py, b.zx, b.z0);
  }
}

// Velocity (dof 3) at one point: the waterfall over the wave's leaves.
template <int D, int MODE>
__device__ __forceinline__ void eval_vec3(const TreeView &t, double x,
                                          double y, double z, double &v0,
                                          double &v1, double &v2, int hint = -1) {
  PointBasis<D, MODE> b;
  b.build(t, x, y, z, hint);
  // the waves of a workgroup start at different components: after a
  // regrouping they reach the same leaf together, and in the same order
  // they would all miss on the same coefficient lines at once
  const int l0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) % 3;
  for_each_wave_leaf(b.leaf, [&](int lf) {
    // one component at a time: unrolled side by side, the three
    // contractions' scalar loads exhaust the SGPRs and spill
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
      const int l = l0 + k < 3 ? l0 + k : l0 + k - 3;
      const double u = leaf_value<D, MODE>(t, lf, l, b);
      v0 = l == 0 ? u : v0;
      v1 = l == 1 ? u : v1;
      v2 = l == 2 ? u : v2;
    }
  });
}

// All dof components at one point, each written through store(l, value).
template <int D, int MODE, typename Store>
__device__ __forceinline__ void eval_dof(const TreeView &t, double x, double y,
                                         double z, Store store, int hint = -1) {
  PointBasis<D, MODE> b;
  b.build(t, x, y, z, hint);
  const int dof = t.dof;
  for_each_wave_leaf(b.leaf, [&](int lf) {
    for (int l = 0; l < dof; ++l) store(l, leaf_value<D, MODE>(t, lf, l, b));
  });
}

// CubicInterpPolicy::InterpCubic1D (cubic.h:42-56) with its tangent
// (cubic.h:36-39) and Hermite basis (cubic.h:28-34), in the reference's
// operation order.
__device__ __forceinline__ double cubic1d(double x, const double *xx, double p0,
                                          double p1, double p2, double p3) {
  const double mk = (p2 - p1) * 0.5 / (xx[2] - xx[1]) + (p1 - p0) * 0.5 / (xx[1] - xx[0]);
  const double mk1 = (p3 - p2) * 0.5 / (xx[3] - xx[2]) + (p2 - p1) * 0.5 / (xx[2] - xx[1]);
  const double t = (x - xx[1]) / (xx[2] - xx[1]);
  const double h00 = 2 * t * t * t - 3 * t * t + 1;
  const double h10 = t * t * t - 2 * t * t + t;
  const double h01 = -2 * t * t * t + 3 * t * t;
  const double h11 = t * t * t - t * t;
  return h00 * p1 + h10 * (xx[2] - xx[1]) * mk + h01 * p2 + h11 * (xx[2] - xx[1]) * mk1;
}

// The velocity functor at one point for RK stage `stage` (1 or 2) at time
// ts.  A single tree is contracted inline; several trees go through one
// rolled loop (one contraction body) and are combined as the reference's
// FieldExtrapFunctor (1.5 c - 0.5 p, tree_extrap_functor.h:71-77) or
// FieldSetFunctor (cubic Hermite in time, tree_set_functor.h:66-72).
// Those functors evaluate their trees one NodeFieldFunctor after another on
// the same array, and each EvalTree wraps that array in place
// (tree_functor.h:442-449): the point is wrapped again before every tree
// after the first (the caller wrapped it for the first), and the caller's
// point ends up wrapped once per tree.  This matters only for points more
// than one period outside the domain.
// KIND is the launch's vs.kind as a compile-time constant: the composite
// kinds keep up to four trees' values live, which a single-tree kernel must
// not pay for in registers.
template <int D, int MODE, int KIND>
__device__ __forceinline__ void eval_velocity(const VelSource &vs, int stage, double ts,
                                              int scaled, double g, int bc, double &x,
                                              double &y, double &z, double &v0, double &v1,
                                              double &v2, int hint0 = -1) {
  if constexpr (KIND == kVelTree) {
    eval_vec3<D, MODE>(vs.t[0], x, y, z, v0, v1, v2, hint0);
    if (scaled) {
      v0 = v0 * g;
      v1 = v1 * g;
      v2 = v2 * g;
    }
  } else {
    const int nt = KIND == kVelExtrap ? (stage == 1 ? 1 : 2) : 4;
    double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    double c0 = 0, c1 = 0, c2 = 0, d0 = 0, d1 = 0, d2 = 0;
#pragma unroll 1
    for (int k = 0; k < nt; ++k) {
      if (k > 0 && bc == TBX_BC_PERIODIC) {
        x = wrap1(x);
        y = wrap1(y);
        z = wrap1(z);
      }
      const TreeView &tk = k == 0 ? vs.t[0] : k == 1 ? vs.t[1] : k == 2 ? vs.t[2] : vs.t[3];
      double r0, r1, r2;
      eval_vec3<D, MODE>(tk, x, y, z, r0, r1, r2, k == 0 ? hint0 : -1);
      if (k == 0) {
        a0 = r0; a1 = r1; a2 = r2;
      } else if (k == 1) {
        b0 = r0; b1 = r1; b2 = r2;
      } else if (k == 2) {
        c0 = r0; c1 = r1; c2 = r2;
      } else {
        d0 = r0; d1 = r1; d2 = r2;
      }
    }
    if constexpr (KIND == kVelExtrap) {
      if (stage == 1) {
        v0 = a0; v1 = a1; v2 = a2;
      } else {
        const double ccoeff = 3.0 / 2, pcoeff = 0.5;
        v0 = ccoeff * a0 - pcoeff * b0;
        v1 = ccoeff * a1 - pcoeff * b1;
        v2 = ccoeff * a2 - pcoeff * b2;
      }
    } else {
      v0 = cubic1d(ts, vs.tt, a0, b0, c0, d0);
      v1 = cubic1d(ts, vs.tt, a1, b1, c1, d1);
      v2 = cubic1d(ts, vs.tt, a2, b2, c2, d2);
    }
  }
}

// One IntegrateRK2 substep (traj.inc:20-46) of one point, in place.  The
// first evaluation wraps x itself (EvalTree mutates the caller's array,
// and ComputeTrajRK2 passes xsol as xinit); the midpoint is wrapped only
// for its evaluation.
template <int D, int MODE, int KIND>
__device__ __forceinline__ void rk2_finish(const VelSource &vel, int bc, double htau,
                                           double tau, int scaled, double g2, double ts2,
                                           double &x, double &y, double &z, double v0,
                                           double v1, double v2);

template <int D, int MODE, int KIND>
__device__ __forceinline__ void rk2_substep(const VelSource &vel, int bc,
                                            double htau, double tau, int scaled,
                                            double g1, double g2, double ts1,
                                            double ts2, double &x, double &y,
                                            double &z) {
  if (bc == TBX_BC_PERIODIC) {
    x = wrap1(x);
    y = wrap1(y);
    z = wrap1(z);
  }
  double v0, v1, v2;
  eval_velocity<D, MODE, KIND>(vel, 1, ts1, scaled, g1, bc, x, y, z, v0, v1, v2);
  rk2_finish<D, MODE, KIND>(vel, bc, htau, tau, scaled, g2, ts2, x, y, z, v0, v1, v2);
}

// The substep's second stage, given the first stage's velocity at (x, y, z).
template <int D, int MODE, int KIND>
__device__ __forceinline__ void rk2_finish(const VelSource &vel, int bc, double htau,
                                           double tau, int scaled, double g2, double ts2,
                                           double &x, double &y, double &z, double v0,
                                           double v1, double v2) {
  // xtmp = xinit + (0.5 * dt) * vtmp   (traj.inc:36)
  double hx = x + htau * v0, hy = y + htau * v1, hz = z + htau * v2;
  if (bc == TBX_BC_PERIODIC) {
    hx = wrap1(hx);
    hy = wrap1(hy);
    hz = wrap1(hz);
  }
  eval_velocity<D, MODE, KIND>(vel, 2, ts2, scaled, g2, bc, hx, hy, hz, v0, v1, v2);
  // xsol = xinit + dt * vtmp   (traj.inc:41-42)
  x = x + tau * v0;
  y = y + tau * v1;
  z = z + tau * v2;
}

// ---------------------------------------------------------------------------
// The first RK2 stage at an arrival grid node.
//
// A step's first substep starts at the arrival points, which are the grid
// nodes corner + len * node[a] of the concentration's leaves; the first
// stage evaluates the velocity there.  Where the velocity leaf that the
// (wrapped) node is located in has the arrival leaf's depth, the node is one
// of that leaf's own grid nodes, and its velocity is the velocity tree's
// polynomial expanded at its nodes (tensor-product Chebyshev values, computed
// once per step for all leaves) instead of a contraction.  That covers every
// node of a leaf whose velocity leaf is the same cell, and the upper-face
// nodes (node[D-1] = 1 up to rounding) that lie on the lower face of an
// equal-sized neighbour (its node 0).
// ---------------------------------------------------------------------------

// The precision of the nodal velocities a step of arithmetic mode MODE reads.
template <int MODE>
using NodalT = typename std::conditional<MODE == kFp32 || MODE == kFp32h, float, double>::type;

// Node index of arrival node m of a leaf `own` (D nodes per axis, the first
// at the leaf's lower face and the last at its upper face up to rounding)
// in the leaf `o` it is located in, or -1 when o is of another depth.  Per axis, the anchors at
// the common depth differ by 0 (same node), +1 (own's upper face: o's node
// 0) or -1 (own's lower face rounded below: o's node D-1), modulo the
// periodic wrap.
__device__ __forceinline__ int node_of(uint4 own, uint4 o, int m, int D) {
  if (o.w != own.w) return -1;
  const uint32_t mask = (1u << own.w) - 1u;
  const uint32_t ov[3] = {o.x, o.y, o.z}, nv[3] = {own.x, own.y, own.z};
  int r = 0, mul = 1, rest = m;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ak = rest % D;
    rest /= D;
    const uint32_t d = (ov[k] - nv[k]) & mask;
    // a single leaf (depth 0) is its own neighbour across the periodic
    // wrap, so an upper-face node may stand for node 0: left to evaluation
    if (mask == 0 && ak == D - 1) return -1;
    int bk;
    if (d == 0)
      bk = ak;
    else if (d == 1 && ak == D - 1)
      bk = 0;
    else if (d == mask && ak == 0)
      bk = D - 1;
    else
      return -1;
    r += bk * mul;
    mul *= D;
  }
  return r;
}

// v = g * (velocity leaf lf's values at node node_of(own, lf, m)) from
// vnode [L][3][P] (the scaling as in eval_velocity); false when m is not a
// node of lf.
template <int D, typename VT>
__device__ __forceinline__ bool nodal_velocity(const VT *__restrict__ vnode,
                                               const TreeView &vt, uint4 own, int m, int lf,
                                               int P, int scaled, double g, double &v0,
                                               double &v1, double &v2) {
  const int mm = node_of(own, vt.info[lf], m, D);
  if (mm < 0) return false;
  const VT *p = vnode + (int64_t)lf * 3 * P + mm;
  v0 = (double)p[0];
  v1 = (double)p[P];
  v2 = (double)p[2 * P];
  if (scaled) {
    v0 = v0 * g;
    v1 = v1 * g;
    v2 = v2 * g;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Regrouping a workgroup's points by leaf.
//
// Consecutive arrival points lie mostly in one leaf, but the grid points on
// a leaf's upper faces belong to the neighbours (half-open boxes) and the
// departure points near any face scatter across it, so a wavefront of 64
// consecutive points touches about 2.7 distinct leaves and the waterfall
// repeats the contraction that often.  Before each evaluation the
// workgroup therefore permutes its points through LDS so that the points of
// one leaf are contiguous: leaves get slots in a small open-addressing hash
// table, slots get counts and offsets, and each point moves to offset +
// rank.  A wavefront then spans one leaf except where two slots meet.  The
// order inside a slot is arbitrary; results do not depend on it, because
// each point's arithmetic is independent of the lane that does it.
// ---------------------------------------------------------------------------
constexpr int kRgSlots = 64;  // hash slots; one more collects any overflow

template <int BLOCK, int NF>
struct RegroupLds {
  int key[kRgSlots];
  int cnt[kRgSlots + 1];
  int orig[BLOCK];
  int leaf[BLOCK];
  double f[NF][BLOCK];
};

// New position of this thread's item, items of equal key contiguous.
// Every thread of the workgroup must call it (it synchronises the block).
template <int BLOCK, int NF>
__device__ __forceinline__ int regroup_pos(RegroupLds<BLOCK, NF> &s, int key) {
  const int tid = threadIdx.x;
  if (tid < kRgSlots) s.key[tid] = -1;
  if (tid <= kRgSlots) s.cnt[tid] = 0;
  __syncthreads();
  uint32_t h = ((uint32_t)key * 0x9E3779B1u) >> 26;  // 6 bits: kRgSlots == 64
  int slot = kRgSlots;
#pragma unroll 1
  for (int p = 0; p < kRgSlots; ++p) {
    const int prev = atomicCAS(&s.key[h], -1, key);
    if (prev == -1 || prev == key) {
      slot = (int)h;
      break;
    }
    h = (h + 1) & (kRgSlots - 1);
  }
  const int rank = atomicAdd(&s.cnt[slot], 1);
  __syncthreads();
  if (tid < 64) {  // exclusive scan of the 64 slot counts by wave 0
    const int c = s.cnt[tid];
    int incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(incl, d, 64);
      if (tid >= d) incl += o;
    }
    s.cnt[tid] = incl - c;
    if (tid == 63) s.cnt[kRgSlots] = incl;  // the overflow slot goes last
  }
  __syncthreads();
  return s.cnt[slot] + rank;
}

// Moves (orig, leaf, f[0..nf)) of every thread to its regrouped position
// and returns the item now at this thread's position.
template <int BLOCK, int NF, int NUSE>
__device__ __forceinline__ void regroup(RegroupLds<BLOCK, NF> &s, int &orig, int &leaf,
                                        double (&f)[NUSE]) {
  static_assert(NUSE <= NF, "payload too wide");
  const int pos = regroup_pos(s, leaf);
  s.orig[pos] = orig;
  s.leaf[pos] = leaf;
#pragma unroll
  for (int k = 0; k < NUSE; ++k) s.f[k][pos] = f[k];
  __syncthreads();
  const int tid = threadIdx.x;
  orig = s.orig[tid];
  leaf = s.leaf[tid];
#pragma unroll
  for (int k = 0; k < NUSE; ++k) f[k] = s.f[k][tid];
}

}  // namespace tbx
