// Adjoint of one SolveSemilagInSitu step for one Chebyshev degree
// (compiled once per degree with -DTBX_Q=<q>, like tbx_kernels.hip).
//
// The forward step is c' = R E_D(V) c: E_D interpolates c at the departure
// points D(V) of the grid points (RK2 through the velocity tree V), R is the
// per-leaf refit.  For a cotangent lam' of c':
//   dc = E_D^T R^T lam'          transposed interpolation: every departure
//                                point adds u_p B_m(D_p) into the
//                                coefficients of its leaf (u = R^T lam')
//   dV = sum_p a_p . dD_p/dV     a_p = sum_c u_{p,c} grad c_c(D_p), chained
//                                through the midpoint RK2 substep
// R^T lam' is a GEMM on the host side.  The reference has no adjoint; these
// kernels differentiate exactly what the forward kernels compute (the same
// trajectory arithmetic, the same leaf assignment), for one RK2 substep and
// a single velocity tree.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tbx_device.h"

#ifndef TBX_Q
#error "compile with -DTBX_Q=<degree>"
#endif

namespace tbx {
namespace {

constexpr int D = TBX_Q + 1;
constexpr int NC = D * (D + 1) * (D + 2) / 6;

// T_0..T_{D-1} and dT/dxi by differentiating the three-term recurrence,
// dT_{k+1} = 2 T_k + 2 xi dT_k - dT_{k-1}; zeros outside [-1,1] like
// cheb_poly.
__device__ __forceinline__ void cheb_poly_d(double x, double *p, double *dp) {
  const bool in = fabs(x) <= 1.0;
  const double xx = in ? x : 0.0;
  p[0] = in ? 1.0 : 0.0;
  dp[0] = 0.0;
  if constexpr (D > 1) {
    p[1] = xx;
    dp[1] = p[0];
  }
#pragma unroll
  for (int k = 2; k < D; ++k) {
    p[k] = 2.0 * xx * p[k - 1] - p[k - 2];
    dp[k] = 2.0 * p[k - 1] + 2.0 * xx * dp[k - 1] - dp[k - 2];
  }
}

// Basis and its derivatives at one point, in the point's leaf.
struct AdjBasis {
  int leaf;
  double s2;  // d xi / dx = 2 * 2^depth
  double tx[D], ty[D], tz[D], dx[D], dy[D], dz[D];

  __device__ __forceinline__ void build(const TreeView &t, double x, double y, double z,
                                        int hint = -1) {
    leaf = hint >= 0 ? hint : locate(t, x, y, z);
    double ex, ey, ez;
    leaf_xi(t, leaf, x, y, z, ex, ey, ez);
    s2 = 2.0 * (double)(1u << t.info[leaf].w);
    cheb_poly_d(ex, tx, dx);
    cheb_poly_d(ey, ty, dy);
    cheb_poly_d(ez, tz, dz);
  }
};

// Gradient (d/dx, d/dy, d/dz) of one component at the point of b, c the
// component's coefficients in b's leaf.  One pass over the packed triangle:
// every coefficient feeds the x-derivative and the value sums along k, and
// the (i, j) partial sums are then combined with the y and z bases and
// derivatives.
__device__ __forceinline__ void grad3(const double *c, const AdjBasis &b, double &gx,
                                      double &gy, double &gz) {
  double ux = 0.0, uy = 0.0, uz = 0.0;
  int idx = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double vx = 0.0, vy = 0.0, v = 0.0;
#pragma unroll
    for (int j = 0; i + j < D; ++j) {
      double wx = 0.0, w = 0.0;
#pragma unroll
      for (int k = 0; i + j + k < D; ++k) {
        const double cm = load_uniform(c + idx++);
        wx = __builtin_fma(b.dx[k], cm, wx);
        w = __builtin_fma(b.tx[k], cm, w);
      }
      vx = __builtin_fma(b.ty[j], wx, vx);
      vy = __builtin_fma(b.dy[j], w, vy);
      v = __builtin_fma(b.ty[j], w, v);
    }
    ux = __builtin_fma(b.tz[i], vx, ux);
    uy = __builtin_fma(b.tz[i], vy, uy);
    uz = __builtin_fma(b.dz[i], v, uz);
  }
  gx = ux * b.s2;
  gy = uy * b.s2;
  gz = uz * b.s2;
}

// Cross-lane exchanges of one double.  swap_add32 / swap_add16: with lo
// and hi two blocks of values, lanes whose bit 5 (bit 4) is clear return
// lo + lo of lane l ^ 32 (l ^ 16) and the others hi + hi of that lane
// (v_permlane32_swap / v_permlane16_swap exchange the upper lanes of the
// first operand with the lower lanes of the second).  dpp_pair: the value
// of the partner lane of a DPP pattern (row_ror:8 = l ^ 8 in a row,
// row_half_mirror = 7 - l in each eight, quad_perm = l ^ 2, l ^ 1).
__device__ __forceinline__ double swap_add32(double lo, double hi) {
  const auto a = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(lo),
                                                  (unsigned)__double2loint(hi), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(lo),
                                                  (unsigned)__double2hiint(hi), false, false);
  return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
}

__device__ __forceinline__ double swap_add16(double lo, double hi) {
  const auto a = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(lo),
                                                  (unsigned)__double2loint(hi), false, false);
  const auto b = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(lo),
                                                  (unsigned)__double2hiint(hi), false, false);
  return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
}

template <int CTRL>
__device__ __forceinline__ double dpp_pair(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// One halving step over b[0, 2H): the lane keeps the upper half when
// `upper`, sends the other half to its partner and adds what it receives.
template <int CTRL, int H, int N>
__device__ __forceinline__ void halve(double (&b)[N], bool upper) {
#pragma unroll
  for (int t = 0; t < H; ++t) {
    const double send = upper ? b[t] : b[t + H];
    const double keep = upper ? b[t + H] : b[t];
    b[t] = keep + dpp_pair<CTRL>(send);
  }
}

// Reduce-scatter over the wave (all 64 lanes active): on return lane l
// holds the sum over all lanes of b[l] (N = 64), or, for N = 32, lanes l
// and l ^ 1 both hold the sum of b[l >> 1].  Recursive halving: at every
// step the two partners hold the same block of indices and each keeps one
// half, so N - 1 exchanges replace N full reductions.
template <int N>
__device__ __forceinline__ double wave_reduce_scatter(double (&b)[N]) {
  static_assert(N == 64 || N == 32, "chunk of 32 or 64");
  const unsigned l = threadIdx.x & 63u;
  constexpr int H1 = N / 2, H2 = N / 4, H3 = N / 8, H4 = N / 16, H5 = N / 32;
#pragma unroll
  for (int t = 0; t < H1; ++t) b[t] = swap_add32(b[t], b[t + H1]);
#pragma unroll
  for (int t = 0; t < H2; ++t) b[t] = swap_add16(b[t], b[t + H2]);
  halve<0x128, H3>(b, (l & 8u) != 0);
  halve<0x141, H4>(b, (l & 4u) != 0);
  halve<0x4e, H5>(b, (l & 2u) != 0);
  if constexpr (N == 64)
    halve<0xb1, 1>(b, (l & 1u) != 0);
  else
    b[0] = b[0] + dpp_pair<0xb1>(b[0]);
  return b[0];
}

// (i, j, k) of packed triangle index m (order i(z) -> j(y) -> k(x)).
struct Tri {
  int i[NC], j[NC], k[NC];
  constexpr Tri() : i(), j(), k() {
    int m = 0;
    for (int a = 0; a < D; ++a)
      for (int b = 0; a + b < D; ++b)
        for (int c = 0; a + b + c < D; ++c) {
          i[m] = a;
          j[m] = b;
          k[m] = c;
          ++m;
        }
  }
};
constexpr Tri kTri{};

// Transposed interpolation of ncomp weighted components: for every leaf
// held by a lane of the wave, G[leaf][c][m] += sum over its lanes of
// w(c) B_m(point).  All lanes of the wave must be active; each lane's basis
// b is in its own leaf.  The coefficients go in chunks of CH (64 or 32:
// CH doubles of registers per lane) through the wave reduce-scatter, one
// coalesced atomic add per lane and chunk.
template <int CH, typename W>
__device__ __forceinline__ void scatter_transposed(double *G, int ncomp, const AdjBasis &b,
                                                   W weight) {
  constexpr int PER = 64 / CH;  // lanes per reduced coefficient
  const int lane = (int)(threadIdx.x & 63u);
  const bool writer = (lane % PER) == 0;
  const int slot = lane / PER;
  uint64_t pending = __builtin_amdgcn_ballot_w64(true);
  while (pending) {
    const int lf = __builtin_amdgcn_readlane(b.leaf, __builtin_ctzll(pending));
    const bool mine = b.leaf == lf;
    pending &= ~__builtin_amdgcn_ballot_w64(mine);
#pragma unroll 1
    for (int c = 0; c < ncomp; ++c) {
      const double wc = mine ? weight(c) : 0.0;
      double *g = G + ((size_t)lf * ncomp + c) * NC;
#pragma unroll
      for (int m0 = 0; m0 < NC; m0 += CH) {
        double buf[CH];
#pragma unroll
        for (int r = 0; r < CH; ++r) {
          const int m = m0 + r;
          buf[r] = m < NC ? wc * b.tz[kTri.i[m < NC ? m : 0]] * b.ty[kTri.j[m < NC ? m : 0]] *
                                b.tx[kTri.k[m < NC ? m : 0]]
                          : 0.0;
        }
        const double s = wave_reduce_scatter(buf);
        if (writer && m0 + slot < NC) atomicAdd(g + m0 + slot, s);
      }
    }
  }
}

// Chunk of the concentration kernel's scatter: 64 while the basis arrays
// leave room for it, 32 above degree 10.
constexpr int kConcChunk = D <= 11 ? 64 : 32;

// Arrival point of global grid index gi (tree_utils.h:480-491).
__device__ __forceinline__ void arrival_point(const TreeView &conc, const StepArgs &a,
                                              int64_t gi, double &x, double &y, double &z) {
  const int64_t leaf = gi / a.P;
  const int m = (int)(gi - leaf * a.P);
  const uint4 inf = conc.info[leaf];
  const double len = 1.0 / (double)(1u << inf.w);
  x = (double)inf.x * len + len * a.node[m % D];
  y = (double)inf.y * len + len * a.node[(m / D) % D];
  z = (double)inf.z * len + len * a.node[m / (D * D)];
}

// Cotangent of c and, when A is given, the per-point velocity cotangent
// factor a_p += sum_c u_{p,c} grad c_c(D_p).  ua[al][c][am] = (R^T lam')
// at arrival point (al, am).  MODE is the forward's trajectory arithmetic;
// with STORED the departure points are read from dep[3 gi] (the forward
// step's own output, wrapped as the concentration saw them) instead of
// being integrated again.
template <int MODE, int BLOCK, bool STORED>
__global__ __launch_bounds__(BLOCK) void k_adj_conc(VelSource vel, TreeView conc, StepArgs a,
                                                    const double *__restrict__ dep,
                                                    const double *__restrict__ ua,
                                                    double *__restrict__ lam,
                                                    double *__restrict__ A) {
  __shared__ RegroupLds<BLOCK, 3> sh;
  const uint32_t b = blockIdx.x;
  const int64_t lb = (int64_t)(b & 7u) * a.per_xcd + (int64_t)(b >> 3);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  const int64_t i0 = lb * BLOCK;
  int orig = (int)threadIdx.x;
  double p3[3];
  {
    const int64_t i = i0 + orig < a.npts ? i0 + orig : a.npts - 1;
    if constexpr (STORED) {
      p3[0] = dep[3 * i + 0];
      p3[1] = dep[3 * i + 1];
      p3[2] = dep[3 * i + 2];
    } else {
      double x, y, z;
      arrival_point(conc, a, i, x, y, z);
#pragma unroll 1
      for (int s = 0; s < a.nsub; ++s)
        rk2_substep<D, MODE, kVelTree>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[2 * s],
                                       a.g[2 * s + 1], a.ts[2 * s], a.ts[2 * s + 1], x, y, z);
      if (a.bc == TBX_BC_PERIODIC) {
        x = wrap1(x);
        y = wrap1(y);
        z = wrap1(z);
      }
      p3[0] = x;
      p3[1] = y;
      p3[2] = z;
    }
  }
  int lf = locate(conc, p3[0], p3[1], p3[2]);
  regroup(sh, orig, lf, p3);
  const int64_t gi = i0 + orig;
  const bool live = gi < a.npts;
  const int64_t al = live ? gi / a.P : 0;
  const int am = live ? (int)(gi - al * a.P) : 0;
  const int dof = conc.dof;
  AdjBasis bs;
  bs.build(conc, p3[0], p3[1], p3[2], lf);
  auto weight = [&](int c) { return live ? ua[(al * dof + c) * a.P + am] : 0.0; };
  if (A) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    for_each_wave_leaf(bs.leaf, [&](int leaf) {
      for (int c = 0; c < dof; ++c) {
        double gx, gy, gz;
        grad3(conc.coeff + ((size_t)leaf * dof + c) * NC, bs, gx, gy, gz);
        const double u = weight(c);
        ax = __builtin_fma(u, gx, ax);
        ay = __builtin_fma(u, gy, ay);
        az = __builtin_fma(u, gz, az);
      }
    });
    if (live) {
      A[3 * gi + 0] += ax;
      A[3 * gi + 1] += ay;
      A[3 * gi + 2] += az;
    }
  }
  scatter_transposed<kConcChunk>(lam, dof, bs, weight);
}

// The first-stage velocity at arrival point i (its wrapped position x, y,
// z) as the forward step takes it from the velocity's nodal values
// (a.vnode, FP64; nodal_velocity); false where the forward evaluates it.
__device__ __forceinline__ bool adj_nodal_stage1(const VelSource &vel, const TreeView &conc,
                                                 const StepArgs &a, int64_t i, double x,
                                                 double y, double z, double &v0, double &v1,
                                                 double &v2) {
  if (!a.vnode) return false;
  const int64_t jl = i / a.P;
  const int lf = locate(vel.t[0], x, y, z);
  return nodal_velocity<D>((const double *)a.vnode, vel.t[0], conc.info[jl],
                           (int)(i - jl * a.P), lf, a.P, a.scaled, a.g[0], v0, v1, v2);
}

// Waves per SIMD asked of the velocity kernel: two fit without spills up
// to degree 8 (the register budget halves to 256).
constexpr int kVelWaves = D <= 9 ? 2 : 1;

// Velocity cotangent through the num_rk_step substeps.  Substep s maps x_s
// to x_{s+1} = x_s + tau g2_s V(h_s), h_s = x_s + (tau/2) g1_s V(x_s)
// (traj.inc:20-46; the periodic wraps are translations), so with a_{s+1}
// the cotangent of x_{s+1} (a_S = A, the departure point's):
//   a . dx_{s+1} gives   tau g2_s a_{s+1} . dV(h_s)
//                      + w_s . dV(x_s),  w_s = tau g2_s (tau/2) g1_s J_V(h_s)^T a_{s+1}
//   a_s = a_{s+1} + tau g2_s J_V(h_s)^T a_{s+1} + J_V(x_s)^T w_s,
// each dV scattered into G[leaf][k][m] at its point (in its leaf).  The
// substeps run backwards; x_s is integrated again from the arrival point
// (s substeps) for each one.  Before the work at h_s and again before the
// work at x_s the workgroup's points are regrouped by the leaf they fall in
// (as in the concentration kernel), carrying their state along, so a
// wavefront's scatter reads and adds one leaf's coefficients at a time.
template <int MODE>
__global__ __launch_bounds__(kAdjBlock, kVelWaves) void k_adj_vel(VelSource vel, TreeView conc,
                                                                  StepArgs a,
                                                                  const double *__restrict__ A,
                                                                  double *__restrict__ G) {
  __shared__ RegroupLds<kAdjBlock, 12> sh;
  const uint32_t b = blockIdx.x;
  const int64_t lb = (int64_t)(b & 7u) * a.per_xcd + (int64_t)(b >> 3);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  const int64_t i0 = lb * kAdjBlock;
  const TreeView &V = vel.t[0];
  int orig = (int)threadIdx.x;
  double ax, ay, az;
  {
    const int64_t gi = i0 + orig;
    const bool live = gi < a.npts;
    ax = live ? A[3 * gi + 0] : 0.0;  // idle lanes carry zero weights
    ay = live ? A[3 * gi + 1] : 0.0;
    az = live ? A[3 * gi + 2] : 0.0;
  }
#pragma unroll 1
  for (int s = a.nsub - 1; s >= 0; --s) {
    const int64_t i = i0 + orig < a.npts ? i0 + orig : a.npts - 1;
    double x, y, z;
    arrival_point(conc, a, i, x, y, z);
#pragma unroll 1
    for (int r = 0; r < s; ++r) {
      double w0, w1, w2;
      if (r == 0 && a.vnode) {  // the forward's first stage from the nodal values
        if (a.bc == TBX_BC_PERIODIC) {
          x = wrap1(x);
          y = wrap1(y);
          z = wrap1(z);
        }
        if (!adj_nodal_stage1(vel, conc, a, i, x, y, z, w0, w1, w2))
          eval_velocity<D, MODE, kVelTree>(vel, 1, a.ts[0], a.scaled, a.g[0], a.bc, x, y, z, w0,
                                           w1, w2);
        rk2_finish<D, MODE, kVelTree>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[1], a.ts[1], x, y,
                                      z, w0, w1, w2);
      } else {
        rk2_substep<D, MODE, kVelTree>(vel, a.bc, a.htau, a.tau, a.scaled, a.g[2 * r],
                                       a.g[2 * r + 1], a.ts[2 * r], a.ts[2 * r + 1], x, y, z);
      }
    }
    if (a.bc == TBX_BC_PERIODIC) {
      x = wrap1(x);
      y = wrap1(y);
      z = wrap1(z);
    }
    // the forward's midpoint of substep s (rk2_substep, stage 1)
    double v0, v1, v2;
    double xs = x, ys = y, zs = z;
    if (s != 0 || !adj_nodal_stage1(vel, conc, a, i, x, y, z, v0, v1, v2))
      eval_velocity<D, MODE, kVelTree>(vel, 1, a.ts[2 * s], a.scaled, a.g[2 * s], a.bc, xs, ys,
                                       zs, v0, v1, v2);
    double hx = x + a.htau * v0, hy = y + a.htau * v1, hz = z + a.htau * v2;
    if (a.bc == TBX_BC_PERIODIC) {
      hx = wrap1(hx);
      hy = wrap1(hy);
      hz = wrap1(hz);
    }
    const double g1 = a.scaled ? a.g[2 * s] : 1.0, g2 = a.scaled ? a.g[2 * s + 1] : 1.0;
    const double fh = a.tau * g2, fx = a.tau * g2 * a.htau * g1;
    // work at h: regroup by h's leaf
    int lh = locate(V, hx, hy, hz);
    double ph[9] = {hx, hy, hz, ax, ay, az, x, y, z};
    regroup(sh, orig, lh, ph);
    ax = ph[3];
    ay = ph[4];
    az = ph[5];
    double jx = 0.0, jy = 0.0, jz = 0.0;
    {
      AdjBasis bh;
      bh.build(V, ph[0], ph[1], ph[2], lh);
      // J^T a at h: (J^T a)_k = sum_c a_c dV_c/dy_k
      for_each_wave_leaf(bh.leaf, [&](int leaf) {
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
          double gx, gy, gz;
          grad3(V.coeff + ((size_t)leaf * 3 + c) * NC, bh, gx, gy, gz);
          const double ac = c == 0 ? ax : c == 1 ? ay : az;
          jx = __builtin_fma(ac, gx, jx);
          jy = __builtin_fma(ac, gy, jy);
          jz = __builtin_fma(ac, gz, jz);
        }
      });
      scatter_transposed<32>(G, 3, bh, [&](int c) {
        return fh * (c == 0 ? ax : c == 1 ? ay : az);
      });
    }
    // work at x: regroup by x's leaf
    int lx = locate(V, ph[6], ph[7], ph[8]);
    double px[12] = {ph[6], ph[7], ph[8], fx * jx, fx * jy, fx * jz, ax, ay, az, jx, jy, jz};
    if (s == 0 && a.wnode) {
      // x is the point's arrival node: where that is a grid node of its
      // velocity leaf, its weights go to the node's slot (the slots are then
      // projected, G += B^T slots, in one pass per leaf); only the others
      // scatter, and only workgroups holding one regroup
      const int64_t gi = i0 + orig;
      bool rest = false;
      if (gi < a.npts) {
        const int64_t jl = gi / a.P;
        const int mm = node_of(conc.info[jl], V.info[lx], (int)(gi - jl * a.P), D);
        if (mm >= 0) {
          double *wn = a.wnode + (int64_t)lx * 3 * a.P + mm;
          atomicAdd(wn, px[3]);
          atomicAdd(wn + a.P, px[4]);
          atomicAdd(wn + 2 * a.P, px[5]);
          px[3] = 0.0;
          px[4] = 0.0;
          px[5] = 0.0;
        } else {
          rest = true;
        }
      }
      if (!__syncthreads_or(rest)) continue;  // s == 0: the last substep
    }
    regroup(sh, orig, lx, px);
    const double wx = px[3], wy = px[4], wz = px[5];
    AdjBasis bx;
    bx.build(V, px[0], px[1], px[2], lx);
    scatter_transposed<32>(G, 3, bx, [&](int c) { return c == 0 ? wx : c == 1 ? wy : wz; });
    if (s > 0) {
      // a_s = a_{s+1} + tau g2 J(h)^T a_{s+1} + J(x)^T w
      double kx = 0.0, ky = 0.0, kz = 0.0;
      for_each_wave_leaf(bx.leaf, [&](int leaf) {
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {
          double gx, gy, gz;
          grad3(V.coeff + ((size_t)leaf * 3 + c) * NC, bx, gx, gy, gz);
          const double wc = c == 0 ? wx : c == 1 ? wy : wz;
          kx = __builtin_fma(wc, gx, kx);
          ky = __builtin_fma(wc, gy, ky);
          kz = __builtin_fma(wc, gz, kz);
        }
      });
      ax = px[6] + fh * px[9] + kx;
      ay = px[7] + fh * px[10] + ky;
      az = px[8] + fh * px[11] + kz;
    }
  }
}

}  // namespace

#define TBX_CAT2(a, b) a##b
#define TBX_CAT(a, b) TBX_CAT2(a, b)

hipError_t TBX_CAT(launch_adj_conc_q, TBX_Q)(int mode, const VelSource &vel,
                                             const TreeView &conc, const StepArgs &a,
                                             const double *dep, const double *ua, double *lam,
                                             double *A, hipStream_t s) {
  if (a.npts <= 0) return hipSuccess;
  const dim3 g((unsigned)(8 * a.per_xcd)), blk(kAdjBlock);
  if (dep)  // the stored trajectories carry the forward's arithmetic
    hipLaunchKernelGGL((k_adj_conc<kFma64, kAdjBlock, true>), g, blk, 0, s, vel, conc, a, dep,
                       ua, lam, A);
  else if (mode == kFma64)
    hipLaunchKernelGGL((k_adj_conc<kFma64, kAdjBlock, false>), g, blk, 0, s, vel, conc, a, dep,
                       ua, lam, A);
  else if (mode == kExact64)
    hipLaunchKernelGGL((k_adj_conc<kExact64, kAdjBlock, false>), g, blk, 0, s, vel, conc, a,
                       dep, ua, lam, A);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t TBX_CAT(launch_adj_vel_q, TBX_Q)(int mode, const VelSource &vel,
                                            const TreeView &conc, const StepArgs &a,
                                            const double *A, double *G, hipStream_t s) {
  if (a.npts <= 0) return hipSuccess;
  const dim3 g((unsigned)(8 * a.per_xcd)), blk(kAdjBlock);
  if (mode == kFma64)
    hipLaunchKernelGGL((k_adj_vel<kFma64>), g, blk, 0, s, vel, conc, a, A, G);
  else if (mode == kExact64)
    hipLaunchKernelGGL((k_adj_vel<kExact64>), g, blk, 0, s, vel, conc, a, A, G);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace tbx
