// Host side of libtbx: the C ABI of include/tbx.h.
//
// Owns the device copies of each tree (Morton keys, leaf info, coarse-cell
// directory, coefficients and their FP32 mirror), checks arguments, picks
// the per-degree kernels and runs the refit GEMM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tbx.h"
#include "tbx_internal.h"

using namespace tbx;

namespace {

thread_local std::string g_err;
std::atomic<int> g_step_shape{kAutoStepShape};
std::atomic<int> g_tracer_mfma{1};
std::atomic<int> g_grid_lds{1};
std::atomic<int> g_node_velocity{1};

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                   \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess)                                               \
      return fail(TBX_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                  \
  } while (0)

inline int ncoef(int q) { return (q + 1) * (q + 2) * (q + 3) / 6; }

// The contraction streams coefficients in 64-byte scalar chunks, two per
// block, and may read up to 124 bytes past the last triangle; coefficient
// buffers carry this much slack.
constexpr size_t kCoeffPad = 256;

unsigned host_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return std::min(h, 16u);
}

// Run f(i) for i in [0, n) on a few host threads.
template <class F>
void parallel_for(int64_t n, F f) {
  const unsigned nt = (unsigned)std::min<int64_t>(host_threads(), std::max<int64_t>(n, 1));
  if (nt <= 1 || n < 64) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t chunk = std::max<int64_t>(1, n / (nt * 8));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (;;) {
        const int64_t a = next.fetch_add(chunk);
        if (a >= n) break;
        const int64_t b = std::min(n, a + chunk);
        for (int64_t i = a; i < b; ++i) f(i);
      }
    });
  for (auto &x : th) x.join();
}

uint64_t spread3_host(uint32_t v) {
  uint64_t x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

uint64_t leaf_key(uint32_t ax, uint32_t ay, uint32_t az, int depth) {
  const int sh = kKeyBits - depth;
  return spread3_host(ax << sh) | (spread3_host(ay << sh) << 1) |
         (spread3_host(az << sh) << 2);
}

uint32_t compact3(uint64_t k) {
  uint64_t x = k & 0x1249249249249249ull;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x | (x >> 4)) & 0x100f00f00f00f00full;
  x = (x | (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x | (x >> 16)) & 0x1f00000000ffffull;
  x = (x | (x >> 32)) & 0x1fffffull;
  return (uint32_t)x;
}

// T_0..T_q at x by the recurrence the kernels use.
void cheb_poly_host(int q, double x, double *p) {
  const bool in = std::fabs(x) <= 1.0;
  const double xx = in ? x : 0.0;
  double y0 = in ? 1.0 : 0.0;
  p[0] = y0;
  if (q >= 1) p[1] = xx;
  double y1 = xx;
  for (int k = 2; k <= q; ++k) {
    const double y2 = 2.0 * xx * y1 - y0;
    p[k] = y2;
    y0 = y1;
    y1 = y2;
  }
}

// n first-kind Chebyshev nodes on [0, 1] (the cheb_nodes InitTree samples
// at, tree_utils.h:144).
void first_kind_nodes_host(int n, double *x) {
  for (int i = 0; i < n; ++i) x[i] = -std::cos((i + 0.5) * M_PI / n) * 0.5 + 0.5;
}

void new_nodes_host(int q, double *x) {
  const unsigned d = (unsigned)q + 1;
  const double scal = 1.0 / std::cos(0.5 * M_PI / d);
  for (unsigned i = 0; i < d; ++i)
    x[i] = -std::cos((i + 0.5) * M_PI / d) * scal * 0.5 + 0.5;
}

// Householder QR of B (P x nc, column-major in A on entry, in long double),
// B[p][m] = basis m (packed i,j,k) at stretched node p mapped to [-1,1]
// (cheb.h:170-180): returns R (nc x nc, row-major upper triangle) and the
// reflectors (V, beta).
typedef long double R_t;
struct Householder {
  int P = 0, nc = 0;
  std::vector<R_t> R, V, beta;
};

std::vector<double> stretched_basis_1d(int q) {
  // T[k * d + a] = T_k(node a mapped to [-1, 1])
  const int d = q + 1;
  std::vector<double> x(d), T((size_t)d * d);
  new_nodes_host(q, x.data());
  for (int a = 0; a < d; ++a) {
    std::vector<double> p(d);
    cheb_poly_host(q, x[a] * 2.0 - 1.0, p.data());
    for (int k = 0; k < d; ++k) T[(size_t)k * d + a] = p[k];
  }
  return T;
}

Householder householder_basis(int q) {
  const int d = q + 1, P = d * d * d, nc = ncoef(q);
  const std::vector<double> T = stretched_basis_1d(q);
  // A column-major P x nc
  std::vector<R_t> A((size_t)P * nc);
  {
    int m = 0;
    for (int i = 0; i < d; ++i)
      for (int j = 0; i + j < d; ++j)
        for (int k = 0; i + j + k < d; ++k, ++m) {
          R_t *col = &A[(size_t)m * P];
          for (int p = 0; p < P; ++p) {
            const int ix = p % d, iy = (p / d) % d, iz = p / (d * d);
            col[p] = (R_t)T[(size_t)k * d + ix] * T[(size_t)j * d + iy] * T[(size_t)i * d + iz];
          }
        }
  }
  Householder h;
  h.P = P;
  h.nc = nc;
  h.R.assign((size_t)nc * nc, 0.0);
  h.V.assign((size_t)P * nc, 0.0);
  h.beta.assign(nc, 0.0);
  for (int c = 0; c < nc; ++c) {
    R_t *a = &A[(size_t)c * P];
    R_t nrm = 0;
    for (int p = c; p < P; ++p) nrm += a[p] * a[p];
    nrm = std::sqrt(nrm);
    const R_t alpha = a[c] > 0 ? -nrm : nrm;
    R_t *v = &h.V[(size_t)c * P];
    for (int p = c; p < P; ++p) v[p] = a[p];
    v[c] -= alpha;
    R_t vv = 0;
    for (int p = c; p < P; ++p) vv += v[p] * v[p];
    h.beta[c] = vv > 0 ? 2.0L / vv : 0.0L;
    h.R[(size_t)c * nc + c] = alpha;
    parallel_for(nc - c - 1, [&](int64_t jj) {
      const int j = c + 1 + (int)jj;
      R_t *aj = &A[(size_t)j * P];
      R_t s = 0;
      for (int p = c; p < P; ++p) s += v[p] * aj[p];
      s *= h.beta[c];
      for (int p = c; p < P; ++p) aj[p] -= s * v[p];
    });
    for (int j = c + 1; j < nc; ++j) h.R[(size_t)c * nc + j] = A[(size_t)j * P + c];
  }
  return h;
}

// M = pinv(B)^T = Q R^-T (GetPt2CoeffMatrix, cheb.h:166-196), M[P][nc].
std::vector<double> build_pt2coeff(int q) {
  const Householder h = householder_basis(q);
  const int P = h.P, nc = h.nc;
  // Q (P x nc, column-major) = H_0 ... H_{nc-1} [I; 0]
  std::vector<R_t> Q((size_t)P * nc, 0.0);
  for (int c = 0; c < nc; ++c) Q[(size_t)c * P + c] = 1.0;
  parallel_for(nc, [&](int64_t col) {
    R_t *qc = &Q[(size_t)col * P];
    for (int c = nc - 1; c >= 0; --c) {
      const R_t *v = &h.V[(size_t)c * P];
      R_t s = 0;
      for (int p = c; p < P; ++p) s += v[p] * qc[p];
      s *= h.beta[c];
      if (s != 0)
        for (int p = c; p < P; ++p) qc[p] -= s * v[p];
    }
  });
  // row p of M solves R m = Q[p,:]^T
  std::vector<double> M((size_t)P * nc);
  parallel_for(P, [&](int64_t p) {
    std::vector<R_t> r(nc);
    for (int a = 0; a < nc; ++a) r[a] = Q[(size_t)a * P + p];
    for (int a = nc - 1; a >= 0; --a) {
      R_t s = r[a];
      for (int k = a + 1; k < nc; ++k) s -= h.R[(size_t)a * nc + k] * r[k];
      r[a] = s / h.R[(size_t)a * nc + a];
    }
    for (int a = 0; a < nc; ++a) M[(size_t)p * nc + a] = (double)r[a];
  });
  return M;
}

// The refit operator in factored form: pinv(B) = H B^T with
// H = (B^T B)^-1.  H is stored
// zero-padded to [KP][NP] (KP a multiple of 16, NP of 64) for the GEMM, and
// the one-axis basis T_k(node a) for the tensor-product contractions.
struct RefitOperator {
  int KP = 0, NP = 0;
  std::vector<double> H;  // [KP][NP]
  RefitTable tab{};
};

RefitOperator build_refit_operator(int q) {
  // B = Tx (x) Ty (x) Tz on the triangle's columns, so the Gram matrix is
  // B^T B [m][m'] = G[k][k'] G[j][j'] G[i][i'], G = T T^T (d x d); H is its
  // inverse by Cholesky in long double (B^T B is symmetric positive
  // definite with condition number ~14 at every degree built here).
  const int d = q + 1, nc = ncoef(q);
  const std::vector<double> T = stretched_basis_1d(q);
  std::vector<R_t> G((size_t)d * d);
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) {
      R_t s = 0;
      for (int n = 0; n < d; ++n) s += (R_t)T[(size_t)a * d + n] * T[(size_t)b * d + n];
      G[(size_t)a * d + b] = s;
    }
  std::vector<int> ti(nc), tj(nc), tk(nc);
  {
    int m = 0;
    for (int i = 0; i < d; ++i)
      for (int j = 0; i + j < d; ++j)
        for (int k = 0; i + j + k < d; ++k, ++m) {
          ti[m] = i;
          tj[m] = j;
          tk[m] = k;
        }
  }
  // Cholesky A = L L^T, L lower triangular (row-major, in place in A)
  std::vector<R_t> A((size_t)nc * nc);
  parallel_for(nc, [&](int64_t a) {
    for (int b = 0; b < nc; ++b)
      A[(size_t)a * nc + b] = G[(size_t)tk[a] * d + tk[b]] * G[(size_t)tj[a] * d + tj[b]] *
                              G[(size_t)ti[a] * d + ti[b]];
  });
  for (int c = 0; c < nc; ++c) {
    R_t s = A[(size_t)c * nc + c];
    for (int k = 0; k < c; ++k) s -= A[(size_t)c * nc + k] * A[(size_t)c * nc + k];
    const R_t lcc = std::sqrt(s);
    A[(size_t)c * nc + c] = lcc;
    parallel_for(nc - c - 1, [&](int64_t rr) {
      const int r = c + 1 + (int)rr;
      R_t t = A[(size_t)r * nc + c];
      for (int k = 0; k < c; ++k) t -= A[(size_t)r * nc + k] * A[(size_t)c * nc + k];
      A[(size_t)r * nc + c] = t / lcc;
    });
  }
  // Linv (lower triangular), column by column: L Linv = I
  std::vector<R_t> Li((size_t)nc * nc, 0.0);
  parallel_for(nc, [&](int64_t col) {
    for (int r = (int)col; r < nc; ++r) {
      R_t s = r == col ? 1.0L : 0.0L;
      for (int k = (int)col; k < r; ++k) s -= A[(size_t)r * nc + k] * Li[(size_t)k * nc + col];
      Li[(size_t)r * nc + col] = s / A[(size_t)r * nc + r];
    }
  });
  RefitOperator op;
  op.KP = (nc + 15) / 16 * 16;
  op.NP = (nc + 63) / 64 * 64;
  op.H.assign((size_t)op.KP * op.NP, 0.0);
  // H = L^-T L^-1: H[a][b] = sum_k Linv[k][a] Linv[k][b], k >= max(a, b)
  parallel_for(nc, [&](int64_t a) {
    for (int b = 0; b < nc; ++b) {
      R_t s = 0;
      for (int k = std::max((int)a, b); k < nc; ++k)
        s += Li[(size_t)k * nc + a] * Li[(size_t)k * nc + b];
      op.H[(size_t)a * op.NP + b] = (double)s;
    }
  });
  for (int k = 0; k < d; ++k)
    for (int a = 0; a < d; ++a) op.tab.T[k][a] = T[(size_t)k * d + a];
  return op;
}

const RefitOperator &refit_operator_cached(int q) {
  static std::mutex mu;
  static std::map<int, RefitOperator> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(q);
  if (it == cache.end()) it = cache.emplace(q, build_refit_operator(q)).first;
  return it->second;
}

const std::vector<double> &pt2coeff_cached(int q) {
  static std::mutex mu;
  static std::map<int, std::vector<double>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(q);
  if (it == cache.end()) it = cache.emplace(q, build_pt2coeff(q)).first;
  return it->second;
}

}  // namespace

struct tbx_tree_s {
  int device = 0;
  int64_t L = 0;
  int q = 0, dof = 0, nc = 0, max_depth = 0, dir_level = 0;
  std::vector<uint64_t> h_key;
  std::vector<uint4> h_info;
  uint64_t *d_key = nullptr;
  uint4 *d_info = nullptr;
  int32_t *d_dir = nullptr;
  double *d_coeff = nullptr;
  float *d_coeff32 = nullptr;
  bool f32_stale = true;
  double *d_vals = nullptr;  // [L][dof][P] values at the grid points
  double *d_H = nullptr;     // [KP][NP] refit operator (B^T B)^-1, zero-padded
  double *d_W = nullptr;     // [L*dof][KP] refit workspace (B^T values)
  int KP = 0, NP = 0;
  RefitTable tab{};
  double *d_adjA = nullptr;  // [L*P][3] per-point velocity cotangent factors
  double *d_nodal = nullptr;  // [L][dof][P] values at the own grid nodes (velocity trees)
  double *d_adjW = nullptr;   // [L][dof][P] the velocity adjoint's per-node weights

  TreeView view() const {
    TreeView v{};
    v.key = d_key;
    v.info = d_info;
    v.dir = d_dir;
    v.coeff = d_coeff;
    v.coeff32 = d_coeff32;
    v.L = L;
    v.dof = dof;
    v.nc = nc;
    v.q = q;
    v.dir_shift = 3 * (kKeyBits - dir_level);
    v.dir_exact = max_depth <= dir_level ? 1 : 0;
    return v;
  }
  int64_t ncoeff_total() const { return L * dof * nc; }
  int P() const { return (q + 1) * (q + 1) * (q + 1); }
};

namespace {

void free_tree(tbx_tree *t) {
  if (!t) return;
  hipFree(t->d_key);
  hipFree(t->d_info);
  hipFree(t->d_dir);
  hipFree(t->d_coeff);
  hipFree(t->d_coeff32);
  hipFree(t->d_vals);
  hipFree(t->d_H);
  hipFree(t->d_W);
  hipFree(t->d_adjA);
  hipFree(t->d_nodal);
  hipFree(t->d_adjW);
  delete t;
}

int ensure_f32(tbx_tree *t, hipStream_t s) {
  if (!t->d_coeff32) {
    HIP_TRY(hipMalloc(&t->d_coeff32, sizeof(float) * (size_t)t->ncoeff_total() + kCoeffPad));
    t->f32_stale = true;
  }
  if (t->f32_stale) {
    HIP_TRY(launch_to_f32(t->d_coeff, t->d_coeff32, t->ncoeff_total(), s));
    t->f32_stale = false;
  }
  return TBX_OK;
}

// The launch shape a step of this mode and degree runs with under the
// current option (a shape not built for q falls back to the default).
int resolve_step_shape(int mode, int q) {
  int want = g_step_shape.load();
  if (want == kAutoStepShape) want = auto_step_shape(mode, q);
  return step_shape_built(want, q) ? want : kDefaultStepShape;
}

int mode_of(uint32_t flags, int *mode) {
  const uint32_t m = flags & TBX_MODE_MASK;
  if (m == TBX_FP64)
    *mode = kFma64;
  else if (m == TBX_EXACT)
    *mode = kExact64;
  else if (m == TBX_FP32)
    *mode = kFp32;
  else if (m == TBX_FP32_HEAD64)
    *mode = kFp32h;
  else
    return fail(TBX_EINVAL, "unknown arithmetic mode 0x%x", m);
  return TBX_OK;
}

int prepare_mode(tbx_tree *t, int mode, hipStream_t s) {
  if (mode == kFp32 || mode == kFp32h) return ensure_f32(t, s);
  return TBX_OK;
}

// A bump allocator over caller-independent scratch memory (the
// allocation-free entry point's static device array): regions are 256-byte
// aligned and live for one call.
struct ScratchArena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  bool nodal = true;  // room for the velocity's nodal values
  void *take(size_t bytes) {
    const size_t off = (used + 255) & ~(size_t)255;
    if (off > size || bytes > size - off) return nullptr;
    used = off + bytes;
    return base + off;
  }
};

// The tree's device view for a mode.  FP32 modes read an FP32 mirror of the
// coefficients: the tree's own (made on first use), or, with an arena and a
// tree that has none, one converted into the arena for this call.
int prepare_view(tbx_tree *t, int mode, hipStream_t s, ScratchArena *arena, TreeView *v,
                 const char *who) {
  if ((mode == kFp32 || mode == kFp32h) && arena && !t->d_coeff32) {
    *v = t->view();
    float *m = (float *)arena->take(sizeof(float) * (size_t)t->ncoeff_total() + kCoeffPad);
    if (!m) return fail(TBX_EOVERFLOW, "%s: FP32 coefficients exceed the scratch", who);
    HIP_TRY(launch_to_f32(t->d_coeff, m, t->ncoeff_total(), s));
    v->coeff32 = m;
    return TBX_OK;
  }
  int rc = prepare_mode(t, mode, s);
  if (rc) return rc;
  *v = t->view();
  return TBX_OK;
}

#define TBX_Q_CASES(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

hipError_t dispatch_eval(int q, int mode, const TreeView &t, double *pts,
                         int64_t n, int bc, int wb, double *out, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_eval_q##Q(mode, t, pts, n, bc, wb, out, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_contract(int q, int mode, const TreeView &t, int64_t n, int wpl, double *out,
                             double *pts, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_contract_q##Q(mode, t, n, wpl, out, pts, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_veval(int q, int mode, const VelSource &vel, double *pts, int64_t n,
                          int bc, int wb, double t, int scaled, double g, double *out,
                          hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_veval_q##Q(mode, vel, pts, n, bc, wb, t, scaled, g, out, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_traj(int q, int mode, const VelSource &vel, double *x,
                         int64_t n, const SubArgs &a, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_traj_q##Q(mode, vel, x, n, a, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_step(int q, int mode, const VelSource &vel,
                         const TreeView &conc, const StepArgs &a, double *vals,
                         double *dep, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_step_q##Q(mode, vel, conc, a, vals, dep, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_adj_conc(int q, int mode, const VelSource &vel, const TreeView &conc,
                             const StepArgs &a, const double *dep, const double *ua,
                             double *lam, double *A, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_adj_conc_q##Q(mode, vel, conc, a, dep, ua, lam, A, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

hipError_t dispatch_adj_vel(int q, int mode, const VelSource &vel, const TreeView &conc,
                            const StepArgs &a, const double *A, double *G, hipStream_t s) {
  switch (q) {
#define X(Q) \
  case Q:    \
    return launch_adj_vel_q##Q(mode, vel, conc, a, A, G, s);
    TBX_Q_CASES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

int check_points(const void *p, int64_t n, const char *what) {
  if (n < 0) return fail(TBX_EINVAL, "%s: negative point count", what);
  if (n > 0 && !p) return fail(TBX_EINVAL, "%s: null buffer", what);
  if (n > ((int64_t)1 << 40)) return fail(TBX_EOVERFLOW, "%s: %lld points", what, (long long)n);
  return TBX_OK;
}

int check_bc(int bc) {
  if (bc != TBX_BC_FREESPACE && bc != TBX_BC_PERIODIC)
    return fail(TBX_EINVAL, "unknown boundary condition %d", bc);
  return TBX_OK;
}

// Stage factors of one substep at current time tcur (traj.inc:33-40).
SubArgs sub_args(double tcur, double tau, tbx_tfac_fn g, void *gctx, int bc) {
  SubArgs a{};
  a.tau = tau;
  a.htau = 0.5 * tau;
  a.scaled = g ? 1 : 0;
  a.ts1 = tcur;
  a.ts2 = tcur + 0.5 * tau;
  a.g1 = g ? g(a.ts1, gctx) : 1.0;
  a.g2 = g ? g(a.ts2, gctx) : 1.0;
  a.bc = bc;
  return a;
}

// Checks a velocity functor and prepares its trees for the mode.
int prepare_velocity(const tbx_velocity *v, int mode, hipStream_t s, VelSource *out,
                     int *q, const char *who, ScratchArena *arena = nullptr) {
  if (!v) return fail(TBX_EINVAL, "%s: null velocity", who);
  int nt;
  switch (v->kind) {
    case TBX_VEL_TREE: nt = 1; break;
    case TBX_VEL_EXTRAP: nt = 2; break;
    case TBX_VEL_SET: nt = 4; break;
    default: return fail(TBX_EINVAL, "%s: unknown velocity kind %d", who, v->kind);
  }
  if (v->kind != TBX_VEL_TREE && v->g)
    return fail(TBX_EINVAL, "%s: a time factor applies to a single velocity tree only", who);
  *out = VelSource{};
  for (int k = 0; k < nt; ++k) {
    tbx_tree *t = const_cast<tbx_tree *>(v->trees[k]);
    if (!t) return fail(TBX_EINVAL, "%s: velocity tree %d is null", who, k);
    if (t->dof != 3) return fail(TBX_EINVAL, "%s: velocity tree %d has dof %d", who, k, t->dof);
    if (k > 0 && t->q != v->trees[0]->q)
      return fail(TBX_EUNSUPPORTED, "%s: velocity trees differ in degree", who);
    int rc = prepare_view(t, mode, s, arena, &out->t[k], who);
    if (rc) return rc;
  }
  if (v->kind == TBX_VEL_SET)
    for (int k = 0; k < 4; ++k) out->tt[k] = v->times[k];
  out->kind = v->kind;
  out->ntree = nt;
  *q = v->trees[0]->q;
  return TBX_OK;
}

tbx_velocity single_tree(const tbx_tree *t, tbx_tfac_fn g, void *gctx) {
  tbx_velocity v{};
  v.kind = TBX_VEL_TREE;
  v.trees[0] = t;
  v.g = g;
  v.gctx = gctx;
  return v;
}

// The refit operator on the device (uploaded when the tree is created).
int ensure_refit_operator(tbx_tree *t) {
  if (!t->d_H) {
    const RefitOperator &op = refit_operator_cached(t->q);
    HIP_TRY(hipMalloc(&t->d_H, sizeof(double) * op.H.size()));
    HIP_TRY(hipMemcpy(t->d_H, op.H.data(), sizeof(double) * op.H.size(), hipMemcpyHostToDevice));
    t->KP = op.KP;
    t->NP = op.NP;
    t->tab = op.tab;
  }
  return TBX_OK;
}

// The operator and the tree's own refit workspace.
int ensure_refit(tbx_tree *t, hipStream_t s) {
  int rc = ensure_refit_operator(t);
  if (rc) return rc;
  if (!t->d_W) HIP_TRY(hipMalloc(&t->d_W, sizeof(double) * (size_t)t->L * t->dof * t->KP));
  (void)s;
  return TBX_OK;
}

// coeff = pinv(B) vals for every leaf component (SetTreeGridValues,
// tree_utils.h:534-546): W = B^T vals by the one-axis contractions into the
// workspace W ([L*dof][KP]), then coeff = W H on the matrix cores, written
// in place together with the FP32 mirror when the tree has one.  Allocates
// nothing.
int refit_with(tbx_tree *t, const double *vals, double *W, hipStream_t s) {
  const int64_t rows = t->L * t->dof;
  if (rows > INT32_MAX) return fail(TBX_EOVERFLOW, "refit: %lld rows", (long long)rows);
  if (!t->d_H) return fail(TBX_EINVAL, "refit: the tree has no refit operator");
  HIP_TRY(launch_refit_project(t->q + 1, vals, rows, t->dof, t->P(), 1, t->tab, W, t->KP, s));
  HIP_TRY(launch_refit_gemm(W, t->KP, rows, t->KP, t->KP, t->d_H, t->NP, t->d_coeff, t->nc,
                            t->nc, t->d_coeff32, s));
  t->f32_stale = t->d_coeff32 == nullptr;
  return TBX_OK;
}

int refit(tbx_tree *t, const double *vals, hipStream_t s) {
  int rc = ensure_refit(t, s);
  if (rc) return rc;
  return refit_with(t, vals, t->d_W, s);
}

// The leaf structure of a tree from a host leaf list: validated (depths,
// anchors, strict Morton order, an exact tiling of [0,1)^3), host mirrors
// of keys and (anchor, depth), and their device copies with the location
// directory, replacing any previous ones.  Coefficients are untouched.
int set_leaves(tbx_tree *t, int64_t L, const uint32_t *anchor, const uint8_t *depth) {
  if (L <= 0 || L > INT32_MAX) return fail(TBX_EOVERFLOW, "%lld leaves", (long long)L);
  std::vector<uint64_t> key(L);
  std::vector<uint4> info(L);
  int maxd = 0;
  for (int64_t j = 0; j < L; ++j) {
    const int dj = depth[j];
    const uint32_t ax = anchor[3 * j], ay = anchor[3 * j + 1], az = anchor[3 * j + 2];
    if (dj > kKeyBits)
      return fail(TBX_EUNSUPPORTED, "leaf %lld: depth %d > %d", (long long)j, dj, kKeyBits);
    const uint32_t lim = 1u << dj;
    if (ax >= lim || ay >= lim || az >= lim)
      return fail(TBX_EINVAL, "leaf %lld: anchor outside depth-%d grid", (long long)j, dj);
    key[j] = leaf_key(ax, ay, az, dj);
    info[j] = make_uint4(ax, ay, az, (uint32_t)dj);
    maxd = std::max(maxd, dj);
  }
  // strictly increasing keys whose ranges tile [0, 8^21) without gaps
  const unsigned __int128 total = (unsigned __int128)1 << (3 * kKeyBits);
  if (key[0] != 0) return fail(TBX_ECOVER, "first leaf does not start at the origin");
  for (int64_t j = 0; j < L; ++j) {
    const unsigned __int128 end =
        (unsigned __int128)key[j] + ((unsigned __int128)1 << (3 * (kKeyBits - depth[j])));
    if (j + 1 < L) {
      if (key[j + 1] <= key[j])
        return fail(TBX_EORDER, "leaves %lld and %lld are not in Morton order", (long long)j,
                    (long long)(j + 1));
      if (end != key[j + 1])
        return fail(TBX_ECOVER, "leaf %lld and the next leaf overlap or leave a gap", (long long)j);
    } else if (end != total) {
      return fail(TBX_ECOVER, "last leaf does not reach the far corner");
    }
  }
  // Directory level g <= 6: 8^6 + 1 int32 entries (1 MiB) stay resident in
  // one XCD's L2.  On uniformly random points a level-7 directory (8 MiB)
  // located a depth-7 tree at 1.8 ms per 2^27 points, level 6 plus a search
  // of at most three steps at 1.0 ms.
  const int dir_level = std::min(maxd, 6);
  const int64_t ncell = (int64_t)1 << (3 * dir_level);
  std::vector<int32_t> dir(ncell + 1);
  {
    const int sh = 3 * (kKeyBits - dir_level);
    int64_t j = 0;
    for (int64_t c = 0; c < ncell; ++c) {
      const uint64_t ck = (uint64_t)c << sh;
      while (j + 1 < L && key[j + 1] <= ck) ++j;
      dir[c] = (int32_t)j;
    }
    dir[ncell] = (int32_t)(L - 1);
  }
  uint64_t *d_key = nullptr;
  uint4 *d_info = nullptr;
  int32_t *d_dir = nullptr;
  hipError_t e = hipMalloc(&d_key, sizeof(uint64_t) * L);
  if (e == hipSuccess) e = hipMalloc(&d_info, sizeof(uint4) * L);
  if (e == hipSuccess) e = hipMalloc(&d_dir, sizeof(int32_t) * dir.size());
  if (e == hipSuccess) e = hipMemcpy(d_key, key.data(), sizeof(uint64_t) * L, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_info, info.data(), sizeof(uint4) * L, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d_dir, dir.data(), sizeof(int32_t) * dir.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(d_key);
    hipFree(d_info);
    hipFree(d_dir);
    return fail(e == hipErrorOutOfMemory ? TBX_ENOMEM : TBX_EHIP, "leaf structure upload: %s",
                hipGetErrorString(e));
  }
  hipFree(t->d_key);
  hipFree(t->d_info);
  hipFree(t->d_dir);
  t->d_key = d_key;
  t->d_info = d_info;
  t->d_dir = d_dir;
  t->h_key.swap(key);
  t->h_info.swap(info);
  t->L = L;
  t->max_depth = maxd;
  t->dir_level = dir_level;
  return TBX_OK;
}

}  // namespace

extern "C" {

const char *tbx_last_error(void) { return g_err.c_str(); }
const char *tbx_version(void) { return "tbx 0.1 (gfx950)"; }

int tbx_tree_create(const tbx_tree_desc *desc, tbx_tree **out) {
  if (!desc || !out) return fail(TBX_EINVAL, "tbx_tree_create: null argument");
  *out = nullptr;
  const int64_t L = desc->num_leaves;
  if (L <= 0 || !desc->anchor || !desc->depth)
    return fail(TBX_EINVAL, "tbx_tree_create: empty leaf list");
  if (L > INT32_MAX) return fail(TBX_EOVERFLOW, "tbx_tree_create: %lld leaves", (long long)L);
  if (desc->q < 1 || desc->q > TBX_MAX_Q)
    return fail(TBX_EUNSUPPORTED, "tbx_tree_create: degree %d outside 1..%d", desc->q, TBX_MAX_Q);
  if (desc->dof < 1) return fail(TBX_EINVAL, "tbx_tree_create: dof %d", desc->dof);
  const int nc = ncoef(desc->q);
  if ((double)L * desc->dof * nc > 4.0e10)
    return fail(TBX_EOVERFLOW, "tbx_tree_create: coefficient array too large");

  tbx_tree *t = new tbx_tree_s();
  t->q = desc->q;
  t->dof = desc->dof;
  t->nc = nc;
  hipGetDevice(&t->device);
  if (int rc = set_leaves(t, L, desc->anchor, desc->depth)) {
    free_tree(t);
    return rc;
  }
  hipError_t e;
  if ((e = hipMalloc(&t->d_coeff, sizeof(double) * (size_t)t->ncoeff_total() + kCoeffPad)) !=
      hipSuccess) {
    free_tree(t);
    return fail(e == hipErrorOutOfMemory ? TBX_ENOMEM : TBX_EHIP, "hipMalloc: %s",
                hipGetErrorString(e));
  }
  if (desc->coeff)
    e = hipMemcpy(t->d_coeff, desc->coeff, sizeof(double) * (size_t)t->ncoeff_total(), hipMemcpyHostToDevice);
  else
    e = hipMemset(t->d_coeff, 0, sizeof(double) * (size_t)t->ncoeff_total());
  if (e != hipSuccess) {
    free_tree(t);
    return fail(TBX_EHIP, "coefficient upload: %s", hipGetErrorString(e));
  }
  // the refit operator goes to the device with the tree, so a step never
  // has to allocate it
  if (int rc = ensure_refit_operator(t)) {
    free_tree(t);
    return rc;
  }
  *out = t;
  return TBX_OK;
}

int tbx_tree_destroy(tbx_tree *t) {
  if (!t) return fail(TBX_EINVAL, "tbx_tree_destroy: null tree");
  hipDeviceSynchronize();
  free_tree(t);
  return TBX_OK;
}

int tbx_tree_info(const tbx_tree *t, int64_t *num_leaves, int32_t *q, int32_t *dof,
                  int32_t *max_depth) {
  if (!t) return fail(TBX_EINVAL, "tbx_tree_info: null tree");
  if (num_leaves) *num_leaves = t->L;
  if (q) *q = t->q;
  if (dof) *dof = t->dof;
  if (max_depth) *max_depth = t->max_depth;
  return TBX_OK;
}

int tbx_tree_set_coeffs(tbx_tree *t, const double *coeff, int on_device, tbx_stream s) {
  if (!t || !coeff) return fail(TBX_EINVAL, "tbx_tree_set_coeffs: null argument");
  HIP_TRY(hipMemcpyAsync(t->d_coeff, coeff, sizeof(double) * (size_t)t->ncoeff_total(),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                         (hipStream_t)s));
  if (!on_device) HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  t->f32_stale = true;
  return TBX_OK;
}

int tbx_tree_get_coeffs(const tbx_tree *t, double *coeff, int on_device, tbx_stream s) {
  if (!t || !coeff) return fail(TBX_EINVAL, "tbx_tree_get_coeffs: null argument");
  HIP_TRY(hipMemcpyAsync(coeff, t->d_coeff, sizeof(double) * (size_t)t->ncoeff_total(),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                         (hipStream_t)s));
  if (!on_device) HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  return TBX_OK;
}

int tbx_locate(const tbx_tree *t, const double *pts, int64_t n, int bc, int32_t *leaf,
               tbx_stream s) {
  if (!t) return fail(TBX_EINVAL, "tbx_locate: null tree");
  int rc = check_points(pts, n, "tbx_locate");
  if (rc) return rc;
  if ((rc = check_bc(bc))) return rc;
  if (n > 0 && !leaf) return fail(TBX_EINVAL, "tbx_locate: null output");
  HIP_TRY(launch_locate(t->view(), pts, n, bc, leaf, (hipStream_t)s));
  return TBX_OK;
}

int tbx_eval(const tbx_tree *tc, double *pts, int64_t n, int bc, uint32_t flags, double *out,
             tbx_stream s) {
  tbx_tree *t = const_cast<tbx_tree *>(tc);
  if (!t) return fail(TBX_EINVAL, "tbx_eval: null tree");
  int rc = check_points(pts, n, "tbx_eval");
  if (rc) return rc;
  if ((rc = check_bc(bc))) return rc;
  if (n > 0 && !out) return fail(TBX_EINVAL, "tbx_eval: null output");
  int mode;
  if ((rc = mode_of(flags, &mode))) return rc;
  if ((rc = prepare_mode(t, mode, (hipStream_t)s))) return rc;
  HIP_TRY(dispatch_eval(t->q, mode, t->view(), pts, n, bc,
                        (flags & TBX_WRITE_BACK_WRAP) ? 1 : 0, out, (hipStream_t)s));
  return TBX_OK;
}

int tbx_contraction_only(const tbx_tree *tc, int64_t n, int waves_per_leaf, uint32_t flags,
                         double *out, double *pts, tbx_stream s) {
  tbx_tree *t = const_cast<tbx_tree *>(tc);
  if (!t) return fail(TBX_EINVAL, "tbx_contraction_only: null tree");
  if (n < 0) return fail(TBX_EINVAL, "tbx_contraction_only: n = %lld", (long long)n);
  if (waves_per_leaf < 1)
    return fail(TBX_EINVAL, "tbx_contraction_only: waves_per_leaf %d < 1", waves_per_leaf);
  if (n > 0 && !out) return fail(TBX_EINVAL, "tbx_contraction_only: null output");
  if (n > (int64_t)INT32_MAX * kBlock) return fail(TBX_EOVERFLOW, "tbx_contraction_only: n too large");
  int mode, rc;
  if ((rc = mode_of(flags, &mode))) return rc;
  if ((rc = prepare_mode(t, mode, (hipStream_t)s))) return rc;
  HIP_TRY(dispatch_contract(t->q, mode, t->view(), n, waves_per_leaf, out, pts, (hipStream_t)s));
  return TBX_OK;
}

int tbx_velocity_eval(const tbx_velocity *vel, double *pts, int64_t n, double t, int bc,
                      uint32_t flags, double *out, tbx_stream s) {
  int rc = check_points(pts, n, "tbx_velocity_eval");
  if (rc) return rc;
  if ((rc = check_bc(bc))) return rc;
  if (n > 0 && !out) return fail(TBX_EINVAL, "tbx_velocity_eval: null output");
  int mode, q;
  if ((rc = mode_of(flags, &mode))) return rc;
  VelSource vs;
  if ((rc = prepare_velocity(vel, mode, (hipStream_t)s, &vs, &q, "tbx_velocity_eval"))) return rc;
  if (n == 0) return TBX_OK;
  // The extrapolating functor's value is its second-stage form, the
  // FieldExtrapFunctor itself (tree_extrap_functor.h:47-78).
  HIP_TRY(dispatch_veval(q, mode, vs, pts, n, bc, (flags & TBX_WRITE_BACK_WRAP) ? 1 : 0, t,
                         vel->g ? 1 : 0, vel->g ? vel->g(t, vel->gctx) : 1.0, out,
                         (hipStream_t)s));
  return TBX_OK;
}

int tbx_traj_rk2_v(const tbx_velocity *vel, const double *xinit, int64_t n, double tinit,
                   double tfinal, int num_rk_step, int bc, uint32_t flags, double *xsol,
                   tbx_stream s) {
  int rc = check_points(xinit, n, "tbx_traj_rk2");
  if (rc) return rc;
  if ((rc = check_bc(bc))) return rc;
  if (num_rk_step <= 0) return fail(TBX_EINVAL, "tbx_traj_rk2: num_rk_step %d", num_rk_step);
  if (n > 0 && !xsol) return fail(TBX_EINVAL, "tbx_traj_rk2: null output");
  int mode, q;
  if ((rc = mode_of(flags, &mode))) return rc;
  VelSource vs;
  if ((rc = prepare_velocity(vel, mode, (hipStream_t)s, &vs, &q, "tbx_traj_rk2"))) return rc;
  if (n == 0) return TBX_OK;
  if (xsol != xinit)
    HIP_TRY(hipMemcpyAsync(xsol, xinit, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice,
                           (hipStream_t)s));
  const double tau = (tfinal - tinit) / num_rk_step;  // traj.inc:55
  double tcur = tinit;
  for (int k = 0; k < num_rk_step; ++k) {
    const SubArgs a = sub_args(tcur, tau, vel->g, vel->gctx, bc);
    HIP_TRY(dispatch_traj(q, mode, vs, xsol, n, a, (hipStream_t)s));
    tcur = tcur + tau;  // traj.inc:65
  }
  return TBX_OK;
}

int tbx_traj_rk2(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, const double *xinit, int64_t n,
                 double tinit, double tfinal, int num_rk_step, int bc, uint32_t flags,
                 double *xsol, tbx_stream s) {
  if (!vel) return fail(TBX_EINVAL, "tbx_traj_rk2: null tree");
  const tbx_velocity v = single_tree(vel, g, gctx);
  return tbx_traj_rk2_v(&v, xinit, n, tinit, tfinal, num_rk_step, bc, flags, xsol, s);
}

int tbx_semilag_rk2(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, const tbx_tree *concc,
                    const double *pts, int64_t n, int timestep, double dt, int num_rk_step,
                    int bc, uint32_t flags, double *vals, tbx_stream s) {
  tbx_tree *conc = const_cast<tbx_tree *>(concc);
  if (!vel || !conc) return fail(TBX_EINVAL, "tbx_semilag_rk2: null tree");
  int rc = check_points(pts, n, "tbx_semilag_rk2");
  if (rc) return rc;
  if (n > 0 && !vals) return fail(TBX_EINVAL, "tbx_semilag_rk2: null output");
  if (n == 0) return TBX_OK;
  const double tinit = timestep * dt;  // semilag.inc:34-35
  const double tfinal = tinit - dt;
  double *dep = nullptr;
  HIP_TRY(hipMallocAsync((void **)&dep, sizeof(double) * 3 * (size_t)n, (hipStream_t)s));
  rc = tbx_traj_rk2(vel, g, gctx, pts, n, tinit, tfinal, num_rk_step, bc, flags, dep, s);
  if (!rc) rc = tbx_eval(conc, dep, n, bc, flags & TBX_MODE_MASK, vals, s);
  hipFreeAsync(dep, (hipStream_t)s);
  return rc;
}

int tbx_collect_grid_points(const tbx_tree *t, double *out, tbx_stream s) {
  if (!t || !out) return fail(TBX_EINVAL, "tbx_collect_grid_points: null argument");
  double node[TBX_MAX_Q + 1];
  new_nodes_host(t->q, node);
  HIP_TRY(launch_grid_points(t->view(), node, t->q + 1, out, (hipStream_t)s));
  return TBX_OK;
}

int tbx_set_grid_values(tbx_tree *t, const double *vals, tbx_stream s) {
  if (!t || !vals) return fail(TBX_EINVAL, "tbx_set_grid_values: null argument");
  return refit(t, vals, (hipStream_t)s);
}

namespace {

// vals [L][dof][P] = B c per leaf component (k_refit_expand, the refit's
// transpose), the leaf's polynomial at its own grid nodes, into vals (FP64)
// or vals32 (rounded to FP32).
int expand_grid_values(tbx_tree *t, double *vals, float *vals32, hipStream_t s,
                       const char *who) {
  int rc = ensure_refit_operator(t);
  if (rc) return rc;
  const int64_t rows = t->L * t->dof;
  if (rows > INT32_MAX) return fail(TBX_EOVERFLOW, "%s: %lld rows", who, (long long)rows);
  if (vals32)
    HIP_TRY(launch_refit_expand32(t->q + 1, t->d_coeff, rows, t->tab, t->nc, vals32, s));
  else
    HIP_TRY(launch_refit_expand(t->q + 1, t->d_coeff, rows, t->tab, t->nc, vals, s));
  return TBX_OK;
}

}  // namespace

int tbx_tree_grid_values(const tbx_tree *t, double *vals, tbx_stream s) {
  if (!t || !vals) return fail(TBX_EINVAL, "tbx_tree_grid_values: null argument");
  return expand_grid_values(const_cast<tbx_tree *>(t), vals, nullptr, (hipStream_t)s,
                            "tbx_tree_grid_values");
}

int tbx_tree_grid_values_f32(const tbx_tree *t, float *vals, tbx_stream s) {
  if (!t || !vals) return fail(TBX_EINVAL, "tbx_tree_grid_values_f32: null argument");
  return expand_grid_values(const_cast<tbx_tree *>(t), nullptr, vals, (hipStream_t)s,
                            "tbx_tree_grid_values_f32");
}

namespace {

// Whether a step's first stage reads the velocity tree's nodal values
// (TBX_OPT_NODE_VELOCITY): a single velocity tree, or the extrapolating
// functor (whose first stage is its current tree alone, traj.inc:80); not
// in TBX_EXACT (which evaluates every stage as the reference does, bit for
// bit).
bool uses_node_velocity(int vel_kind, int mode) {
  return g_node_velocity.load() && (vel_kind == TBX_VEL_TREE || vel_kind == TBX_VEL_EXTRAP) &&
         mode != kExact64;
}

// Bytes of a tree's nodal values in a mode (FP32 in the FP32 modes); the
// tree's own buffer is sized for FP64.
size_t nodal_bytes(const tbx_tree *t, int mode = kFma64) {
  const size_t w = (mode == kFp32 || mode == kFp32h) ? sizeof(float) : sizeof(double);
  return w * (size_t)t->L * t->dof * t->P();
}

// The velocity tree's values at its own grid nodes, [L][3][P], expanded
// from its coefficients by the three one-axis contractions (the refit's
// transpose, k_refit_expand; rounded to FP32 in the FP32 modes) into the
// tree's buffer or the arena.
int node_velocity(tbx_tree *vt, int mode, hipStream_t s, ScratchArena *arena, const void **out,
                  const char *who) {
  void *nodal;
  if (arena) {
    nodal = arena->take(nodal_bytes(vt, mode));
    if (!nodal) return fail(TBX_EOVERFLOW, "%s: nodal velocities exceed the scratch", who);
  } else {
    if (!vt->d_nodal) HIP_TRY(hipMalloc(&vt->d_nodal, nodal_bytes(vt)));
    nodal = vt->d_nodal;
  }
  const bool f32 = mode == kFp32 || mode == kFp32h;
  int rc = expand_grid_values(vt, f32 ? nullptr : (double *)nodal, f32 ? (float *)nodal : nullptr,
                              s, who);
  if (rc) return rc;
  *out = nodal;
  return TBX_OK;
}

// Fused arrival -> RK2 -> conc evaluation into vals [L][dof][P].
int step_values(const tbx_velocity *vel, tbx_tree *conc, int timestep, double dt,
                int num_rk_step, int bc, uint32_t flags, double *vals, double *dep,
                hipStream_t st, const char *who, ScratchArena *arena = nullptr) {
  if (!conc) return fail(TBX_EINVAL, "%s: null tree", who);
  if (!vals) return fail(TBX_EINVAL, "%s: null value buffer", who);
  if (num_rk_step <= 0 || num_rk_step > kMaxSub)
    return fail(TBX_EUNSUPPORTED, "%s: num_rk_step %d outside 1..%d", who, num_rk_step,
                kMaxSub);
  int rc = check_bc(bc);
  if (rc) return rc;
  int mode, q;
  if ((rc = mode_of(flags, &mode))) return rc;
  VelSource vs;
  if ((rc = prepare_velocity(vel, mode, st, &vs, &q, who, arena))) return rc;
  if (q != conc->q)
    return fail(TBX_EUNSUPPORTED, "%s: velocity q %d != concentration q %d", who, q, conc->q);
  TreeView cv;
  if ((rc = prepare_view(conc, mode, st, arena, &cv, who))) return rc;
  const int P = conc->P();
  StepArgs a{};
  const int want = resolve_step_shape(mode, conc->q);
  a.shape = vs.kind == kVelTree ? want : kDefaultStepShape;
  // many components (tracers sharing the trajectories) on the matrix cores
  a.mfma = g_tracer_mfma.load() && vs.kind == kVelTree && (mode == kFma64 || mode == kFp32) &&
           conc->dof >= 16 && conc->dof % 16 == 0;
  if (a.mfma) a.shape = kMfmaStepShape;
  a.npts = conc->L * P;
  const int block = a.mfma ? mfma_step_block(mode) : kStepShapeBlock[a.shape];
  a.nblocks = (a.npts + block - 1) / block;
  a.per_xcd = (a.nblocks + 7) / 8;
  if (8 * a.per_xcd > INT32_MAX) return fail(TBX_EOVERFLOW, "%s: grid too large", who);
  a.P = P;
  a.nsub = num_rk_step;
  a.bc = bc;
  a.scaled = vel->g ? 1 : 0;
  const double tinit = timestep * dt, tfinal = tinit - dt;  // semilag.inc:34-35
  a.tau = (tfinal - tinit) / num_rk_step;                    // traj.inc:55
  a.htau = 0.5 * a.tau;
  double tcur = tinit;
  for (int k = 0; k < num_rk_step; ++k) {
    a.ts[2 * k] = tcur;
    a.ts[2 * k + 1] = tcur + 0.5 * a.tau;
    a.g[2 * k] = vel->g ? vel->g(a.ts[2 * k], vel->gctx) : 1.0;
    a.g[2 * k + 1] = vel->g ? vel->g(a.ts[2 * k + 1], vel->gctx) : 1.0;
    tcur = tcur + a.tau;
  }
  new_nodes_host(conc->q, a.node);
  a.vnode = nullptr;
  if (uses_node_velocity(vel->kind, mode) && (!arena || arena->nodal) &&
      (rc = node_velocity(const_cast<tbx_tree *>(vel->trees[0]), mode, st, arena, &a.vnode,
                          who)))
    return rc;
  HIP_TRY(dispatch_step(conc->q, mode, vs, cv, a, vals, dep, st));
  return TBX_OK;
}

int step_insitu(const tbx_velocity *vel, tbx_tree *conc, int timestep, double dt,
                int num_rk_step, int bc, uint32_t flags, double *dep, hipStream_t st,
                const char *who) {
  if (!conc) return fail(TBX_EINVAL, "%s: null tree", who);
  if (!conc->d_vals)
    HIP_TRY(hipMalloc(&conc->d_vals, sizeof(double) * (size_t)conc->L * conc->dof * conc->P()));
  int rc = step_values(vel, conc, timestep, dt, num_rk_step, bc, flags, conc->d_vals, dep, st,
                       who);
  if (rc) return rc;
  return refit(conc, conc->d_vals, st);
}

}  // namespace

int tbx_semilag_values_v(const tbx_velocity *vel, const tbx_tree *conc, int timestep, double dt,
                         int num_rk_step, int bc, uint32_t flags, double *vals, double *dep,
                         tbx_stream s) {
  return step_values(vel, const_cast<tbx_tree *>(conc), timestep, dt, num_rk_step, bc, flags,
                     vals, dep, (hipStream_t)s, "tbx_semilag_values");
}

int tbx_semilag_values(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, const tbx_tree *conc,
                       int timestep, double dt, int num_rk_step, int bc, uint32_t flags,
                       double *vals, double *dep, tbx_stream s) {
  if (!vel) return fail(TBX_EINVAL, "tbx_semilag_values: null tree");
  const tbx_velocity v = single_tree(vel, g, gctx);
  return tbx_semilag_values_v(&v, conc, timestep, dt, num_rk_step, bc, flags, vals, dep, s);
}

int tbx_semilag_insitu_v(const tbx_velocity *vel, tbx_tree *conc, int timestep, double dt,
                         int num_rk_step, int bc, uint32_t flags, double *dep, tbx_stream s) {
  return step_insitu(vel, conc, timestep, dt, num_rk_step, bc, flags, dep, (hipStream_t)s,
                     "tbx_semilag_insitu");
}

int tbx_semilag_insitu(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, tbx_tree *conc,
                       int timestep, double dt, int num_rk_step, int bc, uint32_t flags,
                       double *dep, tbx_stream s) {
  if (!vel) return fail(TBX_EINVAL, "tbx_semilag_insitu: null tree");
  const tbx_velocity v = single_tree(vel, g, gctx);
  return tbx_semilag_insitu_v(&v, conc, timestep, dt, num_rk_step, bc, flags, dep, s);
}

int tbx_semilag_insitu_noalloc(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, tbx_tree *conc,
                               int timestep, double dt, int num_rk_step, int bc,
                               uint32_t flags, double *dep, size_t *scratch_bytes,
                               tbx_stream s) {
  const char *who = "tbx_semilag_insitu_noalloc";
  if (!vel || !conc) return fail(TBX_EINVAL, "%s: null tree", who);
  hipStream_t st = (hipStream_t)s;
  int mode, rc;
  if ((rc = mode_of(flags, &mode))) return rc;
  // the regions this call needs, each 256-byte aligned: values [L][dof][P],
  // the refit workspace [L*dof][KP] and FP32 mirrors of trees that have
  // none (FP32 modes); then, if they fit too, the velocity's nodal values
  // [L][3][P] (every mode but TBX_EXACT, with TBX_OPT_NODE_VELOCITY on)
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t rows = (size_t)conc->L * conc->dof;
  size_t need = up(sizeof(double) * rows * conc->P()) + up(sizeof(double) * rows * conc->KP);
  if (mode == kFp32 || mode == kFp32h) {
    if (!vel->d_coeff32) need += up(sizeof(float) * (size_t)vel->ncoeff_total() + kCoeffPad);
    if (!conc->d_coeff32) need += up(sizeof(float) * (size_t)conc->ncoeff_total() + kCoeffPad);
  }
  if (scratch_bytes) *scratch_bytes = need;
  void *base = nullptr;
  size_t cap = 0;
  HIP_TRY(noalloc_scratch(&base, &cap));
  if (need > cap)
    return fail(TBX_EOVERFLOW, "%s: the step needs %zu bytes of scratch, the maximum is %zu", who,
                need, cap);
  ScratchArena arena;
  arena.base = (char *)base;
  arena.size = cap;
  arena.nodal = false;
  if (uses_node_velocity(TBX_VEL_TREE, mode) && up(nodal_bytes(vel, mode)) <= cap - need) {
    need += up(nodal_bytes(vel, mode));
    arena.nodal = true;
  }
  if (scratch_bytes) *scratch_bytes = need;
  double *vals = (double *)arena.take(sizeof(double) * rows * conc->P());
  double *W = (double *)arena.take(sizeof(double) * rows * conc->KP);
  const tbx_velocity v = single_tree(vel, g, gctx);
  if ((rc = step_values(&v, conc, timestep, dt, num_rk_step, bc, flags, vals, dep, st, who,
                        &arena)))
    return rc;
  return refit_with(conc, vals, W, st);
}

namespace {

// The adjoint's view of one forward step: one RK2 substep of a single
// velocity tree in FP64 or TBX_EXACT arithmetic.
struct AdjCtx {
  tbx_velocity v;
  VelSource vs;
  int mode = kFma64, q = 0;
  int nsub = 1;  // RK substeps of the forward step
};

int adj_prepare(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, tbx_tree *conc,
                int num_rk_step, int bc, uint32_t flags, hipStream_t st, AdjCtx *c,
                const char *who) {
  if (!vel || !conc) return fail(TBX_EINVAL, "%s: null tree", who);
  if (num_rk_step < 1 || num_rk_step > kMaxSub)
    return fail(TBX_EUNSUPPORTED, "%s: num_rk_step %d outside 1..%d", who, num_rk_step,
                kMaxSub);
  c->nsub = num_rk_step;
  int rc = check_bc(bc);
  if (rc) return rc;
  if ((rc = mode_of(flags, &c->mode))) return rc;
  if (c->mode != kFma64 && c->mode != kExact64)
    return fail(TBX_EUNSUPPORTED, "%s: the adjoint runs in FP64 (TBX_FP64 or TBX_EXACT)", who);
  c->v = single_tree(vel, g, gctx);
  if ((rc = prepare_velocity(&c->v, c->mode, st, &c->vs, &c->q, who))) return rc;
  if (c->q != conc->q)
    return fail(TBX_EUNSUPPORTED, "%s: velocity q %d != concentration q %d", who, c->q, conc->q);
  if ((rc = ensure_refit(conc, st))) return rc;
  const size_t nvals = (size_t)conc->L * conc->dof * conc->P();
  if (!conc->d_vals) HIP_TRY(hipMalloc(&conc->d_vals, sizeof(double) * nvals));
  return TBX_OK;
}

// StepArgs of step `timestep` for the adjoint kernels (workgroup `block`).
int adj_args(const AdjCtx &c, const tbx_tree *conc, int timestep, double dt, int bc, int block,
             StepArgs *a) {
  *a = StepArgs{};
  const int P = conc->P();
  a->npts = conc->L * P;
  a->nblocks = (a->npts + block - 1) / block;
  a->per_xcd = (a->nblocks + 7) / 8;
  if (8 * a->per_xcd > INT32_MAX) return fail(TBX_EOVERFLOW, "adjoint: grid too large");
  a->P = P;
  a->nsub = c.nsub;
  a->bc = bc;
  a->scaled = c.v.g ? 1 : 0;
  const double tinit = timestep * dt, tfinal = tinit - dt;  // semilag.inc:34-35
  a->tau = (tfinal - tinit) / c.nsub;                        // traj.inc:55
  a->htau = 0.5 * a->tau;
  double tcur = tinit;
  for (int k = 0; k < c.nsub; ++k) {  // the forward's stage times (step_values)
    a->ts[2 * k] = tcur;
    a->ts[2 * k + 1] = tcur + 0.5 * a->tau;
    a->g[2 * k] = c.v.g ? c.v.g(a->ts[2 * k], c.v.gctx) : 1.0;
    a->g[2 * k + 1] = c.v.g ? c.v.g(a->ts[2 * k + 1], c.v.gctx) : 1.0;
    tcur = tcur + a->tau;
  }
  new_nodes_host(conc->q, a->node);
  return TBX_OK;
}

// lam_out = E_D^T R^T lam_next for the step `timestep` from conc's current
// coefficients; A (may be NULL) += the per-point velocity cotangent factors.
// dep (may be NULL): the step's departure points as its forward wrote them.
int adj_conc_step(const AdjCtx &c, tbx_tree *conc, int timestep, double dt, int bc,
                  const double *lam_next, double *lam_out, double *A, const double *dep,
                  hipStream_t st) {
  const int P = conc->P(), nc = conc->nc;
  const int64_t rows = conc->L * conc->dof;
  if (rows > INT32_MAX) return fail(TBX_EOVERFLOW, "adjoint: %lld rows", (long long)rows);
  // R^T: ua = pinv(B)^T lam = B (H lam) per row
  HIP_TRY(launch_refit_gemm(lam_next, nc, rows, nc, conc->KP, conc->d_H, conc->NP, conc->d_W,
                            conc->KP, nc, nullptr, st));
  HIP_TRY(launch_refit_expand(conc->q + 1, conc->d_W, rows, conc->tab, conc->KP, conc->d_vals,
                              st));
  (void)P;
  HIP_TRY(hipMemsetAsync(lam_out, 0, sizeof(double) * (size_t)conc->ncoeff_total(), st));
  StepArgs a;
  int rc = adj_args(c, conc, timestep, dt, bc, kAdjBlock, &a);
  if (rc) return rc;
  HIP_TRY(dispatch_adj_conc(conc->q, c.mode, c.vs, conc->view(), a, dep, conc->d_vals, lam_out,
                            A, st));
  return TBX_OK;
}

// grad_vel += the velocity cotangent of step `timestep` from A.
int adj_vel_step(const AdjCtx &c, tbx_tree *conc, int timestep, double dt, int bc,
                 const double *A, double *grad_vel, hipStream_t st) {
  StepArgs a;
  int rc = adj_args(c, conc, timestep, dt, bc, kAdjBlock, &a);
  if (rc) return rc;
  // TBX_FP64 with the nodal first stage (as the forward step takes it): the
  // first substep's arrival nodes add their weights to per-node slots of
  // the velocity tree, projected into grad_vel afterwards (B^T per leaf
  // component, k_refit_project) instead of each point scattering
  tbx_tree *vt = const_cast<tbx_tree *>(c.v.trees[0]);
  const bool nodal = uses_node_velocity(TBX_VEL_TREE, c.mode);
  if (nodal) {
    if ((rc = node_velocity(vt, c.mode, st, nullptr, &a.vnode, "tbx_semilag_insitu_vjp")))
      return rc;
    if (!vt->d_adjW) HIP_TRY(hipMalloc(&vt->d_adjW, nodal_bytes(vt)));
    HIP_TRY(hipMemsetAsync(vt->d_adjW, 0, nodal_bytes(vt), st));
    a.wnode = vt->d_adjW;
  }
  HIP_TRY(dispatch_adj_vel(conc->q, c.mode, c.vs, conc->view(), a, A, grad_vel, st));
  if (nodal)
    HIP_TRY(launch_refit_project_add(vt->q + 1, vt->d_adjW, vt->L * 3, 3, vt->P(), 1, vt->tab,
                                     grad_vel, vt->nc, st));
  return TBX_OK;
}

int ensure_adjA(tbx_tree *conc, hipStream_t st) {
  const size_t n = (size_t)conc->L * conc->P() * 3;
  if (!conc->d_adjA) HIP_TRY(hipMalloc(&conc->d_adjA, sizeof(double) * n));
  HIP_TRY(hipMemsetAsync(conc->d_adjA, 0, sizeof(double) * n, st));
  return TBX_OK;
}

}  // namespace

int tbx_semilag_insitu_vjp(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, const tbx_tree *concc,
                           int timestep, double dt, int num_rk_step, int bc, uint32_t flags,
                           const double *dep, const double *lam_next, double *lam_out,
                           double *grad_vel, tbx_stream s) {
  tbx_tree *conc = const_cast<tbx_tree *>(concc);
  hipStream_t st = (hipStream_t)s;
  AdjCtx c;
  int rc = adj_prepare(vel, g, gctx, conc, num_rk_step, bc, flags, st, &c,
                       "tbx_semilag_insitu_vjp");
  if (rc) return rc;
  if (!lam_next || !lam_out) return fail(TBX_EINVAL, "tbx_semilag_insitu_vjp: null cotangent");
  if (grad_vel && (rc = ensure_adjA(conc, st))) return rc;
  if ((rc = adj_conc_step(c, conc, timestep, dt, bc, lam_next, lam_out,
                          grad_vel ? conc->d_adjA : nullptr, dep, st)))
    return rc;
  if (grad_vel) return adj_vel_step(c, conc, timestep, dt, bc, conc->d_adjA, grad_vel, st);
  return TBX_OK;
}

int tbx_adjoint_run(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, tbx_tree *conc, int nsteps,
                    double dt, int bc, uint32_t flags, const double *c_target,
                    int checkpoint_every, double *grad_c0, double *grad_vel, double *loss,
                    tbx_stream s) {
  hipStream_t st = (hipStream_t)s;
  AdjCtx c;
  int rc = adj_prepare(vel, g, gctx, conc, 1, bc, flags, st, &c, "tbx_adjoint_run");
  if (rc) return rc;
  if (nsteps < 1 || checkpoint_every < 1 || !c_target || !grad_c0)
    return fail(TBX_EINVAL, "tbx_adjoint_run: bad argument");
  const int K = checkpoint_every;
  const size_t nco = (size_t)conc->ncoeff_total(), bytes = sizeof(double) * nco;
  const int nck = nsteps / K + 1;
  const bool stationary = (c.v.g == nullptr);
  // Trajectories kept on the device: a stationary velocity has the same
  // departure points every step (one array); otherwise every recomputed
  // step of a segment keeps its own.  Kept only while they fit in half the
  // free device memory; the adjoint kernel integrates them again otherwise.
  const size_t dep_bytes = sizeof(double) * 3 * (size_t)conc->L * conc->P();
  size_t free_mem = 0, total_mem = 0;
  HIP_TRY(hipMemGetInfo(&free_mem, &total_mem));
  const size_t ndep = stationary ? 1 : (size_t)(K > 1 ? K - 1 : 0);
  const bool keep_dep = ndep > 0 && ndep * dep_bytes <= free_mem / 2;
  std::vector<double *> ck(nck, nullptr), seg(K, nullptr), deps(keep_dep ? ndep : 0, nullptr);
  double *lam_a = nullptr, *lam_b = nullptr, *cT = nullptr, *partial = nullptr;
  auto release = [&]() {
    for (double *p : ck) hipFree(p);
    for (double *p : seg) hipFree(p);
    for (double *p : deps) hipFree(p);
    hipFree(lam_a);
    hipFree(lam_b);
    hipFree(cT);
    hipFree(partial);
  };
  auto body = [&]() -> int {
    for (auto &p : ck) HIP_TRY(hipMalloc(&p, bytes));
    for (auto &p : seg) HIP_TRY(hipMalloc(&p, bytes));
    for (auto &p : deps) HIP_TRY(hipMalloc(&p, dep_bytes));
    HIP_TRY(hipMalloc(&lam_a, bytes));
    HIP_TRY(hipMalloc(&lam_b, bytes));
    HIP_TRY(hipMalloc(&cT, bytes));
    HIP_TRY(hipMalloc(&partial, sizeof(double) * diff_norm2_blocks((int64_t)nco)));
    const tbx_velocity &v = c.v;
    // One forward step.  With a stationary velocity every step has step 1's
    // departure points: once they are kept, a one-component step is the
    // concentration at them plus the refit (the evaluation kernel's output
    // layout is the step's value layout for one component, and its
    // per-point arithmetic the step kernel's: the same coefficients).
    bool have_dep = false;
    auto forward = [&](int n, double *dep) -> int {
      if (stationary && keep_dep && have_dep && conc->dof == 1) {
        HIP_TRY(dispatch_eval(conc->q, c.mode, conc->view(), deps[0], (int64_t)conc->L * conc->P(),
                              bc, 0, conc->d_vals, st));
        return refit(conc, conc->d_vals, st);
      }
      const int r = step_insitu(&v, conc, n, dt, 1, bc, flags, dep, st, "tbx_adjoint_run");
      if (!r && dep && stationary) have_dep = true;
      return r;
    };
    // forward sweep, keeping c_0, c_K, c_2K, ... (and, for a stationary
    // velocity, the trajectories of step 1)
    HIP_TRY(hipMemcpyAsync(ck[0], conc->d_coeff, bytes, hipMemcpyDeviceToDevice, st));
    for (int n = 1; n <= nsteps; ++n) {
      double *dep = (stationary && keep_dep && n == 1) ? deps[0] : nullptr;
      int r = forward(n, dep);
      if (r) return r;
      if (n % K == 0 && n / K < nck)
        HIP_TRY(hipMemcpyAsync(ck[n / K], conc->d_coeff, bytes, hipMemcpyDeviceToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(cT, conc->d_coeff, bytes, hipMemcpyDeviceToDevice, st));
    // lam_T = c_T - c*, loss = 0.5 |lam_T|^2
    {
      const int nb = diff_norm2_blocks(nco);
      HIP_TRY(launch_diff_norm2(cT, c_target, nco, lam_a, partial, st));
      std::vector<double> h(nb);
      HIP_TRY(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      double dot = 0.0;
      for (double v : h) dot += v;
      if (loss) *loss = 0.5 * dot;
    }
    // backward sweep by segments, recomputing c_n from the checkpoints
    if (grad_vel) {
      HIP_TRY(hipMemsetAsync(grad_vel, 0, sizeof(double) * (size_t)vel->L * 3 * vel->nc, st));
      int r = ensure_adjA(conc, st);
      if (r) return r;
    }
    for (int s0 = ((nsteps - 1) / K) * K; s0 >= 0; s0 -= K) {
      const int s1 = std::min(s0 + K, nsteps);  // steps s0+1 .. s1
      HIP_TRY(hipMemcpyAsync(seg[0], ck[s0 / K], bytes, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(conc->d_coeff, seg[0], bytes, hipMemcpyDeviceToDevice, st));
      conc->f32_stale = true;
      for (int n = s0 + 1; n < s1; ++n) {
        double *dep = (!stationary && keep_dep) ? deps[n - s0 - 1] : nullptr;
        int r = forward(n, dep);
        if (r) return r;
        HIP_TRY(hipMemcpyAsync(seg[n - s0], conc->d_coeff, bytes, hipMemcpyDeviceToDevice, st));
      }
      for (int n = s1; n > s0; --n) {  // step n maps c_{n-1} -> c_n
        HIP_TRY(hipMemcpyAsync(conc->d_coeff, seg[n - 1 - s0], bytes, hipMemcpyDeviceToDevice,
                               st));
        conc->f32_stale = true;
        double *A = nullptr;
        if (grad_vel) {
          if (!stationary) {
            int r = ensure_adjA(conc, st);
            if (r) return r;
          }
          A = conc->d_adjA;
        }
        // the segment's last step was not recomputed: its trajectories are
        // integrated again unless the velocity is stationary
        const double *dep = !keep_dep      ? nullptr
                            : stationary   ? deps[0]
                            : n < s1       ? deps[n - s0 - 1]
                                           : nullptr;
        int r = adj_conc_step(c, conc, n, dt, bc, lam_a, lam_b, A, dep, st);
        if (r) return r;
        if (grad_vel && !stationary &&
            (r = adj_vel_step(c, conc, n, dt, bc, conc->d_adjA, grad_vel, st)))
          return r;
        std::swap(lam_a, lam_b);
      }
    }
    // a stationary velocity has the same trajectories every step: one
    // scatter of the accumulated factors
    if (grad_vel && stationary) {
      int r = adj_vel_step(c, conc, 1, dt, bc, conc->d_adjA, grad_vel, st);
      if (r) return r;
    }
    HIP_TRY(hipMemcpyAsync(grad_c0, lam_a, bytes, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(conc->d_coeff, cT, bytes, hipMemcpyDeviceToDevice, st));
    conc->f32_stale = true;
    HIP_TRY(hipStreamSynchronize(st));
    return TBX_OK;
  };
  rc = body();
  hipStreamSynchronize(st);
  release();
  return rc;
}

struct tbx_grid_s {
  tbx_tree *tree = nullptr;  // leaf structure (keys, directory); its coefficients are unused
  int n = 0, dof = 0;
  int64_t L = 0;
  double *d_s = nullptr, *d_s_next = nullptr;
  float *d_s32 = nullptr;
  bool s32_stale = true;
  int64_t nsamples() const { return L * dof * (int64_t)n * n * n; }
};

namespace {

void free_grid(tbx_grid *g) {
  if (!g) return;
  if (g->tree) tbx_tree_destroy(g->tree);
  hipFree(g->d_s);
  hipFree(g->d_s_next);
  hipFree(g->d_s32);
  delete g;
}

int grid_mode(uint32_t flags, int *mode, const char *who) {
  int rc = mode_of(flags, mode);
  if (rc) return rc;
  if (*mode == kFp32h) return fail(TBX_EUNSUPPORTED, "%s: no FP64-head mode for grid leaves", who);
  return TBX_OK;
}

int grid_prepare(tbx_grid *g, int mode, hipStream_t st) {
  if (mode != kFp32) return TBX_OK;
  if (!g->d_s32) {
    HIP_TRY(hipMalloc(&g->d_s32, sizeof(float) * (size_t)g->nsamples()));
    g->s32_stale = true;
  }
  if (g->s32_stale) {
    HIP_TRY(launch_to_f32(g->d_s, g->d_s32, g->nsamples(), st));
    g->s32_stale = false;
  }
  return TBX_OK;
}

}  // namespace

int tbx_grid_create(const tbx_grid_desc *desc, tbx_grid **out) {
  if (!desc || !out) return fail(TBX_EINVAL, "tbx_grid_create: null argument");
  *out = nullptr;
  if (desc->n < 4 || desc->n > 16)
    return fail(TBX_EUNSUPPORTED, "tbx_grid_create: %d samples per axis outside 4..16", desc->n);
  if (desc->dof < 1) return fail(TBX_EINVAL, "tbx_grid_create: dof %d", desc->dof);
  tbx_tree_desc td{desc->num_leaves, desc->anchor, desc->depth, 1, 1, nullptr};
  tbx_grid *g = new tbx_grid_s();
  int rc = tbx_tree_create(&td, &g->tree);
  if (rc) {
    delete g;
    return rc;
  }
  g->n = desc->n;
  g->dof = desc->dof;
  g->L = desc->num_leaves;
  if ((double)g->nsamples() > 4.0e10) {
    free_grid(g);
    return fail(TBX_EOVERFLOW, "tbx_grid_create: sample array too large");
  }
  const size_t bytes = sizeof(double) * (size_t)g->nsamples();
  hipError_t e;
  if ((e = hipMalloc(&g->d_s, bytes)) != hipSuccess ||
      (e = hipMalloc(&g->d_s_next, bytes)) != hipSuccess) {
    free_grid(g);
    return fail(TBX_ENOMEM, "tbx_grid_create: hipMalloc: %s", hipGetErrorString(e));
  }
  e = desc->samples ? hipMemcpy(g->d_s, desc->samples, bytes, hipMemcpyHostToDevice)
                    : hipMemset(g->d_s, 0, bytes);
  if (e != hipSuccess) {
    free_grid(g);
    return fail(TBX_EHIP, "tbx_grid_create: upload: %s", hipGetErrorString(e));
  }
  *out = g;
  return TBX_OK;
}

int tbx_grid_destroy(tbx_grid *g) {
  if (!g) return fail(TBX_EINVAL, "tbx_grid_destroy: null grid");
  hipDeviceSynchronize();
  free_grid(g);
  return TBX_OK;
}

int tbx_grid_set_samples(tbx_grid *g, const double *samples, int on_device, tbx_stream s) {
  if (!g || !samples) return fail(TBX_EINVAL, "tbx_grid_set_samples: null argument");
  HIP_TRY(hipMemcpyAsync(g->d_s, samples, sizeof(double) * (size_t)g->nsamples(),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                         (hipStream_t)s));
  if (!on_device) HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  g->s32_stale = true;
  return TBX_OK;
}

int tbx_grid_get_samples(const tbx_grid *g, double *samples, int on_device, tbx_stream s) {
  if (!g || !samples) return fail(TBX_EINVAL, "tbx_grid_get_samples: null argument");
  HIP_TRY(hipMemcpyAsync(samples, g->d_s, sizeof(double) * (size_t)g->nsamples(),
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                         (hipStream_t)s));
  if (!on_device) HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  return TBX_OK;
}

int tbx_grid_eval(const tbx_grid *gc, double *pts, int64_t n, int bc, uint32_t flags,
                  double *out, tbx_stream s) {
  tbx_grid *g = const_cast<tbx_grid *>(gc);
  if (!g) return fail(TBX_EINVAL, "tbx_grid_eval: null grid");
  int rc = check_points(pts, n, "tbx_grid_eval");
  if (rc) return rc;
  if ((rc = check_bc(bc))) return rc;
  if (n > 0 && !out) return fail(TBX_EINVAL, "tbx_grid_eval: null output");
  int mode;
  if ((rc = grid_mode(flags, &mode, "tbx_grid_eval"))) return rc;
  if ((rc = grid_prepare(g, mode, (hipStream_t)s))) return rc;
  HIP_TRY(launch_grid_eval(mode, g->tree->view(), g->d_s, g->d_s32, g->n, g->dof, pts, n, bc,
                           (flags & TBX_WRITE_BACK_WRAP) ? 1 : 0, out, (hipStream_t)s));
  return TBX_OK;
}

int tbx_grid_semilag_insitu(const tbx_grid *velc, tbx_tfac_fn g, void *gctx, tbx_grid *conc,
                            int timestep, double dt, int num_rk_step, int bc, uint32_t flags,
                            double *dep, tbx_stream s) {
  tbx_grid *vel = const_cast<tbx_grid *>(velc);
  const char *who = "tbx_grid_semilag_insitu";
  if (!vel || !conc) return fail(TBX_EINVAL, "%s: null grid", who);
  if (vel->dof != 3) return fail(TBX_EINVAL, "%s: velocity dof %d", who, vel->dof);
  if (num_rk_step <= 0 || num_rk_step > kMaxSub)
    return fail(TBX_EUNSUPPORTED, "%s: num_rk_step %d outside 1..%d", who, num_rk_step, kMaxSub);
  int rc = check_bc(bc);
  if (rc) return rc;
  int mode;
  if ((rc = grid_mode(flags, &mode, who))) return rc;
  hipStream_t st = (hipStream_t)s;
  if ((rc = grid_prepare(vel, mode, st)) || (rc = grid_prepare(conc, mode, st))) return rc;
  GridStepArgs a{};
  const int N = conc->n;
  a.npts = conc->L * (int64_t)N * N * N;
  a.nsub = num_rk_step;
  a.bc = bc;
  a.scaled = g ? 1 : 0;
  const double tinit = timestep * dt, tfinal = tinit - dt;  // semilag.inc:34-35
  a.tau = (tfinal - tinit) / num_rk_step;                    // traj.inc:55
  a.htau = 0.5 * a.tau;
  double tcur = tinit;
  for (int k = 0; k < num_rk_step; ++k) {
    a.g[2 * k] = g ? g(tcur, gctx) : 1.0;
    a.g[2 * k + 1] = g ? g(tcur + 0.5 * a.tau, gctx) : 1.0;
    tcur = tcur + a.tau;
  }
  for (int j = 0; j < N; ++j) a.node[j] = (double)j / (N - 1);
  // FP32 steps on uniform trees of 8^3-sample leaves shared by velocity
  // and concentration: sibling blocks of 8 leaves staged in LDS.  (In FP64
  // the staged block takes 128 KiB, one workgroup per CU at two waves per
  // SIMD, and the global-memory kernel is faster: 18.9 against 22.7 ms per
  // step of the 262,144-leaf workload, against 11.2 and 12.8 ms in FP32.)
  {
    const tbx_tree *tv = vel->tree, *tc = conc->tree;
    const int lds = g_grid_lds.load();
    bool blk = (lds == 2 || (lds == 1 && mode == kFp32)) && vel->n == 8 && N == 8 &&
               conc->dof == 1 &&
               tv->L == tc->L &&
               tc->L % 8 == 0 && tc->max_depth >= 1 && tv->h_key == tc->h_key;
    for (int64_t j = 0; blk && j < tc->L; ++j) blk = (int)tc->h_info[j].w == tc->max_depth;
    a.blk = blk ? 1 : 0;
  }
  // the first stage at a sample point from the velocity sample there
  // (TBX_OPT_NODE_VELOCITY; sample points of equal-sized leaves coincide)
  a.nodal = uses_node_velocity(TBX_VEL_TREE, mode) && vel->n == N ? 1 : 0;
  HIP_TRY(launch_grid_step(mode, vel->tree->view(), vel->d_s, vel->d_s32, vel->n,
                           conc->tree->view(), conc->d_s, conc->d_s32, N, conc->dof, a,
                           conc->d_s_next, dep, st));
  std::swap(conc->d_s, conc->d_s_next);
  conc->s32_stale = true;
  return TBX_OK;
}

struct tbx_batch_s {
  int nprob = 0, bc = 0, cfg = -1, parity = 0, grid = 0;
  std::vector<const tbx_tree *> vel, conc;
  std::vector<double> dt;
  std::vector<BatchProblem> host;  // table for the current configuration
  float *d_vel32 = nullptr, *d_conc32[2] = {nullptr, nullptr};
  float *d_vnode = nullptr;        // the velocities at their grid nodes, or null
  std::map<int, double *> d_M;     // refit operator per degree
  BatchProblem *d_table = nullptr;
  unsigned long long *d_counter = nullptr;
  int64_t total_items = 0;
  std::vector<int64_t> conc_off;   // first FP32 coefficient of each problem
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  // in-situ steps carried by the same graphs ahead of the persistent kernel
  struct Attached {
    const tbx_tree *vel;
    tbx_tree *conc;
    double dt;
    int nsub;
    uint32_t flags;
  };
  std::vector<Attached> attached;
};

namespace {

void free_graphs(tbx_batch *b) {
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
}

void free_batch(tbx_batch *b) {
  if (!b) return;
  hipDeviceSynchronize();
  free_graphs(b);
  if (b->cap) hipStreamDestroy(b->cap);
  hipFree(b->d_vel32);
  hipFree(b->d_vnode);
  hipFree(b->d_conc32[0]);
  hipFree(b->d_conc32[1]);
  for (auto &kv : b->d_M) hipFree(kv.second);
  hipFree(b->d_table);
  hipFree(b->d_counter);
  delete b;
}

// Problem table and persistent grid of configuration cfg; two graphs (one
// per parity) of counter reset + persistent kernel.
int batch_configure(tbx_batch *b, int cfg) {
  if (cfg < 0 || cfg >= kBatchConfigs) return fail(TBX_EINVAL, "batch: configuration %d", cfg);
  HIP_TRY(hipDeviceSynchronize());
  free_graphs(b);
  const int lpi = kBatchLpi[cfg];
  int64_t item = 0;
  for (int p = 0; p < b->nprob; ++p) {
    BatchProblem &pb = b->host[p];
    pb.item0 = item;
    item += (pb.L + lpi - 1) / lpi;
  }
  b->total_items = item;
  HIP_TRY(hipMemcpy(b->d_table, b->host.data(), sizeof(BatchProblem) * b->nprob,
                    hipMemcpyHostToDevice));
  int dev = 0, ncu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  b->grid = ncu * batch_blocks_per_cu(cfg);
  // the attached steps' nodal velocity buffers (the options in force now
  // decide whether the captured steps use them)
  for (const auto &at : b->attached) {
    int md;
    if (mode_of(at.flags, &md) == TBX_OK && uses_node_velocity(TBX_VEL_TREE, md) &&
        !at.vel->d_nodal) {
      tbx_tree *vt = const_cast<tbx_tree *>(at.vel);
      int rc = ensure_refit_operator(vt);
      if (rc) return rc;
      HIP_TRY(hipMalloc(&vt->d_nodal, nodal_bytes(vt)));
    }
  }
  for (int k = 0; k < 2; ++k) {
    HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
    // the attached steps' kernels (their buffers were made when attached,
    // so nothing is allocated or cleared inside the capture)
    int arc = TBX_OK;
    for (const auto &at : b->attached) {
      if (arc) break;
      const tbx_velocity v = single_tree(at.vel, nullptr, nullptr);
      arc = step_values(&v, at.conc, 1, at.dt, at.nsub, b->bc, at.flags, at.conc->d_vals,
                        nullptr, b->cap, "tbx_batch_attach_step");
      if (!arc) arc = refit_with(at.conc, at.conc->d_vals, at.conc->d_W, b->cap);
    }
    hipError_t e = arc ? hipSuccess
                       : launch_batch(cfg, b->d_table, b->nprob, k, b->total_items, b->d_counter,
                                      b->grid, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->cap, &g);
    if (arc) {
      if (g) (void)hipGraphDestroy(g);
      return arc;
    }
    if (e != hipSuccess) return fail(TBX_EHIP, "batch capture: %s", hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(TBX_EHIP, "batch capture: %s", hipGetErrorString(e2));
    b->graph[k] = g;
    HIP_TRY(hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0));
  }
  b->cfg = cfg;
  return TBX_OK;
}

}  // namespace

int tbx_batch_create(int nprob, const tbx_tree *const *vel, const tbx_tree *const *conc,
                     const double *dt, int bc, tbx_batch **out) {
  if (!out || nprob < 1 || !vel || !conc || !dt)
    return fail(TBX_EINVAL, "tbx_batch_create: bad argument");
  *out = nullptr;
  int rc = check_bc(bc);
  if (rc) return rc;
  for (int p = 0; p < nprob; ++p) {
    if (!vel[p] || !conc[p]) return fail(TBX_EINVAL, "tbx_batch_create: problem %d: null tree", p);
    if (vel[p]->dof != 3 || conc[p]->dof != 1)
      return fail(TBX_EINVAL, "tbx_batch_create: problem %d: dof %d/%d (want 3/1)", p,
                  vel[p]->dof, conc[p]->dof);
    const int q = conc[p]->q;
    if (vel[p]->q != q) return fail(TBX_EUNSUPPORTED, "tbx_batch_create: problem %d: degrees", p);
    if (q != 6 && q != 8 && q != 10)
      return fail(TBX_EUNSUPPORTED, "tbx_batch_create: problem %d: degree %d (6, 8 or 10)", p, q);
  }
  tbx_batch *b = new tbx_batch_s();
  b->nprob = nprob;
  b->bc = bc;
  b->vel.assign(vel, vel + nprob);
  b->conc.assign(conc, conc + nprob);
  b->dt.assign(dt, dt + nprob);
  auto body = [&]() -> int {
    int64_t nv = 0, nc = 0;
    b->conc_off.resize(nprob);
    for (int p = 0; p < nprob; ++p) {
      nv += vel[p]->ncoeff_total();
      b->conc_off[p] = nc;
      nc += conc[p]->ncoeff_total();
    }
    HIP_TRY(hipMalloc(&b->d_vel32, sizeof(float) * nv + kCoeffPad));
    HIP_TRY(hipMalloc(&b->d_conc32[0], sizeof(float) * nc + kCoeffPad));
    HIP_TRY(hipMalloc(&b->d_conc32[1], sizeof(float) * nc + kCoeffPad));
    HIP_TRY(hipMalloc(&b->d_table, sizeof(BatchProblem) * nprob));
    HIP_TRY(hipMalloc(&b->d_counter, sizeof(unsigned long long)));
    HIP_TRY(hipStreamCreate(&b->cap));
    for (int q : {6, 8, 10}) {
      const std::vector<double> &M = pt2coeff_cached(q);
      double *d = nullptr;
      HIP_TRY(hipMalloc(&d, sizeof(double) * M.size()));
      b->d_M[q] = d;
      HIP_TRY(hipMemcpy(d, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice));
    }
    b->host.resize(nprob);
    // the stationary velocities at their grid nodes for the first RK2
    // stage (TBX_OPT_NODE_VELOCITY when the batch is made), if they take at
    // most half of the free device memory; evaluated otherwise
    size_t nodal = 0;
    for (int p = 0; p < nprob; ++p) nodal += nodal_bytes(vel[p], kFp32);
    if (uses_node_velocity(TBX_VEL_TREE, kFp32)) {
      size_t freeb = 0, totalb = 0;
      HIP_TRY(hipMemGetInfo(&freeb, &totalb));
      if (nodal <= freeb / 2) HIP_TRY(hipMalloc(&b->d_vnode, nodal));
    }
    int64_t ov = 0, onode = 0;
    for (int p = 0; p < nprob; ++p) {
      const tbx_tree *tv = vel[p], *tc = conc[p];
      HIP_TRY(launch_to_f32(tv->d_coeff, b->d_vel32 + ov, tv->ncoeff_total(), nullptr));
      HIP_TRY(launch_to_f32(tc->d_coeff, b->d_conc32[0] + b->conc_off[p], tc->ncoeff_total(),
                            nullptr));
      BatchProblem &pb = b->host[p];
      pb = BatchProblem{};
      pb.vel = tv->view();
      pb.vel.coeff32 = b->d_vel32 + ov;
      for (int k = 0; k < 2; ++k) {
        pb.conc[k] = tc->view();
        pb.conc[k].coeff32 = b->d_conc32[k] + b->conc_off[p];
      }
      pb.out32[0] = b->d_conc32[1] + b->conc_off[p];
      pb.out32[1] = b->d_conc32[0] + b->conc_off[p];
      pb.M = b->d_M[tc->q];
      if (b->d_vnode) {
        float *vn = b->d_vnode + onode;
        const int erc = expand_grid_values(const_cast<tbx_tree *>(tv), nullptr, vn, nullptr,
                                           "tbx_batch_create");
        if (erc) return erc;
        pb.vnode = vn;
        onode += (int64_t)tv->L * 3 * tv->P();
      }
      pb.L = tc->L;
      pb.q = tc->q;
      pb.bc = bc;
      // one RK2 substep from t = dt back to 0 (semilag.inc:34-35,
      // traj.inc:55); the velocity is stationary
      const double tinit = 1 * dt[p], tfinal = tinit - dt[p];
      pb.tau = (tfinal - tinit) / 1;
      pb.htau = 0.5 * pb.tau;
      new_nodes_host(tc->q, pb.node);
      ov += tv->ncoeff_total();
    }
    HIP_TRY(hipDeviceSynchronize());
    return batch_configure(b, 2);
  };
  rc = body();
  if (rc) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return TBX_OK;
}

int tbx_batch_destroy(tbx_batch *b) {
  if (!b) return fail(TBX_EINVAL, "tbx_batch_destroy: null batch");
  free_batch(b);
  return TBX_OK;
}

int tbx_batch_set_config(tbx_batch *b, int cfg) {
  if (!b) return fail(TBX_EINVAL, "tbx_batch_set_config: null batch");
  return batch_configure(b, cfg);
}

int tbx_batch_get_config(const tbx_batch *b, int *cfg) {
  if (!b || !cfg) return fail(TBX_EINVAL, "tbx_batch_get_config: null argument");
  *cfg = b->cfg;
  return TBX_OK;
}

int tbx_batch_node_velocity(const tbx_batch *b, int *on) {
  if (!b || !on) return fail(TBX_EINVAL, "tbx_batch_node_velocity: null argument");
  *on = b->d_vnode ? 1 : 0;
  return TBX_OK;
}

int tbx_batch_tune(tbx_batch *b, double *ms, tbx_stream s) {
  if (!b) return fail(TBX_EINVAL, "tbx_batch_tune: null batch");
  hipStream_t st = (hipStream_t)s;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int best = -1;
  float best_ms = 0.f;
  int rc = TBX_OK;
  for (int cfg = 0; cfg < kBatchConfigs && !rc; ++cfg) {
    if ((rc = batch_configure(b, cfg))) break;
    // a step of the current parity reads one buffer and writes the other:
    // repeating it leaves the state unchanged
    hipError_t e = launch_batch(cfg, b->d_table, b->nprob, b->parity, b->total_items,
                                b->d_counter, b->grid, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess)
      e = launch_batch(cfg, b->d_table, b->nprob, b->parity, b->total_items, b->d_counter,
                       b->grid, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) {
      rc = fail(TBX_EHIP, "tbx_batch_tune: %s", hipGetErrorString(e));
      break;
    }
    if (ms) ms[cfg] = t;
    if (best < 0 || t < best_ms) {
      best = cfg;
      best_ms = t;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  return batch_configure(b, best);
}

int tbx_batch_attach_step(tbx_batch *b, const tbx_tree *vel, tbx_tree *conc, double dt,
                          int num_rk_step, uint32_t flags) {
  const char *who = "tbx_batch_attach_step";
  if (!b || !vel || !conc) return fail(TBX_EINVAL, "%s: null argument", who);
  if (num_rk_step < 1 || num_rk_step > kMaxSub)
    return fail(TBX_EUNSUPPORTED, "%s: num_rk_step %d outside 1..%d", who, num_rk_step, kMaxSub);
  int mode, rc;
  if ((rc = mode_of(flags, &mode))) return rc;
  if (vel->dof != 3 || vel->q != conc->q)
    return fail(TBX_EINVAL, "%s: velocity dof %d / degree %d for a degree-%d concentration", who,
                vel->dof, vel->q, conc->q);
  // every buffer the step uses, made now: the capture allocates nothing
  HIP_TRY(hipDeviceSynchronize());
  if ((rc = ensure_refit(conc, nullptr))) return rc;
  if (!conc->d_vals)
    HIP_TRY(hipMalloc(&conc->d_vals, sizeof(double) * (size_t)conc->L * conc->dof * conc->P()));
  if ((rc = prepare_mode(const_cast<tbx_tree *>(vel), mode, nullptr)) ||
      (rc = prepare_mode(conc, mode, nullptr)))
    return rc;
  HIP_TRY(hipDeviceSynchronize());
  b->attached.push_back({vel, conc, dt, num_rk_step, flags});
  return batch_configure(b, b->cfg < 0 ? 2 : b->cfg);
}

int tbx_batch_step(tbx_batch *b, int nsteps, tbx_stream s) {
  if (!b || nsteps < 0) return fail(TBX_EINVAL, "tbx_batch_step: bad argument");
  for (int k = 0; k < nsteps; ++k) {
    HIP_TRY(hipGraphLaunch(b->exec[b->parity], (hipStream_t)s));
    b->parity ^= 1;
  }
  return TBX_OK;
}

int tbx_batch_get_coeffs(const tbx_batch *b, int p, double *coeff, tbx_stream s) {
  if (!b || !coeff || p < 0 || p >= b->nprob)
    return fail(TBX_EINVAL, "tbx_batch_get_coeffs: bad argument");
  const size_t n = (size_t)b->conc[p]->ncoeff_total();
  std::vector<float> tmp(n);
  HIP_TRY(hipMemcpyAsync(tmp.data(), b->d_conc32[b->parity] + b->conc_off[p], sizeof(float) * n,
                         hipMemcpyDeviceToHost, (hipStream_t)s));
  HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  for (size_t i = 0; i < n; ++i) coeff[i] = tmp[i];
  return TBX_OK;
}

int tbx_set_option(int option, int value) {
  if (option == TBX_OPT_STEP_SHAPE) {
    if (value < kAutoStepShape || value >= kStepShapes)
      return fail(TBX_EINVAL, "tbx_set_option: step shape %d outside -1..%d", value,
                  kStepShapes - 1);
    g_step_shape = value;
    return TBX_OK;
  }
  if (option == TBX_OPT_TRACER_MFMA) {
    if (value != 0 && value != 1)
      return fail(TBX_EINVAL, "tbx_set_option: tracer MFMA switch %d is not 0 or 1", value);
    g_tracer_mfma = value;
    return TBX_OK;
  }
  if (option == TBX_OPT_GRID_LDS) {
    if (value < 0 || value > 2)
      return fail(TBX_EINVAL, "tbx_set_option: grid LDS switch %d outside 0..2", value);
    g_grid_lds = value;
    return TBX_OK;
  }
  if (option == TBX_OPT_NODE_VELOCITY) {
    if (value != 0 && value != 1)
      return fail(TBX_EINVAL, "tbx_set_option: node velocity switch %d is not 0 or 1", value);
    g_node_velocity = value;
    return TBX_OK;
  }
  return fail(TBX_EINVAL, "tbx_set_option: unknown option %d", option);
}

int tbx_step_shape(uint32_t flags, int q, int *shape) {
  if (!shape) return fail(TBX_EINVAL, "tbx_step_shape: null output");
  if (q < 1 || q > TBX_MAX_Q) return fail(TBX_EINVAL, "tbx_step_shape: q %d outside 1..%d", q, TBX_MAX_Q);
  int mode, rc;
  if ((rc = mode_of(flags, &mode))) return rc;
  *shape = resolve_step_shape(mode, q);
  return TBX_OK;
}

int tbx_get_option(int option, int *value) {
  if (!value) return fail(TBX_EINVAL, "tbx_get_option: null output");
  if (option == TBX_OPT_STEP_SHAPE) {
    *value = g_step_shape;
    return TBX_OK;
  }
  if (option == TBX_OPT_TRACER_MFMA) {
    *value = g_tracer_mfma;
    return TBX_OK;
  }
  if (option == TBX_OPT_GRID_LDS) {
    *value = g_grid_lds;
    return TBX_OK;
  }
  if (option == TBX_OPT_NODE_VELOCITY) {
    *value = g_node_velocity;
    return TBX_OK;
  }
  return fail(TBX_EINVAL, "tbx_get_option: unknown option %d", option);
}

int tbx_new_nodes(int q, double *x) {
  if (q < 0 || q > 64 || !x) return fail(TBX_EINVAL, "tbx_new_nodes: bad argument");
  new_nodes_host(q, x);
  return TBX_OK;
}

int tbx_pt2coeff_matrix(int q, double *M) {
  if (q < 1 || q > TBX_MAX_Q || !M) return fail(TBX_EINVAL, "tbx_pt2coeff_matrix: bad argument");
  const std::vector<double> &src = pt2coeff_cached(q);
  std::memcpy(M, src.data(), sizeof(double) * src.size());
  return TBX_OK;
}

int tbx_refit_gram_inverse(int q, double *H) {
  if (q < 1 || q > TBX_MAX_Q || !H) return fail(TBX_EINVAL, "tbx_refit_gram_inverse: bad argument");
  const RefitOperator &op = refit_operator_cached(q);
  const int nc = ncoef(q);
  for (int a = 0; a < nc; ++a)
    for (int b = 0; b < nc; ++b) H[(size_t)a * nc + b] = op.H[(size_t)a * op.NP + b];
  return TBX_OK;
}

int tbx_uniform_leaves(int depth, uint32_t *anchor, uint8_t *dep) {
  if (depth < 0 || depth > 10 || !anchor || !dep)
    return fail(TBX_EINVAL, "tbx_uniform_leaves: depth %d", depth);
  const int64_t L = (int64_t)1 << (3 * depth);
  for (int64_t m = 0; m < L; ++m) {
    anchor[3 * m + 0] = compact3((uint64_t)m);
    anchor[3 * m + 1] = compact3((uint64_t)m >> 1);
    anchor[3 * m + 2] = compact3((uint64_t)m >> 2);
    dep[m] = (uint8_t)depth;
  }
  return TBX_OK;
}

int tbx_first_kind_points(int64_t L, const uint32_t *anchor, const uint8_t *depth, int q,
                          double *out) {
  if (L < 0 || q < 0 || q > 64 || (L > 0 && (!anchor || !depth || !out)))
    return fail(TBX_EINVAL, "tbx_first_kind_points: bad argument");
  const int d = q + 1, P = d * d * d;
  std::vector<double> x(d);
  for (int i = 0; i < d; ++i) x[i] = -std::cos((i + 0.5) * M_PI / d) * 0.5 + 0.5;
  parallel_for(L, [&](int64_t j) {
    const double len = std::pow(0.5, depth[j]);
    const double c0 = anchor[3 * j] * len, c1 = anchor[3 * j + 1] * len, c2 = anchor[3 * j + 2] * len;
    double *o = out + (size_t)j * P * 3;
    for (int m = 0; m < P; ++m) {
      o[3 * m + 0] = c0 + len * x[m % d];
      o[3 * m + 1] = c1 + len * x[(m / d) % d];
      o[3 * m + 2] = c2 + len * x[m / (d * d)];
    }
  });
  return TBX_OK;
}

int tbx_fit_leaves(int64_t L, int q, int dof, const double *vals, double *coeff) {
  if (L < 0 || q < 0 || q > 64 || dof < 1 || (L > 0 && (!vals || !coeff)))
    return fail(TBX_EINVAL, "tbx_fit_leaves: bad argument");
  const int d = q + 1, P = d * d * d, nc = ncoef(q);
  // C[k][i] = (2/d) T_k(xi_i), halved for k = 0: the discrete Chebyshev
  // transform at first-kind nodes.
  std::vector<double> C((size_t)d * d);
  for (int i = 0; i < d; ++i) {
    const double xi = (-std::cos((i + 0.5) * M_PI / d) * 0.5 + 0.5) * 2.0 - 1.0;
    std::vector<double> p(d);
    cheb_poly_host(q, xi, p.data());
    for (int k = 0; k < d; ++k) C[(size_t)k * d + i] = p[k] * (2.0 / d) * (k == 0 ? 0.5 : 1.0);
  }
  parallel_for(L * dof, [&](int64_t r) {
    const int64_t j = r / dof;
    const int l = (int)(r % dof);
    std::vector<double> a(P), b(P);
    for (int m = 0; m < P; ++m) a[m] = vals[((size_t)j * P + m) * dof + l];
    // a[z][y][x] -> b[z][y][kx] -> a[z][ky][kx] -> b[kz][ky][kx]
    for (int zy = 0; zy < d * d; ++zy)
      for (int k = 0; k < d; ++k) {
        double s = 0;
        for (int i = 0; i < d; ++i) s += C[(size_t)k * d + i] * a[(size_t)zy * d + i];
        b[(size_t)zy * d + k] = s;
      }
    for (int z = 0; z < d; ++z)
      for (int kx = 0; kx < d; ++kx)
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int i = 0; i < d; ++i) s += C[(size_t)k * d + i] * b[((size_t)z * d + i) * d + kx];
          a[((size_t)z * d + k) * d + kx] = s;
        }
    for (int ky = 0; ky < d; ++ky)
      for (int kx = 0; kx < d; ++kx)
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int i = 0; i < d; ++i) s += C[(size_t)k * d + i] * a[((size_t)i * d + ky) * d + kx];
          b[((size_t)k * d + ky) * d + kx] = s;
        }
    double *c = coeff + ((size_t)j * dof + l) * nc;
    int m = 0;
    for (int i = 0; i < d; ++i)
      for (int jj = 0; i + jj < d; ++jj)
        for (int k = 0; i + jj + k < d; ++k) c[m++] = b[((size_t)i * d + jj) * d + k];
  });
  return TBX_OK;
}

namespace {

// C[k][a] = (2/d) T_k(xi_a), halved for k = 0, xi_a the first-kind nodes on
// [-1, 1]: the one-axis discrete Chebyshev transform of the initial fit
// (tbx_fit_leaves' table).
RefitTable fit_table(int q) {
  const int d = q + 1;
  RefitTable tab{};
  double x[TBX_MAX_Q + 1];
  first_kind_nodes_host(d, x);
  for (int a = 0; a < d; ++a) {
    double p[TBX_MAX_Q + 1];
    cheb_poly_host(q, x[a] * 2.0 - 1.0, p);
    for (int k = 0; k < d; ++k) tab.T[k][a] = p[k] * (2.0 / d) * (k == 0 ? 0.5 : 1.0);
  }
  return tab;
}

}  // namespace

int tbx_tree_first_kind_points(const tbx_tree *t, double *out, tbx_stream s) {
  if (!t || !out) return fail(TBX_EINVAL, "tbx_tree_first_kind_points: null argument");
  double x[TBX_MAX_Q + 1];
  first_kind_nodes_host(t->q + 1, x);
  HIP_TRY(launch_grid_points(t->view(), x, t->q + 1, out, (hipStream_t)s));
  return TBX_OK;
}

int tbx_tree_fit(tbx_tree *t, const double *vals, tbx_stream s) {
  if (!t || !vals) return fail(TBX_EINVAL, "tbx_tree_fit: null argument");
  const int64_t rows = t->L * t->dof;
  if (rows > INT32_MAX) return fail(TBX_EOVERFLOW, "tbx_tree_fit: %lld rows", (long long)rows);
  HIP_TRY(launch_refit_project(t->q + 1, vals, rows, t->dof, 1, t->dof, fit_table(t->q),
                               t->d_coeff, t->nc, (hipStream_t)s));
  t->f32_stale = true;
  return TBX_OK;
}

int tbx_error_points(const tbx_tree *t, double *out, tbx_stream s) {
  if (!t || !out) return fail(TBX_EINVAL, "tbx_error_points: null argument");
  const int n = t->q / 3 + 1;
  double x[TBX_MAX_Q + 1];
  first_kind_nodes_host(n, x);
  HIP_TRY(launch_grid_points(t->view(), x, n, out, (hipStream_t)s));
  return TBX_OK;
}

int tbx_check_cheb_output(const tbx_tree *tc, const double *ref, uint32_t flags, double *res,
                          tbx_stream s) {
  if (!tc || !ref || !res) return fail(TBX_EINVAL, "tbx_check_cheb_output: null argument");
  tbx_tree *t = const_cast<tbx_tree *>(tc);
  if (t->dof > 8) return fail(TBX_EUNSUPPORTED, "tbx_check_cheb_output: dof %d > 8", t->dof);
  hipStream_t st = (hipStream_t)s;
  int mode, rc;
  if ((rc = mode_of(flags, &mode))) return rc;
  if ((rc = prepare_mode(t, mode, st))) return rc;
  const int n = t->q / 3 + 1, dof = t->dof;
  const int64_t npts = t->L * (int64_t)n * n * n;
  const int nb = metric_blocks(npts);
  double *pts = nullptr, *val = nullptr, *part = nullptr;
  auto body = [&]() -> int {
    HIP_TRY(hipMallocAsync((void **)&pts, sizeof(double) * 3 * (size_t)npts, st));
    HIP_TRY(hipMallocAsync((void **)&val, sizeof(double) * dof * (size_t)npts, st));
    HIP_TRY(hipMallocAsync((void **)&part, sizeof(double) * nb * std::max(dof, 4), st));
    double x[TBX_MAX_Q + 1];
    first_kind_nodes_host(n, x);
    HIP_TRY(launch_grid_points(t->view(), x, n, pts, st));
    HIP_TRY(dispatch_eval(t->q, mode, t->view(), pts, npts, TBX_BC_FREESPACE, 0, val, st));
    std::vector<double> h((size_t)nb * std::max(dof, 4));
    // pass 0: the volume-weighted mean error per component (utils.txx:158-187)
    HIP_TRY(launch_metric(t->d_info, n, val, ref, npts, dof, nullptr, 0, part, st));
    HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * nb * dof, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double avg[8] = {};
    for (int b = 0; b < nb; ++b)
      for (int l = 0; l < dof; ++l) avg[l] += h[(size_t)b * dof + l];
    // pass 1: L2 and L-inf of the field and of the centred error (:240-293)
    HIP_TRY(launch_metric(t->d_info, n, val, ref, npts, dof, avg, 1, part, st));
    HIP_TRY(hipMemcpyAsync(h.data(), part, sizeof(double) * nb * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double l2 = 0, l2e = 0, mx = 0, mxe = 0;
    for (int b = 0; b < nb; ++b) {
      l2 += h[(size_t)b * 4 + 0];
      l2e += h[(size_t)b * 4 + 1];
      mx = std::max(mx, h[(size_t)b * 4 + 2]);
      mxe = std::max(mxe, h[(size_t)b * 4 + 3]);
    }
    res[0] = std::sqrt(l2e);
    res[1] = std::sqrt(l2e / l2);
    res[2] = mxe;
    res[3] = mxe / mx;
    return TBX_OK;
  };
  rc = body();
  hipFreeAsync(pts, st);
  hipFreeAsync(val, st);
  hipFreeAsync(part, st);
  return rc;
}

int tbx_tail_norm(int64_t L, int q, int dof, const double *coeff, double *tail) {
  if (L < 0 || q < 0 || dof < 1 || (L > 0 && (!coeff || !tail)))
    return fail(TBX_EINVAL, "tbx_tail_norm: bad argument");
  const int d = q + 1, nc = ncoef(q);
  std::vector<int> shell;
  {
    int m = 0;
    for (int i = 0; i < d; ++i)
      for (int j = 0; i + j < d; ++j)
        for (int k = 0; i + j + k < d; ++k, ++m)
          if (i + j + k == q) shell.push_back(m);
  }
  parallel_for(L, [&](int64_t j) {
    double best = 0;
    for (int l = 0; l < dof; ++l) {
      const double *c = coeff + ((size_t)j * dof + l) * nc;
      double s = 0;
      for (int m : shell) s += std::fabs(c[m]);
      best = std::max(best, s);
    }
    tail[j] = best;
  });
  return TBX_OK;
}

int tbx_balance21(int64_t *num_leaves, int64_t capacity, uint32_t *anchor, uint8_t *depth,
                  int bc) {
  if (!num_leaves || !anchor || !depth) return fail(TBX_EINVAL, "tbx_balance21: null argument");
  int rc = check_bc(bc);
  if (rc) return rc;
  int64_t L = *num_leaves;
  for (int round = 0; round < 64; ++round) {
    std::vector<uint64_t> key(L);
    for (int64_t j = 0; j < L; ++j)
      key[j] = leaf_key(anchor[3 * j], anchor[3 * j + 1], anchor[3 * j + 2], depth[j]);
    for (int64_t j = 1; j < L; ++j)
      if (key[j] <= key[j - 1]) return fail(TBX_EORDER, "tbx_balance21: leaves not in Morton order");
    std::unique_ptr<std::atomic<uint8_t>[]> split(new std::atomic<uint8_t>[L]);
    for (int64_t j = 0; j < L; ++j) split[j].store(0, std::memory_order_relaxed);
    std::atomic<int64_t> nsplit{0};
    parallel_for(L, [&](int64_t j) {
      const int dj = depth[j];
      if (dj < 2) return;
      const int64_t lim = (int64_t)1 << dj;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (!dx && !dy && !dz) continue;
            int64_t n[3] = {(int64_t)anchor[3 * j] + dx, (int64_t)anchor[3 * j + 1] + dy,
                            (int64_t)anchor[3 * j + 2] + dz};
            bool inside = true;
            for (int a = 0; a < 3; ++a) {
              if (n[a] < 0 || n[a] >= lim) {
                if (bc == TBX_BC_PERIODIC)
                  n[a] = (n[a] + lim) % lim;
                else
                  inside = false;
              }
            }
            if (!inside) continue;
            const uint64_t k = leaf_key((uint32_t)n[0], (uint32_t)n[1], (uint32_t)n[2], dj);
            const int64_t jn = std::upper_bound(key.begin(), key.end(), k) - key.begin() - 1;
            if (jn >= 0 && depth[jn] < dj - 1 && split[jn].exchange(1) == 0)
              nsplit.fetch_add(1);
          }
    });
    if (nsplit.load() == 0) break;
    int64_t newL = 0;
    for (int64_t j = 0; j < L; ++j) newL += split[j] ? 8 : 1;
    if (newL > capacity) return fail(TBX_ENOMEM, "tbx_balance21: capacity %lld too small", (long long)capacity);
    // expand back to front so the arrays can be rewritten in place
    int64_t w = newL;
    for (int64_t j = L - 1; j >= 0; --j) {
      const uint32_t ax = anchor[3 * j], ay = anchor[3 * j + 1], az = anchor[3 * j + 2];
      const uint8_t dj = depth[j];
      if (split[j]) {
        for (int c = 7; c >= 0; --c) {
          --w;
          anchor[3 * w + 0] = 2 * ax + (c & 1);
          anchor[3 * w + 1] = 2 * ay + ((c >> 1) & 1);
          anchor[3 * w + 2] = 2 * az + ((c >> 2) & 1);
          depth[w] = dj + 1;
        }
      } else {
        --w;
        anchor[3 * w + 0] = ax;
        anchor[3 * w + 1] = ay;
        anchor[3 * w + 2] = az;
        depth[w] = dj;
      }
    }
    L = newL;
  }
  *num_leaves = L;
  return TBX_OK;
}

}  // extern "C"

// RefineTree + Balance21 on the device.
namespace {

// A device buffer that frees itself (adaptivity's temporaries).
struct DevBuf {
  void *p = nullptr;
  hipStream_t s = nullptr;
  ~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

int dev_alloc(DevBuf &b, size_t bytes, hipStream_t s) {
  b.s = s;
  HIP_TRY(hipMallocAsync(&b.p, bytes ? bytes : 8, s));
  return TBX_OK;
}

// Coefficients [nleaf][dof][nc] of leaves whose grid-node values are vals
// [nleaf][dof][P]: the tree's refit operator, into out.
int refit_rows(tbx_tree *t, const double *vals, int64_t nleaf, double *out, hipStream_t s) {
  const int64_t rows = nleaf * t->dof;
  if (rows <= 0) return TBX_OK;
  DevBuf W;
  if (int rc = dev_alloc(W, sizeof(double) * (size_t)rows * t->KP, s)) return rc;
  HIP_TRY(launch_refit_project(t->q + 1, vals, rows, t->dof, t->P(), 1, t->tab, W.as<double>(),
                               t->KP, s));
  HIP_TRY(launch_refit_gemm(W.as<double>(), t->KP, rows, t->KP, t->KP, t->d_H, t->NP, out, t->nc,
                            t->nc, nullptr, s));
  return TBX_OK;
}

// Coefficients of new leaves resampled from the tree's current leaves:
// leaf n from src[n] (restrict: the 8 children starting at src[n]).
int resample(tbx_tree *t, const std::vector<int32_t> &src, const std::vector<uint4> &info,
             int restrict_children, double *out, hipStream_t s) {
  const int64_t n = (int64_t)src.size();
  if (!n) return TBX_OK;
  DevBuf dsrc, dinfo, vals;
  int rc;
  if ((rc = dev_alloc(dsrc, sizeof(int32_t) * n, s)) || (rc = dev_alloc(dinfo, sizeof(uint4) * n, s)) ||
      (rc = dev_alloc(vals, sizeof(double) * (size_t)n * t->dof * t->P(), s)))
    return rc;
  HIP_TRY(hipMemcpyAsync(dsrc.p, src.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dinfo.p, info.data(), sizeof(uint4) * n, hipMemcpyHostToDevice, s));
  double node[TBX_MAX_Q + 1];
  new_nodes_host(t->q, node);
  HIP_TRY(launch_leaf_resample(t->view(), dsrc.as<int32_t>(), dinfo.as<uint4>(),
                               restrict_children, n, node, vals.as<double>(), s));
  rc = refit_rows(t, vals.as<double>(), n, out, s);
  if (rc) return rc;
  // the host vectors must outlive the copies
  HIP_TRY(hipStreamSynchronize(s));
  return TBX_OK;
}

// One new leaf of an adaptation: its anchor and depth and where its
// coefficients come from (kind 0: old leaf src; 1: prolongation of old leaf
// src; 2: row src of the restricted parents).
struct NewLeaf {
  uint32_t a[3];
  uint8_t d;
  int kind;
  int64_t src;
};

// Replace the tree's leaves by `leaves` (valid by construction), with
// coefficients from the old leaves, from prolongation (computed here) and
// from the restricted parents pcoef.
int rebuild(tbx_tree *t, const std::vector<NewLeaf> &leaves, const double *pcoef, hipStream_t s) {
  const int64_t Ln = (int64_t)leaves.size();
  const size_t row = (size_t)t->dof * t->nc;
  std::vector<int32_t> psrc;
  std::vector<uint4> pinfo;
  std::vector<int64_t> map_old(Ln, -1), map_par(Ln, -1), map_pro(Ln, -1);
  for (int64_t n = 0; n < Ln; ++n) {
    const NewLeaf &l = leaves[n];
    if (l.kind == 0) {
      map_old[n] = l.src;
    } else if (l.kind == 2) {
      map_par[n] = l.src;
    } else {
      map_pro[n] = (int64_t)psrc.size();
      psrc.push_back((int32_t)l.src);
      pinfo.push_back(make_uint4(l.a[0], l.a[1], l.a[2], l.d));
    }
  }
  double *coeff = nullptr;
  HIP_TRY(hipMalloc(&coeff, sizeof(double) * row * Ln + kCoeffPad));
  auto body = [&]() -> int {
    DevBuf pro, maps;
    int rc;
    if ((rc = dev_alloc(pro, sizeof(double) * row * psrc.size(), s)) ||
        (rc = dev_alloc(maps, sizeof(int64_t) * 3 * Ln, s)))
      return rc;
    if ((rc = resample(t, psrc, pinfo, 0, pro.as<double>(), s))) return rc;
    int64_t *dm = maps.as<int64_t>();
    HIP_TRY(hipMemcpyAsync(dm, map_old.data(), sizeof(int64_t) * Ln, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dm + Ln, map_pro.data(), sizeof(int64_t) * Ln, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dm + 2 * Ln, map_par.data(), sizeof(int64_t) * Ln, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_gather_rows(t->d_coeff, dm, Ln, (int)row, coeff, s));
    HIP_TRY(launch_gather_rows(pro.as<double>(), dm + Ln, Ln, (int)row, coeff, s));
    if (pcoef) HIP_TRY(launch_gather_rows(pcoef, dm + 2 * Ln, Ln, (int)row, coeff, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TBX_OK;
  };
  int rc = body();
  if (rc) {
    (void)hipFree(coeff);
    return rc;
  }
  std::vector<uint32_t> an(3 * Ln);
  std::vector<uint8_t> de(Ln);
  for (int64_t n = 0; n < Ln; ++n) {
    for (int k = 0; k < 3; ++k) an[3 * n + k] = leaves[n].a[k];
    de[n] = leaves[n].d;
  }
  if ((rc = set_leaves(t, Ln, an.data(), de.data()))) {
    (void)hipFree(coeff);
    return rc;
  }
  (void)hipFree(t->d_coeff);
  t->d_coeff = coeff;
  // buffers sized by the leaf count are made again on their next use
  (void)hipFree(t->d_coeff32);
  (void)hipFree(t->d_vals);
  (void)hipFree(t->d_W);
  (void)hipFree(t->d_adjA);
  (void)hipFree(t->d_nodal);
  (void)hipFree(t->d_adjW);
  t->d_coeff32 = nullptr;
  t->d_vals = nullptr;
  t->d_W = nullptr;
  t->d_adjA = nullptr;
  t->d_nodal = nullptr;
  t->d_adjW = nullptr;
  t->f32_stale = true;
  return TBX_OK;
}

NewLeaf keep_leaf(const tbx_tree *t, int64_t j) {
  const uint4 f = t->h_info[j];
  return NewLeaf{{f.x, f.y, f.z}, (uint8_t)f.w, 0, j};
}

void push_children(const tbx_tree *t, int64_t j, std::vector<NewLeaf> &out) {
  const uint4 f = t->h_info[j];
  for (uint32_t c = 0; c < 8; ++c)
    out.push_back(NewLeaf{{2 * f.x + (c & 1), 2 * f.y + ((c >> 1) & 1), 2 * f.z + ((c >> 2) & 1)},
                          (uint8_t)(f.w + 1), 1, j});
}

}  // namespace

extern "C" {

int tbx_tree_adapt(tbx_tree *t, double tol, int min_depth, int max_depth, int bc, tbx_stream s) {
  if (!t) return fail(TBX_EINVAL, "tbx_tree_adapt: null tree");
  if (int rc = check_bc(bc)) return rc;
  if (!(tol >= 0.0) || min_depth < 0 || max_depth < min_depth || max_depth > kKeyBits)
    return fail(TBX_EINVAL, "tbx_tree_adapt: tol %g, depths %d..%d", tol, min_depth, max_depth);
  hipStream_t st = (hipStream_t)s;
  const int64_t L = t->L;
  const int q = t->q, dof = t->dof;
  const size_t row = (size_t)dof * t->nc;
  // 1. the criterion per leaf
  std::vector<double> tail(L);
  {
    DevBuf d;
    if (int rc = dev_alloc(d, sizeof(double) * L, st)) return rc;
    HIP_TRY(launch_adapt_tail(t->d_coeff, L, dof, q, d.as<double>(), st));
    HIP_TRY(hipMemcpyAsync(tail.data(), d.p, sizeof(double) * L, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  std::vector<uint8_t> refine(L);
  for (int64_t j = 0; j < L; ++j)
    refine[j] = tail[j] > tol && (int)t->h_info[j].w < max_depth ? 1 : 0;
  // 2. complete sibling groups that may merge: 8 consecutive leaves of one
  //    parent, none to be refined, the parent not above min_depth
  std::vector<int64_t> cand;
  for (int64_t j = 0; j + 8 <= L; ++j) {
    const uint4 f = t->h_info[j];
    if ((int)f.w <= min_depth || (f.x & 1) || (f.y & 1) || (f.z & 1)) continue;
    bool ok = true;
    for (uint32_t c = 0; c < 8 && ok; ++c) {
      const uint4 g = t->h_info[j + c];
      ok = g.w == f.w && g.x == f.x + (c & 1) && g.y == f.y + ((c >> 1) & 1) &&
           g.z == f.z + ((c >> 2) & 1) && !refine[j + c];
    }
    if (ok) {
      cand.push_back(j);
      j += 7;
    }
  }
  // 3. the parents' fits from their children (restriction) and their tails
  DevBuf pcoef;
  std::vector<uint8_t> merge(cand.size(), 0);
  if (!cand.empty()) {
    const int64_t nc = (int64_t)cand.size();
    std::vector<int32_t> src(nc);
    std::vector<uint4> info(nc);
    for (int64_t k = 0; k < nc; ++k) {
      const uint4 f = t->h_info[cand[k]];
      src[k] = (int32_t)cand[k];
      info[k] = make_uint4(f.x / 2, f.y / 2, f.z / 2, f.w - 1);
    }
    if (int rc = dev_alloc(pcoef, sizeof(double) * row * nc, st)) return rc;
    if (int rc = resample(t, src, info, 1, pcoef.as<double>(), st)) return rc;
    std::vector<double> ptail(nc);
    DevBuf d;
    if (int rc = dev_alloc(d, sizeof(double) * nc, st)) return rc;
    HIP_TRY(launch_adapt_tail(pcoef.as<double>(), nc, dof, q, d.as<double>(), st));
    HIP_TRY(hipMemcpyAsync(ptail.data(), d.p, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int64_t k = 0; k < nc; ++k) merge[k] = ptail[k] <= tol ? 1 : 0;
  }
  // 4. refine and merge
  std::vector<NewLeaf> leaves;
  leaves.reserve(L);
  bool changed = false;
  {
    size_t k = 0;
    for (int64_t j = 0; j < L; ++j) {
      if (k < cand.size() && cand[k] == j) {
        if (merge[k]) {
          const uint4 f = t->h_info[j];
          leaves.push_back(NewLeaf{{f.x / 2, f.y / 2, f.z / 2}, (uint8_t)(f.w - 1), 2, (int64_t)k});
          j += 7;
          ++k;
          changed = true;
          continue;
        }
        ++k;
      }
      if (refine[j]) {
        push_children(t, j, leaves);
        changed = true;
      } else {
        leaves.push_back(keep_leaf(t, j));
      }
    }
  }
  if (changed)
    if (int rc = rebuild(t, leaves, pcoef.as<double>(), st)) return rc;
  // 5. 2:1 balance: split every leaf more than one level coarser than a
  //    neighbour, until none is
  for (;;) {
    const int64_t Lb = t->L;
    std::vector<uint8_t> split(Lb);
    {
      DevBuf d;
      if (int rc = dev_alloc(d, Lb, st)) return rc;
      HIP_TRY(hipMemsetAsync(d.p, 0, Lb, st));
      HIP_TRY(launch_balance_probe(t->view(), bc, d.as<uint8_t>(), st));
      HIP_TRY(hipMemcpyAsync(split.data(), d.p, Lb, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    std::vector<NewLeaf> next;
    bool any = false;
    for (int64_t j = 0; j < Lb; ++j) {
      if (split[j]) {
        push_children(t, j, next);
        any = true;
      } else {
        next.push_back(keep_leaf(t, j));
      }
    }
    if (!any) break;
    if (int rc = rebuild(t, next, nullptr, st)) return rc;
  }
  return TBX_OK;
}

int tbx_tree_get_leaves(const tbx_tree *t, uint32_t *anchor, uint8_t *depth) {
  if (!t || !anchor || !depth) return fail(TBX_EINVAL, "tbx_tree_get_leaves: null argument");
  for (int64_t j = 0; j < t->L; ++j) {
    const uint4 f = t->h_info[j];
    anchor[3 * j] = f.x;
    anchor[3 * j + 1] = f.y;
    anchor[3 * j + 2] = f.z;
    depth[j] = (uint8_t)f.w;
  }
  return TBX_OK;
}

// Write2File: the tree on order^3 uniform points per leaf as VTK XML
// hexahedra (one .vtu piece plus the .pvtu that names it).
namespace {

struct AppendedBlock {
  const void *data;
  uint64_t bytes;
};

int write_vtu_files(const std::string &prefix, int dof, int64_t npts, int64_t ncells,
                    const std::vector<double> &xyz, const std::vector<double> &val,
                    const std::vector<int64_t> &conn, const std::vector<int64_t> &offs,
                    const std::vector<uint8_t> &types) {
  const std::string piece = prefix + "_000000.vtu";
  const size_t slash = piece.find_last_of('/');
  const std::string piece_name = slash == std::string::npos ? piece : piece.substr(slash + 1);
  const AppendedBlock blocks[5] = {{val.data(), sizeof(double) * val.size()},
                                   {xyz.data(), sizeof(double) * xyz.size()},
                                   {conn.data(), sizeof(int64_t) * conn.size()},
                                   {offs.data(), sizeof(int64_t) * offs.size()},
                                   {types.data(), types.size()}};
  uint64_t off[5], o = 0;
  for (int b = 0; b < 5; ++b) {
    off[b] = o;
    o += sizeof(uint64_t) + blocks[b].bytes;
  }
  FILE *f = std::fopen(piece.c_str(), "wb");
  if (!f) return fail(TBX_EINVAL, "tbx_tree_write_vtu: cannot open %s", piece.c_str());
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n"
               " <UnstructuredGrid>\n"
               "  <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n"
               "   <PointData Scalars=\"value\">\n"
               "    <DataArray type=\"Float64\" Name=\"value\" NumberOfComponents=\"%d\" "
               "format=\"appended\" offset=\"%llu\"/>\n"
               "   </PointData>\n"
               "   <Points>\n"
               "    <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" "
               "offset=\"%llu\"/>\n"
               "   </Points>\n"
               "   <Cells>\n"
               "    <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" "
               "offset=\"%llu\"/>\n"
               "    <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" "
               "offset=\"%llu\"/>\n"
               "    <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" "
               "offset=\"%llu\"/>\n"
               "   </Cells>\n"
               "  </Piece>\n"
               " </UnstructuredGrid>\n"
               " <AppendedData encoding=\"raw\">\n_",
               (long long)npts, (long long)ncells, dof, (unsigned long long)off[0],
               (unsigned long long)off[1], (unsigned long long)off[2], (unsigned long long)off[3],
               (unsigned long long)off[4]);
  bool ok = true;
  for (int b = 0; b < 5 && ok; ++b) {
    ok = std::fwrite(&blocks[b].bytes, sizeof(uint64_t), 1, f) == 1;
    if (ok && blocks[b].bytes)
      ok = std::fwrite(blocks[b].data, 1, blocks[b].bytes, f) == blocks[b].bytes;
  }
  std::fprintf(f, "\n </AppendedData>\n</VTKFile>\n");
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(TBX_EINVAL, "tbx_tree_write_vtu: write to %s failed", piece.c_str());
  const std::string pv = prefix + ".pvtu";
  f = std::fopen(pv.c_str(), "w");
  if (!f) return fail(TBX_EINVAL, "tbx_tree_write_vtu: cannot open %s", pv.c_str());
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n"
               " <PUnstructuredGrid GhostLevel=\"0\">\n"
               "  <PPointData Scalars=\"value\">\n"
               "   <PDataArray type=\"Float64\" Name=\"value\" NumberOfComponents=\"%d\"/>\n"
               "  </PPointData>\n"
               "  <PPoints>\n"
               "   <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n"
               "  </PPoints>\n"
               "  <Piece Source=\"%s\"/>\n"
               " </PUnstructuredGrid>\n"
               "</VTKFile>\n",
               dof, piece_name.c_str());
  if (std::fclose(f) != 0) return fail(TBX_EINVAL, "tbx_tree_write_vtu: write to %s failed", pv.c_str());
  return TBX_OK;
}

}  // namespace

int tbx_tree_write_vtu(const tbx_tree *tc, const char *prefix, int order, tbx_stream s) {
  if (!tc || !prefix) return fail(TBX_EINVAL, "tbx_tree_write_vtu: null argument");
  if (order < 2 || order > TBX_MAX_Q + 1)
    return fail(TBX_EINVAL, "tbx_tree_write_vtu: order %d outside 2..%d", order, TBX_MAX_Q + 1);
  tbx_tree *t = const_cast<tbx_tree *>(tc);
  hipStream_t st = (hipStream_t)s;
  const int n = order, n3 = n * n * n;
  const int64_t npts = t->L * (int64_t)n3, ncells = t->L * (int64_t)(n - 1) * (n - 1) * (n - 1);
  // the leaf's own polynomial on its faces: evaluation points are pulled
  // inside the leaf by a factor 1 - 2^-40 of the offset from its corner, so
  // an upper-face point is not located in the neighbour; the file holds the
  // uniform points
  double node[TBX_MAX_Q + 1];
  std::vector<double> unode(n);
  for (int a = 0; a < n; ++a) {
    unode[a] = (double)a / (n - 1);
    node[a] = unode[a] * (1.0 - std::ldexp(1.0, -40));
  }
  double *pts = nullptr, *val = nullptr;
  std::vector<double> hval((size_t)npts * t->dof);
  auto body = [&]() -> int {
    HIP_TRY(hipMallocAsync((void **)&pts, sizeof(double) * 3 * (size_t)npts, st));
    HIP_TRY(hipMallocAsync((void **)&val, sizeof(double) * t->dof * (size_t)npts, st));
    HIP_TRY(launch_grid_points(t->view(), node, n, pts, st));
    HIP_TRY(dispatch_eval(t->q, kFma64, t->view(), pts, npts, TBX_BC_FREESPACE, 0, val, st));
    HIP_TRY(hipMemcpyAsync(hval.data(), val, sizeof(double) * hval.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TBX_OK;
  };
  int rc = body();
  hipFreeAsync(pts, st);
  hipFreeAsync(val, st);
  if (rc) return rc;
  std::vector<double> xyz((size_t)npts * 3);
  std::vector<int64_t> conn((size_t)ncells * 8), offs(ncells);
  std::vector<uint8_t> types(ncells, 12);  // VTK_HEXAHEDRON
  parallel_for(t->L, [&](int64_t j) {
    const uint4 inf = t->h_info[j];
    const double len = 1.0 / (double)(1u << inf.w);
    const int64_t p0 = j * n3;
    for (int m = 0; m < n3; ++m) {
      double *o = &xyz[(size_t)(p0 + m) * 3];
      o[0] = (double)inf.x * len + len * unode[m % n];
      o[1] = (double)inf.y * len + len * unode[(m / n) % n];
      o[2] = (double)inf.z * len + len * unode[m / (n * n)];
    }
    const int64_t c0 = j * (int64_t)(n - 1) * (n - 1) * (n - 1);
    int64_t c = c0;
    for (int k = 0; k + 1 < n; ++k)
      for (int b = 0; b + 1 < n; ++b)
        for (int a = 0; a + 1 < n; ++a, ++c) {
          auto id = [&](int da, int db, int dk) {
            return p0 + (a + da) + (int64_t)n * ((b + db) + (int64_t)n * (k + dk));
          };
          int64_t *q = &conn[(size_t)c * 8];
          q[0] = id(0, 0, 0);
          q[1] = id(1, 0, 0);
          q[2] = id(1, 1, 0);
          q[3] = id(0, 1, 0);
          q[4] = id(0, 0, 1);
          q[5] = id(1, 0, 1);
          q[6] = id(1, 1, 1);
          q[7] = id(0, 1, 1);
          offs[c] = 8 * (c + 1);
        }
  });
  return write_vtu_files(prefix, t->dof, npts, ncells, xyz, hval, conn, offs, types);
}

}  // extern "C"
