/*
 * tbx oracle: FP64 CPU restatement of the tbslas hot path.
 * TEST INFRASTRUCTURE ONLY (see tbx_oracle.h for scope and parity status).
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "tbx_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2
#define ORC_EORDER 3

int orc_ncoef(int q) { return (q + 1) * (q + 2) * (q + 3) / 6; }

/* ------------------------------------------------------------------------ */
/* Nodes and the Chebyshev recurrence                                        */
/* ------------------------------------------------------------------------ */

/* utils/cheb.h:50-58: x_i = -cos((i+1/2) pi/d) * scal * 0.5 + 0.5 with
 * scal = 1/cos(pi/(2d)), so the end nodes land on (about) 0 and 1. */
void orc_new_nodes(int q, double *x) {
  unsigned int d = (unsigned int)q + 1;
  double scal = 1.0 / cos(0.5 * M_PI / d);
  for (unsigned int i = 0; i < d; i++)
    x[i] = -cos((i + 0.5) * M_PI / d) * scal * 0.5 + 0.5;
}

/* First-kind nodes on [0,1]: x_i = -cos((i+1/2) pi/d) * 0.5 + 0.5. */
void orc_cheb_nodes_first_kind(int q, double *x) {
  unsigned int d = (unsigned int)q + 1;
  for (unsigned int i = 0; i < d; i++)
    x[i] = -cos((i + 0.5) * M_PI / d) * 0.5 + 0.5;
}

/* T_0..T_q at each x, out[k*n + i].  Three-term recurrence
 * T_{k+1} = (2 x) T_k - T_{k-1}; a point with |x| > 1 (or NaN) gets an
 * all-zero column, which is the rule PVFMM's cheb_poly applies. */
void orc_cheb_poly(int q, const double *x, int64_t n, double *out) {
  for (int64_t i = 0; i < n; i++) {
    int in = fabs(x[i]) <= 1.0;
    double xx = in ? x[i] : 0.0;
    double y0 = in ? 1.0 : 0.0;
    out[i] = y0;
    if (q >= 1) out[n + i] = xx;
    double y1 = xx;
    for (int k = 2; k <= q; k++) {
      double y2 = 2 * xx * y1 - y0;
      out[k * n + i] = y2;
      y0 = y1;
      y1 = y2;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* MortonId (PVFMM) restated                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t x, y, z;
  uint8_t d;
} orc_mid;

/* true iff the most significant set bit of a is below that of b */
static inline int msb_less(uint32_t a, uint32_t b) { return a < b && a < (a ^ b); }

/* MortonId::operator<: equal anchors order by depth (ancestor first);
 * otherwise the highest differing bit decides, z above y above x. */
static inline int mid_less(const orc_mid *a, const orc_mid *b) {
  if (a->x == b->x && a->y == b->y && a->z == b->z) return a->d < b->d;
  uint32_t xx = a->x ^ b->x, yy = a->y ^ b->y, zz = a->z ^ b->z;
  int dim = 2;
  uint32_t best = zz;
  if (msb_less(best, yy)) {
    dim = 1;
    best = yy;
  }
  if (msb_less(best, xx)) dim = 0;
  if (dim == 2) return a->z < b->z;
  if (dim == 1) return a->y < b->y;
  return a->x < b->x;
}

/* Integer cell of one coordinate at ORC_MAXD (tree_functor.h:464-479):
 * c == 1.0 under a non-periodic bc is shifted down by 2^-MAX_DEPTH before
 * floor(c * 2^MAX_DEPTH).  Outside [0,1) the reference is undefined; here
 * the cell is clamped into [0, 2^MAXD - 1] and NaN maps to cell 0. */
static inline uint32_t point_cell(double c, int bc) {
  const double scale = (double)(1u << ORC_MAXD);
  if (c == 1.0 && bc != ORC_BC_PERIODIC) c = c - 1.0 / scale;
  if (!(c >= 0.0)) return 0;
  double f = floor(c * scale);
  if (f > scale - 1.0) return (1u << ORC_MAXD) - 1u;
  return (uint32_t)f;
}

static inline orc_mid point_mid(const double *p, int bc) {
  orc_mid m;
  m.x = point_cell(p[0], bc);
  m.y = point_cell(p[1], bc);
  m.z = point_cell(p[2], bc);
  m.d = ORC_MAXD;
  return m;
}

static inline orc_mid leaf_mid(const orc_tree *t, int64_t j) {
  orc_mid m;
  int sh = ORC_MAXD - t->depth[j];
  m.x = t->anchor[3 * j + 0] << sh;
  m.y = t->anchor[3 * j + 1] << sh;
  m.z = t->anchor[3 * j + 2] << sh;
  m.d = t->depth[j];
  return m;
}

/* ------------------------------------------------------------------------ */
/* Parallel merge sort of (key, index) pairs (omp_par::merge_sort role)      */
/* ------------------------------------------------------------------------ */

typedef struct {
  orc_mid key;
  int64_t idx;
} orc_pair;

static int pair_cmp(const void *a, const void *b) {
  const orc_pair *pa = (const orc_pair *)a, *pb = (const orc_pair *)b;
  if (mid_less(&pa->key, &pb->key)) return -1;
  if (mid_less(&pb->key, &pa->key)) return 1;
  return (pa->idx > pb->idx) - (pa->idx < pb->idx);
}

static void merge_runs(const orc_pair *a, int64_t na, const orc_pair *b,
                       int64_t nb, orc_pair *out) {
  int64_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (pair_cmp(&b[j], &a[i]) < 0)
      out[k++] = b[j++];
    else
      out[k++] = a[i++];
  }
  while (i < na) out[k++] = a[i++];
  while (j < nb) out[k++] = b[j++];
}

static int sort_pairs(orc_pair *p, int64_t n) {
  int nt = omp_get_max_threads();
  if (n < 4096 || nt == 1) {
    qsort(p, (size_t)n, sizeof(orc_pair), pair_cmp);
    return ORC_OK;
  }
  int64_t *bnd = (int64_t *)malloc(sizeof(int64_t) * (nt + 1));
  orc_pair *tmp = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)n);
  if (!bnd || !tmp) {
    free(bnd);
    free(tmp);
    return ORC_ENOMEM;
  }
  for (int i = 0; i <= nt; i++) bnd[i] = (n * i) / nt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nt; i++)
    qsort(p + bnd[i], (size_t)(bnd[i + 1] - bnd[i]), sizeof(orc_pair), pair_cmp);
  orc_pair *src = p, *dst = tmp;
  for (int w = 1; w < nt; w *= 2) {
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nt; i += 2 * w) {
      int m = i + w < nt ? i + w : nt;
      int e = i + 2 * w < nt ? i + 2 * w : nt;
      merge_runs(src + bnd[i], bnd[m] - bnd[i], src + bnd[m], bnd[e] - bnd[m],
                 dst + bnd[i]);
    }
    orc_pair *s = src;
    src = dst;
    dst = s;
  }
  if (src != p) memcpy(p, src, sizeof(orc_pair) * (size_t)n);
  free(tmp);
  free(bnd);
  return ORC_OK;
}

static int check_tree(const orc_tree *t) {
  if (!t || t->num_leaves <= 0 || t->q < 0 || t->q > 30 || t->dof <= 0)
    return ORC_EINVAL;
  for (int64_t j = 0; j < t->num_leaves; j++)
    if (t->depth[j] > ORC_MAXD) return ORC_EINVAL;
  for (int64_t j = 1; j < t->num_leaves; j++) {
    orc_mid a = leaf_mid(t, j - 1), b = leaf_mid(t, j);
    if (!mid_less(&a, &b)) return ORC_EORDER;
  }
  return ORC_OK;
}

/* Periodic wrap, in place (tree_functor.h:442-449). */
static void wrap_periodic(double *pts, int64_t n3) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n3; i++) {
    double c = pts[i];
    if (c < 0.0) c = c + 1.0;
    if (c >= 1.0) c = c - 1.0;
    pts[i] = c;
  }
}

/* Keys, sort and leaf partition (tree_functor.h:464-489, :190-198).
 * On return pairs[] is sorted and part[j] is the first sorted position whose
 * key is not below leaf j's MortonId; part[L] = n. */
static int sort_and_partition(const orc_tree *t, const double *pts, int64_t n,
                              int bc, orc_pair **pairs_out, int64_t **part_out) {
  int64_t L = t->num_leaves;
  orc_pair *pairs = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)(n ? n : 1));
  int64_t *part = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L + 1));
  if (!pairs || !part) {
    free(pairs);
    free(part);
    return ORC_ENOMEM;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    pairs[i].key = point_mid(pts + 3 * i, bc);
    pairs[i].idx = i;
  }
  int rc = sort_pairs(pairs, n);
  if (rc) {
    free(pairs);
    free(part);
    return rc;
  }
  part[L] = n;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < L; j++) {
    orc_mid m = leaf_mid(t, j);
    int64_t lo = 0, hi = n; /* lower_bound */
    while (lo < hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (mid_less(&pairs[mid].key, &m))
        lo = mid + 1;
      else
        hi = mid;
    }
    part[j] = lo;
  }
  *pairs_out = pairs;
  *part_out = part;
  return ORC_OK;
}

int orc_locate(const orc_tree *t, const double *pts, int64_t n, int bc,
               int32_t *leaf) {
  int rc = check_tree(t);
  if (rc) return rc;
  if (n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  double *cp = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  if (!cp) return ORC_ENOMEM;
  memcpy(cp, pts, sizeof(double) * 3 * (size_t)n);
  if (bc == ORC_BC_PERIODIC) wrap_periodic(cp, 3 * n);
  orc_pair *pairs;
  int64_t *part;
  rc = sort_and_partition(t, cp, n, bc, &pairs, &part);
  free(cp);
  if (rc) return rc;
  /* Points sorted below the first leaf own no partition; they keep -1. */
  for (int64_t i = 0; i < part[0]; i++) leaf[pairs[i].idx] = -1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t j = 0; j < t->num_leaves; j++)
    for (int64_t s = part[j]; s < part[j + 1]; s++) leaf[pairs[s].idx] = (int32_t)j;
  free(pairs);
  free(part);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* EvalNodesLocal + vec_eval                                                 */
/* ------------------------------------------------------------------------ */

/* vec_eval (tree_functor.h:27-84) on one block of n points: for each dof,
 * u = sum_i pz_i (sum_j py_j (sum_k px_k c_ijk)), k innermost, separate
 * multiply and add, coefficients in packed-triangle order. */
static void vec_eval_block(int n, int d, int dof, const double *px,
                           const double *py, const double *pz,
                           const double *coeff, int nc, double *tmp_out,
                           double *w, double *v) {
  for (int l0 = 0; l0 < dof; l0++) {
    const double *c = coeff + (size_t)l0 * nc;
    double *u = tmp_out + (size_t)l0 * n;
    int idx = 0;
    for (int p = 0; p < n; p++) u[p] = 0.0;
    for (int i = 0; i < d; i++) {
      for (int p = 0; p < n; p++) v[p] = 0.0;
      for (int j = 0; i + j < d; j++) {
        for (int p = 0; p < n; p++) w[p] = 0.0;
        for (int k = 0; i + j + k < d; k++) {
          const double ck = c[idx++];
          const double *pk = px + (size_t)k * n;
          for (int p = 0; p < n; p++) w[p] = w[p] + pk[p] * ck;
        }
        const double *pj = py + (size_t)j * n;
        for (int p = 0; p < n; p++) v[p] = v[p] + pj[p] * w[p];
      }
      const double *pi = pz + (size_t)i * n;
      for (int p = 0; p < n; p++) u[p] = u[p] + pi[p] * v[p];
    }
  }
}

/* EvalNodesLocal (tree_functor.h:155-395) on points already sorted by key:
 * sorted_pts[s] is the s-th point in key order, part[] the leaf partition.
 * Thread pid owns sorted points [pid*N/p, (pid+1)*N/p); each leaf's points
 * go through blocks of 288. */
static int eval_nodes_local(const orc_tree *t, const double *sorted_pts,
                            int64_t n, const int64_t *part, double *sorted_val) {
  const int q = t->q, d = q + 1, dof = t->dof, nc = orc_ncoef(q);
  const int64_t L = t->num_leaves;
  const int blk = 288;
  int nt = omp_get_max_threads();
  int err = 0;
#pragma omp parallel num_threads(nt)
  {
    int pid = omp_get_thread_num(), np = omp_get_num_threads();
    int nmax = blk + 16;
    double *cx = (double *)malloc(sizeof(double) * nmax * 3);
    double *px = (double *)malloc(sizeof(double) * (size_t)nmax * d * 3);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)nmax * dof);
    double *w = (double *)malloc(sizeof(double) * nmax * 2);
    if (!cx || !px || !tmp || !w) {
#pragma omp atomic write
      err = 1;
    } else {
      double *cy = cx + nmax, *cz = cy + nmax;
      double *py = px + (size_t)nmax * d, *pz = py + (size_t)nmax * d;
      double *v = w + nmax;
      int64_t s0 = (pid * n) / np, s1 = ((pid + 1) * n) / np;
      /* first leaf whose range reaches s0 */
      int64_t lo = 0, hi = L;
      while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (part[mid + 1] <= s0)
          lo = mid + 1;
        else
          hi = mid;
      }
      for (int64_t j = lo; j < L && part[j] < s1; j++) {
        int64_t a = part[j] > s0 ? part[j] : s0;
        int64_t b = part[j + 1] < s1 ? part[j + 1] : s1;
        if (a >= b) continue;
        const double *coeff = t->coeff + (size_t)j * dof * nc;
        const double s = (double)(1ull << t->depth[j]);
        const double len = 1.0 / s;
        const double c0 = t->anchor[3 * j + 0] * len;
        const double c1 = t->anchor[3 * j + 1] * len;
        const double c2 = t->anchor[3 * j + 2] * len;
        for (int64_t pb = a; pb < b; pb += blk) {
          int np_ = (int)((b - pb) < blk ? (b - pb) : blk);
          int nn = np_;
          while (nn % 16) nn++;
          const double *cp = sorted_pts + 3 * pb;
          /* tree_functor.h:286-301: xi = ((x - c) * 2.0) * s - 1.0, pad 0 */
          for (int i = 0; i < np_; i++) {
            cx[i] = (cp[3 * i + 0] - c0) * 2.0 * s - 1.0;
            cy[i] = (cp[3 * i + 1] - c1) * 2.0 * s - 1.0;
            cz[i] = (cp[3 * i + 2] - c2) * 2.0 * s - 1.0;
          }
          for (int i = np_; i < nn; i++) cx[i] = cy[i] = cz[i] = 0.0;
          orc_cheb_poly(q, cx, nn, px);
          orc_cheb_poly(q, cy, nn, py);
          orc_cheb_poly(q, cz, nn, pz);
          vec_eval_block(nn, d, dof, px, py, pz, coeff, nc, tmp, w, v);
          double *o = sorted_val + (size_t)pb * dof;
          for (int i = 0; i < np_; i++)
            for (int l = 0; l < dof; l++) o[(size_t)i * dof + l] = tmp[(size_t)l * nn + i];
        }
      }
    }
    free(cx);
    free(px);
    free(tmp);
    free(w);
  }
  return err ? ORC_ENOMEM : ORC_OK;
}

int orc_eval_tree(const orc_tree *t, double *pts, int64_t n, int bc,
                  double *out) {
  int rc = check_tree(t);
  if (rc) return rc;
  if (n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  const int dof = t->dof;
  if (bc == ORC_BC_PERIODIC) wrap_periodic(pts, 3 * n);
  orc_pair *pairs;
  int64_t *part;
  rc = sort_and_partition(t, pts, n, bc, &pairs, &part);
  if (rc) return rc;
  double *sp = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *sv = (double *)malloc(sizeof(double) * dof * (size_t)n);
  if (!sp || !sv) {
    free(sp);
    free(sv);
    free(pairs);
    free(part);
    return ORC_ENOMEM;
  }
  /* gather in key order (tree_functor.h:659-667) */
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < n; s++) {
    const double *src = pts + 3 * pairs[s].idx;
    sp[3 * s + 0] = src[0];
    sp[3 * s + 1] = src[1];
    sp[3 * s + 2] = src[2];
  }
  /* points below the first leaf belong to no leaf: value 0 */
  for (int64_t s = 0; s < part[0]; s++)
    for (int l = 0; l < dof; l++) sv[(size_t)s * dof + l] = 0.0;
  rc = eval_nodes_local(t, sp, n, part, sv);
  /* scatter back to caller order (tree_functor.h:683-690) */
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < n; s++)
    for (int l = 0; l < dof; l++)
      out[(size_t)pairs[s].idx * dof + l] = sv[(size_t)s * dof + l];
  free(sp);
  free(sv);
  free(pairs);
  free(part);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* RK2 trajectories and the semi-Lagrangian solve                           */
/* ------------------------------------------------------------------------ */

/* NodeFieldFunctor(vel)(pts, n, t, out) with optional scalar modulation. */
static int vel_eval(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                    double *pts, int64_t n, double t, int bc, double *out) {
  int rc = orc_eval_tree(vel, pts, n, bc, out);
  if (rc || !g) return rc;
  double f = g(t, gctx);
  int64_t m = n * vel->dof;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) out[i] = out[i] * f;
  return ORC_OK;
}

/* IntegrateRK2 (traj.inc:20-46), called in place as ComputeTrajRK2 does. */
static int integrate_rk2(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                         double *x, int64_t n, double tinit, double dt, int bc,
                         double *xtmp, double *vtmp) {
  int64_t m = 3 * n;
  int rc = vel_eval(vel, g, gctx, x, n, tinit, bc, vtmp);
  if (rc) return rc;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) xtmp[i] = x[i] + 0.5 * dt * vtmp[i];
  rc = vel_eval(vel, g, gctx, xtmp, n, tinit + 0.5 * dt, bc, vtmp);
  if (rc) return rc;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) x[i] = x[i] + dt * vtmp[i];
  return ORC_OK;
}

int orc_traj_rk2(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                 const double *xinit, int64_t n, double tinit, double tfinal,
                 int num_rk_step, int bc, double *xsol) {
  if (!vel || vel->dof != 3 || n < 0 || num_rk_step <= 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  int64_t m = 3 * n;
  double tau = (tfinal - tinit) / num_rk_step; /* traj.inc:55 */
  if (xsol != xinit) memcpy(xsol, xinit, sizeof(double) * (size_t)m);
  double *xtmp = (double *)malloc(sizeof(double) * (size_t)m);
  double *vtmp = (double *)malloc(sizeof(double) * (size_t)m);
  if (!xtmp || !vtmp) {
    free(xtmp);
    free(vtmp);
    return ORC_ENOMEM;
  }
  double tcur = tinit;
  int rc = ORC_OK;
  for (int s = 0; s < num_rk_step && !rc; s++) {
    rc = integrate_rk2(vel, g, gctx, xsol, n, tcur, tau, bc, xtmp, vtmp);
    tcur = tcur + tau;
  }
  free(xtmp);
  free(vtmp);
  return rc;
}

int orc_semilag_rk2(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                    const orc_tree *conc, const double *pts, int64_t n,
                    int timestep, double dt, int num_rk_step, int bc,
                    double *vals) {
  if (n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  double tinit = timestep * dt; /* semilag.inc:34-35 */
  double tfinal = tinit - dt;
  double *dep = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  if (!dep) return ORC_ENOMEM;
  int rc = orc_traj_rk2(vel, g, gctx, pts, n, tinit, tfinal, num_rk_step, bc, dep);
  if (!rc) rc = orc_eval_tree(conc, dep, n, bc, vals);
  free(dep);
  return rc;
}

int orc_collect_grid_points(const orc_tree *t, double *out) {
  int rc = check_tree(t);
  if (rc) return rc;
  const int d = t->q + 1;
  const int P = d * d * d;
  double *x = (double *)malloc(sizeof(double) * d);
  if (!x) return ORC_ENOMEM;
  orc_new_nodes(t->q, x);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < t->num_leaves; j++) {
    double len = pow(0.5, t->depth[j]);
    double c0 = t->anchor[3 * j + 0] * len;
    double c1 = t->anchor[3 * j + 1] * len;
    double c2 = t->anchor[3 * j + 2] * len;
    double *o = out + (size_t)j * P * 3;
    for (int i = 0; i < P; i++) { /* cheb.h:60-67 x fastest; tree_utils.h:484-491 */
      o[3 * i + 0] = c0 + len * x[i % d];
      o[3 * i + 1] = c1 + len * x[(i / d) % d];
      o[3 * i + 2] = c2 + len * x[i / (d * d)];
    }
  }
  free(x);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Refit operator: M = pinv(B)^T (cheb.h:166-196)                            */
/* ------------------------------------------------------------------------ */

/* B[p][m]: basis m = packed (i,j,k) evaluated at the stretched node p mapped
 * to [-1,1] (cheb.h:178); value T_k(x) T_j(y) T_i(z).  pinv through the
 * normal equations in long double with a Cholesky factorisation (B has full
 * column rank on the tensor node set). */
int orc_pt2coeff(int q, double *M) {
  if (q < 0 || q > 24) return ORC_EINVAL;
  const int d = q + 1, P = d * d * d, nc = orc_ncoef(q);
  double *x = (double *)malloc(sizeof(double) * d);
  double *tx = (double *)malloc(sizeof(double) * d * d);
  long double *B = (long double *)malloc(sizeof(long double) * (size_t)P * nc);
  long double *G = (long double *)malloc(sizeof(long double) * (size_t)nc * nc);
  if (!x || !tx || !B || !G) {
    free(x);
    free(tx);
    free(B);
    free(G);
    return ORC_ENOMEM;
  }
  orc_new_nodes(q, x);
  double *xi = (double *)malloc(sizeof(double) * d);
  for (int i = 0; i < d; i++) xi[i] = x[i] * 2.0 - 1.0;
  orc_cheb_poly(q, xi, d, tx); /* tx[k*d + node] */
  free(xi);
  for (int p = 0; p < P; p++) {
    int ix = p % d, iy = (p / d) % d, iz = p / (d * d);
    int m = 0;
    for (int i = 0; i < d; i++)
      for (int j = 0; i + j < d; j++)
        for (int k = 0; i + j + k < d; k++)
          B[(size_t)p * nc + m++] =
              (long double)tx[k * d + ix] * tx[j * d + iy] * tx[i * d + iz];
  }
  /* G = B^T B */
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < nc; a++)
    for (int b = 0; b <= a; b++) {
      long double s = 0;
      for (int p = 0; p < P; p++) s += B[(size_t)p * nc + a] * B[(size_t)p * nc + b];
      G[(size_t)a * nc + b] = s;
      G[(size_t)b * nc + a] = s;
    }
  /* Cholesky G = L L^T (lower triangle in place) */
  int bad = 0;
  for (int j = 0; j < nc && !bad; j++) {
    long double s = G[(size_t)j * nc + j];
    for (int k = 0; k < j; k++) s -= G[(size_t)j * nc + k] * G[(size_t)j * nc + k];
    if (s <= 0) {
      bad = 1;
      break;
    }
    long double ljj = sqrtl(s);
    G[(size_t)j * nc + j] = ljj;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < nc; i++) {
      long double t = G[(size_t)i * nc + j];
      for (int k = 0; k < j; k++) t -= G[(size_t)i * nc + k] * G[(size_t)j * nc + k];
      G[(size_t)i * nc + j] = t / ljj;
    }
  }
  if (!bad) {
    /* row p of M solves G m = B[p]^T  (M = B G^-1, G symmetric) */
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; p++) {
      long double *r = (long double *)malloc(sizeof(long double) * nc);
      for (int a = 0; a < nc; a++) r[a] = B[(size_t)p * nc + a];
      for (int a = 0; a < nc; a++) { /* forward L y = r */
        long double s = r[a];
        for (int k = 0; k < a; k++) s -= G[(size_t)a * nc + k] * r[k];
        r[a] = s / G[(size_t)a * nc + a];
      }
      for (int a = nc - 1; a >= 0; a--) { /* backward L^T m = y */
        long double s = r[a];
        for (int k = a + 1; k < nc; k++) s -= G[(size_t)k * nc + a] * r[k];
        r[a] = s / G[(size_t)a * nc + a];
      }
      for (int a = 0; a < nc; a++) M[(size_t)p * nc + a] = (double)r[a];
      free(r);
    }
  }
  free(x);
  free(tx);
  free(B);
  free(G);
  return bad ? ORC_EINVAL : ORC_OK;
}

/* SetTreeGridValues (tree_utils.h:530-551): per leaf and dof,
 * coeff[m] = sum_p vals[p] * M[p][m], ascending p. */
int orc_set_grid_values(orc_tree *t, const double *M, const double *vals) {
  int rc = check_tree(t);
  if (rc) return rc;
  const int d = t->q + 1, P = d * d * d, nc = orc_ncoef(t->q);
  const int64_t rows = t->num_leaves * t->dof;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rows; r++) {
    const double *v = vals + (size_t)r * P;
    double *c = t->coeff + (size_t)r * nc;
    for (int m = 0; m < nc; m++) c[m] = 0.0;
    for (int p = 0; p < P; p++) {
      const double vp = v[p];
      const double *Mp = M + (size_t)p * nc;
      for (int m = 0; m < nc; m++) c[m] = c[m] + vp * Mp[m];
    }
  }
  return ORC_OK;
}

int orc_semilag_insitu(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                       orc_tree *conc, const double *M, int timestep, double dt,
                       int num_rk_step, int bc) {
  int rc = check_tree(conc);
  if (rc) return rc;
  const int d = conc->q + 1, P = d * d * d, dof = conc->dof;
  const int64_t n = conc->num_leaves * P;
  double *pts = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *vals = (double *)malloc(sizeof(double) * dof * (size_t)n);
  double *vt = (double *)malloc(sizeof(double) * dof * (size_t)n);
  if (!pts || !vals || !vt) {
    free(pts);
    free(vals);
    free(vt);
    return ORC_ENOMEM;
  }
  rc = orc_collect_grid_points(conc, pts);
  if (!rc)
    rc = orc_semilag_rk2(vel, g, gctx, conc, pts, n, timestep, dt, num_rk_step,
                         bc, vals);
  if (!rc) {
    /* point-major [pt][dof] -> per-leaf dof-major [leaf][dof][P]
     * (the transpose tree_ns.h:223-234 applies before SetTreeGridValues) */
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < conc->num_leaves; j++)
      for (int l = 0; l < dof; l++)
        for (int p = 0; p < P; p++)
          vt[((size_t)j * dof + l) * P + p] = vals[((size_t)j * P + p) * dof + l];
    rc = orc_set_grid_values(conc, M, vt);
  }
  free(pts);
  free(vals);
  free(vt);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Initial fit at first-kind nodes (cheb_approx role)                       */
/* ------------------------------------------------------------------------ */

/* Per axis c_k = (2/d) sum_i f_i T_k(xi_i), c_0 halved (discrete Chebyshev
 * transform at first-kind nodes, exact for degree < d), applied along x,
 * then y, then z, then truncated to total degree <= q in packed order. */
int orc_cheb_approx(int q, int dof, const double *vals, double *coeff) {
  if (q < 0 || dof <= 0) return ORC_EINVAL;
  const int d = q + 1, P = d * d * d, nc = orc_ncoef(q);
  double *x = (double *)malloc(sizeof(double) * d);
  double *T = (double *)malloc(sizeof(double) * d * d);
  double *a = (double *)malloc(sizeof(double) * P);
  double *b = (double *)malloc(sizeof(double) * P);
  if (!x || !T || !a || !b) {
    free(x);
    free(T);
    free(a);
    free(b);
    return ORC_ENOMEM;
  }
  orc_cheb_nodes_first_kind(q, x);
  for (int i = 0; i < d; i++) x[i] = x[i] * 2.0 - 1.0;
  orc_cheb_poly(q, x, d, T); /* T[k*d + i] */
  for (int l = 0; l < dof; l++) {
    memcpy(a, vals + (size_t)l * P, sizeof(double) * P);
    /* three passes, each contracts the fastest axis and moves its degree
     * index to the slowest position */
    for (int pass = 0; pass < 3; pass++) {
      for (int r = 0; r < d * d; r++)
        for (int k = 0; k < d; k++) {
          double s = 0;
          for (int i = 0; i < d; i++) s += a[r * d + i] * T[k * d + i];
          s = s * (2.0 / d);
          if (k == 0) s = s * 0.5;
          b[k * d * d + r] = s;
        }
      memcpy(a, b, sizeof(double) * P);
    }
    /* pass 0 maps [z][y][x] -> [kx][z][y], pass 1 -> [ky][kx][z],
     * pass 2 -> [kz][ky][kx] */
    double *c = coeff + (size_t)l * nc;
    int m = 0;
    for (int i = 0; i < d; i++)
      for (int j = 0; i + j < d; j++)
        for (int k = 0; i + j + k < d; k++) c[m++] = a[(i * d + j) * d + k];
  }
  free(x);
  free(T);
  free(a);
  free(b);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Tricubic regular-grid interpolation (fast_interp)                         */
/* ------------------------------------------------------------------------ */

/* fast_interp for one query point p in [0,1]^3 of one regular N^3 grid
 * (tree_functor.h:89-153); out[dof]. */
static void fast_interp_point(const double *grid, int dof, int N, const double *p,
                              double *out) {
  double lagr_denom[4];
  for (int i = 0; i < 4; i++) { /* tree_functor.h:93-99 */
    lagr_denom[i] = 1;
    for (int j = 0; j < 4; j++)
      if (i != j) lagr_denom[i] /= (double)(i - j);
  }
  const int64_t N3 = (int64_t)N * N * N;
  if (p[0] < 0 || p[0] > 1.0 || p[1] < 0 || p[1] > 1.0 || p[2] < 0 ||
      p[2] > 1.0) { /* :106-114 */
    for (int k = 0; k < dof; k++) out[k] = 0;
    return;
  }
  double point[3];
  int gi[3];
  for (int j = 0; j < 3; j++) { /* :118-124 */
    point[j] = p[j] * (N - 1);
    gi[j] = ((int)point[j]) - 1;
    if (gi[j] < 0) gi[j] = 0;
    if (gi[j] > N - 4) gi[j] = N - 4;
    point[j] -= gi[j];
  }
  double Mw[3][4];
  for (int j = 0; j < 3; j++) { /* :127-135 */
    double x = point[j];
    for (int k = 0; k < 4; k++) {
      Mw[j][k] = lagr_denom[k];
      for (int l = 0; l < 4; l++)
        if (k != l) Mw[j][k] *= (x - l);
    }
  }
  for (int k = 0; k < dof; k++) { /* :137-151 */
    double val = 0;
    for (int j2 = 0; j2 < 4; j2++)
      for (int j1 = 0; j1 < 4; j1++) {
        double M1M2 = Mw[1][j1] * Mw[2][j2];
        int64_t ij = (int64_t)N * ((gi[1] + j1) + (int64_t)N * (gi[2] + j2));
        for (int j0 = 0; j0 < 4; j0++) {
          int64_t idx = (gi[0] + j0) + ij;
          val += Mw[0][j0] * M1M2 * grid[idx + k * N3];
        }
      }
    out[k] = val;
  }
}

int orc_fast_interp(const double *grid, int dof, int N, const double *pts,
                    int64_t n, double *out) {
  if (N < 4 || dof <= 0 || n < 0) return ORC_EINVAL;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) fast_interp_point(grid, dof, N, pts + 3 * i, out + i * dof);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Regular-grid leaves (tricubic)                                            */
/* ------------------------------------------------------------------------ */

static orc_tree grid_as_tree(const orc_grid *g) {
  orc_tree t = {g->num_leaves, g->anchor, g->depth, 1, g->dof, NULL};
  return t;
}

int orc_grid_eval(const orc_grid *g, double *pts, int64_t n, int bc, double *out) {
  if (!g || g->n < 4 || g->dof <= 0 || n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  orc_tree t = grid_as_tree(g);
  if (bc == ORC_BC_PERIODIC) wrap_periodic(pts, 3 * n);
  int32_t *leaf = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!leaf) return ORC_ENOMEM;
  int rc = orc_locate(&t, pts, n, bc, leaf);
  if (rc) {
    free(leaf);
    return rc;
  }
  const int64_t N3 = (int64_t)g->n * g->n * g->n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const int64_t j = leaf[i];
    const double s = ldexp(1.0, g->depth[j]), len = 1.0 / s;
    double u[3];
    for (int a = 0; a < 3; a++) u[a] = (pts[3 * i + a] - g->anchor[3 * j + a] * len) * s;
    fast_interp_point(g->samples + j * g->dof * N3, g->dof, g->n, u, out + i * g->dof);
  }
  free(leaf);
  return ORC_OK;
}

int orc_grid_points(const orc_grid *g, double *out) {
  if (!g || g->n < 2) return ORC_EINVAL;
  const int N = g->n;
  const int64_t N3 = (int64_t)N * N * N;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < g->num_leaves; j++) {
    const double len = 1.0 / ldexp(1.0, g->depth[j]);
    const double cx = g->anchor[3 * j] * len, cy = g->anchor[3 * j + 1] * len,
                 cz = g->anchor[3 * j + 2] * len;
    for (int64_t m = 0; m < N3; m++) {
      const int a = (int)(m % N), b = (int)((m / N) % N), c = (int)(m / ((int64_t)N * N));
      double *o = out + 3 * (j * N3 + m);
      o[0] = cx + len * ((double)a / (N - 1));
      o[1] = cy + len * ((double)b / (N - 1));
      o[2] = cz + len * ((double)c / (N - 1));
    }
  }
  return ORC_OK;
}

static int grid_vel(const orc_grid *v, orc_tfac_fn g, void *gctx, double *pts, int64_t n,
                    double t, int bc, double *out) {
  int rc = orc_grid_eval(v, pts, n, bc, out);
  if (rc || !g) return rc;
  const double f = g(t, gctx);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 3 * n; i++) out[i] = out[i] * f;
  return ORC_OK;
}

int orc_grid_traj_rk2(const orc_grid *vel, orc_tfac_fn g, void *gctx, const double *xinit,
                      int64_t n, double tinit, double tfinal, int num_rk_step, int bc,
                      double *xsol) {
  if (!vel || vel->dof != 3 || num_rk_step <= 0 || n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  const int64_t m = 3 * n;
  const double tau = (tfinal - tinit) / num_rk_step;
  if (xsol != xinit) memcpy(xsol, xinit, sizeof(double) * (size_t)m);
  double *xtmp = (double *)malloc(sizeof(double) * (size_t)m);
  double *vtmp = (double *)malloc(sizeof(double) * (size_t)m);
  if (!xtmp || !vtmp) {
    free(xtmp);
    free(vtmp);
    return ORC_ENOMEM;
  }
  double tcur = tinit;
  int rc = ORC_OK;
  for (int s = 0; s < num_rk_step && !rc; s++) { /* traj.inc:20-68 */
    rc = grid_vel(vel, g, gctx, xsol, n, tcur, bc, vtmp);
    if (rc) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) xtmp[i] = xsol[i] + 0.5 * tau * vtmp[i];
    rc = grid_vel(vel, g, gctx, xtmp, n, tcur + 0.5 * tau, bc, vtmp);
    if (rc) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) xsol[i] = xsol[i] + tau * vtmp[i];
    tcur = tcur + tau;
  }
  free(xtmp);
  free(vtmp);
  return rc;
}

int orc_grid_semilag_insitu(const orc_grid *vel, orc_tfac_fn g, void *gctx, orc_grid *conc,
                            int timestep, double dt, int num_rk_step, int bc) {
  if (!vel || !conc || conc->n < 4) return ORC_EINVAL;
  const int dof = conc->dof;
  const int64_t N3 = (int64_t)conc->n * conc->n * conc->n, n = conc->num_leaves * N3;
  double *pts = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *vals = (double *)malloc(sizeof(double) * dof * (size_t)n);
  if (!pts || !vals) {
    free(pts);
    free(vals);
    return ORC_ENOMEM;
  }
  int rc = orc_grid_points(conc, pts);
  const double tinit = timestep * dt, tfinal = tinit - dt; /* semilag.inc:34-35 */
  if (!rc) rc = orc_grid_traj_rk2(vel, g, gctx, pts, n, tinit, tfinal, num_rk_step, bc, pts);
  if (!rc) rc = orc_grid_eval(conc, pts, n, bc, vals);
  if (!rc) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < conc->num_leaves; j++)
      for (int l = 0; l < dof; l++)
        for (int64_t p = 0; p < N3; p++)
          conc->samples[(j * dof + l) * N3 + p] = vals[(j * N3 + p) * dof + l];
  }
  free(pts);
  free(vals);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Time functors                                                             */
/* ------------------------------------------------------------------------ */

static double h00(double t) { return 2 * t * t * t - 3 * t * t + 1; }
static double h10(double t) { return t * t * t - 2 * t * t + t; }
static double h01(double t) { return -2 * t * t * t + 3 * t * t; }
static double h11(double t) { return t * t * t - t * t; }
static double tangent(double tk_1, double tk, double tk1, double pk_1, double pk,
                      double pk1) {
  return (pk1 - pk) * 0.5 / (tk1 - tk) + (pk - pk_1) * 0.5 / (tk - tk_1);
}

double orc_cubic1d(double x, const double *xx, const double *pp) {
  double mk = tangent(xx[0], xx[1], xx[2], pp[0], pp[1], pp[2]);
  double mk1 = tangent(xx[1], xx[2], xx[3], pp[1], pp[2], pp[3]);
  double t = (x - xx[1]) / (xx[2] - xx[1]);
  return h00(t) * pp[1] + h10(t) * (xx[2] - xx[1]) * mk + h01(t) * pp[2] +
         h11(t) * (xx[2] - xx[1]) * mk1;
}

int orc_field_set_eval(const orc_tree *const *trees, const double *tt,
                       double *pts, int64_t n, double t, int bc, double *out) {
  if (n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  const int dof = trees[0]->dof;
  const int64_t m = n * dof;
  double *v = (double *)malloc(sizeof(double) * 4 * (size_t)m);
  if (!v) return ORC_ENOMEM;
  int rc = ORC_OK;
  for (int s = 0; s < 4 && !rc; s++)
    rc = orc_eval_tree(trees[s], pts, n, bc, v + (size_t)s * m);
  if (!rc) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) {
      double pp[4] = {v[i], v[m + i], v[2 * m + i], v[3 * m + i]};
      out[i] = orc_cubic1d(t, tt, pp);
    }
  }
  free(v);
  return rc;
}

int orc_field_extrap_eval(const orc_tree *tp, const orc_tree *tc, double *pts,
                          int64_t n, int bc, double *out) {
  if (n < 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  const int64_t m = n * tc->dof;
  double *vc = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  if (!vc) return ORC_ENOMEM;
  double *vp = vc + m;
  int rc = orc_eval_tree(tc, pts, n, bc, vc);
  if (!rc) rc = orc_eval_tree(tp, pts, n, bc, vp);
  if (!rc) {
    const double ccoeff = 3.0 / 2, pcoeff = 0.5;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) out[i] = ccoeff * vc[i] - pcoeff * vp[i];
  }
  free(vc);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Velocity functors and the extrapolating RK2                               */
/* ------------------------------------------------------------------------ */

/* stage 1 / 2 of one IntegrateRK2 call: the 4-argument functor call with the
 * stage time (traj.inc:33, 40), or field_fn / extrap_fn (traj.inc:80, 86) */
static int vel_stage(const orc_velocity *v, int stage, double *pts, int64_t n,
                     double t, int bc, double *out) {
  switch (v->kind) {
    case ORC_VEL_TREE:
      return vel_eval(v->t[0], v->g, v->gctx, pts, n, t, bc, out);
    case ORC_VEL_EXTRAP:
      if (stage == 1) return orc_eval_tree(v->t[0], pts, n, bc, out);
      return orc_field_extrap_eval(v->t[1], v->t[0], pts, n, bc, out);
    case ORC_VEL_SET:
      return orc_field_set_eval(v->t, v->tt, pts, n, t, bc, out);
  }
  return ORC_EINVAL;
}

int orc_traj_rk2_v(const orc_velocity *vel, const double *xinit, int64_t n,
                   double tinit, double tfinal, int num_rk_step, int bc,
                   double *xsol) {
  if (!vel || n < 0 || num_rk_step <= 0) return ORC_EINVAL;
  if (n == 0) return ORC_OK;
  const int64_t m = 3 * n;
  const double tau = (tfinal - tinit) / num_rk_step;
  if (xsol != xinit) memcpy(xsol, xinit, sizeof(double) * (size_t)m);
  double *xtmp = (double *)malloc(sizeof(double) * (size_t)m);
  double *vtmp = (double *)malloc(sizeof(double) * (size_t)m);
  if (!xtmp || !vtmp) {
    free(xtmp);
    free(vtmp);
    return ORC_ENOMEM;
  }
  double tcur = tinit;
  int rc = ORC_OK;
  for (int s = 0; s < num_rk_step && !rc; s++) {
    rc = vel_stage(vel, 1, xsol, n, tcur, bc, vtmp);
    if (rc) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) xtmp[i] = xsol[i] + 0.5 * tau * vtmp[i];
    rc = vel_stage(vel, 2, xtmp, n, tcur + 0.5 * tau, bc, vtmp);
    if (rc) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) xsol[i] = xsol[i] + tau * vtmp[i];
    tcur = tcur + tau;
  }
  free(xtmp);
  free(vtmp);
  return rc;
}

int orc_semilag_insitu_v(const orc_velocity *vel, orc_tree *conc,
                         const double *M, int timestep, double dt,
                         int num_rk_step, int bc) {
  int rc = check_tree(conc);
  if (rc) return rc;
  const int d = conc->q + 1, P = d * d * d, dof = conc->dof;
  const int64_t n = conc->num_leaves * P;
  double *pts = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *dep = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *vals = (double *)malloc(sizeof(double) * dof * (size_t)n);
  double *vt = (double *)malloc(sizeof(double) * dof * (size_t)n);
  if (!pts || !dep || !vals || !vt) {
    free(pts);
    free(dep);
    free(vals);
    free(vt);
    return ORC_ENOMEM;
  }
  rc = orc_collect_grid_points(conc, pts);
  const double tinit = timestep * dt, tfinal = tinit - dt;
  if (!rc) rc = orc_traj_rk2_v(vel, pts, n, tinit, tfinal, num_rk_step, bc, dep);
  if (!rc) rc = orc_eval_tree(conc, dep, n, bc, vals);
  if (!rc) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < conc->num_leaves; j++)
      for (int l = 0; l < dof; l++)
        for (int p = 0; p < P; p++)
          vt[((size_t)j * dof + l) * P + p] = vals[((size_t)j * P + p) * dof + l];
    rc = orc_set_grid_values(conc, M, vt);
  }
  free(pts);
  free(dep);
  free(vals);
  free(vt);
  return rc;
}
