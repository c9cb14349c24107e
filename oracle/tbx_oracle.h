/*
 * tbx oracle: FP64 CPU restatement of the tbslas semi-Lagrangian hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library holds the reference results the
 * HIP engine (libtbx.so) is compared against, and the CPU baseline that the
 * benchmark times beside it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The engine never links or calls it.
 *
 * Every function restates the reference algorithm from the source text of
 * arashb/tbslas; the citation next to each prototype names the lines.  The
 * reference's arithmetic order is kept (separate multiply and add, same
 * nesting) and the library is compiled with -ffp-contract=off, so its
 * results are the reference algorithm's results, not merely close to them.
 *
 * Parity status (what pins each function; see tests/test_oracle_*.py):
 *   orc_new_nodes        closed form, symmetry x_i + x_{d-1-i} = 1, endpoints
 *   orc_cheb_poly        T_k(cos t) = cos(k t), T_k(+-1) = (+-1)^k, |T_k| <= 1
 *   orc_locate           brute-force half-open box scan, uniform closed form,
 *                        edge-case tables (faces, corners, 1.0, -0.0, wrap)
 *   orc_eval_tree        long-double brute force, polynomial reproduction
 *   orc_pt2coeff         numpy.linalg.pinv, M^T B = I on the triangle space
 *   orc_traj_rk2         closed form for linear (rotation) velocity fields
 *   orc_cheb_approx      polynomial reproduction, numpy DCT
 *   orc_fast_interp      exact for per-axis cubics, 0 outside [0,1]^3
 *   orc_cubic1d          exact for quadratics on uniform times
 *   orc_grid_eval        per-axis cubics reproduced, leaf-local coordinates
 *                        against locate + orc_fast_interp on the leaf grid
 * The PVFMM routines the reference calls (MortonId, cheb_poly, cheb_eval,
 * Matrix::pinv, cheb_approx) are not in the reference repository; their
 * semantics are defined here as documented per function and pinned by the
 * mathematics above.
 */
#ifndef TBX_ORACLE_H_
#define TBX_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BC_FREESPACE 0
#define ORC_BC_PERIODIC 1

/* MortonId quantisation depth of the oracle (PVFMM's MAX_DEPTH role). */
#define ORC_MAXD 30

/* One rank's leaves, in Morton order (PVFMM postorder over non-ghost leaves). */
typedef struct {
  int64_t num_leaves;
  const uint32_t *anchor; /* [L][3] corner in units of 2^-depth, x,y,z   */
  const uint8_t *depth;   /* [L]                                          */
  int32_t q;              /* Chebyshev degree (ChebDeg)                   */
  int32_t dof;            /* DataDOF                                      */
  double *coeff;          /* [L][dof][Nc] packed triangle i(z),j(y),k(x)  */
} orc_tree;

/* Scalar time modulation of a steady velocity tree: v(x,t) = g(t) * V(x). */
typedef double (*orc_tfac_fn)(double t, void *ctx);

int orc_ncoef(int q);

/* utils/cheb.h:41-67 new_nodes: stretched Chebyshev nodes on [0,1]. */
void orc_new_nodes(int q, double *x);
/* First-kind Chebyshev nodes on [0,1] (PVFMM cheb_nodes, as used by
 * InitTree tree_utils.h:144). */
void orc_cheb_nodes_first_kind(int q, double *x);
/* pvfmm::cheb_poly as called at tree_functor.h:328-330. */
void orc_cheb_poly(int q, const double *x, int64_t n, double *out);

/* Leaf of each point as EvalTree/EvalNodesLocal assign it
 * (tree_functor.h:442-449 wrap, :464-489 keys + sort, :190-198 partition).
 * Wraps a private copy; pts are not modified. */
int orc_locate(const orc_tree *t, const double *pts, int64_t n, int bc,
               int32_t *leaf);

/* EvalTree (tree_functor.h:397-791) + EvalNodesLocal (:155-395) + vec_eval
 * (:27-84).  Wraps pts in place when bc is periodic, as the reference does. */
int orc_eval_tree(const orc_tree *t, double *pts, int64_t n, int bc,
                  double *out);

/* ComputeTrajRK2 + IntegrateRK2 (semilag/traj.inc:20-68) with the
 * NodeFieldFunctor of `vel` (tree_functor.h:793-815), optionally scaled by
 * g(t).  xsol may alias xinit. */
int orc_traj_rk2(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                 const double *xinit, int64_t n, double tinit, double tfinal,
                 int num_rk_step, int bc, double *xsol);

/* SolveSemilagRK2 (semilag/semilag.inc:26-45). */
int orc_semilag_rk2(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                    const orc_tree *conc, const double *pts, int64_t n,
                    int timestep, double dt, int num_rk_step, int bc,
                    double *vals);

/* CollectChebTreeGridPoints (tree_utils.h:442-498). out: [L*P][3]. */
int orc_collect_grid_points(const orc_tree *t, double *out);

/* GetPt2CoeffMatrix (utils/cheb.h:166-196): M = pinv(B)^T, [P][Nc]. */
int orc_pt2coeff(int q, double *M);

/* SetTreeGridValues (tree_utils.h:500-552) with values in dof-major leaf
 * layout [L][dof][P]; M from orc_pt2coeff. */
int orc_set_grid_values(orc_tree *t, const double *M, const double *vals);

/* SolveSemilagInSitu (tree/tree_semilag.h:92-135). */
int orc_semilag_insitu(const orc_tree *vel, orc_tfac_fn g, void *gctx,
                       orc_tree *conc, const double *M, int timestep, double dt,
                       int num_rk_step, int bc);

/* Initial fit of one leaf from values at first-kind nodes (x fastest),
 * [dof][d^3] -> [dof][Nc] (PVFMM cheb_approx, used at tree_utils.h:172). */
int orc_cheb_approx(int q, int dof, const double *vals, double *coeff);

/* fast_interp (tree_functor.h:89-153): tricubic Lagrange on a regular
 * N^3 grid of [0,1]^3, values [dof][N^3], x fastest. */
int orc_fast_interp(const double *grid, int dof, int N, const double *pts,
                    int64_t n, double *out);

/* Leaves holding regular N^3 sample grids (values at corner + 2^-depth *
 * (a, b, c) / (N - 1), x fastest, [L][dof][N^3]) evaluated by fast_interp
 * in the leaf's unit coordinates u = (x - corner) * 2^depth.  The
 * reference's fast_interp (tree_functor.h:89-153) has no caller there; this
 * tree-level use defines the tricubic variant of the step. */
typedef struct {
  int64_t num_leaves;
  const uint32_t *anchor;
  const uint8_t *depth;
  int32_t n;
  int32_t dof;
  double *samples;
} orc_grid;
int orc_grid_eval(const orc_grid *g, double *pts, int64_t n, int bc, double *out);
int orc_grid_points(const orc_grid *g, double *out);
int orc_grid_traj_rk2(const orc_grid *vel, orc_tfac_fn g, void *gctx,
                      const double *xinit, int64_t n, double tinit, double tfinal,
                      int num_rk_step, int bc, double *xsol);
/* SolveSemilagInSitu with tricubic leaves: the new samples are the values
 * at the departure points of the sample points (no refit). */
int orc_grid_semilag_insitu(const orc_grid *vel, orc_tfac_fn g, void *gctx,
                            orc_grid *conc, int timestep, double dt,
                            int num_rk_step, int bc);

/* CubicInterpPolicy::InterpCubic1D (utils/cubic.h:42-56). */
double orc_cubic1d(double x, const double *xx, const double *pp);

/* FieldSetFunctor::operator() (tree/tree_set_functor.h:49-79): 4 trees at
 * times tt[4], cubic Hermite in time.  Wraps pts in place when periodic. */
int orc_field_set_eval(const orc_tree *const *trees, const double *tt,
                       double *pts, int64_t n, double t, int bc, double *out);

/* FieldExtrapFunctor::operator() (tree/tree_extrap_functor.h:47-78). */
int orc_field_extrap_eval(const orc_tree *tp, const orc_tree *tc, double *pts,
                          int64_t n, int bc, double *out);

/* A velocity functor of the kinds the reference's trajectory templates take:
 *  ORC_VEL_TREE   NodeFieldFunctor(t[0]) (tree_functor.h:793-815), optionally
 *                 scaled by g(t);
 *  ORC_VEL_EXTRAP the extrapolating overloads (traj.inc:70-115): field_fn =
 *                 NodeFieldFunctor(t[0]) (current), extrap_fn =
 *                 FieldExtrapFunctor(t[1] previous, t[0] current);
 *  ORC_VEL_SET    FieldSetFunctor(t[0..3], tt[0..3]) (tree_set_functor.h). */
#define ORC_VEL_TREE 0
#define ORC_VEL_EXTRAP 1
#define ORC_VEL_SET 2
typedef struct {
  int32_t kind;
  const orc_tree *t[4];
  double tt[4];
  orc_tfac_fn g;
  void *gctx;
} orc_velocity;

/* ComputeTrajRK2 with either overload (traj.inc:48-68, 97-115). */
int orc_traj_rk2_v(const orc_velocity *vel, const double *xinit, int64_t n,
                   double tinit, double tfinal, int num_rk_step, int bc,
                   double *xsol);
/* SolveSemilagInSitu with either overload (tree_semilag.h:92-181). */
int orc_semilag_insitu_v(const orc_velocity *vel, orc_tree *conc,
                         const double *M, int timestep, double dt,
                         int num_rk_step, int bc);

#ifdef __cplusplus
}
#endif
#endif
