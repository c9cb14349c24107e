/*
 * tbx.h: C ABI of the MI355X semi-Lagrangian advection engine.
 *
 * The engine replaces, for one MPI rank's whole tree on one GPU, the hot path
 * of arashb/tbslas: the RK2 backward-trajectory solve (src/semilag) and the
 * per-point octree leaf location plus Chebyshev interpolation of velocity
 * and concentration at departure points (src/tree).  The header-only C++
 * shim include/tbx/tbslas.hpp restores the reference's functor/step
 * templates over this ABI.
 *
 * Conventions
 *  - Every entry point returns an int status (TBX_OK = 0).  tbx_last_error()
 *    returns the calling thread's message for the last failure.  No C++
 *    exception crosses the ABI.
 *  - Point arrays are AoS xyz FP64 [n][3]; field values are point-major
 *    [n][dof] (the reference's functor layout, tree_functor.h:376-383).
 *  - Pointers passed to compute calls are device pointers (hipMalloc or any
 *    allocator of device memory) owned by the caller.  The tree handle owns
 *    device copies of the tree.
 *  - Calls are ordered on the given stream (NULL = the null stream).  A tree
 *    handle is externally synchronised: calls that use one handle on two
 *    host threads at once race.
 *  - Counts are int64.  Leaf indices are int32 positions in Morton order.
 *
 * Leaf layout (PVFMM Cheb_Node, kept so trees move between the two codes):
 *  leaves are listed in Morton order (PVFMM postorder over non-ghost leaves);
 *  each has an integer anchor (x,y,z) in units of its own edge 2^-depth and a
 *  depth; coefficients per leaf are [dof][Nc], Nc = (q+1)(q+2)(q+3)/6, each a
 *  packed triangle in order i(z) -> j(y) -> k(x), i+j+k <= q, basis
 *  T_k(x) T_j(y) T_i(z) on the leaf mapped to [-1,1]^3 (tree_functor.h:256-266).
 */
#ifndef TBX_H_
#define TBX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbx_tree_s tbx_tree;
typedef void *tbx_stream; /* hipStream_t */

enum {
  TBX_OK = 0,
  TBX_EINVAL = 1,      /* bad argument                                  */
  TBX_ENOMEM = 2,      /* host or device allocation failed              */
  TBX_EORDER = 3,      /* leaves not strictly in Morton order           */
  TBX_ECOVER = 4,      /* leaves overlap or leave a gap in [0,1)^3       */
  TBX_EHIP = 5,        /* HIP runtime or library error                  */
  TBX_EOVERFLOW = 6,   /* size exceeds an int64/int32 limit             */
  TBX_EUNSUPPORTED = 7 /* degree or combination the library does not cover */
};

enum { TBX_BC_FREESPACE = 0, TBX_BC_PERIODIC = 1 }; /* sim_config.h:59 bc */

/* Arithmetic of the coefficient contraction (flags, low byte). */
#define TBX_FP64 0x0u  /* FP64, fused multiply-add                      */
#define TBX_EXACT 0x1u /* FP64, the reference's order with separate     */
                       /* multiply and add (vec_eval, tree_functor.h:56-76); */
                       /* bit-identical to the CPU path                  */
#define TBX_FP32 0x2u  /* FP32 coefficients and accumulation; point      */
                       /* coordinates, RK2 updates and keys stay FP64    */
#define TBX_FP32_HEAD64 0x3u /* TBX_FP32 for the terms of total degree > 1, */
                             /* FP64 for the 4 terms of degree <= 1, which  */
                             /* carry most of a smooth field's magnitude    */
#define TBX_MODE_MASK 0xFu
/* EvalTree wraps the caller's points in place under a periodic bc
 * (tree_functor.h:442-449, const_cast at :803).  tbx_eval does so only when
 * this flag is set. */
#define TBX_WRITE_BACK_WRAP 0x10u

/* Largest Chebyshev degree the kernels are built for. */
#define TBX_MAX_Q 16
/* Deepest leaf the engine's Morton keys resolve. */
#define TBX_MAX_DEPTH 21

/* Host-side description of one tree. */
typedef struct {
  int64_t num_leaves;
  const uint32_t *anchor; /* [L][3] */
  const uint8_t *depth;   /* [L]    */
  int32_t q;              /* Chebyshev degree, 1..TBX_MAX_Q             */
  int32_t dof;            /* field components per point, >= 1           */
  const double *coeff;    /* host [L][dof][Nc], or NULL for zeros       */
} tbx_tree_desc;


/* Scalar time modulation g(t) of a steady velocity tree, v(x,t) = g(t) V(x);
 * NULL means steady.  Evaluated on the host once per RK stage.  It takes the
 * role of the time argument of the functor concept (traj.inc:33, 40), which
 * NodeFieldFunctor ignores (tree_functor.h:808-811). */
typedef double (*tbx_tfac_fn)(double t, void *ctx);

/* A velocity functor (dof 3 trees of equal degree):
 *  TBX_VEL_TREE   NodeFieldFunctor(trees[0]) (tree_functor.h:793-815),
 *                 scaled by g(t) when g is not NULL;
 *  TBX_VEL_EXTRAP the extrapolating RK2 overloads (traj.inc:70-115):
 *                 NodeFieldFunctor(trees[0]) for the first stage,
 *                 FieldExtrapFunctor(trees[1], trees[0]) = 1.5 v(t^n) -
 *                 0.5 v(t^n-1) for the second (tree_extrap_functor.h:47-78);
 *                 trees[0] current, trees[1] previous;
 *  TBX_VEL_SET    FieldSetFunctor(trees[0..3], times[0..3]): cubic Hermite
 *                 in time across four trees (tree_set_functor.h:49-79). */
enum { TBX_VEL_TREE = 0, TBX_VEL_EXTRAP = 1, TBX_VEL_SET = 2 };
typedef struct {
  int32_t kind;
  const struct tbx_tree_s *trees[4];
  double times[4];
  tbx_tfac_fn g;
  void *gctx;
} tbx_velocity;

/* ---- tree handle: replaces the pvfmm::MPI_Tree<Cheb_Node<double>> that
 *      NodeFieldFunctor points at (tree_functor.h:796, 814) -------------- */
int tbx_tree_create(const tbx_tree_desc *desc, tbx_tree **out);
int tbx_tree_destroy(tbx_tree *t);
int tbx_tree_info(const tbx_tree *t, int64_t *num_leaves, int32_t *q,
                  int32_t *dof, int32_t *max_depth);
/* Replace / read the coefficients ([L][dof][Nc] FP64).  *_device selects
 * whether the buffer is device (1) or host (0) memory. */
int tbx_tree_set_coeffs(tbx_tree *t, const double *coeff, int on_device,
                        tbx_stream s);
int tbx_tree_get_coeffs(const tbx_tree *t, double *coeff, int on_device,
                        tbx_stream s);

/* ---- point location: EvalTree's periodic wrap + Morton keys + sort +
 *      lower_bound partition (tree_functor.h:442-489, :190-198).  leaf[i] is
 *      the leaf EvalTree evaluates point i in.  pts are not modified. ---- */
int tbx_locate(const tbx_tree *t, const double *pts, int64_t n, int bc,
               int32_t *leaf, tbx_stream s);

/* ---- EvalTree (tree_functor.h:397-791) with EvalNodesLocal (:155-395) and
 *      vec_eval (:27-84): the NodeFieldFunctor::operator() body (:800-806).
 *      out: [n][dof]. ------------------------------------------------------ */
int tbx_eval(const tbx_tree *t, double *pts, int64_t n, int bc, uint32_t flags,
             double *out, tbx_stream s);

/* ---- The leaf contraction alone (no counterpart in the reference: a
 *      measurement of the evaluation's inner kernel without location, RK2
 *      or regrouping).  Point i lies in leaf (i / 64 / waves_per_leaf) mod L
 *      at a hashed position in its box; every component of that leaf is
 *      contracted there in the mode of flags and out[i] = ((0 + u_0) + u_1)
 *      + ... .  pts, if not null, receives the points ([n][3]). ---------- */
int tbx_contraction_only(const tbx_tree *t, int64_t n, int waves_per_leaf, uint32_t flags,
                         double *out, double *pts, tbx_stream s);

/* ---- ComputeTrajRK2 / IntegrateRK2 (semilag/traj.inc:20-68) through the
 *      velocity tree's functor.  xsol may alias xinit (the reference runs in
 *      place, traj.inc:64).  Like the reference, the first stage of every
 *      substep wraps xsol in place under a periodic bc. ------------------- */
int tbx_traj_rk2(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                 const double *xinit, int64_t n, double tinit, double tfinal,
                 int num_rk_step, int bc, uint32_t flags, double *xsol,
                 tbx_stream s);

/* ---- SolveSemilagRK2 (semilag/semilag.inc:26-45): departure points of pts
 *      over [timestep*dt, timestep*dt - dt], then conc at them.
 *      vals: [n][conc dof]. ------------------------------------------------ */
int tbx_semilag_rk2(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                    const tbx_tree *conc, const double *pts, int64_t n,
                    int timestep, double dt, int num_rk_step, int bc,
                    uint32_t flags, double *vals, tbx_stream s);

/* ---- CollectChebTreeGridPoints (tree_utils.h:442-498): out [L*P][3],
 *      P = (q+1)^3 stretched nodes per leaf, x fastest (cheb.h:41-67). ---- */
int tbx_collect_grid_points(const tbx_tree *t, double *out, tbx_stream s);

/* ---- SetTreeGridValues (tree_utils.h:500-552): refit every leaf from
 *      values at its grid points, vals [L][dof][P] (dof-major per leaf; the
 *      reference's in-situ call passes point-major values, which agrees for
 *      dof = 1 only).  coeff = vals * pinv(B)^T (cheb.h:166-196). -------- */
int tbx_set_grid_values(tbx_tree *t, const double *vals, tbx_stream s);

/* ---- The tree's values at its own grid points, vals [L][dof][P]: each
 *      leaf's polynomial at its (q+1)^3 stretched nodes, upper-face nodes
 *      included (the tensor-product Chebyshev values B c, by three one-axis
 *      contractions).  No counterpart in the reference: it is the inverse of
 *      tbx_set_grid_values on represented data, and the expansion a step's
 *      first RK2 stage reads under TBX_OPT_NODE_VELOCITY. ------------------ */
int tbx_tree_grid_values(const tbx_tree *t, double *vals, tbx_stream s);
/* The same values rounded to FP32 (what the FP32 modes' first stage reads). */
int tbx_tree_grid_values_f32(const tbx_tree *t, float *vals, tbx_stream s);

/* ---- SolveSemilagInSitu (tree/tree_semilag.h:92-135): one fused step.
 *      Arrival points are generated on the fly, RK2 runs through vel, conc
 *      is evaluated at the departure points and refit in place.  vel and
 *      conc must share q; up to 16 RK substeps.  With a conc dof > 1 each
 *      component is an independent tracer carried by the same trajectories.
 *      dep, if not NULL, receives the departure points [L*P][3]. --------- */
int tbx_semilag_insitu(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                       tbx_tree *conc, int timestep, double dt,
                       int num_rk_step, int bc, uint32_t flags, double *dep,
                       tbx_stream s);

/* ---- tbx_semilag_insitu without allocation.  No counterpart in the
 *      reference: it is for callers that cannot allocate device memory
 *      (code under HIP graph capture, tight serving loops, harnesses that
 *      forbid hipMalloc).  The step's value buffer, the refit workspace and
 *      FP32 coefficient mirrors the trees lack live in a static device
 *      scratch of 256 MiB per GPU (held until the process ends); a step
 *      that needs more returns TBX_EOVERFLOW before any work is queued.
 *      The velocity's nodal values (TBX_OPT_NODE_VELOCITY) go there too
 *      when they fit beside those; when they do not, the step evaluates its
 *      first RK2 stage instead (results then agree with
 *      tbx_semilag_insitu's to rounding, as with the option off).
 *      *scratch_bytes (may be NULL) receives the bytes this step uses.
 *      Every call on one GPU shares that GPU's scratch: calls that overlap
 *      on one GPU (two streams, two host threads) race.  Otherwise results
 *      equal tbx_semilag_insitu's. ----------------------------------------- */
int tbx_semilag_insitu_noalloc(const tbx_tree *vel, tbx_tfac_fn g, void *gctx, tbx_tree *conc,
                               int timestep, double dt, int num_rk_step, int bc,
                               uint32_t flags, double *dep, size_t *scratch_bytes,
                               tbx_stream s);

/* ---- the first two stages of SolveSemilagInSitu without the refit:
 *      arrival points of conc's leaves (tree_semilag.h:105-115) through
 *      SolveSemilagRK2 (tree_semilag.h:124-127).  vals receives [L][dof][P]
 *      (the layout tbx_set_grid_values takes); dep, if not NULL, the
 *      departure points [L*P][3].  tbx_semilag_insitu = this call followed
 *      by tbx_set_grid_values(conc, vals). -------------------------------- */
int tbx_semilag_values(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                       const tbx_tree *conc, int timestep, double dt,
                       int num_rk_step, int bc, uint32_t flags, double *vals,
                       double *dep, tbx_stream s);

/* ---- the same entry points for any velocity functor ------------------- */
/* The functor's value at pts at time t: a scaled tree, FieldExtrapFunctor's
 * extrapolation, or FieldSetFunctor's Hermite value.  out [n][3]. */
int tbx_velocity_eval(const tbx_velocity *vel, double *pts, int64_t n, double t,
                      int bc, uint32_t flags, double *out, tbx_stream s);
int tbx_traj_rk2_v(const tbx_velocity *vel, const double *xinit, int64_t n,
                   double tinit, double tfinal, int num_rk_step, int bc,
                   uint32_t flags, double *xsol, tbx_stream s);
int tbx_semilag_values_v(const tbx_velocity *vel, const tbx_tree *conc,
                         int timestep, double dt, int num_rk_step, int bc,
                         uint32_t flags, double *vals, double *dep,
                         tbx_stream s);
/* SolveSemilagInSitu, both overloads (tree_semilag.h:92-181). */
int tbx_semilag_insitu_v(const tbx_velocity *vel, tbx_tree *conc, int timestep,
                         double dt, int num_rk_step, int bc, uint32_t flags,
                         double *dep, tbx_stream s);

/* ---- leaves with regular sample grids (tricubic) ----------------------
 * The reference's fast_interp (tree_functor.h:89-153: 4-point Lagrange per
 * axis, first index (int)(u (N-1)) - 1 clamped to [0, N-4], zero outside
 * [0,1]^3; it has no caller there) applied per leaf: each leaf holds an
 * N^3 grid of samples at corner + 2^-depth (a, b, c) / (N - 1), x fastest,
 * samples [L][dof][N^3], N in 4..16, evaluated in the leaf's unit
 * coordinates u = (x - corner) 2^depth.  Flags: TBX_EXACT (fast_interp's
 * operation order), TBX_FP64, TBX_FP32 (FP32 samples and sums). */
typedef struct tbx_grid_s tbx_grid;
typedef struct {
  int64_t num_leaves;
  const uint32_t *anchor; /* [L][3], Morton order as tbx_tree_desc */
  const uint8_t *depth;
  int32_t n;              /* samples per axis */
  int32_t dof;
  const double *samples;  /* host [L][dof][N^3], or NULL for zeros */
} tbx_grid_desc;

int tbx_grid_create(const tbx_grid_desc *desc, tbx_grid **out);
int tbx_grid_destroy(tbx_grid *g);
int tbx_grid_set_samples(tbx_grid *g, const double *samples, int on_device, tbx_stream s);
int tbx_grid_get_samples(const tbx_grid *g, double *samples, int on_device, tbx_stream s);
/* Values at pts (device [n][3], wrapped in place with TBX_WRITE_BACK_WRAP
 * when periodic), out device [n][dof]. */
int tbx_grid_eval(const tbx_grid *g, double *pts, int64_t n, int bc, uint32_t flags,
                  double *out, tbx_stream s);
/* SolveSemilagInSitu for grid leaves (vel: dof 3, scaled by g(t) when g
 * is given): the sample points' departure points by num_rk_step RK2
 * substeps, the new samples the values there (no refit); dep (device
 * [L*N^3][3]) optional. */
int tbx_grid_semilag_insitu(const tbx_grid *vel, tbx_tfac_fn g, void *gctx, tbx_grid *conc,
                            int timestep, double dt, int num_rk_step, int bc, uint32_t flags,
                            double *dep, tbx_stream s);

/* ---- initial fit and error metric on the device ----------------------
 * First-kind Chebyshev points of every leaf (the cheb_nodes InitTree
 * samples the initial field at, tree_utils.h:144): out [L*(q+1)^3][3],
 * x fastest, leaf after leaf. */
int tbx_tree_first_kind_points(const tbx_tree *t, double *out, tbx_stream s);
/* InitTree's fit (the cheb_approx role, tree_utils.h:132-178) on the
 * device: vals [L*(q+1)^3][dof] point-major at tbx_tree_first_kind_points
 * -> the tree's coefficients (per-axis discrete Chebyshev transform,
 * truncated to the triangle; the same as tbx_fit_leaves). */
int tbx_tree_fit(tbx_tree *t, const double *vals, tbx_stream s);
/* CheckChebOutput (applications/include/utils.txx:118-297) on the device.
 * tbx_error_points: the metric's points, n^3 first-kind nodes per leaf with
 * n = q/3 + 1 (utils.txx:149-153), out [L*n^3][3].  tbx_check_cheb_output:
 * the tree (evaluated in the mode of flags) against ref [L*n^3][dof], the
 * exact field at those points (device): the volume-weighted mean error is
 * removed and res[4] = {absolute L2, relative L2, absolute L-inf,
 * relative L-inf} (utils.txx:158-293).  dof <= 8. */
int tbx_error_points(const tbx_tree *t, double *out, tbx_stream s);
int tbx_check_cheb_output(const tbx_tree *t, const double *ref, uint32_t flags, double *res,
                          tbx_stream s);

/* ---- adaptivity ---------------------------------------------------------
 * RefineTree followed by Balance21 (advection.cpp:305-314,
 * tree_utils.h:117-123), on the device, in place:
 *  1. a leaf whose top-shell tail (tbx_tail_norm's criterion) exceeds tol
 *     and whose depth is below max_depth is split into its 8 children;
 *  2. 8 sibling leaves that all stay, above min_depth, merge into their
 *     parent when the parent's fit of them (each child's polynomial at the
 *     parent's grid nodes, refit) has a tail <= tol;
 *  3. leaves more than one level coarser than a face-, edge- or
 *     corner-adjacent leaf (across the domain when periodic) are split
 *     until none is.
 * A child takes its parent's polynomial (exact: the refit of its values at
 * the child's nodes).  The tree's leaf count changes (tbx_tree_info);
 * buffers sized by it are made again on their next use. */
int tbx_tree_adapt(tbx_tree *t, double tol, int min_depth, int max_depth, int bc,
                   tbx_stream s);
/* The tree's current leaves: anchor [L][3], depth [L] (host). */
int tbx_tree_get_leaves(const tbx_tree *t, uint32_t *anchor, uint8_t *depth);

/* ---- output -------------------------------------------------------------
 * Write2File (the PVFMM tree writer the reference's drivers call,
 * advection.cpp:184-188, 317-324): the tree's values on order^3 uniform
 * points per leaf (corners and faces included; each leaf's own polynomial),
 * as VTK XML hexahedra, order in 2..TBX_MAX_Q + 1, order - 1 cells per axis
 * per leaf.
 * Writes <prefix>_000000.vtu (raw appended data, UInt64 headers, point
 * data "value" with dof components, FP64) and <prefix>.pvtu naming it, the
 * one-rank form of the reference's per-rank pieces. */
int tbx_tree_write_vtu(const tbx_tree *t, const char *prefix, int order, tbx_stream s);

/* ---- many small problems in one persistent launch --------------------
 * A batch advances nprob independent problems (velocity tree vel[p], dof 3;
 * concentration tree conc[p], dof 1; degree 6, 8 or 10; time step dt[p];
 * stationary velocity, one RK2 substep) by one in-situ step each in a
 * single persistent kernel, in FP32: work items of a few leaves are claimed
 * from a global counter, each item evaluates its departure points and
 * refits its leaves at once.  The two steps of each parity are captured as
 * HIP graphs at configuration time and replayed by tbx_batch_step.  The
 * batch keeps its own FP32 coefficients (initialised from conc[p]); the
 * trees provide the leaf structure and must outlive the batch.  The
 * configuration (workgroup size x leaves per item) is chosen by
 * tbx_batch_tune, which times every one on the current state.  With
 * TBX_OPT_NODE_VELOCITY on when the batch is made, the velocities are
 * expanded at their grid nodes once (FP32, 12 bytes per velocity grid
 * point, kept only if they take at most half of the free device memory)
 * and each problem's first RK2 stage reads them where an arrival node is a
 * node of its velocity leaf, as tbx_semilag_insitu does.  No counterpart
 * in the reference, whose ranks each own one tree. */
typedef struct tbx_batch_s tbx_batch;
int tbx_batch_create(int nprob, const tbx_tree *const *vel, const tbx_tree *const *conc,
                     const double *dt, int bc, tbx_batch **out);
int tbx_batch_destroy(tbx_batch *b);
int tbx_batch_set_config(tbx_batch *b, int cfg);       /* 0..5 */
int tbx_batch_get_config(const tbx_batch *b, int *cfg);
/* *on = 1 if the batch's first RK2 stage reads the velocities' nodal values. */
int tbx_batch_node_velocity(const tbx_batch *b, int *on);
/* Times each configuration (ms[6], may be NULL), keeps the fastest. */
int tbx_batch_tune(tbx_batch *b, double *ms, tbx_stream s);
int tbx_batch_step(tbx_batch *b, int nsteps, tbx_stream s);
/* Carry an in-situ step of (vel, conc) -- for example many tracers on one
 * trajectory set -- in the batch's step graphs, ahead of the persistent
 * kernel: each tbx_batch_step replay then advances conc and every problem
 * of the batch in one graph launch.  The velocity is stationary (no time
 * factor), flags give the arithmetic, the batch's bc applies.  The step's
 * buffers are made here (the graphs allocate nothing); conc and vel must
 * outlive the batch. */
int tbx_batch_attach_step(tbx_batch *b, const tbx_tree *vel, tbx_tree *conc, double dt,
                          int num_rk_step, uint32_t flags);
/* Current coefficients of problem p, host [L][1][Nc] (FP32 state). */
int tbx_batch_get_coeffs(const tbx_batch *b, int p, double *coeff, tbx_stream s);

/* ---- adjoint (no counterpart in the reference) --------------------------
 * The step c' = R E_D(V) c of tbx_semilag_insitu (a single velocity tree,
 * optionally scaled by g) is linear in c; its cotangents are
 *   dc = E_D^T R^T lam'                the transposed interpolation, a
 *                                      per-leaf scatter-add into coefficients
 *   dV = sum_p a_p . dD_p/dV           basis derivatives at the departure
 *                                      points chained through the RK2 stage
 * Arithmetic: TBX_FP64 or TBX_EXACT (the forward's trajectory arithmetic);
 * the sums are atomic, so results are reproducible up to summation order. */

/* Cotangent of one tbx_semilag_insitu step taken from conc's current
 * coefficients c (left unchanged): lam_out = dL/dc given lam_next = dL/dc'
 * (device arrays [L][dof][Nc]); grad_vel (device [L_vel][3][Nc_vel], may be
 * NULL) is ADDED to.  dep (device [L*P][3], may be NULL): the departure
 * points that forward step wrote (its dep output); read back instead of
 * integrating the trajectories again.  num_rk_step: the forward step's RK
 * substeps, 1..16 (the velocity cotangent runs back through them). */
int tbx_semilag_insitu_vjp(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                           const tbx_tree *conc, int timestep, double dt,
                           int num_rk_step, int bc, uint32_t flags,
                           const double *dep, const double *lam_next,
                           double *lam_out, double *grad_vel, tbx_stream s);

/* nsteps in-situ steps from conc's coefficients c_0 (timesteps 1..nsteps),
 * loss = 0.5 |c_T - c_target|^2 (returned on the host), and its gradients:
 * grad_c0 [L][dof][Nc] and grad_vel [L_vel][3][Nc_vel] (may be NULL),
 * overwritten.  States are checkpointed every checkpoint_every steps and
 * recomputed segment by segment in the backward sweep; the departure
 * points of each recomputed step (of step 1 alone for a stationary
 * velocity, g NULL, whose trajectories are the same every step) are kept
 * on the device, 24 bytes per grid point and step, while they fit in half
 * the free device memory, and the adjoint steps read them instead of
 * integrating again.  A stationary velocity scatters its cotangent once,
 * and once its departure points are kept, its one-component forward steps
 * (the first sweep and the replays) evaluate the concentration there and
 * refit, with the same results as a full step.  conc holds c_T on return. */
int tbx_adjoint_run(const tbx_tree *vel, tbx_tfac_fn g, void *gctx,
                    tbx_tree *conc, int nsteps, double dt, int bc,
                    uint32_t flags, const double *c_target, int checkpoint_every,
                    double *grad_c0, double *grad_vel, double *loss,
                    tbx_stream s);

/* ---- host helpers (setup; no device work) ----------------------------- */
/* new_nodes (cheb.h:41-58), x[q+1]. */
int tbx_new_nodes(int q, double *x);
/* GetPt2CoeffMatrix (cheb.h:166-196): M[P][Nc] = pinv(B)^T. */
int tbx_pt2coeff_matrix(int q, double *M);
/* The same operator in the factored form the device refit applies:
 * pinv(B) = H B^T with H = (B^T B)^-1, H [Nc][Nc] (symmetric).  B is a
 * tensor product of one-axis bases restricted to the triangle, so B^T runs
 * as three one-axis contractions and H as one GEMM over all leaves. */
int tbx_refit_gram_inverse(int q, double *H);
/* Uniform tree of the given depth in Morton order: anchor [8^depth][3],
 * depth [8^depth]. */
int tbx_uniform_leaves(int depth, uint32_t *anchor, uint8_t *dep);
/* First-kind Chebyshev sample points of each leaf, out [L][(q+1)^3][3],
 * x fastest: where an initial field is sampled for tbx_fit_leaves. */
int tbx_first_kind_points(int64_t num_leaves, const uint32_t *anchor,
                          const uint8_t *depth, int q, double *out);
/* Initial fit (the cheb_approx role in InitTree, tree_utils.h:132-178):
 * vals [L][(q+1)^3][dof] point-major at tbx_first_kind_points ->
 * coeff [L][dof][Nc]. */
int tbx_fit_leaves(int64_t num_leaves, int q, int dof, const double *vals,
                   double *coeff);
/* Largest top-shell coefficient sum per leaf, max over dof of
 * sum_{i+j+k=q} |c_ijk|: the refinement criterion of tbx_refine_*. */
int tbx_tail_norm(int64_t num_leaves, int q, int dof, const double *coeff,
                  double *tail);
/* 2:1 balance (the Balance21 role, tree_utils.h:121-123): splits leaves
 * until face-, edge- and corner-adjacent leaves differ by at most one level
 * (wrapping across the domain when periodic).  In/out arrays hold up to
 * capacity leaves; *num_leaves is updated. */
int tbx_balance21(int64_t *num_leaves, int64_t capacity, uint32_t *anchor,
                  uint8_t *depth, int bc);

/* ---- tuning knobs (process-wide) ---------------------------------------
 * TBX_OPT_STEP_SHAPE: launch shape of the fused step kernel, used by later
 * step calls.  0: 256 threads; 1-3: 256 / 1024 / 512 threads with a
 * register budget of 4 waves per SIMD; 4-7: the workgroup's points
 * regrouped by leaf before each evaluation, 256 / 512 threads (budget 2
 * waves per SIMD) and 1024 / 512 threads (budget 4); 8-10: regrouped,
 * 1024 threads with a budget of 8 waves per SIMD (64 registers), 512 and
 * 768 threads with a budget of 6.  -1 (the default) picks per arithmetic
 * mode and degree the shape measured fastest (q = 14: 8 for TBX_FP32, 10
 * for TBX_FP32_HEAD64; 6 otherwise); shapes other than 0 and 6 exist for
 * q = 8 and 14 only (others run 6).
 * Every shape gives the same results.
 * TBX_OPT_TRACER_MFMA (default 1): a concentration with a multiple of 16
 * components (many tracers on one trajectory set) has its components
 * contracted on the matrix cores in TBX_FP64 and TBX_FP32 steps (regrouped
 * workgroups of 512 threads in FP64 and 256 in FP32, whatever the step
 * shape); 0 keeps the vector contraction.  The two agree to rounding (summation order).
 * TBX_OPT_GRID_LDS (default 1): the grid-leaf step on a uniform tree of 8^3-sample
 * leaves whose velocity and concentration trees share their leaves stages each
 * sibling block of 8 leaves in LDS, in FP32 steps (1) or in every mode (2); 0
 * gathers every stencil from global memory.  All give the same bits.
 * TBX_OPT_NODE_VELOCITY (default 1): a step with a single velocity tree (or
 * the extrapolating functor, whose first stage is its current tree), in
 * any mode but TBX_EXACT, expands the velocity tree at its own grid nodes
 * (three one-axis contractions per leaf, into a buffer of L * 3 * (q+1)^3
 * doubles the tree keeps) and takes the first RK2 stage's velocity at each
 * arrival node from there when the node is a grid node of the velocity leaf
 * it is located in (that leaf has the arrival leaf's depth); other nodes,
 * and every later stage, are evaluated as before.  0 evaluates every stage.
 * The two agree to rounding (the expansion's summation order; in the FP32
 * modes the expansion is FP64).  The grid-leaf step (tbx_grid_semilag_insitu)
 * takes the first stage at a sample point from the velocity's sample there
 * under the same rule (velocity and concentration of the same N), the
 * tricubic interpolation reproducing the sample up to rounding.  Steps
 * captured into a batch graph keep the setting in force when the graph was
 * made. */
enum {
  TBX_OPT_STEP_SHAPE = 1,
  TBX_OPT_TRACER_MFMA = 2,
  TBX_OPT_GRID_LDS = 3,
  TBX_OPT_NODE_VELOCITY = 4
};
int tbx_set_option(int option, int value);
int tbx_get_option(int option, int *value);

/* The step launch shape (TBX_OPT_STEP_SHAPE) that a step with these mode
 * flags at degree q runs with under the current option.  No counterpart in
 * the reference: it reports the launch configuration for profiling. */
int tbx_step_shape(uint32_t flags, int q, int *shape);

const char *tbx_last_error(void);
const char *tbx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TBX_H_ */
