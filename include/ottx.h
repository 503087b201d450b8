/*
 * ottx.h — C ABI of the MI355X (gfx950) log-domain Sinkhorn library.
 *
 * The library replaces the XLA fusions that OTT compiles for
 * ``ott.solvers.linear.solve(PointCloud(x, y), a, b)`` in log-domain mode with
 * hand-written HIP kernels.  OTT itself has no FFI: its boundary is Python.
 * Each entry point below names the reference interface it stands in for
 * (paths relative to ott-jax/ott ``src/ott``); the Python package ``ott_amd``
 * binds them with ctypes and mirrors the reference's classes on top.
 *
 * Conventions
 *   - Every array pointer is device memory owned by the caller (torch tensors
 *     in the Python binding), contiguous, float32 unless stated.
 *   - Points are row-major: x is [n, d], y is [m, d].
 *   - Calls are ordered on the caller's ``stream`` (a hipStream_t passed as
 *     void*; NULL is the default stream).
 *   - Return value: OTTX_OK (0) or a negative OTTX_E* code; nothing throws or
 *     aborts across the ABI.  Numerical failure (NaN, no convergence) is
 *     reported in the outputs exactly as OTT reports it, never as an error code.
 *   - The library allocates no device memory.  Scratch space is a caller
 *     workspace sized by the matching ``*_workspace_bytes`` query.
 *   - Precision: OTTX_PREC_FP32 computes the cross term -2<x,y> in float32
 *     (MFMA f32), OTTX_PREC_BF16_SPLIT with hi/lo bfloat16 operands
 *     (hi*hi + hi*lo + lo*hi, float32 accumulation).  Potentials, exp and
 *     log-sum-exp are float32; marginal-error and cost reductions are float64.
 */
#ifndef OTTX_H_
#define OTTX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OTTX_ABI_VERSION 5

enum ottx_status {
  OTTX_OK = 0,
  OTTX_EINVAL = -1,        /* invalid argument (shape, pointer, option) */
  OTTX_EUNSUPPORTED = -2,  /* valid OTT option this library does not run */
  OTTX_EHIP = -3,          /* a HIP runtime call failed */
  OTTX_EWORKSPACE = -4,    /* workspace too small or misaligned */
  OTTX_ESCRATCH = -5,      /* problem larger than the no-allocation scratch */
};

enum ottx_precision {
  OTTX_PREC_FP32 = 0,
  OTTX_PREC_BF16_SPLIT = 1,
};

enum ottx_init_mode {
  OTTX_INIT_GIVEN = 0,
  OTTX_INIT_F_FROM_G = 1,
  OTTX_INIT_ZEROS = 2,
  OTTX_INIT_F_GIVEN = 3,
};

/* A point cloud geometry with squared-Euclidean cost,
 * C_ij = inv_scale_cost * ||x_i - y_j||^2, and its linear problem
 * (weights a, b and unbalancedness tau_a, tau_b).
 * Reference: geometry/pointcloud.py:54-72 (PointCloud), geometry/costs.py:547-561
 * (SqEuclidean), problems/linear/linear_problem.py:51-63 (LinearProblem). */
typedef struct ottx_problem {
  int64_t n;               /* rows of x */
  int64_t m;               /* rows of y */
  int32_t d;               /* feature dimension (1..256) */
  int32_t precision;       /* enum ottx_precision */
  const float* x;          /* [n, d] */
  const float* y;          /* [m, d] */
  const float* a;          /* [n] weights, >= 0 */
  const float* b;          /* [m] weights, >= 0 */
  double epsilon;          /* target epsilon, > 0, used when relative_epsilon == 0 */
  double inv_scale_cost;   /* s in C = s * ||x - y||^2, > 0 (pointcloud.py:334-340) */
  double tau_a;            /* in (0, 1] (linear_problem.py:51-63) */
  double tau_b;
  double relative_epsilon; /* > 0: epsilon = relative_epsilon * mean(C) or * std(C) (see
                              epsilon_statistic), resolved on the device by ottx_sinkhorn_solve
                              (relative_epsilon="mean" / "std" and the default epsilon
                              0.05 * std(C), geometry.py:147-171); 0: use epsilon */
  int32_t epsilon_statistic; /* enum ottx_eps_statistic: the statistic relative_epsilon
                              multiplies */
  int32_t cost;            /* enum ottx_cost: the translation-invariant cost h(x - y) */
  double cost_p;           /* p of OTTX_COST_PNORMP, _SQPNORM, _EUCLIDEANP (>= 1) */
} ottx_problem;

/* Costs of a point cloud, c(x, y) = h(x - y) (geometry/costs.py):
 * OTTX_COST_SQEUCLIDEAN runs the MFMA half-step on the cross term; the others
 * sum a per-feature term of z = x - y on the vector ALUs, online (the cost is
 * recomputed per tile, never stored).  With them a relative epsilon is
 * resolved by the caller (ottx_cost_summary gives the mean and std of C);
 * ottx_pointcloud_apply and ottx_cost_gradient need the squared-Euclidean
 * cost.  inv_scale_cost multiplies every cost. */
enum ottx_cost {
  OTTX_COST_SQEUCLIDEAN = 0, /* ||z||_2^2 (costs.py:548-573) */
  OTTX_COST_PNORMP = 1,      /* ||z||_p^p / p (costs.py:304-334) */
  OTTX_COST_SQPNORM = 2,     /* ||z||_p^2 / 2 (costs.py:269-301) */
  OTTX_COST_EUCLIDEANP = 3,  /* ||z||_2^p (costs.py:337-361) */
  OTTX_COST_EUCLIDEAN = 4,   /* ||z||_2 (costs.py:528-545) */
};

/* The statistic of the cost matrix a relative epsilon multiplies
 * (geometry/geometry.py:112-136, both in closed form, never forming C). */
enum ottx_eps_statistic {
  OTTX_EPS_STAT_MEAN = 0,  /* mean(C) */
  OTTX_EPS_STAT_STD = 1,   /* std(C) = sqrt(relu(mean(C^2) - mean(C)^2)) */
};

/* Per-block progress callback of a solve (Sinkhorn(progress_fn=...),
 * sinkhorn.py:933-937): called on the host thread of ottx_sinkhorn_solve once
 * each block of inner_iterations has completed, in order, with the block's
 * index and its raw error (before the +inf rule of min_iterations).  The
 * solve waits for every block before queuing the next while a callback is
 * set. */
typedef void (*ottx_progress_fn)(void* user, int32_t block, double error);

/* Options of the Sinkhorn solver: solvers/linear/sinkhorn.py:718-780. */
typedef struct ottx_sinkhorn_opts {
  double threshold;         /* default 1e-3 */
  int32_t norm_error;       /* p of the p-norm, >= 1 (default 1) */
  int32_t inner_iterations; /* default 10 */
  int32_t min_iterations;   /* default 0 */
  int32_t max_iterations;   /* default 2000 */
  double momentum;          /* constant Momentum weight (acceleration.py:161-172), 1 = none */
  int32_t tile_config;      /* -1 = the library's default for the precision, d and size (the
                               operator entry points always use it), else a value listed by
                               ottx_tile_configs */
  int32_t use_graph;        /* 1 = replay each block of iterations as a hipGraph */
  int32_t init_mode;        /* OTTX_INIT_GIVEN: f and g are used as given;
                               OTTX_INIT_F_FROM_G: f starts as the f-update of the given g (a warm
                               start that predicts g only);
                               OTTX_INIT_ZEROS: DefaultInitializer (initializers.py:70-99), f and g
                               start at 0, -inf where the weight is 0 (their contents are ignored);
                               OTTX_INIT_F_GIVEN: f as given, g as OTTX_INIT_ZEROS sets it (a warm
                               start that predicts f: the first g-update discards g but for its
                               -inf entries, geometry.py:297-299) */
  int32_t recenter_potentials; /* 1: SinkhornState.recenter on the outputs (sinkhorn.py:76-114,
                               1008-1009): f -= mean of its finite entries, g += the same, for a
                               balanced problem; no change when tau_a or tau_b is 1 (the errors
                               are unchanged in both cases).  With tau_a < 1 and tau_b < 1 the
                               translation-invariant iteration of lse_step (sinkhorn.py:809-854)
                               runs, the errors are those of the recentred state and the outputs
                               are recentred (the Sejourne shift). */
  int32_t parallel_dual_updates; /* 1: Sinkhorn(parallel_dual_updates=True) (sinkhorn.py:840-846):
                               both updates read the previous state; the error is the row plus
                               the column marginal error (sinkhorn.py:160-189).  On the symmetric
                               problems of a Sinkhorn divergence (sinkhorn_divergence.py:298-318:
                               x and y the same buffer, a and b the same buffer or both NULL,
                               tau_a = tau_b = 1, init_mode OTTX_INIT_ZEROS) f and g stay equal, so
                               each iteration is one pass, g <- (1 - w) g + w g^(f), then f <- g,
                               and the error is twice the column one.  Any other problem runs the
                               f-update first, saving the old f for the g-update.  Not with
                               recenter_potentials, a schedule or adaptive momentum
                               (OTTX_EUNSUPPORTED). */
  double epsilon_init;      /* epsilon schedule (epsilon_scheduler.py:24-60): iteration t uses
                               max(epsilon_init epsilon_decay^t, 1) epsilon; the marginal error
                               and the outputs use epsilon itself.  Values <= 1 (0 included) run
                               at epsilon throughout.  A schedule (epsilon_init > 1) is supported
                               without parallel updates and without the recentred unbalanced
                               iteration; those return OTTX_EUNSUPPORTED, and so does the batched
                               solve. */
  double epsilon_decay;     /* in (0, 1]; 0 is read as 1 */
  int32_t momentum_start;   /* 0: constant momentum (the weight above).  > 0: Momentum(start)
                               (acceleration.py:138-160): iterations from momentum_start on use
                               Lehmann's weight 2 / (1 + sqrt(1 - r^(1 / inner))),
                               r = min(err[k-1] / err[k-2], 0.99), k = momentum_start /
                               inner_iterations, if err[k-1] < momentum_error_threshold, with the
                               errors of the requested norm (as the reference records them).
                               Needs momentum_start >= 2 inner_iterations; not with parallel
                               updates or in the batched solve (OTTX_EUNSUPPORTED). */
  int32_t _pad0;
  double momentum_error_threshold; /* Momentum.error_threshold (inf when unused) */
  ottx_progress_fn progress_fn; /* NULL: none */
  void* progress_user;      /* passed back to progress_fn */
} ottx_sinkhorn_opts;

/* Outputs of one solve (SinkhornOutput fields, sinkhorn.py:280-323). */
typedef struct ottx_sinkhorn_out {
  float* f;                 /* [n] device: initial potential in, final potential out */
  float* g;                 /* [m] device: initial potential in, final potential out */
  float* errors;            /* [ceil(max_iterations / inner_iterations)] device */
  double reg_ot_cost;       /* host, written on return (sinkhorn.py:224-277) */
  double epsilon;           /* host: the target epsilon the solve used */
  double total_mass;        /* host: sum of the coupling at (f, g) */
  int32_t converged;        /* host, sinkhorn.py:1020-1023 */
  int32_t n_iters;          /* host, sinkhorn.py:431-434 */
  int32_t n_blocks;         /* host: blocks of inner_iterations run */
  int32_t _pad;
} ottx_sinkhorn_out;

/* A batch of independent problems of one feature dimension, concatenated:
 * problem k owns rows x_off[k] .. x_off[k+1]-1 of x and a, and rows
 * y_off[k] .. y_off[k+1]-1 of y and b.
 * Reference: tools/segment_sinkhorn.py:24-138 with geometry/segment.py:23-190,
 * which pad every problem to max_measure_size with zero-weight zero vectors
 * and vmap one solve per problem.  Nothing is padded here; mean_rows_x/y
 * reproduce the padded statistics a relative epsilon sees there (padding
 * rows have zero weight, so they change nothing else). */
typedef struct ottx_batch {
  int32_t count;            /* K problems (>= 1) */
  int32_t d;                /* feature dimension (1..256) */
  int32_t precision;        /* OTTX_PREC_* */
  int32_t flags;            /* OTTX_BATCH_* bits */
  const float* x;           /* [x_off[K], d] device */
  const float* y;           /* [y_off[K], d] device */
  float* a;                 /* [x_off[K]] device; read, or written (OTTX_BATCH_UNIFORM_A) */
  float* b;                 /* [y_off[K]] device; read, or written (OTTX_BATCH_UNIFORM_B) */
  const int64_t* x_off;     /* [K + 1] host, x_off[0] = 0, non-decreasing with n_k >= 1 */
  const int64_t* y_off;     /* [K + 1] host */
  const double* epsilon;    /* [K] host absolute epsilons; NULL when relative_epsilon > 0 */
  double relative_epsilon;  /* > 0: eps_k = relative_epsilon * mean(C_k) */
  double mean_rows_x;       /* > 0: mean(C_k) averages over this many x rows, as over a cloud
                               padded to that size with zero vectors; 0: over n_k */
  double mean_rows_y;       /* the same for y; 0: over m_k */
  double inv_scale_cost;
  double tau_a;
  double tau_b;
} ottx_batch;

/* ottx_batch.flags: the solve writes the default weights 1 / n_k (1 / m_k)
 * into a (b) for the rows of each problem, as segment_point_cloud's default
 * (geometry/segment.py:100-104), instead of reading them. */
#define OTTX_BATCH_UNIFORM_A 1
#define OTTX_BATCH_UNIFORM_B 2

typedef struct ottx_batch_out {
  float* f;                 /* [x_off[K]] device: initial potentials in, final out */
  float* g;                 /* [y_off[K]] device */
  float* errors;            /* [K, ceil(max_iterations / inner_iterations)] device */
  double* reg_ot_cost;      /* [K] host */
  double* epsilon;          /* [K] host, nullable */
  int32_t* converged;       /* [K] host */
  int32_t* n_iters;         /* [K] host */
  int32_t n_blocks;         /* host: blocks of inner_iterations launched */
  int32_t _pad;
} ottx_batch_out;

/* Workspace of ottx_sinkhorn_batched_solve. */
size_t ottx_sinkhorn_batched_workspace_bytes(const ottx_batch* batch, const ottx_sinkhorn_opts* opts);
/* Every problem of the batch solved as by ottx_sinkhorn_solve, each with its
 * own epsilon, errors and convergence, in shared launches: one launch per
 * half-step covers the column tiles and row splits of all problems, and a
 * problem whose loop has finished drops out of later launches.
 * Reference: tools/segment_sinkhorn.py:24-138 (returns reg_ot_cost per
 * segment; this returns every SinkhornOutput field per problem).
 * opts->init_mode OTTX_INIT_F_FROM_G is not supported here; max_iterations must be >= inner_iterations. */
int ottx_sinkhorn_batched_solve(const ottx_batch* batch, const ottx_sinkhorn_opts* opts,
                                ottx_batch_out* out, void* workspace, size_t workspace_bytes,
                                void* stream);

/* Library identity. */
int ottx_abi_version(void);
const char* ottx_strerror(int status);

/* Mean of the cost matrix in closed form, float64:
 * s * (mean ||x||^2 + mean ||y||^2 - 2 <mean x, mean y>).
 * Replaces geometry/geometry.py:112-117 -> pointcloud.py:170-186 (the
 * relative_epsilon="mean" path of geometry.py:147-171).  *out_host is written
 * after the stream has been synchronised. */
size_t ottx_mean_cost_workspace_bytes(int64_t n, int64_t m, int32_t d);
int ottx_mean_cost(int64_t n, int64_t m, int32_t d, const float* x,
                   const float* y, double inv_scale_cost, double* out_host,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Mean and standard deviation of the cost matrix, float64, in closed form
 * from per-cloud moments (row sums of u, |u|^2 u, |u|^2, |u|^4 and the Gram
 * matrix U^T U of the centred points): out_host[0] = mean(C), out_host[1] =
 * std(C) = sqrt(relu(mean(C^2) - mean(C)^2)), both over the n m entries with
 * C = inv_scale_cost |x_i - y_j|^2.  Replaces geometry/geometry.py:112-136
 * (mean_cost_matrix, std_cost_matrix), which forms C^2 for the standard
 * deviation.  Blocks until both values are on the host. */
size_t ottx_cost_moments_workspace_bytes(int64_t n, int64_t m, int32_t d);
int ottx_cost_moments(int64_t n, int64_t m, int32_t d, const float* x, const float* y, double inv_scale_cost,
                      double* out_host, void* workspace, size_t workspace_bytes, void* stream);

/* The Geometry operator protocol of the reference, for one PointCloud.
 * mode selects the output written to out[k] (k over rows if axis == 1,
 * columns if axis == 0):
 *   OTTX_OP_APPLY_LSE   apply_lse_kernel(f, g, eps, axis)[0]  (geometry.py:260-299)
 *   OTTX_OP_UPDATE      update_potential(f, g, log_marginal, axis)   (geometry.py:376-398);
 *                       log_marginal ([m] if axis == 0 else [n]) is required for this op only
 *   OTTX_OP_MARGINAL    marginal_from_potentials(f, g, axis)   (geometry.py:326-349)
 * ``eps`` is the epsilon used in the exponent (the problem's epsilon for
 * OTTX_OP_MARGINAL). */
enum ottx_op {
  OTTX_OP_APPLY_LSE = 0,
  OTTX_OP_UPDATE = 1,
  OTTX_OP_MARGINAL = 2,
};
size_t ottx_op_workspace_bytes(const ottx_problem* problem);
int ottx_pointcloud_op(const ottx_problem* problem, int32_t op, int32_t axis,
                       const float* f, const float* g, const float* log_marginal,
                       double eps, float* out, void* workspace,
                       size_t workspace_bytes, void* stream);

/* Size of the library's static scratch for ottx_pointcloud_op_noalloc. */
#define OTTX_NOALLOC_SCRATCH_BYTES (32u << 20)

/* ottx_pointcloud_op without a workspace argument and without any allocation.
 * No reference counterpart: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The scratch is a file-scope __device__ array of
 * OTTX_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a
 * problem whose scratch need (ottx_pointcloud_op_noalloc_bytes) exceeds it
 * returns OTTX_ESCRATCH.  Every call on a device shares that device's array
 * (each call re-initialises what must start at zero): calls that overlap on
 * one device, from several streams or threads, race.  Results are those of
 * ottx_pointcloud_op on the same inputs. */
int ottx_pointcloud_op_noalloc(const ottx_problem* problem, int32_t op, int32_t axis,
                               const float* f, const float* g, const float* log_marginal,
                               double eps, float* out, void* stream);
/* Scratch bytes ottx_pointcloud_op_noalloc needs for the problem (0: invalid). */
size_t ottx_pointcloud_op_noalloc_bytes(const ottx_problem* problem);

/* Transport applied to vectors, geometry.py:445-486
 * (apply_transport_from_potentials, vmapped over a batch):
 *   axis 0: out[v * m + j] = sum_i P_ij vecs[v * n + i]   (vecs [k, n], out [k, m])
 *   axis 1: out[v * n + i] = sum_j P_ij vecs[v * m + j]   (vecs [k, m], out [k, n])
 * with P_ij = exp((f_i + g_j - C_ij) / eps), signed vectors included (the
 * reference's signed log-sum-exp with b=vec, return_sign=True).  One
 * log-sum-exp pass, then one launch per 16 vectors: the cross term on MFMA as
 * in the half-step, the products with the vectors on fp32 MFMA.  Backs
 * SinkhornOutput.apply (sinkhorn.py:453-467). */
size_t ottx_apply_workspace_bytes(const ottx_problem* problem, int32_t axis);
/* flags: repeated applies of the same transport (iterative solvers) may keep
 * the workspace between calls and skip the preparation of the clouds
 * (OTTX_APPLY_PREPARED: a previous call on this workspace prepared this
 * problem) and the column log-sum-exp of this axis (OTTX_APPLY_LSE_READY: a
 * previous call on this workspace computed it for the same f, g, eps and
 * axis).  A workspace for both axes: the larger of the two sizes. */
enum ottx_apply_flags {
  OTTX_APPLY_PREPARED = 1,
  OTTX_APPLY_LSE_READY = 2,
};
int ottx_pointcloud_apply(const ottx_problem* problem, int32_t axis, const float* f, const float* g,
                          const float* vecs, int32_t k, double eps, int32_t flags, float* out, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Cost maxima behind PointCloud's scale_cost modes (pointcloud.py:342-383):
 * out_host[0] = max_i |x_i|^2, out_host[1] = max_j |y_j|^2 (cost_fn.norm of the
 * squared-Euclidean cost, for "max_norm" and "max_bound") and, when
 * with_max_cost, out_host[2] = max_ij |x_i - y_j|^2 (the unscaled cost, for
 * "max_cost"; one pass over the pairs).  workspace: at least 256 bytes,
 * 256-byte aligned.  Blocks until the values are on the host. */
int ottx_cost_maxima(int64_t n, int64_t m, int32_t d, const float* x, const float* y, int32_t with_max_cost,
                     double* out_host, void* workspace, size_t workspace_bytes, void* stream);

/* Median of the unscaled cost matrix c(x_i, y_j) over its n m entries, the
 * value behind scale_cost="median" (pointcloud.py:351-353, geometry.py:
 * 106-109 median_cost_matrix: jnp.median, the mean of the two middle entries
 * when n m is even; NaN when an entry is NaN).  An exact radix select over
 * the entries' float bits, costs computed from the differences in float32
 * with the problem's cost, four passes over the pairs; the matrix is never
 * formed.  n m <= 2^40.  Uses n, m, d, x, y, cost and cost_p of the problem.
 * workspace: ottx_cost_median_workspace_bytes(), 256-byte aligned.  Blocks
 * until the value is on the host. */
size_t ottx_cost_median_workspace_bytes(void);
int ottx_cost_median(const ottx_problem* problem, double* out_host, void* workspace, size_t workspace_bytes,
                     void* stream);

/* Sum, sum of squares and maximum of the unscaled cost matrix c(x_i, y_j)
 * (out_host[0..2], the sums in float64), one pass over the pairs: the mean
 * and std of the cost matrix (geometry.py:112-136) and scale_cost "mean" /
 * "max_cost" (pointcloud.py:342-383) for costs without the squared
 * Euclidean's closed form (the reference forms the matrix or applies it to a
 * vector).  Uses n, m, d, x, y, cost and cost_p of the problem.  Blocks until
 * the values are on the host. */
size_t ottx_cost_summary_workspace_bytes(int64_t n, int64_t m);
int ottx_cost_summary(const ottx_problem* problem, double* out_host, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Dense transport matrix from potentials, geometry.py:361-365
 * (transport_from_potentials): out[i * m + j] = exp((f_i + g_j - C_ij) / eps)
 * with C_ij = inv_scale_cost * |x_i - y_j|^2 summed from the differences in
 * float32, row-major [n, m].  For SinkhornOutput.matrix (sinkhorn.py:445-451);
 * the caller owns out (n * m floats). */
int ottx_transport_matrix(const ottx_problem* problem, const float* f, const float* g, double eps, float* out,
                          void* stream);

/* <P, C'> = sum_ij P_ij C'_ij, P_ij = exp((f_i + g_j - C_ij) / eps) of the
 * problem's clouds and C'_ij = other_inv_scale_cost |x'_i - y'_j|^2 of another
 * pair of clouds of the same sizes (x_other [n, d_other], y_other [m, d_other]
 * row-major float32), or C' = C when both are NULL.  Replaces
 * SinkhornOutput.transport_cost_at_geom and primal_cost (sinkhorn.py:345-411),
 * which the reference evaluates through the cost's low-rank factors
 * (pointcloud.py:298-319); here one pass over the pairs, neither matrix stored,
 * float32 pair terms summed in float64.  Blocks until *out_host is written. */
size_t ottx_transport_cost_workspace_bytes(int64_t n, int64_t m);
int ottx_transport_cost(const ottx_problem* problem, const float* f, const float* g, double eps,
                        const float* x_other, const float* y_other, int32_t d_other,
                        double other_inv_scale_cost, double* out_host, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Gradient of reg_ot_cost with respect to the points of one cloud, with the
 * potentials held fixed: the Danskin gradient that jax.grad takes through
 * SinkhornOutput.reg_ot_cost, whose f and g are stop_gradient'ed
 * (sinkhorn.py:329-335, compute_kl_reg_cost :224-277).  Only the coupling's
 * total mass depends on the points, so
 *   side 0: grad[i * d + k] = 2 s sum_j P_ij (x_ik - y_jk)   ([n, d])
 *   side 1: grad[j * d + k] = 2 s sum_i P_ij (y_jk - x_ik)   ([m, d])
 * with P_ij = exp((f_i + g_j - s |x_i - y_j|^2) / eps), s = inv_scale_cost,
 * held constant (a relative epsilon or a data-dependent scale_cost adds terms
 * this does not include).  d <= 64 (else OTTX_EUNSUPPORTED).  Asynchronous on
 * the stream. */
size_t ottx_cost_gradient_workspace_bytes(const ottx_problem* problem, int32_t side);
int ottx_cost_gradient(const ottx_problem* problem, const float* f, const float* g, double eps, int32_t side,
                       float* grad, void* workspace, size_t workspace_bytes, void* stream);

/* Full log-domain Sinkhorn solve: Sinkhorn.__call__ -> run -> iterations
 * (sinkhorn.py:782-801, 1052-1099) with the fixed-point loop of
 * math/fixed_point_loop.py:23-93, the error of sinkhorn.py:127-221 fused into
 * the half-steps, and reg_ot_cost of sinkhorn.py:224-277 in float64.
 * out->f and out->g hold the initial potentials on entry (zeros with -inf on
 * zero weights reproduces DefaultInitializer, initializers.py:70-131).
 * Blocks until the loop has stopped; outputs are final on return. */
size_t ottx_sinkhorn_workspace_bytes(const ottx_problem* problem,
                                     const ottx_sinkhorn_opts* opts);
int ottx_sinkhorn_solve(const ottx_problem* problem,
                        const ottx_sinkhorn_opts* opts, ottx_sinkhorn_out* out,
                        void* workspace, size_t workspace_bytes, void* stream);

/* Tile configurations of the fused half-step for (precision, d): the search
 * space of the autotuner (no reference counterpart; OTT leaves tiling to XLA).
 * Returns the number of configurations; when desc is non-NULL writes up to
 * max_configs records of 5 int32: {tile_config, NB (16-column blocks per wave),
 * MT (16-row blocks per streamed tile), W (waves per workgroup), target
 * workgroups per CU}.  tile_config is what ottx_sinkhorn_opts.tile_config and
 * ottx_benchmark_halfstep accept (-1 = default). */
int ottx_tile_configs(int32_t precision, int32_t d, int32_t* desc,
                      int32_t max_configs);

/* Measurement helper (no reference counterpart; used by bench.py for the
 * roofline figure and by the autotuner).  Prepares the problem once, then
 * runs ``reps`` real half-steps back to back on ``stream`` -- the g-update
 * (axis 0: new g from f) or the f-update (axis 1) -- with the given
 * tile_config, bracketed by hipEvents, and writes the mean duration of one
 * half-step launch in milliseconds to *avg_ms_host.  f and g are updated in
 * place.  Uses the workspace of ottx_sinkhorn_workspace_bytes. */
int ottx_benchmark_halfstep(const ottx_problem* problem, int32_t axis,
                            int32_t tile_config, int32_t reps, float* f,
                            float* g, double* avg_ms_host, void* workspace,
                            size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Learned warm start: a DeepSets network predicting g from the two clouds.
 * No direct reference counterpart: OTT's learned initializer
 * (initializers/neural/meta_initializer.py:35-213) maps the weights of a fixed
 * geometry to f with flax and optax.adam; this network reads the points, is
 * trained with the kernels below (forward, backward, Adam), and its output is
 * turned into a warm start (f0 from one f-update of g) by the Python layer.
 * Parameters are one flat float32 buffer of ottx_ds_param_count(d) values
 * (layout in ott_amd/csrc/ottx_deepsets.hip: param_layout); bf16 copies of the
 * weight matrices (ottx_ds_weight_bf16_count(d) values) feed the MFMA GEMMs.
 * ------------------------------------------------------------------------ */
typedef struct ottx_ds_model {
  int32_t d;               /* feature dimension of the clouds */
  int32_t _pad;
  float* params;           /* [ottx_ds_param_count(d)] device, float32 master weights */
  void* weights_bf16;      /* [ottx_ds_weight_bf16_count(d)] device, bf16 copies */
} ottx_ds_model;

/* A batch of problems; all arrays are device memory.  Encoder rows are the
 * points ordered x_0, y_0, x_1, y_1, ... (segment 2 b + side); head rows are
 * the y points in y order. */
typedef struct ottx_ds_batch {
  int32_t count;           /* problems B */
  int32_t d;
  int32_t n_total;         /* sum of n_b */
  int32_t m_total;         /* sum of m_b */
  const float* x;          /* [n_total, d] problems concatenated */
  const float* y;          /* [m_total, d] */
  const float* eps;        /* [B] target epsilon of each problem */
  const int32_t* x_off;    /* [B + 1] row offsets into x */
  const int32_t* y_off;    /* [B + 1] row offsets into y */
  const int32_t* y_prob;   /* [roundup(m_total, 64)] problem of each head row, 0 on padding */
  const int32_t* enc_src;  /* [enc_rows] row in x (side 0) or y (side 1) of encoder row r */
  const int32_t* seg_of_row; /* [enc_rows] segment 2 b + side of encoder row r */
  const int32_t* seg_off;  /* [2 B + 1] encoder row offsets of the segments */
  const float* eps_rel;    /* [B] epsilon / mean cost of each problem, or NULL (then eps / sigma^2 / 2) */
  int32_t enc_rows;        /* encoder rows = seg_off[2 B]: every point, or a subsample of each cloud
                              (the pooled embedding is a mean over the set) */
  int32_t _pad1;
} ottx_ds_batch;

int64_t ottx_ds_param_count(int32_t d);
int64_t ottx_ds_weight_bf16_count(int32_t d);
/* Workspace of ottx_ds_forward (training = 0) or ottx_ds_loss_backward (training = 1). */
size_t ottx_ds_workspace_bytes(const ottx_ds_batch* batch, int32_t training);
/* Recompute the bf16 weight copies from the float32 parameters. */
int ottx_ds_refresh_bf16(const ottx_ds_model* model, void* stream);
/* g_out[m_total]: predicted potentials g_j = sigma_b^2 * psi(...) (up to a constant), with
   sigma_b^2 half the mean squared-Euclidean cost of problem b. */
int ottx_ds_forward(const ottx_ds_model* model, const ottx_ds_batch* batch,
                    float* g_out, void* workspace, size_t workspace_bytes,
                    void* stream);
/* Centred mean-squared loss against target[m_total] (potentials divided by
 * sigma_b^2, any per-problem constant), its gradient into grad[param_count] and the
 * per-problem losses into loss_out[B] (nullable). */
int ottx_ds_loss_backward(const ottx_ds_model* model,
                          const ottx_ds_batch* batch, const float* target,
                          float* grad, float* loss_out, void* workspace,
                          size_t workspace_bytes, void* stream);
/* One Adam step (bias-corrected, optax.adam semantics) on the float32
 * parameters, then a refresh of the bf16 copies. */
int ottx_ds_adam_step(const ottx_ds_model* model, const float* grad, float* m,
                      float* v, int32_t step, float lr, float beta1,
                      float beta2, float eps, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* OTTX_H_ */
