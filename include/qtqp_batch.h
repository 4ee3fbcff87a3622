/* C ABI of the batched device interior-point engine (MI355X, gfx950).
 *
 * Solves many independent dense QPs of uneven sizes with qtqp's algorithm
 * (Mehrotra predictor-corrector on the homogeneous embedding), every
 * per-iteration step on the GPU: one workgroup runs one whole IPM iteration
 * of one QP (KKT update, MFMA Gram product and LDL', three refined KKT
 * solves, tau, step, sigma, termination), a persistent grid pulls QPs from
 * a work queue in longest-first order, and the per-iteration launch pair is
 * replayed from a HIP graph until every QP has a status.
 *
 * Reference behaviour restated per QP (qtqp v0.0.5, paths relative to its
 * checkout): src/qtqp/__init__.py:351-614 (QTQP._solve_impl) and 671-1008;
 * src/qtqp/direct.py:183-307 (regularised diagonal, iterative refinement);
 * src/qtqp/solvers_dense.py:99-142 (Gram/Schur reduction); Ruiz
 * equilibration (__init__.py:616-669) also runs on the device.  Presolve and
 * the final status-dependent scaling stay on the host (qtqp_amd/batch.py),
 * as in __init__.py:248-297 and 580-614.
 *
 * Host pointers are caller-owned C-contiguous float64 arrays, copied during
 * the call.  Calls are synchronous.  A handle is not thread-safe.
 */
#ifndef QTQP_BATCH_H_
#define QTQP_BATCH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QTQP_BATCH_UNFINISHED 0
#define QTQP_BATCH_SOLVED 1
#define QTQP_BATCH_INFEASIBLE 2
#define QTQP_BATCH_UNBOUNDED 3
#define QTQP_BATCH_ERROR 5      /* FP64 factor breakdown (the reference raises LinAlgError) */

typedef struct qtqp_batch qtqp_batch;

/* Solve options: the keyword arguments of QTQP.solve (__init__.py:309-326)
 * plus the engine's precision mode (QTQP_PREC_FP64 / QTQP_PREC_MIXED). */
typedef struct qtqp_batch_opts {
  double atol, rtol, atol_infeas, rtol_infeas;
  double step_size_scale, min_static_regularization;
  double linear_solver_atol, linear_solver_rtol;
  int32_t max_iter, max_iterative_refinement_steps;
  int32_t precision;
  int32_t use_graph;    /* launch from a hipGraph (the solve, or each iteration) */
  int32_t equilibrate;  /* Ruiz scaling on the device (__init__.py:616-663) */
  int32_t per_iteration;/* 0: one persistent launch for the whole solve, each
                           workgroup iterating a QP to completion; 1: one launch
                           per IPM iteration of every unfinished QP */
} qtqp_batch_opts;

/* Engine counters of one qtqp_batch_solve call. */
typedef struct qtqp_batch_counters {
  int64_t backend_solves;  /* K_reg^{-1} applications over all QPs */
  int64_t inner_steps;     /* FP32 refinement steps over all QPs */
  int64_t fp64_fallbacks;  /* QPs whose factor fell back to FP64 */
  int64_t ir_not_converged;/* KKT solves that ended stalled / non-converged */
  int32_t iterations;      /* IPM iterations launched (max over QPs) */
  double device_ms;        /* device time of the whole solve (HIP events) */
  int64_t jitter_retries;  /* FP64 factors redone with a larger diagonal shift */
} qtqp_batch_counters;

/* Packs `count` QPs onto the device (no compute).  QP i has sizes n[i],
 * m[i], z[i] (1 <= n[i] <= 1024, 0 <= z[i] < m[i]) and its unscaled,
 * presolved data at a + a_off[i] (m x n row-major), p + p_off[i] (n x n
 * symmetric), b + m_off[i], c + n_off[i].  Packed vectors x/y/s used by
 * set_start and results follow n_off / m_off. */
int qtqp_batch_create(int32_t count, const int32_t* n, const int32_t* m, const int32_t* z,
                      const int64_t* a_off, const int64_t* p_off, const int64_t* n_off,
                      const int64_t* m_off, const double* a, const double* p,
                      const double* b, const double* c, qtqp_batch** out);

/* qtqp_batch_create with flags (QTQP_BATCH_NO_ROW_PAIRS or 0).  By default
 * rows that are exact negations of another row (A[q] = -A[p], as in the
 * two-sided bounds lo <= G x <= hi written as two inequality blocks) are
 * stored once: every pass over A reads the stored rows only and forms both
 * constraints' products from them (A[q] x = -A[p] x; the pair's terms of
 * A'y and A'WA are summed as (y_p - y_q) A[p] and (w_p + w_q) A[p]'A[p]), so
 * results agree with the unpaired layout to rounding, not bit for bit.
 * No reference counterpart: the reference keeps A as given. */
#define QTQP_BATCH_NO_ROW_PAIRS 1
/* By default each stored row's passes read only its nonzero column extent
 * (the zero columns before the first and after the last nonzero read as
 * zero without memory traffic: the bound rows of [I; -I] read one 16-byte
 * piece).  Results are those of the full rows.  This flag reads full rows. */
#define QTQP_BATCH_NO_ROW_EXTENTS 2
/* By default, once a QP factors in FP64 and n <= 256, each IPM iteration's
 * two solves that need only the factor -- K^{-1} [c; b] and the
 * predictor's (__init__.py:482-503) -- refine together: each keeps its own
 * residuals, tolerance and step count, and every refinement pass while both
 * still step reads A and P once for both.  Results are bit for bit those of
 * the two solves one after the other, which this flag runs instead.  No
 * reference counterpart (the reference solves them in turn). */
#define QTQP_BATCH_NO_FUSED_SOLVES 4
int qtqp_batch_create_ex(int32_t count, const int32_t* n, const int32_t* m, const int32_t* z,
                         const int64_t* a_off, const int64_t* p_off, const int64_t* n_off,
                         const int64_t* m_off, const double* a, const double* p,
                         const double* b, const double* c, int32_t flags, qtqp_batch** out);

/* A batch of `count` QPs of one parametric family: every QP has the same
 * n, m, z, A (m x n row-major) and P (n x n), and its own b (row q of the
 * count x m row-major b) and c (row q of the count x n c).  A and P are
 * uploaded, paired and equilibrated once (Ruiz scaling depends on A and P
 * only, __init__.py:616-663) and every QP reads that one device copy;
 * results are those of qtqp_batch_create_ex on `count` copies of A and P.
 * Packed x/y/s follow n_off[q] = q n, m_off[q] = q m.  Serves the same
 * QTQP(...).solve() per QP as the reference (__init__.py:309-614); the
 * sharing itself has no reference counterpart. */
int qtqp_batch_create_family(int32_t count, int32_t n, int32_t m, int32_t z, const double* a,
                             const double* p, const double* b, const double* c, int32_t flags,
                             qtqp_batch** out);

/* Replaces b and c of every QP (packed by m_off / n_off, as at creation) for
 * the next solve: a re-solve of the same A and P with new right-hand sides,
 * as QTQP(a, b', c', z, p).solve() would (the caller keeps the presolve's
 * dropped rows the same).  No reference counterpart (the reference builds a
 * new QTQP per problem). */
int qtqp_batch_set_bc(qtqp_batch* h, const double* b, const double* c);

/* Stored rows of A per QP (m[i] less the number of row pairs), in creation
 * order.  No reference counterpart (batch diagnostics). */
int qtqp_batch_stored_rows(qtqp_batch* h, int32_t* mc);

/* FP64 bytes one pass over each QP's stored rows reads: the 16-byte pieces
 * inside each row's nonzero column extent, in creation order.  No reference
 * counterpart (batch diagnostics: the executed-byte count of a pass). */
int qtqp_batch_row_bytes(qtqp_batch* h, int64_t* bytes);

/* Sets the starting point of the next solve for every QP: x, y, s in the
 * unscaled coordinates (equilibrated on the device like _init_variables,
 * __init__.py:299-307 and 668-669) and tau.  Without this call the default
 * start x = 0, y[z:] = s[z:] = 1, tau = 1 is used. */
int qtqp_batch_set_start(qtqp_batch* h, const double* x, const double* y, const double* s,
                         const double* tau);

/* qtqp_batch_set_start from device memory of the handle's device: x, y, s
 * and tau packed as for qtqp_batch_set_start (for example filled by
 * qtqp_mlp_propose).  Same meaning, same __init__.py lines. */
int qtqp_batch_set_start_device(qtqp_batch* h, const double* x, const double* y,
                                const double* s, const double* tau);

/* The handle's own packed device buffers of x (sum of n), y and s (sum of m)
 * and tau (count): a producer on the device may fill them and pass them to
 * qtqp_batch_set_start_device.  qtqp_batch_results and qtqp_batch_solution
 * overwrite them.  No reference counterpart. */
int qtqp_batch_device_buffers(qtqp_batch* h, double** x, double** y, double** s, double** tau);

/* Equilibrates on the device and runs the IPM on every QP until each has a
 * status or max_iter iterations.  Inputs stay on the device, so a handle can
 * be solved repeatedly. */
int qtqp_batch_solve(qtqp_batch* h, const qtqp_batch_opts* opts, qtqp_batch_counters* counters);

/* Copies back the final iterates in the unscaled coordinates (x = e x_eq,
 * y = d y_eq, s = s_eq / d; not yet divided by tau), tau, status
 * (QTQP_BATCH_*; UNFINISHED after max_iter), IPM iteration count and the
 * final pres/dres/gap.  Any output pointer may be NULL. */
int qtqp_batch_results(qtqp_batch* h, double* x, double* y, double* s, double* tau,
                       int32_t* status, int32_t* iters, double* pres, double* dres,
                       double* gap);

/* The solution of every QP as QTQP.solve returns it before the presolve's
 * dropped rows are restored (__init__.py:580-614): x, y, s divided by tau
 * (solved, max_iter, error), or the infeasibility certificate (x = s = NaN,
 * y / |b'y|), or the unboundedness certificate (y = NaN, x and s divided by
 * |c'x|), formed on the device; plus status and iteration count.  Packed as
 * qtqp_batch_results; any output pointer may be NULL. */
int qtqp_batch_solution(qtqp_batch* h, double* x, double* y, double* s, int32_t* status,
                        int32_t* iters);

/* Device time of each IPM launch of the last solve (HIP events on the launch
 * stream): one value for the persistent schedule, one per iteration with
 * per_iteration.  Writes up to `cap` values; returns how many exist. */
int qtqp_batch_launch_times(qtqp_batch* h, double* ms, int32_t cap);

/* Completion time of each QP in the last solve, in creation order: ms of the
 * device wall clock from the start of the first IPM launch to the end of the
 * QP's last iteration (NaN if it did not run).  Writes up to `cap` values;
 * returns how many exist.  No reference counterpart (batch diagnostics). */
int qtqp_batch_instance_ms(qtqp_batch* h, double* ms, int32_t cap);

/* Per-iteration trace of QP q for the last solve: 24 doubles per IPM
 * iteration (mu, sigma, alpha, tau, tau_predictor, tau_corrector,
 * alpha_predictor, then solves and final residual of the q / predictor /
 * corrector KKT solves (solves + 1000 when stalled), pres, dres, gap) --
 * the scalars of the reference's per-iteration stats (__init__.py:967-990)
 * -- then, for the predictor and the corrector, the coefficients a, b, c of
 * the tau quadratic and max |K^{-1} r|.  Copies min(rows, 128) rows. */
int qtqp_batch_trace(qtqp_batch* h, int32_t q, double* out, int32_t rows);

/* Executed work of the last solve per QP, 5 counts each: FP32 and FP64
 * factorisations, refinement passes (one sweep of A each: the rest of a
 * correction, the FP64 K_reg residual and the next right-hand side), FP32
 * and FP64 triangular factor solves. */
int qtqp_batch_work(qtqp_batch* h, int32_t* counts);

/* Tile shapes of the Gram product G = P + diag + A'WA, the MFMA-heavy part
 * of each factorisation.  Each tile row of G is computed pass by pass; a
 * pass reads the tile row's strip of A once for all its tiles.  More tiles
 * per pass read A fewer times but hold more accumulators in registers;
 * which wins depends on n and the precision. */
#define QTQP_BATCH_TILE_64 0   /* 64 x 64 MFMA tiles, two per pass (the default) */
#define QTQP_BATCH_TILE_128 1  /* 128 x 128 MFMA tiles (FP32 only) */
#define QTQP_BATCH_TILE_64X4 2 /* 64 x 64 tiles, four per pass */
#define QTQP_BATCH_TILE_64X1 3 /* 64 x 64 tiles, one per pass */
#define QTQP_BATCH_NTILES 4
/* Autotuner size buckets by n: n <= 64, 128, 192, 256, 384, 512, 768, 1024. */
#define QTQP_BATCH_NBUCKETS 8

/* Tile autotuner (no reference counterpart).  For every size bucket that
 * holds QPs of this batch and for each precision (FP32, then FP64), times
 * `iters` Gram products of up to `sample` of its QPs (spread over the
 * bucket) in each tile shape, keeps the fastest and applies it to every QP
 * of the bucket.  table (2 x QTQP_BATCH_NBUCKETS: the FP32 choices, then the
 * FP64 ones; -1 for empty buckets) and ms (2 x QTQP_BATCH_NBUCKETS x
 * QTQP_BATCH_NTILES, milliseconds, NaN where not measured) may be NULL.
 * opts is accepted for the solve options of later calls and not used here.
 * Discards a start set with qtqp_batch_set_start.  Tile shapes change no
 * result: each G entry sums the same products in the same order. */
int qtqp_batch_autotune(qtqp_batch* h, const qtqp_batch_opts* opts, int32_t sample,
                        int32_t iters, int32_t* table, double* ms);

/* Applies a tile table (2 x QTQP_BATCH_NBUCKETS: one QTQP_BATCH_TILE_* per
 * bucket for the FP32 Gram product, then for the FP64 one, which may not be
 * QTQP_BATCH_TILE_128), e.g. a saved autotune result. */
int qtqp_batch_set_tiles(qtqp_batch* h, const int32_t* table);

void qtqp_batch_free(qtqp_batch* h);

/* Message of the calling thread's last error ("" if none). */
const char* qtqp_batch_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QTQP_BATCH_H_ */
