/* C ABI of the MI355X dense KKT engine for qtqp's interior-point method.
 *
 * Plain C types only (pointers, sizes, doubles).  All pointers are host
 * pointers to C-contiguous float64 arrays owned by the caller; the library
 * copies them during the call and never keeps them.  Every call is
 * synchronous on return.  A handle owns its device memory and one HIP stream
 * and is not thread-safe.  Functions return QTQP_OK (0) or an error code;
 * qtqp_last_error() gives the message of the calling thread's last failure.
 *
 * Reference interfaces replaced (qtqp v0.0.5, paths relative to its checkout):
 *   src/qtqp/direct.py:49-109          LinearSolver protocol (the plug-in API)
 *   src/qtqp/solvers_dense.py:25-145   ScipyDenseSolver (dense Gram backend)
 *   src/qtqp/direct.py:183-307         DirectKktSolver.update / .solve (IR)
 */
#ifndef QTQP_KKT_H_
#define QTQP_KKT_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QTQP_OK 0
#define QTQP_ERR_ARG 1     /* bad argument (maps to ValueError) */
#define QTQP_ERR_HIP 2     /* HIP runtime failure (maps to RuntimeError/OSError) */
#define QTQP_ERR_FACTOR 3  /* factorisation breakdown (maps to numpy.linalg.LinAlgError) */
#define QTQP_ERR_NAN 4     /* NaN in a solution (maps to ValueError) */
#define QTQP_ERR_STATE 5   /* call out of order, e.g. solve before factorize */

/* Factor precision of a handle. */
#define QTQP_PREC_FP64 0   /* plain: FP64 Gram SYRK, FP64 LDL', FP64 solves */
#define QTQP_PREC_MIXED 1  /* FP32 MFMA factor, FP64 refinement, FP64 fallback */

typedef struct qtqp_kkt qtqp_kkt;

/* Result of one refined KKT solve; mirrors the stats dict returned by
 * DirectKktSolver.solve (src/qtqp/direct.py:300-307). */
typedef struct qtqp_ir_stats {
  int32_t solves;   /* backend solves performed (1..max_steps) */
  int32_t status;   /* 0 converged, 1 non-converged, 2 stalled */
  double final_residual_norm;
  double rhs_norm;
  double tolerance;
} qtqp_ir_stats;

/* Engine counters (no reference counterpart; diagnostics for benchmarks). */
typedef struct qtqp_kkt_info {
  int32_t precision;          /* QTQP_PREC_* the handle was created with */
  int32_t factor_precision;   /* precision of the current factor: 64 or 32 */
  int64_t factorizations;     /* factorize() calls */
  int64_t fp64_fallbacks;     /* FP32 factors replaced by an FP64 refactor */
  int64_t backend_solves;     /* K_reg^{-1} applications */
  int64_t inner_steps;        /* FP32 refinement steps inside backend solves */
  double factor_ms;           /* device time in factorisations (HIP events) */
  double solve_ms;            /* device time in solves and matvecs */
  int64_t jitter_retries;     /* FP64 factors redone with a larger diagonal shift */
  int64_t graph_redos;        /* device-loop iterations redone host-driven (qtqp_ipm) */
} qtqp_kkt_info;

/* Creates a handle for an (n+m)x(n+m) KKT system.  Replaces
 * LinearSolver.set_dims + set_kkt (direct.py:57-73) as specialised by
 * ScipyDenseSolver.set_dims/set_kkt (solvers_dense.py:68-93): `a` is the dense
 * m x n constraint block (row-major), `p` the dense symmetric n x n cost
 * block (row-major; its diagonal is ignored because update_diag supplies the
 * whole KKT diagonal). */
int qtqp_kkt_create(int64_t n, int64_t m, const double* a, const double* p,
                    int32_t precision, qtqp_kkt** out);

/* Sets the regularised KKT diagonal, y-block negated (N = n+m values).
 * Replaces LinearSolver.update_diag (direct.py:75-83, solvers_dense.py:95-97). */
int qtqp_kkt_update_diag(qtqp_kkt* h, const double* diag);

/* Factorises K_reg.  Replaces LinearSolver.factorize (direct.py:85-92,
 * solvers_dense.py:99-116).  Returns QTQP_ERR_FACTOR on breakdown of the
 * FP64 factor (the FP32 factor falls back to FP64 instead of failing). */
int qtqp_kkt_factorize(qtqp_kkt* h);

/* out = K_reg^{-1} rhs to FP64 accuracy.  Replaces LinearSolver.solve
 * (direct.py:94-96, solvers_dense.py:130-142). */
int qtqp_kkt_solve(qtqp_kkt* h, const double* rhs, double* out);

/* out = K_reg x in FP64.  Replaces LinearSolver.__matmul__ (direct.py:98-102,
 * solvers_dense.py:118-128). */
int qtqp_kkt_matvec(qtqp_kkt* h, const double* x, double* out);

/* out = K_reg x, as qtqp_kkt_matvec, on caller-owned device memory and with
 * no allocation of any kind.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc).  a: m x n row-major A, poff: n x n
 * row-major P with a zero diagonal (NULL when P = 0), diag: the N = n+m KKT
 * diagonal (y-block negated), x and out: N values; all device pointers.  The
 * kernels are enqueued on `stream` (a hipStream_t; NULL is the null stream)
 * and the call returns without waiting: out is ready when that stream's work
 * is.  Their partial sums live in a fixed 16 MiB scratch array per GPU, held
 * until the process ends; a size whose partials need more,
 *   (n + m) * ceil(n / 256) + ceil(m / 64) * n + n  >  2^21 doubles,
 * is rejected with QTQP_ERR_ARG (n = 4000, m = 6000 needs 0.54e6).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (two streams,
 * two threads) race and must be serialised by the caller. */
int qtqp_kkt_matvec_noalloc(int64_t n, int64_t m, const double* a, const double* poff,
                            const double* diag, const double* x, double* out, void* stream);

/* Stores diag_correction = reg - true (N values), used by solve_ir.
 * Replaces the host-side buffer of DirectKktSolver.update (direct.py:216). */
int qtqp_kkt_set_correction(qtqp_kkt* h, const double* corr);

/* Device-resident DirectKktSolver.solve (direct.py:218-307): negates the
 * y-block of rhs, refines sol from `warm` against K_true = K_reg - diag(corr)
 * until ||res||_inf < atol + rtol*||rhs||_inf, a stall, or max_steps backend
 * solves.  Returns QTQP_ERR_NAN if the solution has NaNs. */
int qtqp_kkt_solve_ir(qtqp_kkt* h, const double* rhs, const double* warm,
                      int32_t max_steps, double atol, double rtol, double* sol,
                      qtqp_ir_stats* stats);

/* Engine counters. */
int qtqp_kkt_get_info(qtqp_kkt* h, qtqp_kkt_info* info);

/* Releases the handle; NULL is accepted.  Replaces LinearSolver.free
 * (direct.py:108-109, idempotence as solvers_gpu.py:109-115). */
void qtqp_kkt_free(qtqp_kkt* h);

/* Message of the calling thread's last error ("" if none). */
const char* qtqp_last_error(void);

/* Number of visible HIP devices (0 when no GPU is usable). */
int qtqp_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* QTQP_KKT_H_ */
