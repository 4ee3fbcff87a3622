/* C ABI of the device-resident interior-point driver for one dense QP
 * (MI355X, gfx950).
 *
 * Runs qtqp's Mehrotra predictor-corrector method on the homogeneous
 * embedding with every per-iteration vector operation, reduction and scalar
 * decision on the GPU.  The problem data stays in HBM for the life of the
 * handle.  The host calls qtqp_ipm_iterate once per IPM iteration; each call
 * returns a small record of that iteration's statistics and status.  No
 * N-vector crosses PCIe inside the loop.  The KKT factorisation, solves and
 * refinement are those of include/qtqp_kkt.h (same kernels, same precision
 * modes).
 *
 * Reference code replaced (qtqp v0.0.5, paths relative to its checkout):
 *   src/qtqp/__init__.py:616-669   QTQP._equilibrate, _(un)equilibrate_iterates
 *   src/qtqp/__init__.py:299-307   QTQP._init_variables (default start)
 *   src/qtqp/__init__.py:473-578   one pass of the iteration loop of _solve_impl
 *   src/qtqp/__init__.py:671-1008  step size, sigma, Newton step, tau,
 *                                  normalisation, termination
 *   src/qtqp/direct.py:183-307     DirectKktSolver.update / solve
 *
 * Host pointers are caller-owned C-contiguous float64 arrays, copied during
 * the call.  Calls are synchronous on return.  A handle is not thread-safe.
 * Return codes and qtqp_last_error() are those of include/qtqp_kkt.h.
 */
#ifndef QTQP_IPM_H_
#define QTQP_IPM_H_

#include <stdint.h>

#include "qtqp_kkt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qtqp_ipm qtqp_ipm;

/* Solver options; the keyword arguments of QTQP.solve (__init__.py:309-320). */
typedef struct qtqp_ipm_opts {
  double atol, rtol;                 /* optimality tolerances */
  double atol_infeas, rtol_infeas;   /* certificate tolerances */
  double step_size_scale;            /* fraction of the step to the boundary */
  double min_static_regularization;  /* floor of the KKT diagonal */
  double linear_solver_atol, linear_solver_rtol;  /* refinement tolerances */
  int32_t max_iterative_refinement_steps;
  int32_t collect_stats;             /* nonzero: fill the collect_stats fields */
  int32_t use_graph;    /* nonzero: replay the iteration from a HIP graph with the
                           refinement and precision decisions made on the device
                           (0: host-driven launches, one readback per refinement step) */
  int32_t graph_steps;  /* refinement steps per refined solve held in the graph (0:
                           1 with an FP64 factor, 2 with FP32, one more after each
                           iteration that needed more, which is redone host-driven) */
} qtqp_ipm_opts;

/* Status after an iteration (SolutionStatus, __init__.py:129-137). */
#define QTQP_IPM_UNFINISHED 0
#define QTQP_IPM_SOLVED 1
#define QTQP_IPM_INFEASIBLE 2
#define QTQP_IPM_UNBOUNDED 3

/* One iteration's statistics: the keys _check_termination stores
 * (__init__.py:967-1008) and the three refined solves' stats
 * (direct.py:300-307, tau_method from __init__.py:741-752). */
typedef struct qtqp_ipm_record {
  int32_t status;         /* QTQP_IPM_* */
  int32_t tau_method[2];  /* predictor, corrector: 1 quadratic, 0 linearized */
  int32_t reserved;
  qtqp_ir_stats lin[3];   /* kinv_q, predictor, corrector solves */
  double ctx, bty, pcost, dcost, pres, dres, gap, pinfeas, dinfeas, dinfeas_a, dinfeas_p;
  double mu, sigma, alpha, tau, norm_x, norm_y, prelrhs, drelrhs;
  /* filled when opts.collect_stats is nonzero */
  double complementarity, norm_s, max_sy, min_sy, std_sy;
  double max_s_over_y, min_s_over_y, mean_s_over_y, std_s_over_y;
} qtqp_ipm_record;

/* Uploads one QP (after presolve):  min 1/2 x'Px + c'x  s.t.  Ax + s = b,
 * s[:z] = 0, s[z:] >= 0.  `a` is the dense m x n A and `p` the dense
 * symmetric n x n P (NULL when P = 0), both row-major, or both column-major
 * when col_major is nonzero (the layout scipy's CSC toarray() returns; the
 * device transposes them).  has_p says whether P has stored entries (the
 * reference's p.nnz > 0).  With equilibrate, runs the 10 Ruiz passes of
 * _equilibrate on the device.  Sets the default start.  precision is
 * QTQP_PREC_FP64 or QTQP_PREC_MIXED. */
int qtqp_ipm_create(int64_t n, int64_t m, int64_t z, const double* a, const double* p,
                    int32_t col_major, int32_t has_p, const double* b, const double* c,
                    int32_t precision, int32_t equilibrate, qtqp_ipm** out);

/* Runs one IPM iteration (__init__.py:473-566) and its termination check
 * (888-1008); fills *rec.  With opts->use_graph the iteration is one launch
 * of a captured HIP graph and one readback of the record (the graph is
 * captured on the first call and again when the options change).  Returns QTQP_ERR_FACTOR on a breakdown of the FP64
 * factor and QTQP_ERR_NAN when a refined solve has NaNs, as the reference's
 * backends raise LinAlgError and ValueError. */
int qtqp_ipm_iterate(qtqp_ipm* h, const qtqp_ipm_opts* opts, qtqp_ipm_record* rec);

/* The current iterate, unequilibrated (_unequilibrate_iterates,
 * __init__.py:665-666): x (n), y (m), s (m) and tau. */
int qtqp_ipm_get_iterates(qtqp_ipm* h, double* x, double* y, double* s, double* tau);

/* Engine counters of the handle's KKT engine. */
int qtqp_ipm_get_info(qtqp_ipm* h, qtqp_kkt_info* info);

/* Releases the handle; NULL is accepted. */
void qtqp_ipm_free(qtqp_ipm* h);

#ifdef __cplusplus
}
#endif

#endif /* QTQP_IPM_H_ */
