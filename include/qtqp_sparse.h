/* C ABI of the MI355X sparse KKT engine: supernodal LDL' of the quasi-definite
 * KKT matrix of qtqp's interior-point method, with the reference's sparse
 * problems in mind (A and P stored sparse, the factor sparse).
 *
 * Plain C types only.  Index and value arrays are host pointers to
 * C-contiguous int64 / float64 arrays owned by the caller; the library copies
 * them during the call and never keeps them.  Every call is synchronous on
 * return.  A handle owns its device memory and one HIP stream and is not
 * thread-safe.  Functions return the QTQP_* codes of include/qtqp_kkt.h;
 * qtqp_sparse_last_error() gives the message of the calling thread's last
 * failure.
 *
 * Reference interfaces replaced (qtqp v0.0.5, paths relative to its checkout):
 *   src/qtqp/direct.py:49-109            LinearSolver protocol (the plug-in API)
 *   src/qtqp/solvers_gpu.py:27-115       CuDssSolver (sparse LDL' on a GPU)
 *   src/qtqp/solvers_sparse.py:74-93     QdldlSolver (sparse quasi-definite LDL')
 */
#ifndef QTQP_SPARSE_H_
#define QTQP_SPARSE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Storage of the upper triangle handed to create/analyze. */
#define QTQP_SPARSE_CSC 0  /* column pointers; row indices <= column */
#define QTQP_SPARSE_CSR 1  /* row pointers; column indices >= row */

typedef struct qtqp_sparse qtqp_sparse;

/* Symbolic statistics and counters (no reference counterpart). */
typedef struct qtqp_sparse_info {
  int64_t n;                /* dimension N = n + m of the KKT matrix */
  int64_t nnz_kkt;          /* stored entries of the upper triangle of K */
  int64_t nnz_l;            /* entries of L, diagonal included */
  int64_t supernodes;       /* supernodes after splitting to at most 64 columns */
  int64_t levels;           /* levels of the supernodal elimination tree */
  int64_t max_rows;         /* rows of the tallest supernode */
  int64_t update_tiles;     /* 64x64 Schur-update tiles per factorisation */
  int64_t factorizations;   /* factorize() calls */
  int64_t solves;           /* solve() calls */
  double factor_flops;      /* flops of one numeric factorisation */
  double analyze_ms;        /* host time of the symbolic analysis */
  double factor_ms;         /* device time in factorisations (HIP events) */
  double solve_ms;          /* device time in solves and matvecs */
  int64_t phases;           /* launches of Schur updates: levels split into classes of
                               supernodes that share no ancestor row (deterministic
                               assembly without atomics) */
} qtqp_sparse_info;

/* Symbolic analysis only, on the host (no device is touched): fill-reducing
 * ordering, elimination tree, supernodes, level schedule.  perm_out (n
 * entries, may be NULL) receives the elimination order: perm_out[k] is the
 * original index of the k-th eliminated row/column.  Counterpart of the
 * ordering/analysis step of CuDssSolver.factorize (solvers_gpu.py:64-91,
 * `DirectSolver.plan`). */
int qtqp_sparse_analyze(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                        int64_t* perm_out, qtqp_sparse_info* info);

/* qtqp_sparse_analyze with a choice of fill-reducing ordering: 0 (what
 * qtqp_sparse_analyze and create use) exact minimum degree up to N = 40000 and
 * approximate minimum degree on a quotient graph above; 1 exact minimum degree
 * (approximate above N = 40000, where the explicit elimination graph would not
 * fit); 2 approximate minimum degree at every size.  A factor whose panels
 * would exceed 64 GiB is refused with QTQP_ERR_ARG.  No reference counterpart
 * (the reference's sparse backends choose their orderings internally). */
int qtqp_sparse_analyze_ex(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                           int32_t ordering, int64_t* perm_out, qtqp_sparse_info* info);

/* Creates a handle for the N x N KKT matrix given by its upper triangle
 * (ptr[N+1], ind[nnz], val[nnz] in `format`).  Runs the symbolic analysis
 * and keeps the pattern, the off-diagonal values and the stored diagonal on
 * the device.  Replaces LinearSolver.set_kkt (direct.py:65-73) as
 * specialised by CuDssSolver.set_kkt (solvers_gpu.py:52-57). */
int qtqp_sparse_create(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                       const double* val, qtqp_sparse** out);

/* Replaces the diagonal of K (diag[N], original order; the y-block already
 * negated by the caller).  Replaces LinearSolver.update_diag
 * (direct.py:75-77; solvers_gpu.py:59-62). */
int qtqp_sparse_update_diag(qtqp_sparse* h, const double* diag);

/* Numeric supernodal LDL' of the current K on the device.  Returns
 * QTQP_ERR_FACTOR on a zero or non-finite pivot (the caller raises
 * numpy.linalg.LinAlgError).  Replaces CuDssSolver.factorize
 * (solvers_gpu.py:64-91) and QdldlSolver.factorize (solvers_sparse.py:83-87). */
int qtqp_sparse_factorize(qtqp_sparse* h);

/* out = K^{-1} rhs with the current factor (both N-vectors, original order).
 * Replaces CuDssSolver.solve (solvers_gpu.py:101-104). */
int qtqp_sparse_solve(qtqp_sparse* h, const double* rhs, double* out);

/* out = K x, the symmetric product with the current diagonal.  Replaces
 * CuDssSolver.__matmul__ (solvers_gpu.py:93-99) and the base product
 * LinearSolver.__matmul__ (direct.py:98-102). */
int qtqp_sparse_matvec(qtqp_sparse* h, const double* x, double* out);

int qtqp_sparse_get_info(qtqp_sparse* h, qtqp_sparse_info* info);

/* Releases the handle (NULL is a no-op).  Replaces CuDssSolver.free
 * (solvers_gpu.py:109-115). */
void qtqp_sparse_free(qtqp_sparse* h);

const char* qtqp_sparse_last_error(void);

/* Number of visible HIP devices (0 when none). */
int qtqp_sparse_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* QTQP_SPARSE_H_ */
