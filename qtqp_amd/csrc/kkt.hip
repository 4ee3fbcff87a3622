// LinearSolver backend of the single-instance dense KKT engine: the C ABI of
// include/qtqp_kkt.h over the kernels and helpers of kkt_engine.h.
#include <mutex>

#include "kkt_engine.h"

using namespace qk;

thread_local std::string g_qtqp_last_error;

namespace {

// Scratch of qtqp_kkt_matvec_noalloc: the row partials of A x and P_off x and
// the column partials and total of A'y.  A file-scope device array, so the
// entry point allocates nothing; every GPU holds its own copy until the
// process ends, and its address is taken once per device with
// hipGetSymbolAddress (the array's name is not a device address on the host).
constexpr int64_t kNoAllocScratchDoubles = (int64_t)1 << 21;  // 16 MiB

}  // namespace

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array.
int scratch_ptr(double** out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(QTQP_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return QTQP_OK;
}

// Doubles of scratch a product with K of this size takes: row partials of
// A x (m per 256-column tile) and of P_off x (n per tile), column partials of
// A'y (n per 32-row tile), and A'y itself.
int64_t noalloc_scratch_doubles(int64_t n, int64_t m) {
  return (m + n) * tiles_c(n) + (int64_t)tiles_rx(m) * n + n;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* qtqp_last_error(void) { return g_qtqp_last_error.c_str(); }

int qtqp_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int qtqp_kkt_create(int64_t n, int64_t m, const double* a, const double* p, int32_t precision,
                    qtqp_kkt** out) {
  return engine_create(n, m, a, p, precision, true, out);
}

int qtqp_kkt_update_diag(qtqp_kkt* h, const double* diag) {
  if (!h || !diag) return fail(QTQP_ERR_ARG, "NULL argument");
  const int64_t N = h->n + h->m;
  for (int64_t i = 0; i < h->n; ++i)
    if (!(diag[i] > 0.0)) return fail(QTQP_ERR_ARG, "x-block diagonal must be positive");
  for (int64_t i = h->n; i < N; ++i)
    if (!(diag[i] < 0.0)) return fail(QTQP_ERR_ARG, "y-block diagonal must be negative");
  HIPC(hipMemcpyAsync(h->diag, diag, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  h->have_diag = true;
  h->have_factor = false;
  return QTQP_OK;
}

int qtqp_kkt_set_correction(qtqp_kkt* h, const double* corr) {
  if (!h || !corr) return fail(QTQP_ERR_ARG, "NULL argument");
  HIPC(hipMemcpyAsync(h->corr, corr, sizeof(double) * (h->n + h->m), hipMemcpyHostToDevice,
                      h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return QTQP_OK;
}

int qtqp_kkt_factorize(qtqp_kkt* h) {
  if (!h) return fail(QTQP_ERR_ARG, "NULL handle");
  if (!h->have_diag) return fail(QTQP_ERR_STATE, "factorize called before update_diag");
  return engine_factorize(h);
}

int qtqp_kkt_solve(qtqp_kkt* h, const double* rhs, double* out) {
  if (!h || !rhs || !out) return fail(QTQP_ERR_ARG, "NULL argument");
  if (!h->have_factor) return fail(QTQP_ERR_STATE, "solve called before factorize");
  const int64_t N = h->n + h->m;
  HIPC(hipEventRecord(h->ev0, h->stream));
  HIPC(hipMemcpyAsync(h->vin, rhs, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
  int rc = backend_solve(h, h->vin, h->vout);
  if (rc) return rc;
  HIPC(hipMemcpyAsync(out, h->vout, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  h->info_acc.solve_ms += elapsed_ms(h);
  return QTQP_OK;
}

int qtqp_kkt_matvec(qtqp_kkt* h, const double* x, double* out) {
  if (!h || !x || !out) return fail(QTQP_ERR_ARG, "NULL argument");
  if (!h->have_diag) return fail(QTQP_ERR_STATE, "matvec called before update_diag");
  const int64_t N = h->n + h->m;
  HIPC(hipMemcpyAsync(h->vin, x, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
  int rc = kmatvec(h, h->vin, nullptr, nullptr, h->vout, nullptr);
  if (rc) return rc;
  HIPC(hipMemcpyAsync(out, h->vout, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return QTQP_OK;
}

int qtqp_kkt_matvec_noalloc(int64_t n, int64_t m, const double* a, const double* poff,
                            const double* diag, const double* x, double* out, void* stream) {
  if (!a || !diag || !x || !out) return fail(QTQP_ERR_ARG, "NULL argument");
  if (n < 1 || m < 1 || n > (int64_t)1 << 30 || m > (int64_t)1 << 30)
    return fail(QTQP_ERR_ARG, "n and m must be in [1, 2^30]");
  const int64_t need = noalloc_scratch_doubles(n, m);
  if (need > kNoAllocScratchDoubles) {
    char buf[200];
    snprintf(buf, sizeof(buf),
             "problem too large for the fixed scratch: needs %lld doubles, the maximum is %lld "
             "(16 MiB)", (long long)need, (long long)kNoAllocScratchDoubles);
    return fail(QTQP_ERR_ARG, buf);
  }
  double* sc = nullptr;
  int rc = scratch_ptr(&sc);
  if (rc) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  double* rpa = sc;                             // m * tiles_c(n)
  double* rpp = rpa + m * tiles_c(n);           // n * tiles_c(n)
  double* cpart = rpp + n * tiles_c(n);         // tiles_rx(m) * n
  if ((rc = tile_pass_on<double>(st, cpart, a, m, n, x, rpa, x + n, nullptr))) return rc;
  if (poff && (rc = tile_pass_on<double>(st, cpart, poff, n, n, x, rpp, nullptr, nullptr)))
    return rc;
  hipLaunchKernelGGL(k_kmv_finish, dim3(blocks_for(n + m)), dim3(WG), 0, st, (int)n, (int)m, x,
                     diag, poff ? rpp : nullptr, rpa, tiles_c(n), cpart, tiles_rx(m), nullptr,
                     nullptr, nullptr, nullptr, out, nullptr, nullptr, 0);
  HIPC(hipGetLastError());
  return QTQP_OK;
}

int qtqp_kkt_solve_ir(qtqp_kkt* h, const double* rhs, const double* warm, int32_t max_steps,
                      double atol, double rtol, double* sol, qtqp_ir_stats* stats) {
  if (!h || !rhs || !warm || !sol) return fail(QTQP_ERR_ARG, "NULL argument");
  if (max_steps < 1) return fail(QTQP_ERR_ARG, "max_steps must be >= 1");
  if (!h->have_factor) return fail(QTQP_ERR_STATE, "solve_ir called before factorize");
  const int64_t N = h->n + h->m;
  HIPC(hipEventRecord(h->ev0, h->stream));
  HIPC(hipMemcpyAsync(h->rin, rhs, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->vsol, warm, sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
  const int rc = ir_refine(h, h->rin, h->vsol, max_steps, atol, rtol, stats);
  if (rc != QTQP_OK && rc != QTQP_ERR_NAN) return rc;
  HIPC(hipMemcpyAsync(sol, h->vsol, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  h->info_acc.solve_ms += elapsed_ms(h);
  return rc;
}

int qtqp_kkt_get_info(qtqp_kkt* h, qtqp_kkt_info* info) {
  if (!h || !info) return fail(QTQP_ERR_ARG, "NULL argument");
  *info = h->info_acc;
  info->factor_precision = h->fprec;
  return QTQP_OK;
}

void qtqp_kkt_free(qtqp_kkt* h) {
  if (!h) return;
  release(h);
  delete h;
}

}  // extern "C"
