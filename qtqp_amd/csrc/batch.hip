// Batched device interior-point engine for many uneven dense QPs (gfx950).
//
// Layout in HBM (one allocation per array kind, instances packed back to
// back in input order, every per-instance array 256-byte aligned):
//   raw A (m x n row-major, as uploaded), the equilibrated A in FP64 (with
//   read slack after it) and in FP32 (for the FP32 factor path), raw and
//   equilibrated P (n x n) in FP64 (with slack) and FP32, vectors b, c, d,
//   e, the Gram/factor workspace G (n x n FP64, whose front half holds the
//   FP32 factor), iterates and scratch vectors.
// Every pass over A is one coalesced, software-pipelined sweep of its rows
// that yields both A x (row dots, wavefront reductions) and A' u (column
// partials per lane, combined across the four wavefronts in LDS).
//
// Execution: k_prepare (Ruiz scaling, FP32 copies, start) once per solve,
// then either one persistent k_ipm_solve launch that runs every QP to
// completion, or per IPM iteration the pair {k_reset_queue, k_ipm_iter}
// replayed from a HIP graph.  Both kernels, and k_prepare, are persistent
// grids: each workgroup takes the next QP from an atomic queue (QPs sorted
// by descending m*n^2, so large ones start first) and every workgroup exits
// when the queue is empty.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "dense_ops.h"
#include "qtqp_batch.h"
#include "qtqp_kkt.h"

using namespace qk;

namespace {

thread_local std::string g_err;
constexpr double kEps = 1e-15;  // qtqp _EPS (__init__.py:51)
constexpr int kMaxInner = 12;   // FP32 refinement steps inside one backend solve
constexpr int kStatusError = 5; // factorisation breakdown in FP64 (LinAlgError)

enum { S_TAU = 0, S_MU, S_SIGMA, S_ALPHA, S_PRES, S_DRES, S_GAP, S_NSCAL = 8 };
// Per-iteration trace row: the scalars of the reference's per-iteration stats.
enum { T_MU = 0, T_SIGMA, T_ALPHA, T_TAU, T_TAU_P, T_TAU_C, T_ALPHA_P, T_Q_SOLVES, T_Q_RES,
       T_P_SOLVES, T_P_RES, T_C_SOLVES, T_C_RES, T_PRES, T_DRES, T_GAP, T_DBG0,
       kTraceFields = 24 };
constexpr int kTraceRows = 128;
enum { I_STATUS = 0, I_ITERS, I_FORCE64, I_FALLBACK, I_F32, I_F64, I_KMV, I_FS32, I_FS64,
       I_NINT = 12 };

// Rows of A that are exact negations of each other (a two-sided constraint
// l <= a'x <= u posed as a'x <= u and -a'x <= -l, or a bound pair) are
// stored once: stored row r stands for row p(r) and, if it has one, its
// partner q(r) with A[q] = -A[p].  Every pass over A then reads the stored
// rows only.  The map is a few segments of consecutive rows: in segment k
// (stored rows [start[k], start[k+1])) p = r + dp[k] and q = r + dq[k], or
// no partner when dq[k] == kNoPartner.
constexpr int kMaxSeg = 4;
constexpr int kNoPartner = -(1 << 30);
struct PairMap {
  int nseg;
  int start[kMaxSeg + 1];
  int dp[kMaxSeg], dq[kMaxSeg];
};

struct Inst {
  int n, m, z, has_p;
  int mc;       // stored rows of A (m when no row has a partner)
  PairMap pm;
  const uint32_t* ext;  // nonzero column extent lo | hi << 16 of each stored row
  const uint32_t* cext; // the same for each 32-row chunk of stored rows (their union)
  int ld;    // leading dimension of A, A32, P, P32: n rounded up to 4
  int tile;    // FP32 Gram tile shape (QTQP_BATCH_TILE_*), chosen per size bucket
  int tile64;  // FP64 Gram tile shape (not QTQP_BATCH_TILE_128)
  // raw data
  const double *Araw, *Praw, *braw, *craw;
  // equilibrated data
  double *A, *P, *b, *c, *d, *e;
  float *A32, *P32;
  double normb, normc;
  // workspace
  double* G;
  double* LiB;  // per 64-block of the factor: L_bb^{-1} column-major then row-major
  double *x, *y, *s, *kq, *dg, *corr, *krhs, *sol, *res, *iacc, *ires, *r, *xp;
  double *um, *un, *ds, *cm;
  double* pkq;  // P kq_x of this iteration's kinv_q
  // per constraint row, for this iteration: wy = 1 / d_y (d_y the
  // regularised diagonal) and ey = d_y - corr_y = -(s/y + mu), the true one
  double *wy, *ey;
  // products kept for the termination test: A x (stored rows), A'y and
  // P_off x of the iterate at the start of the iteration, of kinv_q and of
  // the corrector's solution, each [m | n | n]
  double* tk;
  double* irb;  // refinement state, two copies: A x (m), A'y, P_off x, g (n each)
  // a second refined solve run together with the first (solve_ir2): its
  // right-hand side, solution, KKT right-hand side, refinement state,
  // residuals, correction and A'(W r_y)
  double *r2, *sol2, *krhs2, *irb2, *res2, *ires2, *iacc2, *um2;
  double* sd;
  int* si;
  double* trace;  // kTraceFields doubles per IPM iteration (max_iter rows)
};

struct Opts {
  double atol, rtol, atol_inf, rtol_inf, step_scale, min_reg, ls_atol, ls_rtol;
  int max_iter, max_ir, precision, equilibrate;
  int fuse;  // the q-solve and the predictor's refine together (solve_ir2)
};

struct Counters {
  unsigned long long backend_solves, inner_steps, fallbacks, ir_bad, jitter_retries;
  int done;
  int max_iters;  // most IPM iterations any QP took
};

struct Smem {
  unsigned char* u;  // union scratch: {S, Li, sinv} or {si, sj}
  double* v;         // triangular-solve vector (n values of T)
  double* red;       // WG + 8 doubles
};

__device__ inline Smem carve(double* base, int nmax) {
  Smem s;
  s.u = reinterpret_cast<unsigned char*>(base);
  s.v = base + (2 * NB * (NB + 1) + NB);
  s.red = s.v + ((nmax + 1) & ~1);
  return s;
}

// The IPM kernel's LDS: one static array (77 KB, 2 workgroups per CU) laid
// out as {union scratch of two 64 x 65 tiles + 64 | vector of n <= 1024 |
// reduction slots}.  Being a named LDS object, every __noinline__ phase can
// re-derive its pointers from it, so their accesses compile to LDS
// instructions: a pointer that arrives through a reference is generic and
// would be accessed with flat instructions, whose completion is counted
// with the global loads (an LDS wait would then drain every prefetch).
constexpr int kLdsU = 2 * NB * (NB + 1) + NB;  // doubles
constexpr int kLdsV = 1024;
constexpr int kLdsRed = WG + 8;
__shared__ double g_lds[kLdsU + kLdsV + kLdsRed];
// The per-QP ints (status, iterations, FP64 switch, work counts) and scalars
// (tau, mu, ...) of the QP a workgroup of the IPM kernels holds: loaded from
// Inst::si / Inst::sd when it takes the QP and written back when it lets it
// go (qp_take / qp_release), so the iteration reads and counts them in LDS
// instead of waiting on a global round trip each time.
__shared__ int g_si[I_NINT];
__shared__ double g_sd[S_NSCAL];
// The batch counters a refinement step bumps (backend solves, FP32 inner
// steps, solves that ended unconverged), counted in LDS while the workgroup
// holds the QP and added to the launch's Counters when it lets the QP go: a
// global atomic per step would make the step's next wait for loads wait for
// the atomic too.
enum { C_SOLVES = 0, C_INNER, C_IRBAD, C_NCNT };
__shared__ unsigned g_cnt[C_NCNT];

__device__ inline void qp_take(const Inst& I) {
  if (threadIdx.x < I_NINT) g_si[threadIdx.x] = I.si[threadIdx.x];
  if (threadIdx.x < S_NSCAL) g_sd[threadIdx.x] = I.sd[threadIdx.x];
  if (threadIdx.x < C_NCNT) g_cnt[threadIdx.x] = 0u;
  __syncthreads();
}
__device__ inline void qp_release(const Inst& I, Counters* cnt) {
  __syncthreads();
  if (threadIdx.x < I_NINT) I.si[threadIdx.x] = g_si[threadIdx.x];
  if (threadIdx.x < S_NSCAL) I.sd[threadIdx.x] = g_sd[threadIdx.x];
  if (threadIdx.x == 0) {
    if (g_cnt[C_SOLVES]) atomicAdd(&cnt->backend_solves, (unsigned long long)g_cnt[C_SOLVES]);
    if (g_cnt[C_INNER]) atomicAdd(&cnt->inner_steps, (unsigned long long)g_cnt[C_INNER]);
    if (g_cnt[C_IRBAD]) atomicAdd(&cnt->ir_bad, (unsigned long long)g_cnt[C_IRBAD]);
  }
  __syncthreads();
}

__device__ inline Smem lds_smem() {
  Smem s;
  s.u = reinterpret_cast<unsigned char*>(g_lds);
  s.v = g_lds + kLdsU;
  s.red = g_lds + kLdsU + kLdsV;
  return s;
}

__device__ inline Smem lds_view(const Smem&) { return lds_smem(); }

// Per-phase device clocks of the IPM kernels, compiled in with
// -DQTQP_PHASE_CLOCKS only: PHASE(k) adds the wall-clock ticks (100 MHz)
// since the previous mark of this workgroup to slot k of g_phase, PHASE(-1)
// restarts the interval.  Each workgroup's thread 0 stands for it
// (every phase ends in a barrier or is uniform).  Read and cleared with
// qtqp_batch_phase_clocks.
#ifdef QTQP_PHASE_CLOCKS
constexpr int kPhaseSlots = 32;
__device__ unsigned long long g_phase[kPhaseSlots];
__shared__ unsigned long long g_phase_t;
__device__ inline void phase_mark(int k) {
  if (threadIdx.x == 0) {
    const unsigned long long t = wall_clock64();
    if (k >= 0) atomicAdd(&g_phase[k], t - g_phase_t);
    g_phase_t = t;
  }
}
#define PHASE(k) phase_mark(k)
#else
#define PHASE(k) ((void)0)
#endif

// The segment table of a QP's row pairs, in scalar registers.  A row's
// offsets are formed by adding the steps between consecutive segments
// selected by comparisons: choosing among the fields themselves lets the
// compiler turn the choice into a load at a chosen address, which keeps the
// table in scratch memory, and each such load waits for every load in
// flight.
struct PairRegs {
  int s1, s2, s3;
  int dp0, dp1, dp2, dp3, dq0, dq1, dq2, dq3;
  __device__ inline explicit PairRegs(const PairMap& pm) {
    static_assert(kMaxSeg == 4, "four segments");
    s1 = wave_uniform(pm.start[1]);
    s2 = wave_uniform(pm.start[2]);
    s3 = wave_uniform(pm.start[3]);
    dp0 = wave_uniform(pm.dp[0]);
    dp1 = wave_uniform(pm.dp[1]);
    dp2 = wave_uniform(pm.dp[2]);
    dp3 = wave_uniform(pm.dp[3]);
    dq0 = wave_uniform(pm.dq[0]);
    dq1 = wave_uniform(pm.dq[1]);
    dq2 = wave_uniform(pm.dq[2]);
    dq3 = wave_uniform(pm.dq[3]);
  }
  // row p(r) and its partner q(r) (-1: none); segments past nseg are empty
  // (their start is the stored row count), so r picks the last one that
  // starts at or before it
  __device__ inline void rows(int r, int* p, int* q) const {
    const bool g1 = r >= s1, g2 = r >= s2, g3 = r >= s3;
    // (steps between kNoPartner and an offset stay within int: |offset| < 2^29)
    const int a = dp0 + (g1 ? dp1 - dp0 : 0) + (g2 ? dp2 - dp1 : 0) + (g3 ? dp3 - dp2 : 0);
    const int b = dq0 + (g1 ? dq1 - dq0 : 0) + (g2 ? dq2 - dq1 : 0) + (g3 ? dq3 - dq2 : 0);
    *p = r + a;
    *q = b == kNoPartner ? -1 : r + b;
  }
};

template <typename T>
__device__ inline T (*tile_ptr(unsigned char* p, int row_off))[NB + 1] {
  return reinterpret_cast<T (*)[NB + 1]>(p) + row_off;
}

// -------------------------------------------------------------------------
// Block-level vector kernels (every thread of the workgroup calls; each ends
// with a barrier so results are visible to the whole workgroup).
// -------------------------------------------------------------------------


template <typename T, int W> struct VecT;
template <> struct VecT<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct VecT<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecT<uint32_t, 2> { typedef uint32_t type __attribute__((ext_vector_type(2))); };
template <> struct VecT<uint32_t, 4> { typedef uint32_t type __attribute__((ext_vector_type(4))); };

// Buffer resource over `bytes` bytes at p (raw, unswizzled).  A buffer access
// at an offset past the end reads zero or writes nothing, without a memory
// request.
__device__ inline __amdgpu_buffer_rsrc_t buffer_of(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
// Stores v at byte offset off of r when ok, else nothing: the store is
// issued either way, so no branch forms around it (whose join would make
// later waits drain every load in flight).
template <typename T>
__device__ inline void store_if(__amdgpu_buffer_rsrc_t r, bool ok, uint32_t off, T v) {
  const uint32_t o = ok ? off : 0x80000000u;
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(VecT<uint32_t, 2>::type, v), r, o, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, o, 0, 0);
}

template <typename T> struct VW;
template <> struct VW<double> { static constexpr int value = 2; };
template <> struct VW<float> { static constexpr int value = 4; };

__device__ inline double bsum(double v, double* red) { return block_sum(v, red + WG); }
__device__ inline double bmax(double v, double* red) { return block_max(v, red + WG); }
__device__ inline double bmin(double v, double* red) { return -block_max(-v, red + WG); }

// A pointwise loop over k = threadIdx.x, + WG, ... < len, U iterations at a
// time: ld(k) issues iteration k's loads (through global-qualified pointers;
// k is clamped to len - 1, so every load is unconditional), the loads of all
// U iterations go out before the first is used, then body(k, loaded) runs
// for the iterations in range in the order of the plain loop, so sums
// accumulate exactly as there.  The plain loop over pointers reached through
// the QP's table waits for each iteration's loads (flat accesses) before it
// issues the next iteration's.
template <int K>
struct Dv {
  double v[K];
};
template <int U = 4, typename L, typename B>
__device__ inline void vpass(int len, L ld, B body) {
  using T = decltype(ld(0));
  for (int k0 = threadIdx.x; k0 < len; k0 += U * WG) {
    T t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = ld(min(k0 + u * WG, len - 1));
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u * WG < len) body(k0 + u * WG, t[u]);
  }
}

__device__ inline double dot(const double* a, const double* b, int len, double* red) {
  double s = 0.0;
  vpass(len, [&](int i) { return Dv<2>{{ldg(a + i), ldg(b + i)}}; },
        [&](int, const Dv<2>& t) { s += t.v[0] * t.v[1]; });
  return bsum(s, red);
}

// -------------------------------------------------------------------------
// Factorisation of G = P_off + diag(d_x) + A' W A (one workgroup).
// -------------------------------------------------------------------------
// NJ 64 x 64 Gram tiles of one tile row, G[i0 block, j0[t] block] +=
// A[:, i0 block]' W A[:, j0[t] block], streamed together so the row strip
// of A is read once for all of them.  DIAG: the last tile is the diagonal
// one (j0[NJ-1] == i0), whose second strip is the row strip itself, so it
// costs no loads of its own.  Rounds of KCH rows are staged in LDS and kept
// two rounds ahead in registers, with unconditional (clamped) loads, as in
// RankKTileG::run_vec; each tile accumulates exactly as a RankKTileG<T, 1, 1>
// tile would.  lds: (1 + NJ) KCH x 65 elements of T.
//  With cext (the nonzero column extent of each 32-row chunk of X, packed
// lo | hi << 16, 0 for a zero chunk) only the rounds whose rows reach into
// both the row strip and the tiles' column strips are read: their indices
// are listed in `list` (LDS, `cap` ints) first.  The rounds left out add
// only exact zeros, so every tile gets the sums it would get from all rows.
template <typename T, int NJ, bool DIAG, int KCH_ = RankKTileG<T, 1, 1>::KCH>
struct GramRow {
  using acc_t = typename Mfma<T>::acc_t;
  static constexpr int KCH = KCH_;
  static_assert(KCH == 16 || KCH % 32 == 0, "rounds of 16 rows or whole 32-row chunks");
  static constexpr int VW = 16 / sizeof(T);
  static constexpr int VI = NB / VW;    // vectors per staged row of a strip
  static constexpr int RI = WG / VI;    // rows per pass of the workgroup
  static constexpr int PI = KCH / RI;   // vectors per thread per strip and round
  static constexpr int NLD = NJ - (DIAG ? 1 : 0);  // strips loaded besides the row strip
  static constexpr int DEPTH = NLD > 1 ? 1 : 2;     // rounds in flight (registers)
  typedef T VT __attribute__((ext_vector_type(VW)));
  acc_t acc[NJ][2][2];

  __device__ inline void run(const T* __restrict__ X, int ldx, int kend, int i0,
                             const int (&j0)[NJ], const double* __restrict__ w, T* lds,
                             const uint32_t* __restrict__ cext = nullptr, int* list = nullptr,
                             int cap = 0) {
    X = wave_uniform(X);
    w = wave_uniform(w);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // rounds to read: all (cnt = nround, list unused) or the listed ones
    const int nround = (kend + KCH - 1) / KCH;
    int cnt = nround;
    const bool use_list = cext != nullptr && nround <= cap;
    if (use_list) {
      // 32-row chunks per round (a 16-row round lies in one chunk)
      constexpr int SUB = KCH >= 32 ? KCH / 32 : 1;
      const int jlo = j0[0], jhi = j0[0] + NB * NJ;
      if (wave == 0) {
        int base = 0;
        for (int c0 = 0; c0 < nround; c0 += 64) {
          const int c = c0 + lane;
          bool on = false;
          if (c < nround) {
#pragma unroll
            for (int u = 0; u < SUB; ++u) {
              const uint32_t e = ldg(cext + (KCH >= 32 ? SUB * c + u : c * KCH / 32));
              const int lo = e & 0xffff, hi = e >> 16;
              on = on || (lo < i0 + NB && hi > i0 && lo < jhi && hi > jlo);
            }
          }
          const uint64_t mask = __ballot(on);
          if (on) list[base + __popcll(mask & ((1ull << lane) - 1))] = c;
          base += __popcll(mask);
        }
        if (lane == 0) list[cap] = base;
      }
      __syncthreads();
      cnt = list[cap];
      __syncthreads();
      if (cnt == 0) {
#pragma unroll
        for (int t = 0; t < NJ; ++t)
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
              for (int r = 0; r < 4; ++r) acc[t][a][b][r] = T(0);
        return;
      }
    }
    // first row of the round at list position p (clamped to the last one)
    auto round_row = [&](int p) {
      p = min(p, cnt - 1);
      return (use_list ? list[p] : p) * KCH;
    };
    const int jb = 32 * (wave >> 1), ib = 32 * (wave & 1);
    const int vi = tid % VI, ri = tid / VI;
#pragma unroll
    for (int t = 0; t < NJ; ++t)
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[t][a][b][r] = T(0);
    // columns past ldx read as the last vector and are masked when staged
    const int coli = min(i0 + vi * VW, ldx - VW);
    const bool ci = i0 + vi * VW < ldx;
    int colj[NLD > 0 ? NLD : 1];
    bool cj[NLD > 0 ? NLD : 1];
#pragma unroll
    for (int t = 0; t < NLD; ++t) {
      colj[t] = min(j0[t] + vi * VW, ldx - VW);
      cj[t] = j0[t] + vi * VW < ldx;
    }
    struct Round {
      VT a[PI], b[NLD > 0 ? NLD : 1][PI];
      double w[PI];
    };
    auto fetch = [&](int pos, Round& rd) {
      const int kc = round_row(pos);
#pragma unroll
      for (int q = 0; q < PI; ++q) {
        const int k = min(kc + ri + RI * q, kend - 1);
        const T* row = X + (int64_t)k * ldx;
        rd.a[q] = ldg(reinterpret_cast<const VT*>(row + coli));
#pragma unroll
        for (int t = 0; t < NLD; ++t) rd.b[t][q] = ldg(reinterpret_cast<const VT*>(row + colj[t]));
        rd.w[q] = ldg(w + k);
      }
    };
    T* si = lds;
    auto sj = [&](int t) { return lds + (1 + t) * KCH * (NB + 1); };
    auto step = [&](int pos, Round& rd, bool active) {
      const int kc = round_row(pos);
      if (active) {
#pragma unroll
        for (int q = 0; q < PI; ++q) {
          const bool kin = kc + ri + RI * q < kend;
          const VT va = (kin && ci) ? rd.a[q] : VT(0);
          const T wk = (T)rd.w[q];
#pragma unroll
          for (int e = 0; e < VW; ++e) si[(ri + RI * q) * (NB + 1) + vi * VW + e] = va[e];
#pragma unroll
          for (int t = 0; t < NJ; ++t) {
            VT vb;
            if (DIAG && t == NJ - 1)
              vb = (kin && ci) ? rd.a[q] * wk : VT(0);
            else
              vb = (kin && cj[t]) ? rd.b[t][q] * wk : VT(0);
#pragma unroll
            for (int e = 0; e < VW; ++e) sj(t)[(ri + RI * q) * (NB + 1) + vi * VW + e] = vb[e];
          }
        }
      }
      __syncthreads();
      fetch(pos + DEPTH, rd);
      if (active) {
#pragma unroll
        for (int kk = 0; kk < KCH; kk += 4) {
          const int kq = kk + (lane >> 4), cl = lane & 15;
          T bv[2];
#pragma unroll
          for (int b = 0; b < 2; ++b) bv[b] = si[kq * (NB + 1) + ib + 16 * b + cl];
#pragma unroll
          for (int t = 0; t < NJ; ++t) {
            T av[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = sj(t)[kq * (NB + 1) + jb + 16 * a + cl];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int b = 0; b < 2; ++b) acc[t][a][b] = Mfma<T>::mac(av[a], bv[b], acc[t][a][b]);
          }
        }
      }
      __syncthreads();
    };
    if (kend <= 0) return;
    wait_mem();
    Round rd[DEPTH];
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) {
      fetch(d, rd[d]);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int p0 = 0; p0 < cnt; p0 += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) step(p0 + d, rd[d], p0 + d < cnt);
    }
  }

  // f(i, j, value) for every accumulated element of tile t of this thread
  template <typename F>
  __device__ inline void for_each(int t, int i0, int j0, F f) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jb = j0 + 32 * (wave >> 1), ib = i0 + 32 * (wave & 1);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          f(ib + 16 * b + (lane & 15), jb + 16 * a + Mfma<T>::row(lane, r), acc[t][a][b][r]);
  }
};

// G tiles (i0, j0[t] = j00 + 64 t) = A' W A + (P_off or diag(d_x)) for one
// GramRow pass (a function of its own, so each pass shape gets its own
// registers).  DIAG: the last tile is the diagonal one (j00 + 64 (NJ-1) = i0).
template <typename T, int NJ, bool DIAG, int KCH>
__device__ __noinline__ void gram_pass(const Inst& I, const T* A, const T* P, T* G,
                                       const double* w, int i0, int j00, T* lds_in) {
  T* lds = in_lds(lds_in);
  const int n = I.n, m = I.mc, ld = I.ld;  // the stored rows of A
  int j0[NJ];
#pragma unroll
  for (int t = 0; t < NJ; ++t) j0[t] = j00 + t * NB;
  using G_ = GramRow<T, NJ, DIAG, KCH>;
  G_ g;
  // the round list goes after the staging rows in the LDS union
  constexpr int kUsed = (1 + NJ) * G_::KCH * (NB + 1) * (int)sizeof(T);
  constexpr int kCap = (kLdsU * (int)sizeof(double) - kUsed) / (int)sizeof(int) - 1;
  static_assert(kCap >= 256, "room for the round list");
  g.run(A, ld, m, i0, j0, w, lds, I.cext,
        in_lds(reinterpret_cast<int*>(reinterpret_cast<char*>(lds) + kUsed)), kCap);
  // P_off and the diagonal d_x of every element, loaded before any is used
  // (clamped addresses, unconditional loads: one memory latency per group
  // of EG tiles; EG divides NJ, so every group is whole)
  const __amdgpu_buffer_rsrc_t gr = buffer_of(G, n * n * (int)sizeof(T));
  constexpr int EG = NJ % 2 == 0 ? 2 : 1;
  static_assert(NJ % EG == 0, "whole groups of tiles");
#pragma unroll
  for (int t0 = 0; t0 < NJ; t0 += EG) {
    T add[EG][16];
#pragma unroll
    for (int u = 0; u < EG; ++u) {
      int e = 0;
      g.for_each(t0 + u, i0, j0[t0 + u], [&](int i, int j, T) {
        const int ic = min(i, n - 1), jc = min(j, n - 1);
        const T pv = ldg(P + (int64_t)jc * ld + ic);
        const T dv = (T)ldg(I.dg + ic);
        add[u][e++] = ic == jc ? dv : pv;
      });
    }
#pragma unroll
    for (int u = 0; u < EG; ++u) {
      int e = 0;
      g.for_each(t0 + u, i0, j0[t0 + u], [&](int i, int j, T v) {
        const T a = add[u][e++];
        store_if(gr, i < n && j < n && i >= j, (uint32_t)((j * n + i) * (int)sizeof(T)), v + a);
      });
    }
  }
}

// The lower tiles of one tile row, NJ off-diagonal tiles per pass and the
// rest with the diagonal tile in a last pass.
template <typename T, int NJ, int KCH>
__device__ inline void gram_rows(const Inst& I, const T* A, const T* P, T* G, const double* w,
                                 T* lds) {
  static_assert((1 + NJ) * KCH * (NB + 1) * sizeof(T) <= kLdsU * sizeof(double),
                "GramRow staging fits the LDS union");
  const int nt = (I.n + NB - 1) / NB;
  for (int ti = 0; ti < nt; ++ti) {
    const int i0 = ti * NB;
    int tj = 0;
    for (; tj + NJ <= ti; tj += NJ) gram_pass<T, NJ, false, KCH>(I, A, P, G, w, i0, tj * NB, lds);
    const int r = ti - tj;  // 0 .. NJ - 1 tiles left before the diagonal
    if (r == 0) gram_pass<T, 1, true, KCH>(I, A, P, G, w, i0, i0, lds);
    if constexpr (NJ > 1)
      if (r == 1) gram_pass<T, 2, true, KCH>(I, A, P, G, w, i0, tj * NB, lds);
    if constexpr (NJ > 2)
      if (r == 2) gram_pass<T, 3, true, KCH>(I, A, P, G, w, i0, tj * NB, lds);
    if constexpr (NJ > 3)
      if (r == 3) gram_pass<T, 4, true, KCH>(I, A, P, G, w, i0, tj * NB, lds);
  }
}

// Lower tiles of the Gram block G = P_off + diag(d_x) + A' W A in the tile
// shape `tile` (QTQP_BATCH_TILE_*):
//  - 64 x 64 tiles a tile row at a time (GramRow), one, two or four per pass
//    on a shared row strip, the diagonal tile reusing the row strip: a tile
//    row of r + 1 tiles reads about (1 + 1/NJ)(r + 1) strips of A instead of
//    2 (r + 1), at the cost of NJ tiles of accumulators in registers;
//  - 128 x 128 rank-m MFMA tiles (RankKTileG; FP32 only, the FP64
//    accumulators would not fit in registers).
// Each G entry sums the same products in the same order in every shape.
template <typename T>
__device__ inline void gram_tiles(const Inst& I, const T* A, const T* P, T* G, const double* w,
                                  Smem& sm, int tile) {
  const int n = I.n, m = I.mc, ld = I.ld;  // the stored rows of A
  T* lds = reinterpret_cast<T*>(sm.u);
  constexpr int K1 = RankKTileG<T, 1, 1>::KCH;  // 32 rows (FP64), 64 (FP32)
  if (tile == QTQP_BATCH_TILE_64X1) {
    gram_rows<T, 1, K1>(I, A, P, G, w, lds);
  } else if (tile == QTQP_BATCH_TILE_64X4) {
    gram_rows<T, 4, K1 / 2>(I, A, P, G, w, lds);
  } else if (sizeof(T) == 4 && tile == QTQP_BATCH_TILE_128) {
    constexpr int MT = 2, TS = 64 * MT;
    const int nt = (n + TS - 1) / TS;
    using Tile = RankKTileG<T, MT, MT>;
    T* si = lds;
    T* sj = si + Tile::KCH * (TS + 1);
    for (int t = 0; t < nt * (nt + 1) / 2; ++t) {
      int ti, tj;
      lower_tile(t, &ti, &tj);
      const int i0 = ti * TS, j0 = tj * TS;
      Tile acc;
      acc.zero();
      // rows of A are 16-byte aligned and zero-padded to ld (a multiple of 4)
      acc.run_vec(A, ld, 0, m, i0, j0, w, [](double v) { return (T)v; }, si, sj);
      acc.for_each(i0, j0, [&](int i, int j, T v) {
        if (i < n && j < n && i >= j)
          stg(G + (int64_t)j * n + i,
              v + ((i == j) ? (T)ldg(I.dg + i) : ldg(P + (int64_t)j * ld + i)));
      });
    }
  } else {
    gram_rows<T, 2, K1>(I, A, P, G, w, lds);
  }
}

template <typename T>
__device__ __noinline__ void gram_syrk(const Inst& I, const T* A, const T* P, T* G, Smem& sm_in,
                                       double jit_rel = 1e-14) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, mc = I.mc;
  // weights of the stored rows: W = -1/d_y of the row plus its partner's
  // (A[q] = -A[p], so both add w a a' with the same a)
  {
    const PairRegs pr(I.pm);
    const double* dy = I.dg + n;
    vpass(mc,
          [&](int r) {
            int p, q;
            pr.rows(r, &p, &q);
            return Dv<2>{{ldg(dy + p), ldg(dy + (q >= 0 ? q : p))}};
          },
          [&](int r, const Dv<2>& in) {
            int p, q;
            pr.rows(r, &p, &q);
            double w = -1.0 / in.v[0];
            if (q >= 0) w += -1.0 / in.v[1];
            stg(I.um + r, w);
          });
  }
  __syncthreads();
  gram_tiles<T>(I, A, P, G, I.um, sm, sizeof(T) == 4 ? I.tile : I.tile64);
  __syncthreads();
  // diag(G) += jit_rel * max(diag(G)), jit_rel = 1e-14 (solvers_dense.py:107-112)
  // or larger on a retry after a breakdown
  double mx = -INFINITY;
  vpass(n, [&](int i) { return (double)ldg(G + (int64_t)i * n + i); },
        [&](int, double g) { mx = fmax(mx, g); });
  const double jit = jit_rel * bmax(mx, sm.red);
  vpass(n, [&](int i) { return (double)ldg(G + (int64_t)i * n + i); },
        [&](int i, double g) { stg(G + (int64_t)i * n + i, (T)(g + jit)); });
  __syncthreads();
}

// Blocked right-looking LDL' (Cholesky form) in place; returns breakdown
// column (1-based) or 0.
template <typename T>
__device__ __noinline__ int chol(T* G, int n, T* LiB, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  __shared__ int info;
  if (threadIdx.x == 0) info = 0;
  __syncthreads();
  T (*S)[NB + 1] = tile_ptr<T>(sm.u, 0);
  T (*Li)[NB + 1] = tile_ptr<T>(sm.u, NB);
  T* sinv = reinterpret_cast<T*>(tile_ptr<T>(sm.u, 2 * NB));
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = min(NB, n - k0);
    load_diag_block(G, n, k0, nb, S);
    __syncthreads();
    potrf_inv_blk(S, Li, sinv, nb, &info, k0);
    store_diag_block(G, n, k0, nb, S);
    // keep L_bb^{-1} for the triangular solves: column-major, then row-major
    T* dst = LiB + (int64_t)(k0 / NB) * 2 * NB * NB;
    for (int idx = threadIdx.x; idx < NB * NB; idx += WG) {
      dst[idx] = Li[idx & (NB - 1)][idx >> 6];
      dst[NB * NB + idx] = Li[idx >> 6][idx & (NB - 1)];
    }
    PHASE(16);
    if (k0 + NB < n) {
      for (int rb = k0 + NB; rb < n; rb += WG) trsm_mfma_rows(G, n, n, k0, nb, rb, Li);
      __syncthreads();
      PHASE(17);
      const int t0 = k0 + NB;
      const int tt = (n - t0 + NB - 1) / NB;
      T (*si)[NB + 1] = tile_ptr<T>(sm.u, 0);
      T (*sj)[NB + 1] = tile_ptr<T>(sm.u, KC);
      for (int t = 0; t < tt * (tt + 1) / 2; ++t) {
        int ti, tj;
        lower_tile(t, &ti, &tj);
        const int i0 = t0 + ti * NB, j0 = t0 + tj * NB;
        RankKTile<T> acc;
        acc.zero();
        acc.run(G + (int64_t)k0 * n, n, 0, NB, i0, j0, n, n, [](int) { return T(1); }, si, sj);
        // the tile's G values, loaded before any is updated (clamped,
        // unconditional: one memory latency per tile)
        T gv[16];
        int e = 0;
        acc.for_each(i0, j0, [&](int i, int j, T) {
          gv[e++] = ldg(G + (int64_t)min(j, n - 1) * n + min(i, n - 1));
        });
        const __amdgpu_buffer_rsrc_t gr = buffer_of(G, n * n * (int)sizeof(T));
        e = 0;
        acc.for_each(i0, j0, [&](int i, int j, T v) {
          const T gij = gv[e++];
          store_if(gr, i < n && j < n && i >= j, (uint32_t)((j * n + i) * (int)sizeof(T)), gij - v);
        });
      }
    }
    __syncthreads();
    PHASE(18);
  }
  const int r = info;
  __syncthreads();
  return r;
}

// Loads of the triangular sweeps below.  None of them depends on v, so
// they are issued ahead of use: a block's inverse while the previous block
// is solved, each 64-row chunk of a block's column panel while the chunk
// before it is consumed.  Every load is unconditional (indices clamped into
// the block storage or into L, out-of-range terms skipped where they are
// used), which lets the waits count them in issue order.
//
// Li16: the 16 entries X[(p + 4q) * NB + lane] (q < 16) of a 64 x 64 block X.
template <typename T>
__device__ inline void load_blk16(const T* X, T (&d)[16]) {
  const int lane = threadIdx.x & 63, p = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 16; ++q) d[q] = ldg(X + (p + 4 * q) * NB + lane);
}

// Panel16: L[col][i] for the columns col = j0 + 16p + cc (cc < 16) of block
// j0 (clamped below j0 + nb) and the row i (clamped below n) of column-major L.
template <typename T>
__device__ inline void load_panel16(const T* L, int n, int j0, int nb, int i, T (&d)[16]) {
  const int p = threadIdx.x >> 6;
  const T* col = L + min(i, n - 1);
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) d[cc] = ldg(col + (int64_t)min(j0 + 16 * p + cc, j0 + nb - 1) * n);
}

// L v = v in place (L column-major n x n, lower) with the stored inverses of
// its 64 x 64 diagonal blocks.  Per block: v_J = Li_J v_J (the four waves
// split the 64 columns, LDS reduce), then v_i -= L[i][J] v_J for the rows
// below, each wave taking 16 of the block's columns for 64 rows at a time
// (loads along a column, contiguous in i) and the four partials meeting in
// `part` (>= 4 * n doubles of LDS).
template <typename T>
__device__ void trsv_lower_inv(const T* L, int n, const T* LiB, T* v, double* red, double* part) {
  const int tid = threadIdx.x, lane = tid & 63, p = tid >> 6;
  L = wave_uniform(L);
  LiB = wave_uniform(LiB);
  n = wave_uniform(n);
  const int nblk = (n + NB - 1) / NB;
  wait_mem();
  T li[16], pa[16], pb[16];
  load_blk16(LiB, li);  // LiT[k*64 + i] = Li[i][k] of block 0
  for (int b = 0; b < nblk; ++b) {
    const int j0 = b * NB, nb = min(NB, n - j0);
    const int i0 = j0 + nb, rows = n - i0;
    load_panel16(L, n, j0, nb, i0 + lane, pa);
    __builtin_amdgcn_sched_barrier(0);
    double acc = 0.0;
    if (lane < nb) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (p + 4 * q < nb) acc += (double)li[q] * (double)v[j0 + p + 4 * q];
    }
    red[p * NB + lane] = acc;
    __syncthreads();
    if (tid < nb) v[j0 + tid] = (T)((red[tid] + red[NB + tid]) + (red[2 * NB + tid] + red[3 * NB + tid]));
    __syncthreads();
    load_blk16(LiB + (int64_t)min(b + 1, nblk - 1) * 2 * NB * NB, li);
    __builtin_amdgcn_sched_barrier(0);
    if (rows <= 0) break;
    const int c0 = 16 * p;
    auto chunk = [&](int ib, const T (&pv)[16]) {
      const int i = i0 + ib + lane;
      double a = 0.0;
      if (i < n) {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          if (c0 + cc < nb) a += (double)pv[cc] * (double)v[j0 + c0 + cc];
      }
      part[p * n + ib + lane] = a;
    };
    // Whole pairs of chunks, the second skipped (not its refill) past the
    // end, so the loop head is reached with the same loads in flight.
    for (int ib = 0; ib < rows; ib += 128) {
      load_panel16(L, n, j0, nb, i0 + ib + 64 + lane, pb);
      __builtin_amdgcn_sched_barrier(0);
      chunk(ib, pa);
      load_panel16(L, n, j0, nb, i0 + ib + 128 + lane, pa);
      __builtin_amdgcn_sched_barrier(0);
      if (ib + 64 < rows) chunk(ib + 64, pb);
    }
    __syncthreads();
    for (int r = tid; r < rows; r += WG)
      v[i0 + r] = (T)((double)v[i0 + r] -
                      ((part[r] + part[n + r]) + (part[2 * n + r] + part[3 * n + r])));
    __syncthreads();
  }
}

// L' v = v in place with stored diagonal-block inverses.  Per block (last
// first): rb = v_J - L[below][J]' v_below, read down the block's columns (wave
// p takes columns 16p .. 16p+15, lanes consecutive rows, completed across
// the lanes by a transpose-reduce), then v_J = Li_J' rb.  rb: 64 doubles.
template <typename T>
__device__ void trsv_upper_inv(const T* L, int n, const T* LiB, T* v, double* red, double* rb) {
  const int tid = threadIdx.x, lane = tid & 63, p = tid >> 6;
  L = wave_uniform(L);
  LiB = wave_uniform(LiB);
  n = wave_uniform(n);
  const int nblk = (n + NB - 1) / NB;
  wait_mem();
  T li[16], pa[16], pb[16];
  // Li[k*64 + c] (row-major) of the last block
  load_blk16(LiB + (int64_t)(nblk - 1) * 2 * NB * NB + NB * NB, li);
  for (int b = nblk - 1; b >= 0; --b) {
    const int j0 = b * NB, nb = min(NB, n - j0);
    const int c0 = 16 * p;
    double acc[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) acc[cc] = 0.0;
    auto chunk = [&](int i, const T (&pv)[16]) {
      if (i < n) {
        const double vi = (double)v[i];
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          if (c0 + cc < nb) acc[cc] += (double)pv[cc] * vi;
      }
    };
    load_panel16(L, n, j0, nb, j0 + nb + lane, pa);
    __builtin_amdgcn_sched_barrier(0);
    for (int i = j0 + nb; i < n; i += 128) {  // whole pairs of chunks, as in the forward sweep
      load_panel16(L, n, j0, nb, i + 64 + lane, pb);
      __builtin_amdgcn_sched_barrier(0);
      chunk(i + lane, pa);
      load_panel16(L, n, j0, nb, i + 128 + lane, pa);
      __builtin_amdgcn_sched_barrier(0);
      chunk(i + 64 + lane, pb);
    }
    reduce_rows<16>(acc);  // lane l: column c0 + (l >> 2)
    if ((lane & 3) == 0) red[c0 + (lane >> 2)] = acc[0];
    __syncthreads();
    if (tid < nb) rb[tid] = (double)v[j0 + tid] - red[tid];
    __syncthreads();
    double a2 = 0.0;
    if (lane < nb) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (p + 4 * q < nb) a2 += (double)li[q] * rb[p + 4 * q];
    }
    load_blk16(LiB + (int64_t)max(b - 1, 0) * 2 * NB * NB + NB * NB, li);
    __builtin_amdgcn_sched_barrier(0);
    red[p * NB + lane] = a2;
    __syncthreads();
    if (tid < nb) v[j0 + tid] = (T)((red[tid] + red[NB + tid]) + (red[2 * NB + tid] + red[3 * NB + tid]));
    __syncthreads();
  }
}

// FP64 factor with the reference's diagonal shift of 1e-14 max(diag(G))
// (solvers_dense.py:107-112), redone with 1e-12 and 1e-10 after a breakdown.
// Near the end of an IPM on an unbounded or infeasible problem G's condition
// number reaches 1e18 and the sign of its last pivots depends on rounding
// (the reference's LAPACK Gram and Cholesky can land on either side too); a
// larger shift keeps the factor usable and the refinement against the exact
// K removes its effect.  Returns false if all three break down.
__device__ bool factor64_retry(const Inst& I, Smem& sm_in, Counters* cnt) {
  Smem sm = lds_view(sm_in);
  const double rel[3] = {1e-14, 1e-12, 1e-10};
  for (int t = 0; t < 3; ++t) {
    gram_syrk<double>(I, I.A, I.P, I.G, sm, rel[t]);
    PHASE(1);
    const bool ok = chol<double>(I.G, I.n, I.LiB, sm) == 0;
    PHASE(2);
    if (ok) return true;
    if (t < 2 && threadIdx.x == 0) atomicAdd(&cnt->jitter_retries, 1ull);
  }
  return false;
}

// Factor in the instance's current precision; FP32 breakdown falls back to
// FP64.  Returns false on an FP64 breakdown.
__device__ __noinline__ bool factor(const Inst& I, Smem& sm_in, Counters* cnt) {
  Smem sm = lds_view(sm_in);
  if (!g_si[I_FORCE64]) {
    float* G = reinterpret_cast<float*>(I.G);
    if (threadIdx.x == 0) g_si[I_F32] += 1;
    gram_syrk<float>(I, I.A32, I.P32, G, sm);
    PHASE(1);
    const bool ok = chol<float>(G, I.n, reinterpret_cast<float*>(I.LiB), sm) == 0;
    PHASE(2);
    if (ok) return true;
    if (threadIdx.x == 0) {
      g_si[I_FORCE64] = 1;
      g_si[I_FALLBACK] += 1;
      atomicAdd(&cnt->fallbacks, 1ull);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) g_si[I_F64] += 1;
  return factor64_retry(I, sm, cnt);
}

// -------------------------------------------------------------------------
// KKT solves with iterative refinement, one pass over A per step.
// -------------------------------------------------------------------------

// Slack allocated after each instance's FP64 A and P (kPadRows rows of ld
// plus kPadCols values), so that the pipelined row pass below can issue its
// loads unconditionally.
constexpr int kPadRows = 16;
constexpr int kPadCols = 1024;
__host__ __device__ constexpr int64_t a_pad(int ld) { return (int64_t)kPadRows * ld + kPadCols; }

// Rows past the last row of M from which rowpass2<NS, R, D> may start a
// read: a stage starts at row nrow - 1 at the latest and reads R rows.
// Each read then spans the columns up to 128 NS <= kPadCols from its row
// start.
template <int NS, int R, int D>
constexpr int row_overrun() { return R - 1; }

// Rows a pipelined row pass takes between two flushes of its staged row
// outputs: a multiple of every stage rotation (D + 1) * 4R used below, and
// small enough for three doubles per row in the LDS scratch.
constexpr int kPassChunk = 1984;
// The same for passes that stage five outputs per row (a row pair's).
constexpr int kPairChunk = 1280;

// Row extents of one stage of R rows (R packed words) and the word of row r.
template <int R>
struct ExtVec;
template <>
struct ExtVec<2> {
  using type = VecT<uint32_t, 2>::type;
};
template <>
struct ExtVec<4> {
  using type = VecT<uint32_t, 4>::type;
};
struct Ext8 {
  VecT<uint32_t, 4>::type a, b;
};
template <>
struct ExtVec<8> {
  using type = Ext8;
};
template <int R>
__device__ inline typename ExtVec<R>::type load_ext(const uint32_t* p) {
  using U4 = VecT<uint32_t, 4>::type;
  if constexpr (R == 8)
    return Ext8{ldg(reinterpret_cast<const U4*>(p)), ldg(reinterpret_cast<const U4*>(p + 4))};
  else
    return ldg(reinterpret_cast<const typename ExtVec<R>::type*>(p));
}
template <int R>
__device__ inline uint32_t ext_word(const typename ExtVec<R>::type& e, int r) {
  if constexpr (R == 8)
    return r < 4 ? e.a[r] : e.b[r - 4];
  else
    return e[r];
}
// Extent words past the last stored row (read by the stages that prefetch
// beyond it): covers (2 D + 2) stages of 4 R rows for every pass shape.
constexpr int kExtPad = 256;

// One pass over the rows of row-major M (nrow x ncol, leading dimension ld,
// zero-padded to even; ld <= 128 NS): for every row i,
//   rowepi(i, M[i,:] . xs, in_i, &c1_i, &c2_i, out_i)   with in_i = rowload(i)
// then rowflush(i, out_i) and finally
//   colepi(j, sum_i M[i][j] c1_i, sum_i M[i][j] c2_i).
// A wavefront takes R rows at a time and keeps them in registers: it forms
// the dot products, completes them with reduce_rows, evaluates the epilogue
// (every lane holding a row's total computes it), broadcasts the
// coefficients and accumulates the column sums from the same registers.
// Loads are software-pipelined D stages deep: a stage is a wavefront's R
// rows together with the epilogue inputs of those rows, and the loads of
// stage k + D are issued before stage k is processed.  The compiler waits
// for a stage only as long as it can count the memory operations issued
// after it, so the loop body has none that it cannot count:
//  - every load is unconditional: a stage starts at row min(i0, nrow - 1),
//    rows past nrow and columns past ld are read from the slack that
//    follows M (a_pad) and masked to zero where they are used, and the
//    epilogue inputs are read at a clamped row;
//  - nothing is stored to global memory (a pending store makes every later
//    wait drain the whole queue): the NOUT row outputs go to LDS and are
//    written out by rowflush once per kPassChunk rows.
// COLS (0, 1 or 2) column sums are formed; with COLS = 1, c2 is ignored and
// colepi receives 0 for it, with COLS = 0 colepi is not called.  MAXM
// replaces the sums by maxima of absolute values: the row value is
// max_j |M[i][j]| xs[j] and the column value max_i |M[i][j]| c1_i (for
// non-negative xs and c1; equilibration).  NOUT may be 0.
// With EXT, ext[i] = lo | hi << 16 bounds the nonzero columns [lo, hi) of
// row i (0 past the last row): the row is read through a buffer resource
// and the 16-byte pieces outside its extent are given an out-of-range
// offset, so they read as zero without a memory request.  The values left
// out are exact zeros, so every dot product and column sum is the one the
// full row gives.  A stage's extents are loaded D stages ahead of its rows,
// before the loads of the stage then issued, so waiting for them never waits
// for rows not yet needed.
// With NX = 2 the pass forms the dot products with two vectors (xs, xs2) from
// the same rows: rowepi(i, t1, t2, in, c, out) with c[COLS], COLS <= 4
// column sums, colepi(j, s[COLS]); each product and sum is formed as the
// one-vector pass forms it.
// scratch: >= max(4 COLS ncol, NOUT CHUNK) doubles of LDS.
template <int NS, int R, int D, int NOUT, int COLS, bool MAXM = false, int CHUNK = kPassChunk,
          bool EXT = false, int NX = 1, typename RL, typename RE, typename RF, typename CE>
__device__ inline void rowpass2(const double* __restrict__ M, int nrow, int ncol, int ld,
                                const double* __restrict__ xs, RL rowload, RE rowepi, RF rowflush,
                                CE colepi, double* scratch,
                                const uint32_t* __restrict__ ext = nullptr,
                                const double* __restrict__ xs2 = nullptr) {
  using V = VecT<double, 2>::type;
  using In = decltype(rowload(0));
  constexpr int LR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr int STEP = 4 * R;  // rows a workgroup takes per stage
  static_assert(R == 2 || R == 4 || R == 8, "rows per wavefront");
  static_assert(CHUNK % ((D + 1) * STEP) == 0, "chunks hold whole stage rotations");
  static_assert(row_overrun<NS, R, D>() <= kPadRows && 128 * NS <= kPadCols, "slack after M");
  static_assert(3 * D * STEP + R <= kExtPad, "extent words read past the last row");
  static_assert(NX == 1 || NX == 2, "one or two vectors");
  static_assert(NX == 1 ? COLS <= 2 : (COLS <= 4 && !MAXM), "column sums");
  constexpr int NC = COLS > 0 ? COLS : 1;
  M = wave_uniform(M);
  nrow = wave_uniform(nrow);
  ld = wave_uniform(ld);
  const int lane = threadIdx.x & 63, wave = wave_uniform((int)(threadIdx.x >> 6));
  const int lrow = lane >> (6 - LR);  // row of the stage whose total this lane holds
  double xr[NX][NS][2], ca[NC][NS][2];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int col = s * 128 + lane * 2 + e;
      // (a flat load still pending in the loop would make every wait drain)
#pragma unroll
      for (int v = 0; v < NX; ++v) xr[v][s][e] = col < ncol ? ldg((v == 0 ? xs : xs2) + col) : 0.0;
#pragma unroll
      for (int q = 0; q < NC; ++q) ca[q][s][e] = 0.0;
    }
  using E = typename ExtVec<R>::type;
  struct Stage {
    V a[R][NS];
    In in;
    E e;  // extents of the rows of the stage D stages after this one (EXT)
  };
  // (unused without EXT)
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(M), 0, nrow * ld * 8, 0x00020000);
  auto load = [&](int i0, Stage& st, Stage& ahead) {
    if constexpr (EXT) {
      // the extents of rows i0 .. i0 + R - 1 arrived in st.e; those of the
      // stage D ahead go into its slot before this stage's rows are read
      const E e = st.e;
      ahead.e = load_ext<R>(ext + i0 + D * STEP);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t er = ext_word<R>(e, r);
        const int lo = er & 0xffff, hi = er >> 16;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const int c = s * 128 + lane * 2;
          const uint32_t off = (c + 2 > lo && c < hi)
                                   ? (uint32_t)(((i0 + r) * ld + c) * 8)
                                   : 0x80000000u;
          st.a[r][s] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0));
        }
      }
    } else {
      // A stage past the last row re-reads the rows from nrow - 1 on: they
      // are in cache, where fresh slack would cost HBM traffic.
      const double* base = M + (int64_t)min(i0, nrow - 1) * ld + lane * 2;
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int s = 0; s < NS; ++s)
          st.a[r][s] = ldg(reinterpret_cast<const V*>(base + (int64_t)r * ld + s * 128));
    }
    st.in = rowload(min(i0 + lrow, nrow - 1));
  };
  auto process = [&](int i0, int c0, const Stage& st) {
    // With EXT the loads already read zero wherever the mask below would
    // clear: a row past nrow has extent 0 (and lies past the buffer's end),
    // and every piece at or past column ncol <= ld lies outside its row's
    // extent.  Without EXT, rows past nrow and columns past ld hold slack.
    V a[R][NS];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int s = 0; s < NS; ++s)
        a[r][s] = (EXT || (i0 + r < nrow && s * 128 + lane * 2 < ld)) ? st.a[r][s] : V(0);
    double d[NX][R];
#pragma unroll
    for (int v = 0; v < NX; ++v)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        // (the first slot's term starts the sum: adding it to 0.0 would cost
        // an FP64 add per row and change nothing but the sign of a zero)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const double t = MAXM ? fmax(fabs(a[r][s][0]) * xr[v][s][0], fabs(a[r][s][1]) * xr[v][s][1])
                                : a[r][s][0] * xr[v][s][0] + a[r][s][1] * xr[v][s][1];
          d[v][r] = s == 0 ? t : (MAXM ? fmax(d[v][r], t) : d[v][r] + t);
        }
      }
#pragma unroll
    for (int v = 0; v < NX; ++v) reduce_rows<R, MAXM>(d[v]);
    const int row = i0 + lrow;
    double c[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) c[q] = 0.0;
    if (row < nrow) {
      double out[NOUT > 0 ? NOUT : 1];
      double unused = 0.0;
      if constexpr (NX == 1)
        rowepi(row, d[0][0], st.in, &c[0], NC > 1 ? &c[1] : &unused, out);
      else
        rowepi(row, d[0][0], d[NX - 1][0], st.in, c, out);
      if ((lane & ((64 >> LR) - 1)) == 0)
#pragma unroll
        for (int k = 0; k < NOUT; ++k) scratch[k * CHUNK + row - c0] = out[k];
    }
    if constexpr (COLS > 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        // row r's coefficients, from the first lane of its group
        double kc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) kc[q] = lane_value(c[q], r << (6 - LR));
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
          for (int e = 0; e < 2; ++e) {
            if constexpr (MAXM) {
              ca[0][s][e] = fmax(ca[0][s][e], fabs(a[r][s][e]) * kc[0]);
            } else {
#pragma unroll
              for (int q = 0; q < NC; ++q) ca[q][s][e] += a[r][s][e] * kc[q];
            }
          }
      }
    }
  };
  if (nrow > 0) {
    wait_mem();
    Stage st[D + 1];
    const int first = wave * R;
    if constexpr (EXT) {
#pragma unroll
      for (int k = 0; k < D; ++k) st[k].e = load_ext<R>(ext + first + k * STEP);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) load(first + k * STEP, st[k], st[(k + D) % (D + 1)]);
    for (int c0 = 0; c0 < nrow; c0 += CHUNK) {
      const int cend = min(nrow, c0 + CHUNK);
      // A full chunk holds whole rotations, so the next one starts at st[0].
      // Every stage of a rotation issues its loads, also past cend, so the
      // loads in flight at the loop head are the same whichever way it was
      // reached and the wait there covers one stage.
      for (int i0 = c0 + first; i0 < cend; i0 += (D + 1) * STEP) {
#pragma unroll
        for (int u = 0; u <= D; ++u) {
          const int i = i0 + u * STEP;
          load(i + D * STEP, st[(u + D) % (D + 1)], st[(u + 2 * D) % (D + 1)]);
          if (i < cend) process(i, c0, st[u]);
        }
      }
      if constexpr (NOUT > 0) {
        __syncthreads();
        for (int j = threadIdx.x; j < cend - c0; j += WG) {
          double out[NOUT];
#pragma unroll
          for (int k = 0; k < NOUT; ++k) out[k] = scratch[k * CHUNK + j];
          rowflush(c0 + j, out);
        }
        wait_vmem();
        __syncthreads();
      }
    }
  }
  if constexpr (COLS > 0) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int col = s * 128 + lane * 2 + e;
        if (col < ncol) {
#pragma unroll
          for (int q = 0; q < NC; ++q) scratch[(4 * q + wave) * ncol + col] = ca[q][s][e];
        }
      }
    __syncthreads();
    for (int j = threadIdx.x; j < ncol; j += WG) {
      double cs[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        const double* b = scratch + 4 * q * ncol + j;
        cs[q] = MAXM ? fmax(fmax(b[0], b[ncol]), fmax(b[2 * ncol], b[3 * ncol]))
                     : (b[0] + b[ncol]) + (b[2 * ncol] + b[3 * ncol]);
      }
      if constexpr (NX == 1)
        colepi(j, cs[0], NC > 1 ? cs[1] : 0.0);
      else
        colepi(j, cs);
    }
    __syncthreads();
  }
}

// Row-pass shapes by leading dimension: R rows x NS 128-column slots = 8
// vectors per stage, one stage in flight while the next is processed (with
// two, measured on the batched MPC workload, every pass ran slower).
template <int NS_, int R_, int D_>
struct PassShape {
  static constexpr int NS = NS_, R = R_, D = D_;
};

template <typename F>
__device__ inline auto with_pass_shape(int ld, F f) {
  if (ld <= 128) return f(PassShape<1, 8, 1>());
  if (ld <= 256) return f(PassShape<2, 4, 1>());
  if (ld <= 512) return f(PassShape<4, 2, 1>());
  return f(PassShape<8, 2, 1>());
}

// State of a refined solve for the current iterate sol: the products
// A sol_x (m), A' sol_y, P_off sol_x (n), the residual
// r = krhs - K_reg sol + corr .* sol (N) and the factor-solve right-hand
// side g = r_x + A'(W r_y), W = -1/d_y (n) of the next correction.
struct IrState {
  double *sol, *ax, *aty, *px, *r, *g;
};

// Fills S1 from S1.sol (init), or sets S1 = S0 advanced by the correction
// whose x-part is xc:  y_c = W (A x_c - r0_y), sol1 = sol0 + (x_c, y_c),
// with the products updated incrementally, for the KKT right-hand side
// krhs; um (n) receives A'(W r1_y).  One pass over A (both A x and A' y,
// A' u in it), one over P.  Returns ||r1||_inf (NaN reported as inf).
template <int NS, int R, int D>
__device__ __noinline__ double ir_pass(const Inst& I, const IrState S0, const IrState S1,
                                       const double* xc, bool init, const double* krhs,
                                       double* um, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, mc = I.mc;
  const PairRegs pr(I.pm);
  if (threadIdx.x == 0) g_si[I_KMV] += 1;
  double* scratch = reinterpret_cast<double*>(sm.u);
  const double* wy = wave_uniform(I.wy);
  const double* ky = wave_uniform(krhs + n);
  const double* ey = wave_uniform(I.ey);
  const double* r0y = wave_uniform(S0.r + n);
  const double* ax0 = wave_uniform(S0.ax);
  double* s1y = wave_uniform(S1.sol + n);
  double* s1ax = wave_uniform(S1.ax);
  double* s1ry = wave_uniform(S1.r + n);
  double mx = 0.0;
  // Inputs of a stored row r: its row p and its partner q (A[q] = -A[p];
  // without one the partner slots repeat p and are ignored).  The product
  // with A[q] is the negated product with A[p], so one dot product serves
  // both, and the column coefficients are c(p) - c(q).  ax (A x of the
  // refinement iterate) is kept per stored row.
  struct RowIn {
    double sy[2], r0[2], wy[2], ky[2], ey[2], ax0;
    int has_q;
  };
  // Row inputs are loaded whether or not `init` uses them (S0 holds valid
  // arrays in both modes), so no load depends on a branch.
  const double* sy0 = wave_uniform((init ? S1.sol : S0.sol) + n);
  static_assert(5 * kPairChunk <= kLdsU && 8 * 1024 <= kLdsU, "LDS scratch of the row pass");
  rowpass2<NS, R, D, 5, 2, false, kPairChunk, true>(
      I.A, mc, n, I.ld, init ? S1.sol : xc,
      [&](int r) {
        int p, q;
        pr.rows(r, &p, &q);
        RowIn in;
        in.has_q = q >= 0;
        const int qq = q >= 0 ? q : p;
        in.sy[0] = ldg(sy0 + p);
        in.sy[1] = ldg(sy0 + qq);
        in.r0[0] = ldg(r0y + p);
        in.r0[1] = ldg(r0y + qq);
        in.wy[0] = ldg(wy + p);
        in.wy[1] = ldg(wy + qq);
        in.ky[0] = ldg(ky + p);
        in.ky[1] = ldg(ky + qq);
        in.ey[0] = ldg(ey + p);
        in.ey[1] = ldg(ey + qq);
        in.ax0 = ldg(ax0 + r);
        return in;
      },
      [&](int r, double t, const RowIn& in, double* c1, double* c2, double* out) {
        const double axv = init ? t : in.ax0 + t;
        double cc1 = 0.0, cc2 = 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          // row p (h = 0): A x = axv, A x_c = t; its partner: the negatives
          const double sg = h == 0 ? 1.0 : -1.0;
          double yv, yc = 0.0;
          if (init) {
            yv = in.sy[h];
          } else {
            yc = -(sg * t - in.r0[h]) * in.wy[h];
            yv = in.sy[h] + yc;
          }
          const double ry = in.ky[h] - (sg * axv + in.ey[h] * yv);
          if (h == 0 || in.has_q) {
            cc1 += sg * (init ? yv : yc);
            cc2 += sg * (-ry * in.wy[h]);
          }
          out[h == 0 ? 0 : 3] = yv;
          out[h == 0 ? 2 : 4] = ry;
        }
        *c1 = cc1;
        *c2 = cc2;
        out[1] = axv;
      },
      [&](int r, const double* out) {
        int p, q;
        pr.rows(r, &p, &q);
        if (!init) stg(s1y + p, out[0]);
        stg(s1ax + r, out[1]);
        stg(s1ry + p, out[2]);
        double a = fabs(out[2]);
        mx = fmax(mx, (a == a) ? a : INFINITY);
        if (q >= 0) {
          if (!init) stg(s1y + q, out[3]);
          stg(s1ry + q, out[4]);
          a = fabs(out[4]);
          mx = fmax(mx, (a == a) ? a : INFINITY);
        }
      },
      [&](int j, double s1, double s2) {
        S1.aty[j] = init ? s1 : S0.aty[j] + s1;
        um[j] = s2;  // A'(W r1_y)
      },
      scratch, I.ext);
  PHASE(22);
  if (I.has_p) {
    // P_off u with u = sol_x (init) or x_c, added to P_off sol0_x otherwise.
    struct PIn {
      double u, px0, pjj;
    };
    const double* u = wave_uniform(init ? S1.sol : xc);
    const double* px0 = wave_uniform(S0.px);
    const double* P = wave_uniform(I.P);
    const int ld = I.ld;
    double* s1px = wave_uniform(S1.px);
    rowpass2<NS, R, D, 1, 0>(
        P, n, n, ld, u,
        [&](int j) {
          PIn in;
          in.u = ldg(u + j);
          in.px0 = ldg(px0 + j);
          in.pjj = ldg(P + (int64_t)j * ld + j);
          return in;
        },
        [&](int j, double t, const PIn& in, double* c1, double* c2, double* out) {
          const double v = t - in.pjj * in.u;
          out[0] = init ? v : in.px0 + v;
        },
        [&](int j, const double* out) { stg(s1px + j, out[0]); },
        [](int, double, double) {}, scratch);
  } else {
    for (int j = threadIdx.x; j < n; j += WG) S1.px[j] = 0.0;
    __syncthreads();
  }
  PHASE(23);
  {
    const double* solp = init ? S1.sol : S0.sol;
    const double* xcp = init ? S1.sol : xc;  // (xc is null in init; then unused)
    vpass(n,
          [&](int j) {
            return Dv<8>{{ldg(solp + j), ldg(xcp + j), ldg(krhs + j), ldg(S1.px + j),
                          ldg(I.dg + j), ldg(S1.aty + j), ldg(I.corr + j), ldg(um + j)}};
          },
          [&](int j, const Dv<8>& in) {
            const double xv = init ? in.v[0] : in.v[0] + in.v[1];
            if (!init) stg(S1.sol + j, xv);
            const double rx = in.v[2] - (in.v[3] + in.v[4] * xv + in.v[5]) + in.v[6] * xv;
            stg(S1.r + j, rx);
            stg(S1.g + j, rx + in.v[7]);
            const double a = fabs(rx);
            mx = fmax(mx, (a == a) ? a : INFINITY);
          });
  }
  const double res = bmax(mx, sm.red);
  PHASE(24);
  return res;
}

__device__ inline double ir_pass_any(const Inst& I, const IrState& S0, const IrState& S1,
                                     const double* xc, bool init, const double* krhs, double* um,
                                     Smem& sm) {
  // R rows x NS 128-column slots = 8 vectors per stage, one stage in flight
  if (I.ld <= 128) return ir_pass<1, 8, 1>(I, S0, S1, xc, init, krhs, um, sm);
  if (I.ld <= 256) return ir_pass<2, 4, 1>(I, S0, S1, xc, init, krhs, um, sm);
  if (I.ld <= 512) return ir_pass<4, 2, 1>(I, S0, S1, xc, init, krhs, um, sm);
  return ir_pass<8, 2, 1>(I, S0, S1, xc, init, krhs, um, sm);
}

// Widest rows the two-solve pass takes: wider rows would need more
// registers than a wave has (two solves' x vectors and column sums beside the
// staged rows), and the spills' waits would stall the row pipeline.
constexpr int kPass2MaxLd = 256;
// Rows a chunk of the two-solve pass stages (ten outputs per row).
constexpr int kPair2Chunk = 768;

struct Res2 {
  double a, b;
};

// ir_pass for two refined solves at once: (A0, A1, xca) with the KKT
// right-hand side I.krhs and A'(W r_y) into I.um, and (B0, B1, xcb) with
// I.krhs2 and I.um2.  One stream over A and one over P serve both; each
// solve's arithmetic is ir_pass's, term for term, so each gets ir_pass's
// results bit for bit.  Returns the two ||r1||_inf.
template <int NS, int R, int D>
__device__ __noinline__ Res2 ir_pass2(const Inst& I, const IrState A0, const IrState A1,
                                      const IrState B0, const IrState B1, const double* xca,
                                      const double* xcb, bool init, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, mc = I.mc;
  const PairRegs pr(I.pm);
  if (threadIdx.x == 0) g_si[I_KMV] += 1;
  double* scratch = reinterpret_cast<double*>(sm.u);
  const double* wy = wave_uniform(I.wy);
  const double* ey = wave_uniform(I.ey);
  const double* ky[2] = {wave_uniform(I.krhs + n), wave_uniform(I.krhs2 + n)};
  const double* r0y[2] = {wave_uniform(A0.r + n), wave_uniform(B0.r + n)};
  const double* ax0[2] = {wave_uniform(A0.ax), wave_uniform(B0.ax)};
  const double* sy0[2] = {wave_uniform((init ? A1.sol : A0.sol) + n),
                          wave_uniform((init ? B1.sol : B0.sol) + n)};
  double* s1y[2] = {wave_uniform(A1.sol + n), wave_uniform(B1.sol + n)};
  double* s1ax[2] = {wave_uniform(A1.ax), wave_uniform(B1.ax)};
  double* s1ry[2] = {wave_uniform(A1.r + n), wave_uniform(B1.r + n)};
  double mx[2] = {0.0, 0.0};
  // a stored row's inputs for both solves (row p and its partner q, as in
  // ir_pass); the diagonal terms are shared
  struct RowIn {
    double sy[2][2], r0[2][2], ky[2][2], ax0[2], wy[2], ey[2];
    int has_q;
  };
  static_assert(10 * kPair2Chunk <= kLdsU && 16 * kPass2MaxLd <= kLdsU,
                "LDS scratch of the two-solve pass");
  rowpass2<NS, R, D, 10, 4, false, kPair2Chunk, true, 2>(
      I.A, mc, n, I.ld, init ? A1.sol : xca,
      [&](int r) {
        int p, q;
        pr.rows(r, &p, &q);
        RowIn in;
        in.has_q = q >= 0;
        const int qq = q >= 0 ? q : p;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          in.sy[v][0] = ldg(sy0[v] + p);
          in.sy[v][1] = ldg(sy0[v] + qq);
          in.r0[v][0] = ldg(r0y[v] + p);
          in.r0[v][1] = ldg(r0y[v] + qq);
          in.ky[v][0] = ldg(ky[v] + p);
          in.ky[v][1] = ldg(ky[v] + qq);
          in.ax0[v] = ldg(ax0[v] + r);
        }
        in.wy[0] = ldg(wy + p);
        in.wy[1] = ldg(wy + qq);
        in.ey[0] = ldg(ey + p);
        in.ey[1] = ldg(ey + qq);
        return in;
      },
      [&](int r, double ta, double tb, const RowIn& in, double* c, double* out) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          const double t = v == 0 ? ta : tb;
          const double axv = init ? t : in.ax0[v] + t;
          double cc1 = 0.0, cc2 = 0.0;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const double sg = h == 0 ? 1.0 : -1.0;
            double yv, yc = 0.0;
            if (init) {
              yv = in.sy[v][h];
            } else {
              yc = -(sg * t - in.r0[v][h]) * in.wy[h];
              yv = in.sy[v][h] + yc;
            }
            const double ry = in.ky[v][h] - (sg * axv + in.ey[h] * yv);
            if (h == 0 || in.has_q) {
              cc1 += sg * (init ? yv : yc);
              cc2 += sg * (-ry * in.wy[h]);
            }
            out[5 * v + (h == 0 ? 0 : 3)] = yv;
            out[5 * v + (h == 0 ? 2 : 4)] = ry;
          }
          c[2 * v] = cc1;
          c[2 * v + 1] = cc2;
          out[5 * v + 1] = axv;
        }
      },
      [&](int r, const double* out) {
        int p, q;
        pr.rows(r, &p, &q);
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          const double* o = out + 5 * v;
          if (!init) stg(s1y[v] + p, o[0]);
          stg(s1ax[v] + r, o[1]);
          stg(s1ry[v] + p, o[2]);
          double a = fabs(o[2]);
          mx[v] = fmax(mx[v], (a == a) ? a : INFINITY);
          if (q >= 0) {
            if (!init) stg(s1y[v] + q, o[3]);
            stg(s1ry[v] + q, o[4]);
            a = fabs(o[4]);
            mx[v] = fmax(mx[v], (a == a) ? a : INFINITY);
          }
        }
      },
      [&](int j, const double* cs) {
        A1.aty[j] = init ? cs[0] : A0.aty[j] + cs[0];
        I.um[j] = cs[1];
        B1.aty[j] = init ? cs[2] : B0.aty[j] + cs[2];
        I.um2[j] = cs[3];
      },
      scratch, I.ext, init ? B1.sol : xcb);
  if (I.has_p) {
    struct PIn {
      double u[2], px0[2], pjj;
    };
    const double* u[2] = {wave_uniform(init ? A1.sol : xca), wave_uniform(init ? B1.sol : xcb)};
    const double* px0[2] = {wave_uniform(A0.px), wave_uniform(B0.px)};
    double* s1px[2] = {wave_uniform(A1.px), wave_uniform(B1.px)};
    const double* P = wave_uniform(I.P);
    const int ld = I.ld;
    rowpass2<NS, R, D, 2, 0, false, 2 * kPair2Chunk, false, 2>(
        P, n, n, ld, u[0],
        [&](int j) {
          PIn in;
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            in.u[v] = ldg(u[v] + j);
            in.px0[v] = ldg(px0[v] + j);
          }
          in.pjj = ldg(P + (int64_t)j * ld + j);
          return in;
        },
        [&](int j, double ta, double tb, const PIn& in, double* c, double* out) {
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            const double w = (v == 0 ? ta : tb) - in.pjj * in.u[v];
            out[v] = init ? w : in.px0[v] + w;
          }
        },
        [&](int j, const double* out) {
          stg(s1px[0] + j, out[0]);
          stg(s1px[1] + j, out[1]);
        },
        [](int, const double*) {}, scratch, nullptr, u[1]);
  } else {
    for (int j = threadIdx.x; j < n; j += WG) {
      A1.px[j] = 0.0;
      B1.px[j] = 0.0;
    }
    __syncthreads();
  }
  {
    // each vector's terms as ir_pass forms them
    const double* solp[2] = {init ? A1.sol : A0.sol, init ? B1.sol : B0.sol};
    const double* xcp[2] = {init ? A1.sol : xca, init ? B1.sol : xcb};  // (null in init)
    const double* krhs[2] = {I.krhs, I.krhs2};
    const double* um[2] = {I.um, I.um2};
    const IrState* S1[2] = {&A1, &B1};
    vpass(n,
          [&](int j) {
            Dv<14> t;
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              t.v[6 * v + 0] = ldg(solp[v] + j);
              t.v[6 * v + 1] = ldg(xcp[v] + j);
              t.v[6 * v + 2] = ldg(krhs[v] + j);
              t.v[6 * v + 3] = ldg(S1[v]->px + j);
              t.v[6 * v + 4] = ldg(S1[v]->aty + j);
              t.v[6 * v + 5] = ldg(um[v] + j);
            }
            t.v[12] = ldg(I.dg + j);
            t.v[13] = ldg(I.corr + j);
            return t;
          },
          [&](int j, const Dv<14>& in) {
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              const double* q = in.v + 6 * v;
              const double xv = init ? q[0] : q[0] + q[1];
              if (!init) stg(S1[v]->sol + j, xv);
              const double rx = q[2] - (q[3] + in.v[12] * xv + q[4]) + in.v[13] * xv;
              stg(S1[v]->r + j, rx);
              stg(S1[v]->g + j, rx + q[5]);
              const double a = fabs(rx);
              mx[v] = fmax(mx[v], (a == a) ? a : INFINITY);
            }
          });
  }
  const double ra = bmax(mx[0], sm.red);
  return Res2{ra, bmax(mx[1], sm.red)};
}

__device__ inline Res2 ir_pass2_any(const Inst& I, const IrState& A0, const IrState& A1,
                                    const IrState& B0, const IrState& B1, const double* xca,
                                    const double* xcb, bool init, Smem& sm) {
  // one stage in flight, as in ir_pass (ld <= kPass2MaxLd)
  if (I.ld <= 128) return ir_pass2<1, 8, 1>(I, A0, A1, B0, B1, xca, xcb, init, sm);
  return ir_pass2<2, 4, 1>(I, A0, A1, B0, B1, xca, xcb, init, sm);
}

// xc = G^{-1} g with the factor in precision T.
template <typename T>
__device__ __noinline__ void factor_solve(const Inst& I, const T* L, const T* LiB, const double* g,
                                          double* xc, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int n = I.n;
  if (threadIdx.x == 0) g_si[sizeof(T) == 4 ? I_FS32 : I_FS64] += 1;
  T* v = reinterpret_cast<T*>(sm.v);
  vpass(n, [&](int j) { return ldg(g + j); }, [&](int j, double x) { v[j] = (T)x; });
  __syncthreads();
  PHASE(21);
  trsv_lower_inv(L, n, LiB, v, sm.red, reinterpret_cast<double*>(sm.u));
  PHASE(19);
  trsv_upper_inv(L, n, LiB, v, sm.red, reinterpret_cast<double*>(sm.u));
  PHASE(20);
  for (int j = threadIdx.x; j < n; j += WG) xc[j] = (double)v[j];
  __syncthreads();
}

__device__ inline void factor_solve_current(const Inst& I, const double* g, double* xc, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  if (g_si[I_FORCE64])
    factor_solve<double>(I, I.G, I.LiB, g, xc, sm);
  else
    factor_solve<float>(I, reinterpret_cast<const float*>(I.G),
                        reinterpret_cast<const float*>(I.LiB), g, xc, sm);
}

// Replaces an FP32 factor that refines too slowly by an FP64 one.
__device__ inline bool fallback_fp64(const Inst& I, Smem& sm_in, Counters* cnt) {
  Smem sm = lds_view(sm_in);
  if (threadIdx.x == 0) {
    g_si[I_FORCE64] = 1;
    g_si[I_FALLBACK] += 1;
    g_si[I_F64] += 1;
    atomicAdd(&cnt->fallbacks, 1ull);
  }
  __syncthreads();
  return factor64_retry(I, sm, cnt);
}

// Copies a refinement state's products -- A sol_x over the stored rows,
// A' sol_y and P_off sol_x -- into dst, laid out [m | n | n].
__device__ inline void keep_products(const Inst& I, const IrState& S, double* dst) {
  const int m = I.m, n = I.n;
  const double *ax = S.ax, *aty = S.aty, *px = S.px;
  vpass(I.mc, [&](int k) { return ldg(ax + k); }, [&](int k, double v) { stg(dst + k, v); });
  vpass(n, [&](int j) { return Dv<2>{{ldg(aty + j), ldg(px + j)}}; },
        [&](int j, const Dv<2>& in) {
          stg(dst + m + j, in.v[0]);
          stg(dst + m + n + j, in.v[1]);
        });
  __syncthreads();
}

struct IrResult {
  int solves, status;  // 0 converged, 1 non-converged, 2 stalled
  double res;
  bool ok;
};

// DirectKktSolver.solve (direct.py:243-307) on K_true = K_reg - diag(corr):
// sol from warm, each step sol += K_reg^{-1} res with the current factor and
// res recomputed in FP64, until res < atol + rtol ||rhs||, a stall, or
// max_steps.  A step is: the triangular solves for x_c from g, then one
// ir_pass that completes the correction (y_c), advances sol and the
// products, and forms the next residual and g.  An FP32 step that contracts
// the residual by less than kPoorContraction is dropped and the factor is
// redone in FP64 (it does not count as a step); FP64 steps follow the
// reference rule exactly.  rhs may alias I.r (it is consumed into krhs
// first).  keep_warm / keep_sol, where not null, receive the products of
// warm and of sol (keep_products).
constexpr double kPoorContraction = 3e-2;
constexpr int kF32StepBudget = 3;  // refinement steps per IPM iteration kept on an FP32 factor

__device__ __noinline__ IrResult solve_ir(const Inst& I, const double* rhs, const double* warm,
                                          double* sol, const Opts o, Smem& sm_in, Counters* cnt,
                                          double* keep_warm = nullptr,
                                          double* keep_sol = nullptr) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, m = I.m, N = I.n + I.m;
  double mx = 0.0;
  vpass(N, [&](int k) { return Dv<2>{{ldg(rhs + k), ldg(warm + k)}}; },
        [&](int k, const Dv<2>& in) {
          const double v = k < n ? in.v[0] : -in.v[0];
          stg(I.krhs + k, v);
          stg(sol + k, in.v[1]);
          const double a = fabs(v);
          mx = fmax(mx, (a == a) ? a : INFINITY);
        });
  const double rhs_norm = bmax(mx, sm.red);
  const double tol = o.ls_atol + o.ls_rtol * rhs_norm;
  double* w = I.irb;
  IrState S0{sol, w, w + 2 * m, w + 2 * m + 2 * n, I.res, w + 2 * m + 4 * n};
  IrState S1{I.r, w + m, w + 2 * m + n, w + 2 * m + 3 * n, I.ires, w + 2 * m + 5 * n};
  double* xc = I.iacc;
  double res = ir_pass_any(I, S1, S0, nullptr, true, I.krhs, I.um, sm);
  if (keep_warm) keep_products(I, S0, keep_warm);
  PHASE(3);
  IrResult out{0, 1, res, true};
  int solves = 0;
  while (solves < o.max_ir) {
    const double old = res;
    const bool f32 = g_si[I_FORCE64] == 0;
    if (threadIdx.x == 0) g_cnt[C_SOLVES] += 1;
    factor_solve_current(I, S0.g, xc, sm);
    PHASE(4);
    const double nr = ir_pass_any(I, S0, S1, xc, false, I.krhs, I.um, sm);
    PHASE(5);
    if (f32) {
      if (threadIdx.x == 0) g_cnt[C_INNER] += 1;
      if (!(nr < tol) && !(nr <= kPoorContraction * old)) {
        if (!fallback_fp64(I, sm, cnt)) {
          out.ok = false;
          return out;
        }
        continue;  // S0 is untouched: redo the step with the FP64 factor
      }
    }
    ++solves;
    const IrState t = S0;
    S0 = S1;
    S1 = t;
    res = nr;
    if (res < tol) {
      out.status = 0;
      break;
    }
    if (res >= old) {
      out.status = 2;
      break;
    }
  }
  if (S0.sol != sol) {
    const double* src = S0.sol;
    vpass(N, [&](int k) { return ldg(src + k); }, [&](int k, double v) { stg(sol + k, v); });
    __syncthreads();
  }
  if (keep_sol) keep_products(I, S0, keep_sol);
  PHASE(6);
  out.solves = max(solves, 1);
  out.res = res;
  if (out.status != 0 && threadIdx.x == 0) g_cnt[C_IRBAD] += 1;
  return out;
}

// Two refined solves with the FP64 factor at once: (I.r, warm1) into sol1
// and (I.r2, warm2) into I.sol2, as two calls of solve_ir would run them
// (the FP64 factor takes no fallback, so the two are independent).  Each
// keeps its own residual, tolerance, stall test and step count; while both
// still step, each step's pass serves both (ir_pass2), and a solve that
// has stopped leaves the other to ir_pass.  Each result is solve_ir's bit
// for bit; keep_sol1 receives the products of sol1 and keep_warm2 those of
// warm2 (keep_products).
__device__ __noinline__ void solve_ir2(const Inst& I, const double* warm1, double* sol1,
                                       const double* warm2, const Opts o, Smem& sm_in,
                                       Counters* cnt, IrResult* out1, IrResult* out2,
                                       double* keep_sol1, double* keep_warm2) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, m = I.m, N = I.n + I.m;
  double mx1 = 0.0, mx2 = 0.0;
  double* sol2 = I.sol2;
  vpass(N,
        [&](int k) {
          return Dv<4>{{ldg(I.r + k), ldg(I.r2 + k), ldg(warm1 + k), ldg(warm2 + k)}};
        },
        [&](int k, const Dv<4>& in) {
          const double v1 = k < n ? in.v[0] : -in.v[0];
          const double v2 = k < n ? in.v[1] : -in.v[1];
          stg(I.krhs + k, v1);
          stg(I.krhs2 + k, v2);
          stg(sol1 + k, in.v[2]);
          stg(sol2 + k, in.v[3]);
          double a = fabs(v1);
          mx1 = fmax(mx1, (a == a) ? a : INFINITY);
          a = fabs(v2);
          mx2 = fmax(mx2, (a == a) ? a : INFINITY);
        });
  const double norm1 = bmax(mx1, sm.red);
  const double norm2 = bmax(mx2, sm.red);
  struct Run {
    IrState S0, S1;
    double* xc;
    double tol, res;
    int solves, status;
    bool on;
  };
  double* w = I.irb;
  double* w2 = I.irb2;
  Run q[2] = {
      {IrState{sol1, w, w + 2 * m, w + 2 * m + 2 * n, I.res, w + 2 * m + 4 * n},
       IrState{I.r, w + m, w + 2 * m + n, w + 2 * m + 3 * n, I.ires, w + 2 * m + 5 * n},
       I.iacc, o.ls_atol + o.ls_rtol * norm1, 0.0, 0, 1, true},
      {IrState{sol2, w2, w2 + 2 * m, w2 + 2 * m + 2 * n, I.res2, w2 + 2 * m + 4 * n},
       IrState{I.r2, w2 + m, w2 + 2 * m + n, w2 + 2 * m + 3 * n, I.ires2, w2 + 2 * m + 5 * n},
       I.iacc2, o.ls_atol + o.ls_rtol * norm2, 0.0, 0, 1, true}};
  {
    const Res2 r = ir_pass2_any(I, q[0].S1, q[0].S0, q[1].S1, q[1].S0, nullptr, nullptr, true, sm);
    q[0].res = r.a;
    q[1].res = r.b;
  }
  keep_products(I, q[1].S0, keep_warm2);
  PHASE(3);
  for (;;) {
    bool act[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) act[v] = q[v].on && q[v].solves < o.max_ir;
    if (!act[0] && !act[1]) break;
    double nr[2] = {0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 2; ++v)
      if (act[v]) {
        if (threadIdx.x == 0) g_cnt[C_SOLVES] += 1;
        factor_solve<double>(I, I.G, I.LiB, q[v].S0.g, q[v].xc, sm);
      }
    PHASE(4);
    if (act[0] && act[1]) {
      const Res2 r = ir_pass2_any(I, q[0].S0, q[0].S1, q[1].S0, q[1].S1, q[0].xc, q[1].xc,
                                  false, sm);
      nr[0] = r.a;
      nr[1] = r.b;
    } else if (act[0]) {
      nr[0] = ir_pass_any(I, q[0].S0, q[0].S1, q[0].xc, false, I.krhs, I.um, sm);
    } else {
      nr[1] = ir_pass_any(I, q[1].S0, q[1].S1, q[1].xc, false, I.krhs2, I.um2, sm);
    }
    PHASE(5);
#pragma unroll
    for (int v = 0; v < 2; ++v)
      if (act[v]) {
        Run& u = q[v];
        const double old = u.res;
        ++u.solves;
        const IrState t = u.S0;
        u.S0 = u.S1;
        u.S1 = t;
        u.res = nr[v];
        if (u.res < u.tol) {
          u.status = 0;
          u.on = false;
        } else if (u.res >= old) {
          u.status = 2;
          u.on = false;
        }
      }
  }
  double* sols[2] = {sol1, sol2};
  IrResult* outs[2] = {out1, out2};
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    if (q[v].S0.sol != sols[v]) {
      const double* src = q[v].S0.sol;
      double* dst = sols[v];
      vpass(N, [&](int k) { return ldg(src + k); }, [&](int k, double x) { stg(dst + k, x); });
      __syncthreads();
    }
    if (v == 0) keep_products(I, q[0].S0, keep_sol1);
    *outs[v] = IrResult{max(q[v].solves, 1), q[v].status, q[v].res, true};
    if (q[v].status != 0 && threadIdx.x == 0) g_cnt[C_IRBAD] += 1;
  }
  PHASE(6);
}

// -------------------------------------------------------------------------
// Scalar pieces of the IPM (__init__.py:671-857).
// -------------------------------------------------------------------------
__device__ double max_step(const double* v, const double* dv_a, const double* dv_b, int lo,
                           int hi, double* red) {
  // dv = dv_a - dv_b if dv_b else dv_a
  double mn = INFINITY;
  const double* dvb = dv_b ? dv_b : dv_a;
  vpass(hi - lo,
        [&](int k) { return Dv<3>{{ldg(v + lo + k), ldg(dv_a + lo + k), ldg(dvb + lo + k)}}; },
        [&](int, const Dv<3>& t) {
          const double dv = dv_b ? t.v[1] - t.v[2] : t.v[1];
          if (dv < -kEps) mn = fmin(mn, -t.v[0] / dv);
        });
  mn = bmin(mn, red);
  return mn == INFINITY ? 1.0 : fmin(1.0, mn);
}

// P u for the n-vector u into out (has_p only).
__device__ inline void pmul(const Inst& I, const double* u, double* out, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  out = wave_uniform(out);
  with_pass_shape(I.ld, [&](auto sh) {
    using S = decltype(sh);
    rowpass2<S::NS, S::R, S::D, 1, 0>(
        I.P, I.n, I.n, I.ld, u, [](int) { return 0; },
        [](int, double t, int, double*, double*, double* o) { o[0] = t; },
        [&](int j, const double* o) { stg(out + j, o[0]); }, [](int, double, double) {},
        reinterpret_cast<double*>(sm.u));
  });
}

// Quadratic tau (__init__.py:759-817); returns false where it raises.
// P kq_x is the iteration's pkq.
__device__ __noinline__ bool tau_quadratic(const Inst& I, const double* kr, double mu, double mu_t,
                              double r_tau, double* tau_out, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, m = I.m;
  if (I.has_p) pmul(I, kr, I.un, sm);  // P kr_x
  // the dot products of the tau quadratic and max|kr|, reduced together:
  // slots kq.c, kq_y.b, kr.c, kr_y.b, kq.Pkq, kr.Pkq, kq.Pkr, kr.Pkr, max|kr|
  double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  vpass(n + m, [&](int k) { return ldg(kr + k); },
        [&](int, double x) { v[8] = fmax(v[8], fabs(x)); });
  vpass(n, [&](int j) { return Dv<3>{{ldg(I.kq + j), ldg(kr + j), ldg(I.c + j)}}; },
        [&](int, const Dv<3>& in) {
          v[0] += in.v[0] * in.v[2];
          v[2] += in.v[1] * in.v[2];
        });
  vpass(m, [&](int i) { return Dv<3>{{ldg(I.kq + n + i), ldg(kr + n + i), ldg(I.b + i)}}; },
        [&](int, const Dv<3>& in) {
          v[1] += in.v[0] * in.v[2];
          v[3] += in.v[1] * in.v[2];
        });
  if (I.has_p) {
    vpass(n,
          [&](int j) {
            return Dv<4>{{ldg(I.kq + j), ldg(kr + j), ldg(I.pkq + j), ldg(I.un + j)}};
          },
          [&](int, const Dv<4>& in) {
            v[4] += in.v[0] * in.v[2];
            v[5] += in.v[1] * in.v[2];
            v[6] += in.v[0] * in.v[3];
            v[7] += in.v[1] * in.v[3];
          });
  }
  block_reduce_n<9, 0x100u>(v, sm.red);
  const double dot_dbg_kq = v[8];
  double t_a = mu + (v[0] + v[1]);
  double t_b = -r_tau - (v[2] + v[3]);
  double t_c = -mu_t;
  if (I.has_p) {
    t_a -= v[4];
    t_b += v[5] + v[6];
    t_c -= v[7];
  }
  if (threadIdx.x == 0 && I.trace) {
    const int row = g_si[I_ITERS];
    if (row < kTraceRows) {
      double* tr = I.trace + (int64_t)row * kTraceFields + T_DBG0 + (mu_t == 0.0 ? 0 : 4);
      tr[0] = t_a;
      tr[1] = t_b;
      tr[2] = t_c;
      tr[3] = dot_dbg_kq;
    }
  }
  if (fabs(t_a) < kEps) {
    if (fabs(t_b) < kEps) return false;
    const double t = -t_c / t_b;
    if (!isfinite(t) || t < -1e-10) return false;
    *tau_out = fmax(0.0, t);
    return true;
  }
  double disc = t_b * t_b - 4 * t_a * t_c;
  if (disc < -1e-9) return false;
  disc = fmax(0.0, disc);
  const double sq = sqrt(disc);
  const double t = t_b > 0 ? t_c / (-0.5 * (t_b + sq)) : (-0.5 * (t_b - sq)) / t_a;
  if (!isfinite(t) || t < -1e-10) return false;
  *tau_out = fmax(0.0, t);
  return true;
}

// Linearised tau with trust region (__init__.py:819-857).
__device__ __noinline__ double tau_linearized(const Inst& I, const double* kr, double mu, double mu_t,
                                 double tau_c, double tau_anchor, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, m = I.m;
  double x_px = 0.0, px_kq = 0.0, px_kr = 0.0;
  if (I.has_p) {
    pmul(I, I.x, I.un, sm);
    x_px = dot(I.x, I.un, n, sm.red);
    px_kq = dot(I.un, I.kq, n, sm.red);
    px_kr = dot(I.un, kr, n, sm.red);
  }
  const double q_z = dot(I.c, I.x, n, sm.red) + dot(I.b, I.y, m, sm.red);
  const double q_kq = dot(I.c, I.kq, n, sm.red) + dot(I.b, I.kq + n, m, sm.red);
  const double q_kr = dot(I.c, kr, n, sm.red) + dot(I.b, kr + n, m, sm.red);
  const double q_rz = q_kr - tau_c * q_kq - q_z;
  const double px_rz = px_kr - tau_c * px_kq - x_px;
  const double g = mu * tau_c * tau_c + (mu_t - mu) * tau_anchor * tau_c - tau_c * q_z - mu_t - x_px;
  const double num = g - tau_c * q_rz - 2.0 * px_rz;
  const double den = 2.0 * mu * tau_c + (mu_t - mu) * tau_anchor - q_z + tau_c * q_kq + 2.0 * px_kq;
  const double t = tau_c + (fabs(den) < 1e-16 ? 0.0 : -num / den);
  if (!isfinite(t)) return tau_c;
  return fmin(fmax(t, 0.1 * tau_c), 10.0 * tau_c);
}

// Right-hand side of a Newton step (__init__.py:708-724) into r.
__device__ void newton_rhs(const Inst& I, double mu, double mu_t, const double* anchor,
                           const double* corr_m, double* r) {
  const int n = I.n, z = I.z, N = I.n + I.m;
  // (a plain loop: this function is inlined at two call sites whose results
  // must agree bit for bit, and with its loads batched the multiply and the
  // conditional adds below were contracted into FMAs differently in the two)
  for (int k = threadIdx.x; k < N; k += WG) {
    double v = (mu - mu_t) * anchor[k];
    if (k >= n + z) {
      const int i = k - n;
      if (mu_t != 0.0) v += mu_t / I.y[i];
      v += I.s[i];
      if (corr_m) v += corr_m[i];
    }
    r[k] = v;
  }
  __syncthreads();
}

// The rest of a Newton step from its solve sol (__init__.py:726-757):
// tau+ and [x+; y+] = sol - kinv_q tau+ into outxy.
__device__ __noinline__ void newton_finish(const Inst& I, const IrResult ir, const double* sol,
                                           double mu, double mu_t, double tau_anchor, double tau,
                                           double* outxy, double* tau_plus, Smem& sm_in) {
  Smem sm = lds_view(sm_in);
  const int N = I.n + I.m;
  double t = 0.0;
  bool have = false;
  if (ir.status == 0 || ir.res < 1e-7)
    have = tau_quadratic(I, sol, mu, mu_t, (mu - mu_t) * tau_anchor, &t, sm);
  if (!have) t = tau_linearized(I, sol, mu, mu_t, tau, tau_anchor, sm);
  PHASE(8);
  vpass(N, [&](int k) { return Dv<2>{{ldg(sol + k), ldg(I.kq + k)}}; },
        [&](int k, const Dv<2>& in) { stg(outxy + k, in.v[0] - in.v[1] * t); });
  __syncthreads();
  PHASE(9);
  *tau_plus = t;
}

// Newton step (__init__.py:708-757): writes [x+; y+] into outxy, returns tau+.
__device__ __noinline__ bool newton(const Inst& I, double mu, double mu_t, const double* anchor,
                       double tau_anchor, const double* corr_m, double tau, double* outxy,
                       double* tau_plus, const Opts o, Smem& sm_in, Counters* cnt,
                       IrResult* irout, double* keep_warm, double* keep_sol) {
  Smem sm = lds_view(sm_in);
  newton_rhs(I, mu, mu_t, anchor, corr_m, I.r);
  PHASE(10);
  const IrResult ir = solve_ir(I, I.r, anchor, I.sol, o, sm, cnt, keep_warm, keep_sol);
  *irout = ir;
  if (!ir.ok) return false;
  newton_finish(I, ir, I.sol, mu, mu_t, tau_anchor, tau, outxy, tau_plus, sm);
  return true;
}

// One IPM iteration of one QP (__init__.py:473-578).
__device__ __noinline__ void ipm_iteration(const Inst& I, const Opts o, Smem& sm_in, Counters* cnt) {
  Smem sm = lds_view(sm_in);
  const int n = I.n, m = I.m, z = I.z, N = n + m;
  const int tid = threadIdx.x;
  double* red = sm.red;
  double tau = g_sd[S_TAU];
  PHASE(-1);

  // normalise so ||(x, y, tau)||^2 = m - z + 1
  double nx2 = 0.0;
  vpass(n, [&](int k) { return ldg(I.x + k); }, [&](int, double v) { nx2 += v * v; });
  double ny2 = 0.0;
  vpass(m, [&](int k) { return ldg(I.y + k); }, [&](int, double v) { ny2 += v * v; });
  const double nrm = sqrt(bsum(nx2, red) + bsum(ny2, red) + tau * tau);
  const double scale = sqrt((double)(m - z + 1)) / fmax(kEps, nrm);
  vpass(n, [&](int k) { return ldg(I.x + k); }, [&](int k, double v) { stg(I.x + k, v * scale); });
  vpass(m, [&](int k) { return Dv<2>{{ldg(I.y + k), ldg(I.s + k)}}; },
        [&](int k, const Dv<2>& t) {
          stg(I.y + k, t.v[0] * scale);
          stg(I.s + k, t.v[1] * scale);
        });
  tau *= scale;
  __syncthreads();
  const double mu = dot(I.y, I.s, m, red) / (m - z);


  // regularised diagonal and correction (direct.py:196-216)
  vpass(N,
        [&](int k) {
          const int j = min(k, n - 1), i = max(k - n, 0);
          return Dv<3>{{ldg(I.P + (int64_t)j * I.ld + j), ldg(I.s + i), ldg(I.y + i)}};
        },
        [&](int k, const Dv<3>& in) {
          double t;
          if (k < n) t = in.v[0] + mu;
          else if (k < n + z) t = mu;
          else t = in.v[1] / in.v[2] + mu;
          const double rg = fmax(t, o.min_reg);
          if (k < n) {
            stg(I.dg + k, rg);
            stg(I.corr + k, rg - t);
          } else {
            stg(I.dg + k, -rg);
            stg(I.corr + k, -rg + t);
            stg(I.wy + (k - n), -1.0 / rg);
            stg(I.ey + (k - n), -t);
          }
        });
  __syncthreads();
  PHASE(0);
  if (!factor(I, sm, cnt)) {
    if (tid == 0) g_si[I_STATUS] = kStatusError;
    return;
  }

  IrResult irq{}, irp{}, irc{};
  double tau_p;
  // products for the termination test (keep_products): of [x; y] as the
  // iteration starts, of kinv_q, of the corrector's solution
  const int64_t tks = (int64_t)m + 2 * n;
  double* tk_start = I.tk;
  double* tk_q = I.tk + tks;
  double* tk_c = I.tk + 2 * tks;
  if (o.fuse && g_si[I_FORCE64] && I.ld <= kPass2MaxLd) {
    // With the FP64 factor, kinv_q and the predictor's solve refine together
    // (solve_ir2): the same two solves, sharing each pass over A and P.
    vpass(N,
          [&](int k) {
            const int j = min(k, n - 1), i = max(k - n, 0);
            return Dv<4>{{ldg(I.c + j), ldg(I.b + i), ldg(I.x + j), ldg(I.y + i)}};
          },
          [&](int k, const Dv<4>& in) {
            stg(I.r + k, k < n ? in.v[0] : in.v[1]);
            stg(I.xp + k, k < n ? in.v[2] : in.v[3]);
          });
    __syncthreads();
    newton_rhs(I, mu, 0.0, I.xp, nullptr, I.r2);
    PHASE(15);
    solve_ir2(I, I.kq, I.sol, I.xp, o, sm, cnt, &irq, &irp, tk_q, tk_start);
    vpass(N, [&](int k) { return ldg(I.sol + k); }, [&](int k, double v) { stg(I.kq + k, v); });
    __syncthreads();
    if (I.has_p) pmul(I, I.kq, I.pkq, sm);
    PHASE(7);
    newton_finish(I, irp, I.sol2, mu, 0.0, tau, tau, I.xp, &tau_p, sm);
  } else {
    // kinv_q = K^{-1} [c; b], warm-started from the previous kinv_q
    vpass(N, [&](int k) { return ldg(I.kq + k); }, [&](int k, double v) { stg(I.xp + k, v); });
    __syncthreads();
    {
      vpass(N, [&](int k) { return Dv<2>{{ldg(I.c + min(k, n - 1)), ldg(I.b + max(k - n, 0))}}; },
            [&](int k, const Dv<2>& in) { stg(I.r + k, k < n ? in.v[0] : in.v[1]); });
      __syncthreads();
      const IrResult ir = solve_ir(I, I.r, I.xp, I.sol, o, sm, cnt, nullptr, tk_q);
      irq = ir;
      if (!ir.ok) {
        if (tid == 0) g_si[I_STATUS] = kStatusError;
        return;
      }
      vpass(N, [&](int k) { return ldg(I.sol + k); }, [&](int k, double v) { stg(I.kq + k, v); });
      __syncthreads();
      // P kq_x, which both tau quadratics of the iteration use
      if (I.has_p) pmul(I, I.kq, I.pkq, sm);
      PHASE(7);
    }

    // predictor (mu_target = 0) anchored at [x; y]
    vpass(N, [&](int k) { return Dv<2>{{ldg(I.x + min(k, n - 1)), ldg(I.y + max(k - n, 0))}}; },
          [&](int k, const Dv<2>& in) { stg(I.xp + k, k < n ? in.v[0] : in.v[1]); });
    __syncthreads();
    PHASE(15);
    if (!newton(I, mu, 0.0, I.xp, tau, nullptr, tau, I.xp, &tau_p, o, sm, cnt, &irp, tk_start,
                nullptr)) {
      if (tid == 0) g_si[I_STATUS] = kStatusError;
      return;
    }
  }
  // d_s = -y_p s / y ; step, sigma
  vpass(m, [&](int i) { return Dv<3>{{ldg(I.xp + n + i), ldg(I.s + i), ldg(I.y + i)}}; },
        [&](int i, const Dv<3>& in) { stg(I.ds + i, i < z ? 0.0 : -in.v[0] * in.v[1] / in.v[2]); });
  __syncthreads();
  const double a_s = max_step(I.s, I.ds, nullptr, z, m, red);
  const double a_y = max_step(I.y, I.xp + n, I.y, z, m, red);
  const double alpha_p = fmin(a_s, a_y);
  double sx = 0.0, sy = 0.0, sys = 0.0;
  vpass(n, [&](int k) { return Dv<2>{{ldg(I.x + k), ldg(I.xp + k)}}; },
        [&](int, const Dv<2>& in) {
          const double xa = in.v[0] + alpha_p * (in.v[1] - in.v[0]);
          sx += xa * xa;
        });
  vpass(m,
        [&](int i) { return Dv<4>{{ldg(I.y + i), ldg(I.xp + n + i), ldg(I.s + i), ldg(I.ds + i)}}; },
        [&](int, const Dv<4>& in) {
          const double ya = in.v[0] + alpha_p * (in.v[1] - in.v[0]);
          const double sa = in.v[2] + alpha_p * in.v[3];
          sy += ya * ya;
          sys += ya * sa;
        });
  const double ta = tau + alpha_p * (tau_p - tau);
  double sums[3] = {sx, sy, sys};
  block_reduce_n<3>(sums, red);
  const double nrm2 = sums[0] + sums[1] + ta * ta;
  const double mu_aff = (double)(m - z + 1) / fmax(kEps * kEps, nrm2) * sums[2] / (m - z);
  const double base = mu_aff / fmax(kEps, mu);
  const double sigma = fmin(fmax(base * base * base, 0.0), 1.0);

  // corrector: Mehrotra term, mu_target = sigma mu, anchored at [x_p; y_p]
  PHASE(11);
  vpass(m, [&](int i) { return Dv<3>{{ldg(I.ds + i), ldg(I.xp + n + i), ldg(I.y + i)}}; },
        [&](int i, const Dv<3>& in) {
          stg(I.cm + i, i < z ? 0.0 : -in.v[0] * (in.v[1] - in.v[2]) / in.v[2]);
        });
  __syncthreads();
  PHASE(12);
  double tau_c;
  if (!newton(I, mu, sigma * mu, I.xp, tau_p, I.cm, tau, I.xp, &tau_c, o, sm, cnt, &irc, nullptr,
              tk_c)) {
    if (tid == 0) g_si[I_STATUS] = kStatusError;
    return;
  }
  vpass(m,
        [&](int i) { return Dv<4>{{ldg(I.y + i), ldg(I.cm + i), ldg(I.xp + n + i), ldg(I.s + i)}}; },
        [&](int i, const Dv<4>& in) {
          stg(I.ds + i,
              i < z ? 0.0 : sigma * mu / in.v[0] + in.v[1] - in.v[2] * in.v[3] / in.v[0]);
        });
  __syncthreads();
  const double alpha = fmin(max_step(I.s, I.ds, nullptr, z, m, red),
                            max_step(I.y, I.xp + n, I.y, z, m, red));
  const double step = o.step_scale * alpha;
  vpass(n, [&](int k) { return Dv<2>{{ldg(I.x + k), ldg(I.xp + k)}}; },
        [&](int k, const Dv<2>& in) { stg(I.x + k, in.v[0] + step * (in.v[1] - in.v[0])); });
  vpass(m,
        [&](int i) { return Dv<4>{{ldg(I.y + i), ldg(I.xp + n + i), ldg(I.s + i), ldg(I.ds + i)}}; },
        [&](int i, const Dv<4>& in) {
          double yv = in.v[0] + step * (in.v[1] - in.v[0]);
          double sv = in.v[2] + step * in.v[3];
          if (i >= z) {
            yv = fmax(yv, 1e-30);
            sv = fmax(sv, 1e-30);
          }
          stg(I.y + i, yv);
          stg(I.s + i, sv);
        });
  tau = fmax(tau + step * (tau_c - tau), 1e-30);
  PHASE(13);
  // An FP32 factor whose three solves took more than kF32StepBudget
  // refinement steps costs more than an FP64 one: every extra step streams A
  // in FP64 once more, more than the FP64 Gram product adds.  Such a QP
  // factors in FP64 from its next iteration on.
  if (tid == 0 && !g_si[I_FORCE64] && irq.solves + irp.solves + irc.solves > kF32StepBudget) {
    g_si[I_FORCE64] = 1;
    g_si[I_FALLBACK] += 1;
    atomicAdd(&cnt->fallbacks, 1ull);
  }
  __syncthreads();

  // termination on the unscaled problem (__init__.py:888-965):
  // A_orig x_orig = (A x)/d, A_orig' y_orig = (A' y)/e, P_orig x_orig = (P x)/e.
  // The products of the updated iterate follow from those the refinement
  // already formed, by linearity of the update: x+ = x + step (x_c - x) with
  // x_c = sol_c - tau_c kinv_q, so A x+ = A x + step ((A sol_c - tau_c
  // A kinv_q) - A x), and likewise A'y+ and P_off x+ (P x+ adds the diagonal
  // of P times x+).  No pass over A or P is needed.  (The floor of 1e-30 on
  // y+ changes A'y+ by at most 1e-30 |A|.)
  const double inv_tau = 1.0 / fmax(tau, kEps);
  const bool has_p = I.has_p;
  double* ax = I.um;  // m
  double* aty = I.un; // n
  double* px = I.iacc;  // n
  {
    const PairRegs pr(I.pm);
    vpass(I.mc,
          [&](int r) {
            int p, q;
            pr.rows(r, &p, &q);
            return Dv<4>{{ldg(tk_start + r), ldg(tk_c + r), ldg(tk_q + r), ldg(I.d + p)}};
          },
          [&](int r, const Dv<4>& in) {
            const double t0 = in.v[0];
            const double t = t0 + step * ((in.v[1] - tau_c * in.v[2]) - t0);
            int p, q;
            pr.rows(r, &p, &q);
            const double v = t / in.v[3];
            stg(ax + p, v);
            if (q >= 0) stg(ax + q, -v);  // the partner's d equals p's
          });
    vpass(n,
          [&](int j) {
            return Dv<9>{{ldg(tk_start + m + j), ldg(tk_c + m + j), ldg(tk_q + m + j), ldg(I.e + j),
                          ldg(tk_start + m + n + j), ldg(tk_c + m + n + j), ldg(tk_q + m + n + j),
                          ldg(I.P + (int64_t)j * I.ld + j), ldg(I.x + j)}};
          },
          [&](int j, const Dv<9>& in) {
            const double a0 = in.v[0];
            stg(aty + j, (a0 + step * ((in.v[1] - tau_c * in.v[2]) - a0)) / in.v[3]);
            if (has_p) {
              const double p0 = in.v[4];
              stg(px + j, (p0 + step * ((in.v[5] - tau_c * in.v[6]) - p0)) + in.v[7] * in.v[8]);
            }
          });
    __syncthreads();
  }
  double l_xpx = 0.0, l_ctx = 0.0, l_px = 0.0, l_aty = 0.0, l_dres = 0.0, l_nx = 0.0;
  vpass(n,
        [&](int j) {
          return Dv<5>{{ldg(px + j), ldg(I.x + j), ldg(I.c + j), ldg(I.e + j), ldg(aty + j)}};
        },
        [&](int, const Dv<5>& in) {
          const double pxj = has_p ? in.v[0] : 0.0;
          const double xj = in.v[1], cj = in.v[2], ej = in.v[3], atyj = in.v[4];
          l_xpx += xj * pxj;
          l_ctx += cj * xj;
          const double pxo = pxj / ej;
          l_px = fmax(l_px, fabs(pxo));
          l_aty = fmax(l_aty, fabs(atyj));
          l_dres = fmax(l_dres, fabs((pxo + atyj) * inv_tau + cj / ej));
          l_nx = fmax(l_nx, fabs(ej * xj));
        });
  double l_bty = 0.0, l_ax = 0.0, l_s = 0.0, l_pres = 0.0, l_axs = 0.0, l_ny = 0.0;
  vpass(m,
        [&](int i) {
          return Dv<5>{{ldg(I.b + i), ldg(I.y + i), ldg(I.s + i), ldg(I.d + i), ldg(ax + i)}};
        },
        [&](int, const Dv<5>& in) {
          const double bi = in.v[0], yi = in.v[1], di = in.v[3], axi = in.v[4];
          l_bty += bi * yi;
          const double so = in.v[2] / di;
          l_ax = fmax(l_ax, fabs(axi));
          l_s = fmax(l_s, fabs(so));
          l_pres = fmax(l_pres, fabs((axi + so) * inv_tau - bi / di));
          l_axs = fmax(l_axs, fabs(axi + so));
          l_ny = fmax(l_ny, fabs(di * yi));
        });
  // three sums (slots 0-2) and eleven maxima, reduced together
  double v[14] = {l_xpx, l_ctx, l_bty, l_px, l_aty, l_dres, l_pres, l_ax,
                  l_s,   l_axs, l_nx,  l_ny, 0.0,  0.0};
  block_reduce_n<14, 0x3FF8u>(v, red);
  const double xpx = I.has_p ? v[0] : 0.0;
  const double ctx = v[1], bty = v[2];
  const double norm_px = v[3], norm_aty = v[4];
  const double dres = v[5], pres = v[6];
  const double norm_ax = v[7], norm_s = v[8], norm_axs = v[9];
  const double norm_x = v[10], norm_y = v[11];
  const double pcost = (ctx + 0.5 * xpx * inv_tau) * inv_tau;
  const double dcost = (-bty - 0.5 * xpx * inv_tau) * inv_tau;
  const double gap = fabs((ctx + bty + xpx * inv_tau) * inv_tau);
  const double dinf = fmax(norm_axs / (fabs(ctx) + kEps), norm_px / (fabs(ctx) + kEps));
  const double pinf = norm_aty / (fabs(bty) + kEps);
  const double prel = fmax(fmax(norm_ax * inv_tau, norm_s * inv_tau), I.normb);
  const double drel = fmax(fmax(norm_px * inv_tau, norm_aty * inv_tau), I.normc);
  int status = QTQP_BATCH_UNFINISHED;
  if (gap < o.atol + o.rtol * fmin(fabs(pcost), fabs(dcost)) && pres < o.atol + o.rtol * prel &&
      dres < o.atol + o.rtol * drel)
    status = QTQP_BATCH_SOLVED;
  else if (ctx < -kEps && dinf < o.atol_inf + o.rtol_inf * norm_x / fabs(ctx))
    status = QTQP_BATCH_UNBOUNDED;
  else if (bty < -kEps && pinf < o.atol_inf + o.rtol_inf * norm_y / fabs(bty))
    status = QTQP_BATCH_INFEASIBLE;
  if (tid == 0) {
    const int row = g_si[I_ITERS];
    if (row < kTraceRows) {
      double* tr = I.trace + (int64_t)row * kTraceFields;
      tr[T_MU] = mu;
      tr[T_SIGMA] = sigma;
      tr[T_ALPHA] = alpha;
      tr[T_TAU] = tau;
      tr[T_TAU_P] = tau_p;
      tr[T_TAU_C] = tau_c;
      tr[T_ALPHA_P] = alpha_p;
      tr[T_Q_SOLVES] = irq.solves + (irq.status == 2 ? 1000 : 0);
      tr[T_Q_RES] = irq.res;
      tr[T_P_SOLVES] = irp.solves + (irp.status == 2 ? 1000 : 0);
      tr[T_P_RES] = irp.res;
      tr[T_C_SOLVES] = irc.solves + (irc.status == 2 ? 1000 : 0);
      tr[T_C_RES] = irc.res;
      tr[T_PRES] = pres;
      tr[T_DRES] = dres;
      tr[T_GAP] = gap;
    }
    g_sd[S_TAU] = tau;
    g_sd[S_MU] = mu;
    g_sd[S_SIGMA] = sigma;
    g_sd[S_ALPHA] = alpha;
    g_sd[S_PRES] = pres;
    g_sd[S_DRES] = dres;
    g_sd[S_GAP] = gap;
    g_si[I_STATUS] = status;
    g_si[I_ITERS] += 1;
  }
  __syncthreads();
  PHASE(14);
}

// -------------------------------------------------------------------------
// Kernels
// -------------------------------------------------------------------------

// Per QP: transpose, Ruiz equilibration (__init__.py:616-663), FP32 copies,
// norms of the unscaled b/c, default start (__init__.py:299-307).
// Ruiz equilibration (__init__.py:616-669): 10 passes, each taking the row
// and column inf-norms of D A E and E P E (one pass over each) and scaling
// d by 1/sqrt(row norm), e by 1/sqrt(max(column norms)), clipped to
// [1e-3, 1e3].  Reads the unscaled padded copies I.A and I.P.
// Ruiz equilibration, 10 passes (__init__.py:129-137 via the reference's
// equilibrate): row maxima of |D A E| give D, column maxima of |D A E| and
// |E P E| give E, each factor clamped to [1e-3, 1e3].  Each pass streams A
// and P once with the pipelined row pass in its max-abs form.
//  The passes run over the stored rows of A; a row and its negated partner
//  have equal norms, so they share every factor.
template <typename S>
__device__ inline void ruiz(const Inst& I, double* scratch) {
  const int n = I.n, mc = I.mc, ld = I.ld, tid = threadIdx.x;
  const PairRegs pr(I.pm);
  const double* dv = wave_uniform(I.d);
  const double* ev = wave_uniform(I.e);
  double* um = wave_uniform(I.um);
  for (int pass = 0; pass < 10; ++pass) {
    rowpass2<S::NS, S::R, S::D, 1, 1, true, kPassChunk, true>(
        I.A, mc, n, ld, ev,
        [&](int r) {
          int p, q;
          pr.rows(r, &p, &q);
          return ldg(dv + p);
        },
        [](int, double v, double di, double* c1, double*, double* o) {
          *c1 = di;
          const double r = v * di;
          o[0] = fmin(fmax(1.0 / sqrt(r == 0.0 ? 1.0 : r), 1e-3), 1e3);
        },
        [&](int i, const double* o) { stg(um + i, o[0]); },
        [&](int j, double v, double) { I.un[j] = v * I.e[j]; }, scratch, I.ext);
    rowpass2<S::NS, S::R, S::D, 0, 1, true>(
        I.P, n, n, ld, ev, [&](int i) { return ldg(ev + i); },
        [](int, double, double ei, double* c1, double*, double*) { *c1 = ei; },
        [](int, const double*) {},
        [&](int j, double v, double) {
          const double cn = fmax(I.un[j], v * I.e[j]);
          I.un[j] = fmin(fmax(1.0 / sqrt(cn == 0.0 ? 1.0 : cn), 1e-3), 1e3);
        },
        scratch);
    vpass(mc,
          [&](int r) {
            int p, q;
            pr.rows(r, &p, &q);
            return Dv<3>{{ldg(um + r), ldg(dv + p), ldg(dv + (q >= 0 ? q : p))}};
          },
          [&](int r, const Dv<3>& in) {
            int p, q;
            pr.rows(r, &p, &q);
            stg(I.d + p, in.v[1] * in.v[0]);
            if (q >= 0) stg(I.d + q, in.v[2] * in.v[0]);
          });
    vpass(n, [&](int j) { return Dv<2>{{ldg(ev + j), ldg(I.un + j)}}; },
          [&](int j, const Dv<2>& in) { stg(I.e + j, in.v[0] * in.v[1]); });
    __syncthreads();
  }
}

// rows x n row-major src into dst with leading dimension ld (zeros in the
// pad): each wave copies 4 rows per round, each lane two columns of each
// per 128 columns, so 8 loads of a round are in flight together.
//  With pr, dst row r is src row p(r) (the stored rows of A).
__device__ inline void copy_padded(const double* __restrict__ src, double* __restrict__ dst,
                                   int rows, int n, int ld, const PairRegs* pr = nullptr) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i0 = 4 * wave; i0 < rows; i0 += 16)
    for (int j0 = 0; j0 < ld; j0 += 128) {
      double v[4][2];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int si = i0 + r, q;
        if (pr && si < rows) pr->rows(si, &si, &q);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = i0 + r, j = j0 + lane + 64 * h;
          v[r][h] = (i < rows && j < n) ? ldg(src + (int64_t)si * n + j) : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int i = i0 + r, j = j0 + lane + 64 * h;
          if (i < rows && j < ld) stg(dst + (int64_t)i * ld + j, v[r][h]);
        }
    }
}

// M[i][j] *= sl[i] * sr[j] in place (rows x ld, zero pad kept) and its FP32
// copy: 16-byte vectors, 4 rows per wave per round.
__device__ inline void scale_rows_cols(double* __restrict__ M, float* __restrict__ M32,
                                       const double* __restrict__ sl,
                                       const double* __restrict__ sr, int rows, int n, int ld) {
  using V = VecT<double, 2>::type;
  typedef float F2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < ld; j0 += 128) {
    const int j = j0 + 2 * lane;
    const double e0 = j < n ? ldg(sr + j) : 0.0, e1 = j + 1 < n ? ldg(sr + j + 1) : 0.0;
    for (int i0 = 4 * wave; i0 < rows; i0 += 16) {
      V a[4];
      double di[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = min(i0 + r, rows - 1);
        a[r] = j < ld ? ldg(reinterpret_cast<const V*>(M + (int64_t)i * ld + j)) : V(0);
        di[r] = ldg(sl + i);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r;
        if (i < rows && j < ld) {
          V v;
          v[0] = j < n ? a[r][0] * (di[r] * e0) : 0.0;
          v[1] = j + 1 < n ? a[r][1] * (di[r] * e1) : 0.0;
          stg(reinterpret_cast<V*>(M + (int64_t)i * ld + j), v);
          F2 f;
          f[0] = (float)v[0];
          f[1] = (float)v[1];
          stg(reinterpret_cast<F2*>(M32 + (int64_t)i * ld + j), f);
        }
      }
    }
  }
}

// Equilibration, scaled copies and start of every QP: a persistent grid
// whose workgroups take QPs from an atomic queue in `order` (longest first;
// null: index order), as k_ipm_solve does, so one workgroup does not end up
// with several of the largest.  `phase` selects the matrix work (PREP_MATRIX:
// d, e, the padded, equilibrated and FP32 copies of A and P), the vector work
// (PREP_VECTORS: scaled b and c, their norms, counters, start) or both.  A
// family batch, whose QPs share one A and P, runs the matrix work once for
// QP 0 and then the vector work for every QP.
enum { PREP_MATRIX = 1, PREP_VECTORS = 2, PREP_ALL = 3 };
__global__ __launch_bounds__(WG) void k_prepare(Inst* insts, const int* order, int count,
                                                int* queue, int equilibrate, int default_start,
                                                int force64, int phase) {
  // dynamic LDS: prepare_lds_bytes(nmax)
  extern __shared__ double smem[];
  double* red = smem;               // WG + 8
  double* scratch = smem + WG + 8;  // max(4 nmax, kPassChunk)
  __shared__ int s_q;
  for (;;) {
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int t = s_q;
    __syncthreads();
    if (t >= count) break;
    const int q = order ? order[t] : t;
    const Inst& I = insts[q];
    const int n = I.n, m = I.m, tid = threadIdx.x;
    if (phase & PREP_MATRIX) {
      for (int i = tid; i < m; i += WG) I.d[i] = 1.0;
      for (int j = tid; j < n; j += WG) I.e[j] = 1.0;
      // unscaled copies with the padded leading dimension (zeros in the pad)
      const int ld = I.ld, mc = I.mc;
      const PairRegs pr(I.pm);
      copy_padded(I.Araw, I.A, mc, n, ld, &pr);
      copy_padded(I.Praw, I.P, n, n, ld);
      __syncthreads();
      if (equilibrate) with_pass_shape(ld, [&](auto sh) { ruiz<decltype(sh)>(I, scratch); });
      __syncthreads();
      // D A E and E P E in place, and their FP32 copies (um: d of the stored rows)
      for (int r = tid; r < mc; r += WG) {
        int p, q;
        pr.rows(r, &p, &q);
        I.um[r] = I.d[p];
      }
      __syncthreads();
      scale_rows_cols(I.A, I.A32, I.um, I.e, mc, n, ld);
      scale_rows_cols(I.P, I.P32, I.e, I.e, n, n, ld);
      __syncthreads();
    }
    if (!(phase & PREP_VECTORS)) continue;
    double lb = 0.0, lc = 0.0;
    for (int i = tid; i < m; i += WG) {
      I.b[i] = I.braw[i] * I.d[i];
      lb = fmax(lb, fabs(I.braw[i]));
    }
    for (int j = tid; j < n; j += WG) {
      I.c[j] = I.craw[j] * I.e[j];
      lc = fmax(lc, fabs(I.craw[j]));
    }
    lb = block_max(lb, red + WG);
    lc = block_max(lc, red + WG);
    if (tid == 0) {
      insts[q].normb = lb;
      insts[q].normc = lc;
      for (int k = 0; k < S_NSCAL; ++k) I.sd[k] = (k == S_TAU) ? I.sd[k] : 0.0;
      I.si[I_STATUS] = QTQP_BATCH_UNFINISHED;
      I.si[I_ITERS] = 0;
      for (int k = 0; k < I_NINT; ++k) I.si[k] = 0;
      I.si[I_STATUS] = QTQP_BATCH_UNFINISHED;
      I.si[I_FORCE64] = force64;
    }
    if (default_start) {
      for (int j = tid; j < n; j += WG) I.x[j] = 0.0;
      for (int i = tid; i < m; i += WG) {
        I.y[i] = i < I.z ? 0.0 : 1.0 / I.d[i];
        I.s[i] = i < I.z ? 0.0 : 1.0 * I.d[i];
      }
      if (tid == 0) I.sd[S_TAU] = 1.0;
    } else {
      // caller's start in unscaled coordinates -> equilibrated (__init__.py:668-669)
      for (int j = tid; j < n; j += WG) I.x[j] = I.x[j] / I.e[j];
      for (int i = tid; i < m; i += WG) {
        I.y[i] = I.y[i] / I.d[i];
        I.s[i] = I.s[i] * I.d[i];
      }
    }
    for (int k = tid; k < n + m; k += WG) I.kq[k] = 0.0;
    __syncthreads();
  }
}

// Packed copies between the per-QP iterates and caller layouts.
// Dynamic LDS of k_prepare: reduction slots and the row pass's scratch
// (column maxima of nmax columns per wavefront, or the staged row outputs).
inline size_t prepare_lds_bytes(int nmax) {
  return sizeof(double) * (WG + 8 + std::max(4 * nmax, kPassChunk));
}

__global__ __launch_bounds__(WG) void k_gather(const Inst* insts, int count, const int64_t* n_off,
                                               const int64_t* m_off, double* x, double* y,
                                               double* s, double* scal, int* ints) {
  for (int q = blockIdx.x; q < count; q += gridDim.x) {
    const Inst& I = insts[q];
    for (int j = threadIdx.x; j < I.n; j += WG) x[n_off[q] + j] = I.e[j] * I.x[j];
    for (int i = threadIdx.x; i < I.m; i += WG) {
      y[m_off[q] + i] = I.d[i] * I.y[i];
      s[m_off[q] + i] = I.s[i] / I.d[i];
    }
    if (threadIdx.x < S_NSCAL) scal[(int64_t)q * S_NSCAL + threadIdx.x] = I.sd[threadIdx.x];
    if (threadIdx.x < I_NINT) ints[(int64_t)q * I_NINT + threadIdx.x] = I.si[threadIdx.x];
  }
}

__global__ __launch_bounds__(WG) void k_scatter_start(const Inst* insts, int count,
                                                      const int64_t* n_off, const int64_t* m_off,
                                                      const double* x, const double* y,
                                                      const double* s, const double* tau) {
  for (int q = blockIdx.x; q < count; q += gridDim.x) {
    const Inst& I = insts[q];
    for (int j = threadIdx.x; j < I.n; j += WG) I.x[j] = x[n_off[q] + j];
    for (int i = threadIdx.x; i < I.m; i += WG) {
      I.y[i] = y[m_off[q] + i];
      I.s[i] = s[m_off[q] + i];
    }
    if (threadIdx.x == 0) I.sd[S_TAU] = tau[q];
  }
}

// The solution of each QP as the reference returns it (__init__.py:580-614),
// from the packed unscaled iterates k_gather wrote (x = e x_eq, y = d y_eq,
// s = s_eq / d, not yet divided by tau): infeasible: x = s = NaN and y
// divided by |b'y|; unbounded: y = NaN, x and s divided by |c'x|; otherwise
// (solved, max_iter, error) x, y, s divided by tau.  b and c are the raw
// ones of the presolved QP.
__global__ __launch_bounds__(WG) void k_finish(const Inst* insts, int count, const int64_t* n_off,
                                               const int64_t* m_off, double* x, double* y,
                                               double* s, const double* scal, const int* ints) {
  __shared__ double red[8];
  for (int q = blockIdx.x; q < count; q += gridDim.x) {
    const Inst& I = insts[q];
    const int st = ints[(int64_t)q * I_NINT + I_STATUS];
    double* xq = x + n_off[q];
    double* yq = y + m_off[q];
    double* sq = s + m_off[q];
    const int n = I.n, m = I.m, tid = threadIdx.x;
    if (st == QTQP_BATCH_INFEASIBLE) {
      double acc = 0.0;
      for (int i = tid; i < m; i += WG) acc += I.braw[i] * yq[i];
      const double k = fabs(block_sum(acc, red));
      for (int j = tid; j < n; j += WG) xq[j] = NAN;
      for (int i = tid; i < m; i += WG) {
        yq[i] = yq[i] / k;
        sq[i] = NAN;
      }
    } else if (st == QTQP_BATCH_UNBOUNDED) {
      double acc = 0.0;
      for (int j = tid; j < n; j += WG) acc += I.craw[j] * xq[j];
      const double k = fabs(block_sum(acc, red));
      for (int j = tid; j < n; j += WG) xq[j] = xq[j] / k;
      for (int i = tid; i < m; i += WG) {
        yq[i] = NAN;
        sq[i] = sq[i] / k;
      }
    } else {
      const double tau = scal[(int64_t)q * S_NSCAL + S_TAU];
      for (int j = tid; j < n; j += WG) xq[j] = xq[j] / tau;
      for (int i = tid; i < m; i += WG) {
        yq[i] = yq[i] / tau;
        sq[i] = sq[i] / tau;
      }
    }
    __syncthreads();
  }
}

__global__ void k_reset_queue(int* queue) {
  if (threadIdx.x == 0) *queue = 0;
}

// clk[count] = the earliest workgroup start of the solve (set to the maximum
// here, lowered with atomicMin); clk[q] = wall clock when QP q finished.
__global__ void k_reset_clock(unsigned long long* clk, int count) {
  const int i = blockIdx.x * WG + threadIdx.x;
  if (i < count) clk[i] = 0;
  else if (i == count) clk[count] = ~0ull;
}

// The whole solve in one persistent launch: each workgroup takes the next QP
// in longest-first order and runs its IPM iterations to completion (a final
// status or max_iter), then takes the next.  No workgroup waits on another,
// and with the largest QPs first the launch ends on small ones.
__global__ __launch_bounds__(WG, 2) void k_ipm_solve(const Inst* insts, const int* order, int count,
                                                  int* queue, Opts o, Counters* cnt,
                                                  unsigned long long* clk) {
  Smem sm = lds_smem();
  __shared__ int s_q;
  if (threadIdx.x == 0) atomicMin(&clk[count], (unsigned long long)wall_clock64());
  for (;;) {
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_q;
    __syncthreads();
    if (q >= count) break;
    const int id = order[q];
    const Inst& I = insts[id];
    qp_take(I);
    bool ran = false;
    while (g_si[I_STATUS] == QTQP_BATCH_UNFINISHED && g_si[I_ITERS] < o.max_iter) {
      ipm_iteration(I, o, sm, cnt);
      ran = true;
    }
    if (threadIdx.x == 0 && ran) {
      atomicAdd(&cnt->done, 1);
      atomicMax(&cnt->max_iters, g_si[I_ITERS]);
      clk[id] = wall_clock64();
    }
    qp_release(I, cnt);
  }
}

__global__ __launch_bounds__(WG, 2) void k_ipm_iter(const Inst* insts, const int* order, int count,
                                                 int* queue, Opts o, int nmax, Counters* cnt,
                                                 unsigned long long* clk) {
  Smem sm = lds_smem();
  (void)nmax;
  if (clk && threadIdx.x == 0) atomicMin(&clk[count], (unsigned long long)wall_clock64());
  __shared__ int s_q;
  for (;;) {
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_q;
    __syncthreads();
    if (q >= count) break;
    const Inst& I = insts[order[q]];
    qp_take(I);
    if (g_si[I_STATUS] != QTQP_BATCH_UNFINISHED || g_si[I_ITERS] >= o.max_iter) {
      __syncthreads();  // (everyone has read g_si before the next QP's qp_take)
      continue;
    }
    ipm_iteration(I, o, sm, cnt);
    if (threadIdx.x == 0 &&
        (g_si[I_STATUS] != QTQP_BATCH_UNFINISHED || g_si[I_ITERS] >= o.max_iter)) {
      atomicAdd(&cnt->done, 1);
      atomicMax(&cnt->max_iters, g_si[I_ITERS]);
      if (clk) clk[order[q]] = wall_clock64();
    }
    qp_release(I, cnt);
  }
}

// The autotuner's timing kernel: `reps` Gram products of each QP in the
// queue, in the QP's current tile shape for T.  The KKT diagonal is set to
// d_x = 1, d_y = -1 first; the next solve rewrites it and G before use.
template <typename T>
__global__ __launch_bounds__(WG, 2) void k_gram_bench(const Inst* insts, int count, int* queue,
                                                      int reps) {
  Smem sm = lds_smem();
  __shared__ int s_q;
  for (;;) {
    if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_q;
    __syncthreads();
    if (q >= count) break;
    const Inst& I = insts[q];
    for (int i = threadIdx.x; i < I.n + I.m; i += WG) I.dg[i] = i < I.n ? 1.0 : -1.0;
    __syncthreads();
    for (int r = 0; r < reps; ++r) {
      if (sizeof(T) == 4)
        gram_syrk<float>(I, I.A32, I.P32, reinterpret_cast<float*>(I.G), sm);
      else
        gram_syrk<double>(I, I.A, I.P, I.G, sm);
    }
    __syncthreads();
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct qtqp_batch {
  int count = 0, nmax = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> allocs;
  std::vector<Inst> hinst;
  Inst* dinst = nullptr;
  int* dorder = nullptr;
  int* dqueue = nullptr;
  Counters* dcnt = nullptr;
  bool have_start = false;
  std::vector<int64_t> n_off, m_off;
  std::vector<int> n, m;
  std::vector<int> mc;  // stored rows of A per QP (row pairs kept once)
  std::vector<int64_t> row_bytes;  // FP64 bytes a pass over the stored rows' extents reads
  int64_t tot_n = 0, tot_m = 0;
  int64_t *d_noff = nullptr, *d_moff = nullptr;
  double *d_px = nullptr, *d_py = nullptr, *d_ps = nullptr, *d_scal = nullptr;
  double* d_tau = nullptr;  // packed tau of a start (count values)
  int* d_ints = nullptr;
  size_t smem_iter = 0;
  std::vector<hipEvent_t> iter_ev;  // per-iteration launch timing
  std::vector<int> order;           // longest-first queue order
  std::vector<double> iter_ms;
  unsigned long long* dclk = nullptr;  // per-QP completion clock, [count] = solve start
  std::vector<double> inst_ms;        // per-QP completion time of the last solve
  bool family = false;  // every QP shares QP 0's A and P (qtqp_batch_create_family)
  bool fuse_solves = true;  // see QTQP_BATCH_NO_FUSED_SOLVES
  double *d_braw = nullptr, *d_craw = nullptr;  // raw b and c of all QPs, packed by m_off / n_off
};

namespace {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPC(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(QTQP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

Opts to_opts(const qtqp_batch_opts* opts) {
  Opts o{};
  o.atol = opts->atol;
  o.rtol = opts->rtol;
  o.atol_inf = opts->atol_infeas;
  o.rtol_inf = opts->rtol_infeas;
  o.step_scale = opts->step_size_scale;
  o.min_reg = opts->min_static_regularization;
  o.ls_atol = opts->linear_solver_atol;
  o.ls_rtol = opts->linear_solver_rtol;
  o.max_iter = opts->max_iter;
  o.max_ir = opts->max_iterative_refinement_steps;
  o.precision = opts->precision;
  return o;
}

// Autotuner size buckets by n (upper edges; the last covers n <= 1024).
int size_bucket(int n) {
  static const int edge[QTQP_BATCH_NBUCKETS] = {64, 128, 192, 256, 384, 512, 768, 1024};
  for (int b = 0; b < QTQP_BATCH_NBUCKETS; ++b)
    if (n <= edge[b]) return b;
  return QTQP_BATCH_NBUCKETS - 1;
}

// The row pairs of one QP's A (m x n row-major): rows that are exact
// negations of each other, found by hashing each row with its sign fixed by
// its first nonzero and comparing candidates exactly.  The stored rows are the
// rows without a partner and the first of each pair, in row order; pm maps
// them back as a few segments of consecutive rows.  Returns the number of
// stored rows (m, with the identity map, when pairing is off or the pairs do
// not fit in kMaxSeg segments).
int find_row_pairs(const double* a, int m, int n, bool enable, PairMap* pm) {
  auto identity = [&]() {
    pm->nseg = 1;
    for (int k = 0; k <= kMaxSeg; ++k) pm->start[k] = k == 0 ? 0 : m;
    for (int k = 0; k < kMaxSeg; ++k) {
      pm->dp[k] = 0;
      pm->dq[k] = kNoPartner;
    }
    return m;
  };
  if (!enable) return identity();
  struct Key {
    uint64_t h;
    int neg, row;
  };
  std::vector<Key> keys(m);
  for (int i = 0; i < m; ++i) {
    const double* r = a + (int64_t)i * n;
    int f = 0;
    while (f < n && r[f] == 0.0) ++f;
    const double sg = (f < n && r[f] < 0.0) ? -1.0 : 1.0;
    uint64_t h = 1469598103934665603ull;
    for (int j = f; j < n; ++j) {
      const double v = sg * r[j] + 0.0;  // -0 -> +0
      uint64_t bits;
      std::memcpy(&bits, &v, 8);
      h = (h ^ bits) * 1099511628211ull;
      h ^= h >> 29;
    }
    keys[i] = {h, sg < 0.0 ? 1 : 0, i};
  }
  std::sort(keys.begin(), keys.end(), [](const Key& u, const Key& v) {
    return u.h != v.h ? u.h < v.h : u.row < v.row;
  });
  std::vector<int> partner(m, -1);
  for (size_t b0 = 0; b0 < keys.size();) {
    size_t b1 = b0;
    while (b1 < keys.size() && keys[b1].h == keys[b0].h) ++b1;
    for (size_t u = b0; u < b1; ++u) {
      if (keys[u].neg || partner[keys[u].row] >= 0) continue;
      const double* ru = a + (int64_t)keys[u].row * n;
      for (size_t v = b0; v < b1; ++v) {
        if (!keys[v].neg || partner[keys[v].row] >= 0) continue;
        const double* rv = a + (int64_t)keys[v].row * n;
        bool neg = true;
        for (int j = 0; j < n && neg; ++j) neg = ru[j] == -rv[j];
        if (neg) {
          partner[keys[u].row] = keys[v].row;
          partner[keys[v].row] = keys[u].row;
          break;
        }
      }
    }
    b0 = b1;
  }
  // stored rows and their segments
  int nseg = 0, r = 0, pdp = 0, pdq = 0;
  PairMap out{};
  for (int i = 0; i < m; ++i) {
    const int q = partner[i];
    if (q >= 0 && q < i) continue;  // the second row of a pair
    const int dp = i - r, dq = q >= 0 ? q - r : kNoPartner;
    if (nseg == 0 || dp != pdp || dq != pdq) {
      if (nseg == kMaxSeg) return identity();
      out.start[nseg] = r;
      out.dp[nseg] = dp;
      out.dq[nseg] = dq;
      ++nseg;
      pdp = dp;
      pdq = dq;
    }
    ++r;
  }
  out.nseg = nseg;
  for (int k = nseg; k <= kMaxSeg; ++k) out.start[k] = r;
  for (int k = nseg; k < kMaxSeg; ++k) {
    out.dp[k] = 0;
    out.dq[k] = kNoPartner;
  }
  *pm = out;
  return r;
}

// The nonzero column extent [lo, hi) of each stored row (of row p(r) of the
// m x n row-major a), packed as lo | hi << 16 (0 for a zero row), then
// kExtPad zero words.  Without `enable` every stored row gets [0, n).
void row_extents(const double* a, int n, int mc, const PairMap& pm, bool enable,
                 std::vector<uint32_t>* out) {
  out->assign((size_t)mc + kExtPad, 0u);
  for (int r = 0; r < mc; ++r) {
    const int k = (r >= pm.start[1]) + (r >= pm.start[2]) + (r >= pm.start[3]);
    const double* row = a + (int64_t)(r + pm.dp[k]) * n;
    int lo = 0, hi = n;
    if (enable) {
      while (lo < n && row[lo] == 0.0) ++lo;
      while (hi > lo && row[hi - 1] == 0.0) --hi;
      if (lo == hi) lo = hi = 0;
    }
    (*out)[r] = (uint32_t)lo | (uint32_t)hi << 16;
  }
}

// The union of the row extents of each 32-row chunk of the mc stored rows
// (zero rows left out; 0 for a chunk of zero rows), then zero words up to
// mc / 32 + 4 entries.
std::vector<uint32_t> chunk_extents(const std::vector<uint32_t>& ext, int mc) {
  std::vector<uint32_t> out((size_t)mc / 32 + 4, 0u);
  for (int c = 0; c * 32 < mc; ++c) {
    int lo = 1 << 16, hi = 0;
    for (int r = 32 * c; r < std::min(mc, 32 * c + 32); ++r) {
      const int l = ext[r] & 0xffff, h = ext[r] >> 16;
      if (h > l) {
        lo = std::min(lo, l);
        hi = std::max(hi, h);
      }
    }
    out[c] = hi > 0 ? ((uint32_t)lo | (uint32_t)hi << 16) : 0u;
  }
  return out;
}

// Bump allocator over one device allocation.
struct Arena {
  unsigned char* base = nullptr;
  size_t used = 0;
  template <typename T>
  T* take(int64_t count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += align256(sizeof(T) * (size_t)std::max<int64_t>(count, 1));
    return p;
  }
};

// Device bytes of one QP's matrix data (A, P, their copies, extents, d, e),
// held once per family batch, and of its own vectors and workspace.
size_t matrix_bytes(int n, int m, int mc) {
  const int64_t ld = (n + 3) & ~3;
  const int64_t mn = (int64_t)m * ld, mcn = (int64_t)mc * ld, nn = (int64_t)n * ld;
  size_t b = 0;
  b += align256(8 * mn);           // raw A
  b += align256(8 * (mcn + a_pad((int)ld)));  // stored rows of A (FP64) and read slack
  b += align256(4 * mcn);          // A32
  b += align256(4 * ((int64_t)mc + kExtPad));  // row extents
  b += align256(4 * ((int64_t)mc / 32 + 4));     // chunk extents
  b += align256(8 * nn) + align256(8 * (nn + a_pad((int)ld)));  // raw P, P and its read slack
  b += align256(4 * nn);           // P32
  b += align256(8 * m) + align256(8 * n);  // d, e
  return b;
}

size_t vector_bytes(int n, int m) {
  const int64_t N = n + m, nn = (int64_t)n * n;
  size_t b = 0;
  b += align256(8 * nn);           // G
  b += align256(8 * 2 * 64 * 64 * ((n + 63) / 64));  // diagonal-block inverses
  b += align256(8 * m) + align256(8 * n);  // b, c
  b += 3 * align256(8 * std::max<int64_t>(m, n));  // x/y/s slack (x n, y m, s m)
  b += 10 * align256(8 * N);                       // kq dg corr krhs sol res iacc ires r xp
  b += 2 * align256(8 * std::max<int64_t>(m, n)) * 2;  // um un ds cm
  b += align256(8 * n);                                // pkq
  b += 2 * align256(8 * m);                            // wy, ey
  b += align256(8 * 3 * ((int64_t)m + 2 * (int64_t)n));  // tk
  b += align256(8 * (2 * (int64_t)m + 6 * (int64_t)n));  // refinement state
  b += 6 * align256(8 * N) + align256(8 * (2 * (int64_t)m + 6 * (int64_t)n)) +
       align256(8 * n);  // the second solve's r2 sol2 krhs2 res2 ires2 iacc2, irb2, um2
  b += align256(8 * S_NSCAL) + align256(4 * I_NINT);
  b += align256(8 * kTraceFields * kTraceRows);
  return b;
}

// k_prepare over `count` QPs of `insts` (device table) in `order`: the matrix
// and vector work of each QP, or, for a family batch, the shared matrix work
// once (QP 0 of the handle's table) and then the vector work of each QP.
void launch_prepare(const qtqp_batch* h, Inst* insts, const int* order, int count,
                    int equilibrate, int default_start, int force64, int cus) {
  const size_t lds = prepare_lds_bytes(h->nmax);
  const int grid = std::min(count, std::max(1, cus * 2));
  if (h->family) {
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(WG), lds, h->stream, h->dinst, nullptr, 1,
                       h->dqueue, equilibrate, default_start, force64, (int)PREP_MATRIX);
  }
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
  hipLaunchKernelGGL(k_prepare, dim3(grid), dim3(WG), lds, h->stream, insts, order, count,
                     h->dqueue, equilibrate, default_start, force64,
                     (int)(h->family ? PREP_VECTORS : PREP_ALL));
}

// Packs the QPs onto the device.  With `family`, every QP has QP 0's A and
// P (a_off = p_off = 0): they are uploaded, paired and equilibrated once and
// every QP's table entry points at that one copy.
int create_impl(int32_t count, const int32_t* n, const int32_t* m, const int32_t* z,
                const int64_t* a_off, const int64_t* p_off, const int64_t* n_off,
                const int64_t* m_off, const double* a, const double* p, const double* b,
                const double* c, int32_t flags, bool family, qtqp_batch** out) {
  if (!out) return fail(QTQP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (count < 1) return fail(QTQP_ERR_ARG, "count must be >= 1");
  if (!n || !m || !z || !a_off || !p_off || !n_off || !m_off || !a || !p || !b || !c)
    return fail(QTQP_ERR_ARG, "NULL argument");
  int nmax = 0;
  for (int q = 0; q < count; ++q) {
    if (n[q] < 1 || m[q] < 1 || z[q] < 0 || z[q] >= m[q])
      return fail(QTQP_ERR_ARG, "instance " + std::to_string(q) + ": need n>=1, 0<=z<m");
    if (n[q] > 1024) return fail(QTQP_ERR_ARG, "batched engine supports n <= 1024");
    nmax = std::max(nmax, (int)n[q]);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(QTQP_ERR_HIP, "no HIP device available");
  auto* h = new qtqp_batch();
  auto cleanup = [&](int rc) {
    for (void* ptr : h->allocs) (void)hipFree(ptr);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return rc;
  };
  h->count = count;
  h->nmax = nmax;
  h->n.assign(n, n + count);
  h->m.assign(m, m + count);
  h->n_off.assign(n_off, n_off + count);
  h->m_off.assign(m_off, m_off + count);
  if (hipStreamCreate(&h->stream) != hipSuccess) return cleanup(fail(QTQP_ERR_HIP, "stream"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, "event"));
  // row pairs and row extents of each QP, found on host threads
  std::vector<PairMap> pmap(count);
  std::vector<int> mcs(count);
  std::vector<std::vector<uint32_t>> exts(count);
  {
    h->fuse_solves = (flags & QTQP_BATCH_NO_FUSED_SOLVES) == 0;
    const bool pairs = (flags & QTQP_BATCH_NO_ROW_PAIRS) == 0;
    const bool extents = (flags & QTQP_BATCH_NO_ROW_EXTENTS) == 0;
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    const int distinct = family ? 1 : count;
    auto work = [&]() {
      for (int q; (q = next.fetch_add(1)) < distinct;) {
        mcs[q] = find_row_pairs(a + a_off[q], m[q], n[q], pairs, &pmap[q]);
        row_extents(a + a_off[q], n[q], mcs[q], pmap[q], extents, &exts[q]);
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < std::min(nt, distinct); ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (int q = distinct; q < count; ++q) {
      mcs[q] = mcs[0];
      pmap[q] = pmap[0];
    }
  }
  h->mc = mcs;
  h->family = family;
  h->row_bytes.assign(count, 0);
  for (int q = 0; q < count; ++q) {
    const std::vector<uint32_t>& ex = exts[family ? 0 : q];
    for (int r = 0; r < mcs[q]; ++r) {
      const int lo = ex[r] & 0xffff, hi = ex[r] >> 16;
      if (hi > lo) h->row_bytes[q] += 16 * ((hi + 1) / 2 - lo / 2);
    }
  }
  for (int q = 0; q < count; ++q)
    if ((int64_t)mcs[q] * ((n[q] + 3) & ~3) * 8 >= (int64_t)1 << 31)
      return cleanup(fail(QTQP_ERR_ARG, "instance " + std::to_string(q) +
                                            ": A exceeds 2 GiB (batched engine limit)"));
  int64_t tot_mn = 0, tot_nn = 0, tot_n = 0, tot_m = 0;
  for (int q = 0; q < count; ++q) {
    tot_mn = std::max(tot_mn, a_off[q] + (int64_t)m[q] * n[q]);
    tot_nn = std::max(tot_nn, p_off[q] + (int64_t)n[q] * n[q]);
    tot_n = std::max(tot_n, n_off[q] + n[q]);
    tot_m = std::max(tot_m, m_off[q] + m[q]);
  }
  size_t total = align256(8 * (size_t)std::max<int64_t>(tot_m, 1)) +
                 align256(8 * (size_t)std::max<int64_t>(tot_n, 1)) + 4096;
  for (int q = 0; q < count; ++q)
    total += (family && q > 0 ? 0 : matrix_bytes(n[q], m[q], mcs[q])) + vector_bytes(n[q], m[q]);
  void* mem = nullptr;
  hipError_t err = hipMalloc(&mem, total);
  if (err != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, std::string("hipMalloc batch arena: ") + hipGetErrorString(err)));
  h->allocs.push_back(mem);
  Arena ar{static_cast<unsigned char*>(mem), 0};
  h->hinst.resize(count);
  // raw b and c of every QP, packed as the caller packs them
  h->d_braw = ar.take<double>(std::max<int64_t>(tot_m, 1));
  h->d_craw = ar.take<double>(std::max<int64_t>(tot_n, 1));
  if (hipMemcpyAsync(h->d_braw, b, 8 * tot_m, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(h->d_craw, c, 8 * tot_n, hipMemcpyHostToDevice, h->stream) != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, "upload failed"));
  for (int q = 0; q < count; ++q) {
    Inst& I = h->hinst[q];
    const int nq = n[q], mq = m[q];
    const int64_t N = nq + mq, mn = (int64_t)mq * nq, nn = (int64_t)nq * nq;
    const int64_t mx = std::max<int64_t>(mq, nq);
    I.n = nq;
    I.m = mq;
    I.z = z[q];
    I.mc = mcs[q];
    I.pm = pmap[q];
    I.ld = (nq + 3) & ~3;
    I.tile = QTQP_BATCH_TILE_64;
    I.tile64 = QTQP_BATCH_TILE_64;
    if (family && q > 0) {
      // QP 0's matrix data
      const Inst& T = h->hinst[0];
      I.has_p = T.has_p;
      I.Araw = T.Araw;
      I.A = T.A;
      I.A32 = T.A32;
      I.ext = T.ext;
      I.cext = T.cext;
      I.Praw = T.Praw;
      I.P = T.P;
      I.P32 = T.P32;
      I.d = T.d;
      I.e = T.e;
    } else {
      const int64_t mld = (int64_t)I.mc * I.ld, nld = (int64_t)nq * I.ld;
      I.has_p = 0;
      for (int64_t k = 0; k < nn && !I.has_p; ++k) I.has_p = p[p_off[q] + k] != 0.0;
      double* araw = ar.take<double>(mn);
      I.Araw = araw;
      I.A = ar.take<double>(mld + a_pad(I.ld));
      I.A32 = ar.take<float>(mld);
      uint32_t* ext = ar.take<uint32_t>((int64_t)I.mc + kExtPad);
      I.ext = ext;
      const std::vector<uint32_t> cx = chunk_extents(exts[q], I.mc);
      uint32_t* cext = ar.take<uint32_t>((int64_t)cx.size());
      I.cext = cext;
      double* praw = ar.take<double>(nn);
      I.Praw = praw;
      I.P = ar.take<double>(nld + a_pad(I.ld));
      I.P32 = ar.take<float>(nld);
      I.d = ar.take<double>(mq);
      I.e = ar.take<double>(nq);
      if (hipMemcpyAsync(araw, a + a_off[q], 8 * mn, hipMemcpyHostToDevice, h->stream) !=
              hipSuccess ||
          hipMemcpyAsync(praw, p + p_off[q], 8 * nn, hipMemcpyHostToDevice, h->stream) !=
              hipSuccess ||
          hipMemcpyAsync(ext, exts[q].data(), 4 * exts[q].size(), hipMemcpyHostToDevice,
                         h->stream) != hipSuccess ||
          hipMemcpyAsync(cext, cx.data(), 4 * cx.size(), hipMemcpyHostToDevice, h->stream) !=
              hipSuccess)
        return cleanup(fail(QTQP_ERR_HIP, "upload failed"));
    }
    I.G = ar.take<double>(nn);
    I.LiB = ar.take<double>((int64_t)2 * NB * NB * ((nq + NB - 1) / NB));
    I.braw = h->d_braw + m_off[q];
    I.craw = h->d_craw + n_off[q];
    I.b = ar.take<double>(mq);
    I.c = ar.take<double>(nq);
    I.x = ar.take<double>(mx);
    I.y = ar.take<double>(mx);
    I.s = ar.take<double>(mx);
    I.kq = ar.take<double>(N);
    I.dg = ar.take<double>(N);
    I.corr = ar.take<double>(N);
    I.krhs = ar.take<double>(N);
    I.sol = ar.take<double>(N);
    I.res = ar.take<double>(N);
    I.iacc = ar.take<double>(N);
    I.ires = ar.take<double>(N);
    I.r = ar.take<double>(N);
    I.xp = ar.take<double>(N);
    I.um = ar.take<double>(mx);
    I.un = ar.take<double>(mx);
    I.ds = ar.take<double>(mx);
    I.cm = ar.take<double>(mx);
    I.pkq = ar.take<double>(nq);
    I.wy = ar.take<double>(mq);
    I.ey = ar.take<double>(mq);
    I.tk = ar.take<double>(3 * ((int64_t)mq + 2 * (int64_t)nq));
    I.irb = ar.take<double>(2 * (int64_t)mq + 6 * (int64_t)nq);
    I.r2 = ar.take<double>(N);
    I.sol2 = ar.take<double>(N);
    I.krhs2 = ar.take<double>(N);
    I.res2 = ar.take<double>(N);
    I.ires2 = ar.take<double>(N);
    I.iacc2 = ar.take<double>(N);
    I.irb2 = ar.take<double>(2 * (int64_t)mq + 6 * (int64_t)nq);
    I.um2 = ar.take<double>(nq);
    I.sd = ar.take<double>(S_NSCAL);
    I.si = ar.take<int>(I_NINT);
    I.trace = ar.take<double>(kTraceFields * kTraceRows);
    I.normb = I.normc = 0.0;
  }
  if (ar.used > total) return cleanup(fail(QTQP_ERR_HIP, "arena overflow"));
  h->tot_n = tot_n;
  h->tot_m = tot_m;
  if (hipMalloc(&h->dinst, sizeof(Inst) * count) != hipSuccess ||
      hipMalloc(&h->dorder, sizeof(int) * count) != hipSuccess ||
      hipMalloc(&h->dqueue, sizeof(int)) != hipSuccess ||
      hipMalloc(&h->dcnt, sizeof(Counters)) != hipSuccess ||
      hipMalloc(&h->d_noff, sizeof(int64_t) * count) != hipSuccess ||
      hipMalloc(&h->d_moff, sizeof(int64_t) * count) != hipSuccess ||
      hipMalloc(&h->d_px, sizeof(double) * std::max<int64_t>(tot_n, 1)) != hipSuccess ||
      hipMalloc(&h->d_py, sizeof(double) * std::max<int64_t>(tot_m, 1)) != hipSuccess ||
      hipMalloc(&h->d_ps, sizeof(double) * std::max<int64_t>(tot_m, 1)) != hipSuccess ||
      hipMalloc(&h->d_scal, sizeof(double) * S_NSCAL * count) != hipSuccess ||
      hipMalloc(&h->d_tau, sizeof(double) * count) != hipSuccess ||
      hipMalloc(&h->d_ints, sizeof(int) * I_NINT * count) != hipSuccess ||
      hipMalloc(&h->dclk, sizeof(unsigned long long) * (count + 1)) != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, "hipMalloc tables"));
  for (void* ptr : {(void*)h->dinst, (void*)h->dorder, (void*)h->dqueue, (void*)h->dcnt,
                    (void*)h->d_noff, (void*)h->d_moff, (void*)h->d_px, (void*)h->d_py,
                    (void*)h->d_ps, (void*)h->d_scal, (void*)h->d_tau, (void*)h->d_ints,
                    (void*)h->dclk})
    h->allocs.push_back(ptr);
  if (hipMemcpyAsync(h->d_noff, n_off, sizeof(int64_t) * count, hipMemcpyHostToDevice,
                     h->stream) != hipSuccess ||
      hipMemcpyAsync(h->d_moff, m_off, sizeof(int64_t) * count, hipMemcpyHostToDevice,
                     h->stream) != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, "offset upload"));
  // longest-processing-time order: descending estimated iteration cost, the
  // Gram product and factor (m n^2 + n^3/3) plus the ~20 passes over A a
  // bandwidth-bound iteration makes (weighted as 64 m n)
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int u) {
    return (double)m[u] * n[u] * (n[u] + 64.0) + (double)n[u] * n[u] * n[u] / 3.0;
  };
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return cost(u) > cost(v); });
  h->order = order;
  if (hipMemcpyAsync(h->dorder, order.data(), sizeof(int) * count, hipMemcpyHostToDevice,
                     h->stream) != hipSuccess ||
      hipMemcpyAsync(h->dinst, h->hinst.data(), sizeof(Inst) * count, hipMemcpyHostToDevice,
                     h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return cleanup(fail(QTQP_ERR_HIP, "table upload"));
  h->smem_iter = 0;  // the IPM kernel's LDS is static (g_lds)
  *out = h;
  return QTQP_OK;
}


}  // namespace

extern "C" {

const char* qtqp_batch_last_error(void) { return g_err.c_str(); }

int qtqp_batch_create(int32_t count, const int32_t* n, const int32_t* m, const int32_t* z,
                      const int64_t* a_off, const int64_t* p_off, const int64_t* n_off,
                      const int64_t* m_off, const double* a, const double* p, const double* b,
                      const double* c, qtqp_batch** out) {
  return qtqp_batch_create_ex(count, n, m, z, a_off, p_off, n_off, m_off, a, p, b, c, 0, out);
}

int qtqp_batch_create_ex(int32_t count, const int32_t* n, const int32_t* m, const int32_t* z,
                         const int64_t* a_off, const int64_t* p_off, const int64_t* n_off,
                         const int64_t* m_off, const double* a, const double* p,
                         const double* b, const double* c, int32_t flags, qtqp_batch** out) {
  return create_impl(count, n, m, z, a_off, p_off, n_off, m_off, a, p, b, c, flags, false, out);
}

int qtqp_batch_create_family(int32_t count, int32_t n, int32_t m, int32_t z, const double* a,
                             const double* p, const double* b, const double* c, int32_t flags,
                             qtqp_batch** out) {
  if (!out) return fail(QTQP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (count < 1) return fail(QTQP_ERR_ARG, "count must be >= 1");
  std::vector<int32_t> nv(count, n), mv(count, m), zv(count, z);
  std::vector<int64_t> zero(count, 0), no(count), mo(count);
  for (int q = 0; q < count; ++q) {
    no[q] = (int64_t)q * n;
    mo[q] = (int64_t)q * m;
  }
  return create_impl(count, nv.data(), mv.data(), zv.data(), zero.data(), zero.data(), no.data(),
                     mo.data(), a, p, b, c, flags, true, out);
}

int qtqp_batch_set_bc(qtqp_batch* h, const double* b, const double* c) {
  if (!h || !b || !c) return fail(QTQP_ERR_ARG, "NULL argument");
  HIPC(hipMemcpyAsync(h->d_braw, b, 8 * h->tot_m, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->d_craw, c, 8 * h->tot_n, hipMemcpyHostToDevice, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return QTQP_OK;
}

int qtqp_batch_set_start(qtqp_batch* h, const double* x, const double* y, const double* s,
                         const double* tau) {
  if (!h || !x || !y || !s || !tau) return fail(QTQP_ERR_ARG, "NULL argument");
  HIPC(hipMemcpyAsync(h->d_px, x, 8 * h->tot_n, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->d_py, y, 8 * h->tot_m, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->d_ps, s, 8 * h->tot_m, hipMemcpyHostToDevice, h->stream));
  HIPC(hipMemcpyAsync(h->d_tau, tau, 8 * h->count, hipMemcpyHostToDevice, h->stream));
  return qtqp_batch_set_start_device(h, h->d_px, h->d_py, h->d_ps, h->d_tau);
}

int qtqp_batch_set_start_device(qtqp_batch* h, const double* x, const double* y,
                                const double* s, const double* tau) {
  if (!h || !x || !y || !s || !tau) return fail(QTQP_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(k_scatter_start, dim3(std::min(h->count, 2048)), dim3(WG), 0, h->stream,
                     h->dinst, h->count, h->d_noff, h->d_moff, x, y, s, tau);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(h->stream));
  h->have_start = true;
  return QTQP_OK;
}

int qtqp_batch_device_buffers(qtqp_batch* h, double** x, double** y, double** s, double** tau) {
  if (!h || !x || !y || !s || !tau) return fail(QTQP_ERR_ARG, "NULL argument");
  *x = h->d_px;
  *y = h->d_py;
  *s = h->d_ps;
  *tau = h->d_tau;
  return QTQP_OK;
}

int qtqp_batch_solve(qtqp_batch* h, const qtqp_batch_opts* opts, qtqp_batch_counters* counters) {
  if (!h || !opts) return fail(QTQP_ERR_ARG, "NULL argument");
  if (opts->max_iter < 1 || opts->max_iterative_refinement_steps < 1)
    return fail(QTQP_ERR_ARG, "max_iter and max_iterative_refinement_steps must be >= 1");
  Opts o = to_opts(opts);
  o.fuse = h->fuse_solves ? 1 : 0;
  HIPC(hipEventRecord(h->ev0, h->stream));
  HIPC(hipMemsetAsync(h->dcnt, 0, sizeof(Counters), h->stream));
  int dev = 0, cus = 0;
  HIPC(hipGetDevice(&dev));
  HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  launch_prepare(h, h->dinst, h->dorder, h->count, opts->equilibrate ? 1 : 0,
                 h->have_start ? 0 : 1, o.precision == QTQP_PREC_FP64 ? 1 : 0, cus);
  HIPC(hipGetLastError());
  const int grid = std::min(h->count, std::max(1, cus * 2));
  hipLaunchKernelGGL(k_reset_clock, dim3(std::max(1, (h->count + 1 + WG - 1) / WG)), dim3(WG), 0,
                     h->stream, h->dclk, h->count);
  HIPC(hipGetLastError());
  while ((int)h->iter_ev.size() < o.max_iter + 1) {
    hipEvent_t ev;
    HIPC(hipEventCreate(&ev));
    h->iter_ev.push_back(ev);
  }
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int it = 0, launched = 0;
  if (!opts->per_iteration) {
    // the whole solve in one persistent launch (k_ipm_solve)
    if (opts->use_graph) {
      HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
      hipLaunchKernelGGL(k_ipm_solve, dim3(grid), dim3(WG), h->smem_iter, h->stream, h->dinst,
                         h->dorder, h->count, h->dqueue, o, h->dcnt, h->dclk);
      HIPC(hipStreamEndCapture(h->stream, &graph));
      HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    HIPC(hipEventRecord(h->iter_ev[0], h->stream));
    if (exec) {
      HIPC(hipGraphLaunch(exec, h->stream));
    } else {
      hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
      hipLaunchKernelGGL(k_ipm_solve, dim3(grid), dim3(WG), h->smem_iter, h->stream, h->dinst,
                         h->dorder, h->count, h->dqueue, o, h->dcnt, h->dclk);
      HIPC(hipGetLastError());
    }
    HIPC(hipEventRecord(h->iter_ev[1], h->stream));
    launched = 1;
  } else {
    // one launch per IPM iteration, replayed from a graph {reset queue, k_ipm_iter}
    if (opts->use_graph) {
      HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
      hipLaunchKernelGGL(k_ipm_iter, dim3(grid), dim3(WG), h->smem_iter, h->stream, h->dinst,
                         h->dorder, h->count, h->dqueue, o, h->nmax, h->dcnt, h->dclk);
      HIPC(hipStreamEndCapture(h->stream, &graph));
      HIPC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    const int check_every = 2;
    HIPC(hipEventRecord(h->iter_ev[0], h->stream));
    for (it = 0; it < o.max_iter; ++it) {
      if (exec) {
        HIPC(hipGraphLaunch(exec, h->stream));
      } else {
        hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
        hipLaunchKernelGGL(k_ipm_iter, dim3(grid), dim3(WG), h->smem_iter, h->stream, h->dinst,
                           h->dorder, h->count, h->dqueue, o, h->nmax, h->dcnt, h->dclk);
        HIPC(hipGetLastError());
      }
      HIPC(hipEventRecord(h->iter_ev[it + 1], h->stream));
      if ((it + 1) % check_every == 0 || it + 1 == o.max_iter) {
        Counters hc{};
        HIPC(hipMemcpyAsync(&hc, h->dcnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        if (hc.done >= h->count) {
          ++it;
          break;
        }
      }
    }
    launched = std::min(it, o.max_iter);
  }
  HIPC(hipEventRecord(h->ev1, h->stream));
  HIPC(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  h->iter_ms.assign(launched, 0.0);
  for (int k = 0; k < launched; ++k) {
    float t = 0.f;
    HIPC(hipEventElapsedTime(&t, h->iter_ev[k], h->iter_ev[k + 1]));
    h->iter_ms[k] = t;
  }
  // per-QP completion times from the device wall clock
  {
    std::vector<unsigned long long> clk((size_t)h->count + 1);
    HIPC(hipMemcpy(clk.data(), h->dclk, sizeof(unsigned long long) * (h->count + 1),
                   hipMemcpyDeviceToHost));
    int khz = 0;
    HIPC(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    h->inst_ms.assign(h->count, NAN);
    const unsigned long long t0 = clk[h->count];
    for (int q = 0; q < h->count; ++q)
      if (clk[q] != 0 && t0 != ~0ull && clk[q] >= t0 && khz > 0)
        h->inst_ms[q] = (double)(clk[q] - t0) / (double)khz;
  }
  if (counters) {
    Counters hc{};
    HIPC(hipMemcpy(&hc, h->dcnt, sizeof(Counters), hipMemcpyDeviceToHost));
    counters->backend_solves = (int64_t)hc.backend_solves;
    counters->inner_steps = (int64_t)hc.inner_steps;
    counters->fp64_fallbacks = (int64_t)hc.fallbacks;
    counters->ir_not_converged = (int64_t)hc.ir_bad;
    counters->jitter_retries = (int64_t)hc.jitter_retries;
    counters->iterations = hc.max_iters;
    counters->device_ms = ms;
  }
  h->have_start = false;
  return QTQP_OK;
}

int qtqp_batch_autotune(qtqp_batch* h, const qtqp_batch_opts* opts, int32_t sample,
                        int32_t iters, int32_t* table, double* ms) {
  if (!h || !opts) return fail(QTQP_ERR_ARG, "NULL argument");
  if (sample < 1 || iters < 1) return fail(QTQP_ERR_ARG, "sample and iters must be >= 1");
  constexpr int NBK = QTQP_BATCH_NBUCKETS, NT = QTQP_BATCH_NTILES;
  std::vector<int32_t> choice(2 * NBK, -1);
  std::vector<double> times((size_t)2 * NBK * NT, NAN);
  // QPs of each bucket in longest-first order
  std::vector<std::vector<int>> members(NBK);
  for (int q : h->order) members[size_bucket(h->n[q])].push_back(q);
  Inst* tmp = nullptr;
  HIPC(hipMalloc(&tmp, sizeof(Inst) * sample));
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // the matrix data (equilibrated A and P and their FP32 copies) of every
  // QP, without touching its iterates
  if (e == hipSuccess) {
    launch_prepare(h, h->dinst, h->dorder, h->count, opts->equilibrate ? 1 : 0, 1, 0, cus);
    e = hipGetLastError();
  }
  for (int b = 0; b < NBK && e == hipSuccess; ++b) {
    const std::vector<int>& mem = members[b];
    if (mem.empty()) continue;
    // an evenly spread sample of the bucket (it holds the longest QPs first)
    std::vector<Inst> pick;
    const int k = std::min<int>(sample, (int)mem.size());
    for (int j = 0; j < k; ++j) pick.push_back(h->hinst[mem[(size_t)j * mem.size() / k]]);
    const int grid = std::min(k, std::max(1, cus * 2));
    for (int prec = 0; prec < 2 && e == hipSuccess; ++prec) {
      double best = INFINITY;
      for (int v = 0; v < NT && e == hipSuccess; ++v) {
        if (prec == 1 && v == QTQP_BATCH_TILE_128) continue;  // FP32 only
        for (Inst& I : pick) (prec == 0 ? I.tile : I.tile64) = v;
        if ((e = hipMemcpyAsync(tmp, pick.data(), sizeof(Inst) * k, hipMemcpyHostToDevice,
                                h->stream)) != hipSuccess)
          break;
        // one untimed launch (code and data warm), then the timed one
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
          hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, h->stream, h->dqueue);
          if ((e = hipEventRecord(h->ev0, h->stream)) != hipSuccess) break;
          if (prec == 0)
            hipLaunchKernelGGL(k_gram_bench<float>, dim3(grid), dim3(WG), 0, h->stream, tmp, k,
                               h->dqueue, (int)iters);
          else
            hipLaunchKernelGGL(k_gram_bench<double>, dim3(grid), dim3(WG), 0, h->stream, tmp, k,
                               h->dqueue, (int)iters);
          if ((e = hipGetLastError()) != hipSuccess) break;
          if ((e = hipEventRecord(h->ev1, h->stream)) != hipSuccess) break;
          if ((e = hipEventSynchronize(h->ev1)) != hipSuccess) break;
        }
        if (e != hipSuccess) break;
        float t = 0.f;
        if ((e = hipEventElapsedTime(&t, h->ev0, h->ev1)) != hipSuccess) break;
        times[((size_t)prec * NBK + b) * NT + v] = t;
        if (t < best) {
          best = t;
          choice[prec * NBK + b] = v;
        }
      }
    }
  }
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(QTQP_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
  // the timing runs overwrote the KKT diagonal and G of the sampled QPs
  h->have_start = false;
  std::vector<int32_t> apply(choice);
  for (int32_t& c : apply)
    if (c < 0) c = QTQP_BATCH_TILE_64;
  int rc = qtqp_batch_set_tiles(h, apply.data());
  if (rc) return rc;
  if (table) std::copy(choice.begin(), choice.end(), table);
  if (ms) std::copy(times.begin(), times.end(), ms);
  (void)opts;
  return QTQP_OK;
}

int qtqp_batch_set_tiles(qtqp_batch* h, const int32_t* table) {
  if (!h || !table) return fail(QTQP_ERR_ARG, "NULL argument");
  constexpr int NBK = QTQP_BATCH_NBUCKETS;
  for (int b = 0; b < 2 * NBK; ++b)
    if (table[b] < 0 || table[b] >= QTQP_BATCH_NTILES ||
        (b >= NBK && table[b] == QTQP_BATCH_TILE_128))
      return fail(QTQP_ERR_ARG, "unknown tile shape for its precision");
  for (int q = 0; q < h->count; ++q) {
    h->hinst[q].tile = table[size_bucket(h->n[q])];
    h->hinst[q].tile64 = table[NBK + size_bucket(h->n[q])];
  }
  HIPC(hipMemcpy(h->dinst, h->hinst.data(), sizeof(Inst) * h->count, hipMemcpyHostToDevice));
  return QTQP_OK;
}

int qtqp_batch_results(qtqp_batch* h, double* x, double* y, double* s, double* tau,
                       int32_t* status, int32_t* iters, double* pres, double* dres, double* gap) {
  if (!h) return fail(QTQP_ERR_ARG, "NULL handle");
  hipLaunchKernelGGL(k_gather, dim3(std::min(h->count, 2048)), dim3(WG), 0, h->stream, h->dinst,
                     h->count, h->d_noff, h->d_moff, h->d_px, h->d_py, h->d_ps, h->d_scal,
                     h->d_ints);
  HIPC(hipGetLastError());
  std::vector<double> sd((size_t)S_NSCAL * h->count);
  std::vector<int> si((size_t)I_NINT * h->count);
  if (x) HIPC(hipMemcpyAsync(x, h->d_px, 8 * h->tot_n, hipMemcpyDeviceToHost, h->stream));
  if (y) HIPC(hipMemcpyAsync(y, h->d_py, 8 * h->tot_m, hipMemcpyDeviceToHost, h->stream));
  if (s) HIPC(hipMemcpyAsync(s, h->d_ps, 8 * h->tot_m, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipMemcpyAsync(sd.data(), h->d_scal, sd.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipMemcpyAsync(si.data(), h->d_ints, si.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  for (int q = 0; q < h->count; ++q) {
    const double* d = sd.data() + (size_t)q * S_NSCAL;
    const int* iv = si.data() + (size_t)q * I_NINT;
    if (tau) tau[q] = d[S_TAU];
    if (pres) pres[q] = d[S_PRES];
    if (dres) dres[q] = d[S_DRES];
    if (gap) gap[q] = d[S_GAP];
    if (status) status[q] = iv[I_STATUS];
    if (iters) iters[q] = iv[I_ITERS];
  }
  return QTQP_OK;
}

int qtqp_batch_solution(qtqp_batch* h, double* x, double* y, double* s, int32_t* status,
                        int32_t* iters) {
  if (!h) return fail(QTQP_ERR_ARG, "NULL handle");
  hipLaunchKernelGGL(k_gather, dim3(std::min(h->count, 2048)), dim3(WG), 0, h->stream, h->dinst,
                     h->count, h->d_noff, h->d_moff, h->d_px, h->d_py, h->d_ps, h->d_scal,
                     h->d_ints);
  hipLaunchKernelGGL(k_finish, dim3(std::min(h->count, 2048)), dim3(WG), 0, h->stream, h->dinst,
                     h->count, h->d_noff, h->d_moff, h->d_px, h->d_py, h->d_ps, h->d_scal,
                     h->d_ints);
  HIPC(hipGetLastError());
  std::vector<int> si((size_t)I_NINT * h->count);
  if (x) HIPC(hipMemcpyAsync(x, h->d_px, 8 * h->tot_n, hipMemcpyDeviceToHost, h->stream));
  if (y) HIPC(hipMemcpyAsync(y, h->d_py, 8 * h->tot_m, hipMemcpyDeviceToHost, h->stream));
  if (s) HIPC(hipMemcpyAsync(s, h->d_ps, 8 * h->tot_m, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipMemcpyAsync(si.data(), h->d_ints, si.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  for (int q = 0; q < h->count; ++q) {
    if (status) status[q] = si[(size_t)q * I_NINT + I_STATUS];
    if (iters) iters[q] = si[(size_t)q * I_NINT + I_ITERS];
  }
  return QTQP_OK;
}

int qtqp_batch_instance_ms(qtqp_batch* h, double* ms, int32_t cap) {
  if (!h) return -1;
  const int k = (int)h->inst_ms.size();
  if (ms)
    for (int i = 0; i < std::min(k, (int)cap); ++i) ms[i] = h->inst_ms[i];
  return k;
}

int qtqp_batch_launch_times(qtqp_batch* h, double* ms, int32_t cap) {
  if (!h) return -1;
  const int k = (int)h->iter_ms.size();
  for (int i = 0; i < std::min(k, (int)cap); ++i) ms[i] = h->iter_ms[i];
  return k;
}

int qtqp_batch_trace(qtqp_batch* h, int32_t q, double* out, int32_t rows) {
  if (!h || !out || q < 0 || q >= h->count || rows < 0) return fail(QTQP_ERR_ARG, "bad argument");
  const int r = std::min<int>(rows, kTraceRows);
  HIPC(hipMemcpy(out, h->hinst[q].trace, sizeof(double) * kTraceFields * r, hipMemcpyDeviceToHost));
  return QTQP_OK;
}

int qtqp_batch_row_bytes(qtqp_batch* h, int64_t* bytes) {
  if (!h || !bytes) return fail(QTQP_ERR_ARG, "NULL argument");
  for (int q = 0; q < h->count; ++q) bytes[q] = h->row_bytes[q];
  return QTQP_OK;
}

int qtqp_batch_stored_rows(qtqp_batch* h, int32_t* mc) {
  if (!h || !mc) return fail(QTQP_ERR_ARG, "NULL argument");
  for (int q = 0; q < h->count; ++q) mc[q] = h->mc[q];
  return QTQP_OK;
}

int qtqp_batch_work(qtqp_batch* h, int32_t* counts) {
  if (!h || !counts) return fail(QTQP_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(k_gather, dim3(std::min(h->count, 2048)), dim3(WG), 0, h->stream, h->dinst,
                     h->count, h->d_noff, h->d_moff, h->d_px, h->d_py, h->d_ps, h->d_scal,
                     h->d_ints);
  HIPC(hipGetLastError());
  std::vector<int> si((size_t)I_NINT * h->count);
  HIPC(hipMemcpyAsync(si.data(), h->d_ints, si.size() * 4, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  const int slots[5] = {I_F32, I_F64, I_KMV, I_FS32, I_FS64};
  for (int q = 0; q < h->count; ++q)
    for (int k = 0; k < 5; ++k) counts[q * 5 + k] = si[(size_t)q * I_NINT + slots[k]];
  return QTQP_OK;
}

void qtqp_batch_free(qtqp_batch* h) {
  if (!h) return;
  for (hipEvent_t ev : h->iter_ev) (void)hipEventDestroy(ev);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

#ifdef QTQP_PHASE_CLOCKS
// Copies the per-phase clock sums (kPhaseSlots values, 100 MHz ticks summed
// over workgroups) into out and clears them.
int qtqp_batch_phase_clocks(unsigned long long* out) {
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), sizeof(g_phase)) != hipSuccess) return 2;
  static const unsigned long long zero[kPhaseSlots] = {};
  return hipMemcpyToSymbol(HIP_SYMBOL(g_phase), zero, sizeof(g_phase)) == hipSuccess ? 0 : 2;
}
#endif
}  // extern "C"
