// Workgroup-level dense building blocks for the KKT engine on gfx950 (CDNA4).
//
// Conventions: one workgroup = WG = 256 threads = 4 wavefronts of 64 lanes.
// Symmetric n x n matrices (the Gram block G and its factor L) are stored
// column-major with leading dimension ld and only the lower triangle is
// referenced, so column k of L is contiguous in the row index -- the same
// access pattern as a row of the row-major constraint block A.  That makes
// the Gram product A' W A and the LDL' trailing update L21 L21' the same
// symmetric rank-k tile routine, X' diag(w) X with X read row by row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace qk {

constexpr int NB = 64;   // panel width of the blocked factorisation, tile edge
constexpr int WG = 256;  // threads per workgroup
constexpr int KC = 32;   // k-chunk staged through LDS by the rank-k tile

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA tiles.  Operands: A lane l holds A[l%16][l/16] and B lane l
// holds B[l/16][l%16] for both forms.  Results differ: row(l, r) gives the
// row of D held by register r of lane l (the column is always l%16).
template <typename T> struct Mfma;
// v_mfma_f64_16x16x4_f64: register r holds row (l/16) + 4r.
template <> struct Mfma<double> {
  using acc_t = dbl4;
  __device__ static inline acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
// v_mfma_f32_16x16x4_f32: register r holds row 4(l/16) + r.
template <> struct Mfma<float> {
  using acc_t = flt4;
  __device__ static inline acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// Block-wide sum / max; `red` is LDS scratch of at least 4 doubles.  All
// threads of the workgroup must call; every thread gets the result.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0] + red[1] + red[2] + red[3];
  return r;
}
__device__ inline double block_max(double v, double* red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// K block-wide reductions with one pair of barriers: slot k is a sum, or a
// maximum where bit k of MAXMASK is set.  Each slot is combined exactly as
// block_sum / block_max combine a single value.  red: >= 4 K doubles of LDS.
template <int K, unsigned MAXMASK = 0u>
__device__ inline void block_reduce_n(double (&v)[K], double* red) {
  static_assert(K <= 32, "slots");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((MAXMASK >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave * K + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k)
    v[k] = ((MAXMASK >> k) & 1u) ? fmax(fmax(red[k], red[K + k]), fmax(red[2 * K + k], red[3 * K + k]))
                                 : red[k] + red[K + k] + red[2 * K + k] + red[3 * K + k];
}

// The value v of the lane paired with this one across bit BIT of the lane
// index, moved without LDS: lane ^ 32 and lane ^ 16 by the gfx950 lane
// swaps (v_permlane32_swap / v_permlane16_swap: of the pair they return, the
// first holds the lower half's values in both halves, the second the upper
// half's), lane ^ 8 by DPP row_ror:8, lane ^ 2 and lane ^ 1 by DPP quad
// permutes, and for BIT = 4 lane 7 - (lane & 7) of the same 8-lane group
// (DPP row_half_mirror; it also crosses bit 2, which is what a pairwise
// exchange across bit 2 needs).
template <int BIT>
__device__ inline uint32_t lane_partner32(uint32_t v) {
  const int lane = __lane_id();
  if constexpr (BIT == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return lane < 32 ? r[1] : r[0];
  } else if constexpr (BIT == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    constexpr int ctrl = BIT == 8 ? 0x128 : BIT == 4 ? 0x141 : BIT == 2 ? 0x4E : 0xB1;
    static_assert(BIT == 8 || BIT == 4 || BIT == 2 || BIT == 1, "lane bit");
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xf, 0xf, false);
  }
}
template <int BIT>
__device__ inline double lane_partner(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = lane_partner32<BIT>((uint32_t)u), hi = lane_partner32<BIT>((uint32_t)(u >> 32));
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// v of lane `src` (a constant) in every lane, through scalar registers.
__device__ inline float lane_value(float v, int src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(uint32_t, v), src));
}
__device__ inline double lane_value(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// Completes R per-lane partial sums (or maxima, MAX) across the wavefront
// (R a power of two up to 32): a transpose-reduce that halves the rows a lane carries
// at each of log2(R) exchanges, then combines the rest.  Lane l ends with
// the result of row l >> (6 - log2 R) in d[0].  The exchanges use
// lane_partner (no LDS round trips); the partner across bit 4 is the
// mirror lane, which pairs lanes with opposite bit 4 as the exchange needs.
// The rows live in a vector value while they are exchanged: on a plain
// array, the compiler may merge the lane-dependent choice between two
// elements into one load at a lane-dependent index, which sends the array
// to scratch memory.
template <int BIT>
__device__ inline double partner_of(double v) {
  return lane_partner<BIT>(v);
}
template <int R, bool MAX = false>
__device__ inline void reduce_rows(double* d) {
  typedef double Rows __attribute__((ext_vector_type(R < 2 ? 2 : R)));
  const int lane = __lane_id();
  Rows v;
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = d[q];
  auto exchange = [&](auto bitc, int cnt) {
    constexpr int bit = decltype(bitc)::value;
    const int half = cnt / 2;
    const bool hi = lane & bit;
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      if (q < half) {
        const double lo_v = v[q], hi_v = v[q + half];
        const double send = hi ? lo_v : hi_v;
        const double keep = hi ? hi_v : lo_v;
        const double got = partner_of<bit>(send);
        v[q] = MAX ? fmax(keep, got) : keep + got;
      }
    }
  };
  // exchanges across bits 32, 16, ... 64 / R, each halving the rows
  if constexpr (R >= 2) exchange(std::integral_constant<int, 32>(), R);
  if constexpr (R >= 4) exchange(std::integral_constant<int, 16>(), R / 2);
  if constexpr (R >= 8) exchange(std::integral_constant<int, 8>(), R / 4);
  if constexpr (R >= 16) exchange(std::integral_constant<int, 4>(), R / 8);
  if constexpr (R >= 32) exchange(std::integral_constant<int, 2>(), R / 16);
  double r0 = v[0];
  auto combine = [&](auto bitc) {
    const double got = partner_of<decltype(bitc)::value>(r0);
    r0 = MAX ? fmax(r0, got) : r0 + got;
  };
  // the rest across bits 32 / R ... 1 (all lanes of a row's group end equal)
  if constexpr (32 / R >= 16) combine(std::integral_constant<int, 16>());
  if constexpr (32 / R >= 8) combine(std::integral_constant<int, 8>());
  if constexpr (32 / R >= 4) combine(std::integral_constant<int, 4>());
  if constexpr (32 / R >= 2) combine(std::integral_constant<int, 2>());
  if constexpr (32 / R >= 1) combine(std::integral_constant<int, 1>());
  d[0] = r0;
}

// A value the whole wavefront shares, moved to scalar registers.  The
// compiler cannot see that the arguments of a non-inlined function, or
// threadIdx.x / 64, are the same in every lane; once it can, loops over them
// branch on scalars and addresses keep their uniform part in SGPRs.
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename T>
__device__ inline T* wave_uniform(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// Waits until every vector memory operation of this wavefront has completed
// (s_waitcnt vmcnt(0), the other counters left alone: gfx9 encoding, vmcnt
// in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).  After stores, this
// tells the compiler's wait insertion that none are pending any more:
// loads and stores share the counter without a fixed order between them, so
// while a store may be pending every load is waited for with vmcnt(0).
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Waits for every vector memory and LDS/scalar/flat operation (vmcnt(0)
// lgkmcnt(0)).  Placed before a software-pipelined loop, it retires what
// the code before it left in flight (stores, and flat accesses, which count
// on both counters), so the waits inside the loop count its own loads only.
__device__ inline void wait_mem() { __builtin_amdgcn_s_waitcnt(0x0070); }

// Loads through a pointer known to address global memory (device buffers
// reached through tables in memory are otherwise accessed with flat
// instructions, whose completion is also counted with LDS operations, so an
// LDS wait would drain every outstanding prefetch).
template <typename T>
__device__ inline T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

template <typename T>
__device__ inline void stg(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// The same pointer, marked as addressing LDS (a round trip through the LDS
// address space that the compiler keeps, so accesses become ds_* ops).
template <typename T>
__device__ inline T* in_lds(T* p) {
  return (T*)(__attribute__((address_space(3))) T*)p;
}

// Non-negative doubles compare like their bit patterns, so an inf-norm can be
// accumulated with an integer atomic max (order independent, deterministic).
__device__ inline void atomic_max_nonneg(unsigned long long* slot, double v) {
  if (!(v >= 0.0)) v = __longlong_as_double(0x7ff8000000000000ll);  // NaN wins
  atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// Symmetric rank-k tile on MFMA:
//   acc(j, i) = sum_{k in [kbeg,kend)} X[k*ldx + i] * w(k) * X[k*ldx + j]
// for the (64 MI) x (64 MJ) tile with row block i0 and column block j0
// (global indices, rows >= ilim / cols >= jlim read as zero).  Wave v owns
// the (32 MI) x (32 MJ) sub-tile i in [i0 + 32 MI (v&1), ...),
// j in [j0 + 32 MJ (v>>1), ...) as (2 MJ) x (2 MI) MFMA 16x16 blocks:
// acc[a][b][r] is element j = jb + 16a + Mfma<T>::row(lane,r),
// i = ib + 16b + (lane&15), so stores into a column-major matrix (index
// j*ld + i) are contiguous across lanes.  Larger tiles re-read X less
// (each k-row of X is fetched once per tile) at the cost of registers and
// of work wasted on partial tiles; the batched engine picks MI x MJ per
// size bucket with its autotuner.
// ---------------------------------------------------------------------------
// Weight functor of RankKTileG::run_vec for unweighted rank-k updates.
struct UnitWeight {
  __device__ double operator()(double) const { return 1.0; }
};

template <typename T, int MI = 1, int MJ = 1>
struct RankKTileG {
  using acc_t = typename Mfma<T>::acc_t;
  static constexpr int TI = 64 * MI, TJ = 64 * MJ;  // tile rows (i) and columns (j)
  static constexpr int BI = 2 * MI, BJ = 2 * MJ;    // 16x16 blocks per wave
  acc_t acc[BJ][BI];

  __device__ inline void zero() {
#pragma unroll
    for (int a = 0; a < BJ; ++a)
#pragma unroll
      for (int b = 0; b < BI; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] = T(0);
  }

  // si: KC rows of (TI + 1) elements, sj: KC rows of (TJ + 1) elements of
  // LDS staging.  The next k-chunk is fetched into registers while the
  // MFMAs consume the current one.
  template <typename WFun>
  __device__ inline void run(const T* __restrict__ X, int64_t ldx, int kbeg, int kend, int i0,
                             int j0, int ilim, int jlim, WFun wfun, T* si, T* sj) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = tid & 63, kr = tid >> 6;
    const int jb = 32 * MJ * (wave >> 1), ib = 32 * MI * (wave & 1);
    T pi[KC / 4][MI], pj[KC / 4][MJ];
    auto fetch = [&](int kc) {
#pragma unroll
      for (int q = 0; q < KC / 4; ++q) {
        const int k = kc + kr + 4 * q;
        const bool kin = k < kend;
        const T* row = X + (int64_t)(kin ? k : kbeg) * ldx;
        const T wk = kin ? wfun(k) : T(0);
#pragma unroll
        for (int h = 0; h < MI; ++h) pi[q][h] = (kin && i0 + c + 64 * h < ilim) ? ldg(row + i0 + c + 64 * h) : T(0);
#pragma unroll
        for (int h = 0; h < MJ; ++h)
          pj[q][h] = (kin && j0 + c + 64 * h < jlim) ? ldg(row + j0 + c + 64 * h) * wk : T(0);
      }
    };
    if (kbeg < kend) fetch(kbeg);
    for (int kc = kbeg; kc < kend; kc += KC) {
#pragma unroll
      for (int q = 0; q < KC / 4; ++q) {
#pragma unroll
        for (int h = 0; h < MI; ++h) si[(kr + 4 * q) * (TI + 1) + c + 64 * h] = pi[q][h];
#pragma unroll
        for (int h = 0; h < MJ; ++h) sj[(kr + 4 * q) * (TJ + 1) + c + 64 * h] = pj[q][h];
      }
      __syncthreads();
      if (kc + KC < kend) fetch(kc + KC);
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        const int kq = kk + (lane >> 4), cl = lane & 15;
        T av[BJ], bv[BI];
#pragma unroll
        for (int a = 0; a < BJ; ++a) av[a] = sj[kq * (TJ + 1) + jb + 16 * a + cl];
#pragma unroll
        for (int b = 0; b < BI; ++b) bv[b] = si[kq * (TI + 1) + ib + 16 * b + cl];
#pragma unroll
        for (int a = 0; a < BJ; ++a)
#pragma unroll
          for (int b = 0; b < BI; ++b) acc[a][b] = Mfma<T>::mac(av[a], bv[b], acc[a][b]);
      }
      __syncthreads();
    }
  }

  // Rows of X staged per LDS round by run_vec.
  static constexpr int KCH = sizeof(T) == 4 ? 64 : 32;

  // run() for a row-major X whose rows are 16-byte aligned and zero-padded
  // to ldx (a multiple of 16 / sizeof(T) elements; columns >= ilim / jlim
  // must read as zero, which the padding provides): every thread fetches
  // 16-byte vectors, KCH rows per round, DEPTH rounds ahead of the MFMAs.
  // Row k of the j operand is scaled by wt(wsrc[k]) (by 1 for UnitWeight).
  // The loads are unconditional (rows clamped to kend - 1, columns to
  // ldx - VW, the values masked to zero when they are staged), so the wait
  // for one round leaves the rounds behind it in flight.
  // si: KCH rows of (TI + 1), sj: KCH rows of (TJ + 1) elements of LDS.
  template <typename WT>
  __device__ inline void run_vec(const T* __restrict__ X, int64_t ldx, int kbeg, int kend, int i0,
                                 int j0, const double* __restrict__ wsrc, WT wt, T* si, T* sj) {
    constexpr int VW = 16 / sizeof(T);
    typedef T VT __attribute__((ext_vector_type(VW)));
    constexpr int VI = TI / VW, VJ = TJ / VW;  // vectors per staged row
    constexpr int RI = WG / VI, RJ = WG / VJ;  // rows per pass of the workgroup
    constexpr int PI = KCH / RI, PJ = KCH / RJ;
    constexpr bool kUnit = std::is_same<WT, UnitWeight>::value;
    constexpr int DEPTH = MI * MJ == 1 ? 2 : 1;  // larger tiles keep one round in registers
    X = wave_uniform(X);
    kbeg = wave_uniform(kbeg);
    kend = wave_uniform(kend);
    i0 = wave_uniform(i0);
    j0 = wave_uniform(j0);
    ldx = ((int64_t)wave_uniform((int)(ldx >> 32)) << 32) | (uint32_t)wave_uniform((int)ldx);
    if constexpr (!kUnit) wsrc = wave_uniform(wsrc);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int jb = 32 * MJ * (wave >> 1), ib = 32 * MI * (wave & 1);
    const int vi = tid % VI, ri = tid / VI, vj = tid % VJ, rj = tid / VJ;
    const bool ci = i0 + vi * VW < ldx, cj = j0 + vj * VW < ldx;
    const int64_t coli = min<int64_t>(i0 + vi * VW, ldx - VW);
    const int64_t colj = min<int64_t>(j0 + vj * VW, ldx - VW);
    struct Round {
      VT a[PI], b[PJ];
      double w[kUnit ? 1 : PJ];
    };
    auto fetch = [&](int kc, Round& rd) {
#pragma unroll
      for (int q = 0; q < PI; ++q) {
        const int k = min(kc + ri + RI * q, kend - 1);
        rd.a[q] = ldg(reinterpret_cast<const VT*>(X + (int64_t)k * ldx + coli));
      }
#pragma unroll
      for (int q = 0; q < PJ; ++q) {
        const int k = min(kc + rj + RJ * q, kend - 1);
        rd.b[q] = ldg(reinterpret_cast<const VT*>(X + (int64_t)k * ldx + colj));
        if constexpr (!kUnit) rd.w[q] = ldg(wsrc + k);
      }
    };
    // One round: stage it (when it holds rows below kend), refill its
    // registers DEPTH rounds ahead, run its MFMAs.  The refill is issued on
    // every path, so the loads in flight at the loop head are the same
    // whichever way it was reached and the wait there covers one round.
    auto step = [&](int kc, Round& rd, bool active) {
      if (active) {
#pragma unroll
      for (int q = 0; q < PI; ++q) {
        const VT v = (kc + ri + RI * q < kend && ci) ? rd.a[q] : VT(0);
#pragma unroll
        for (int e = 0; e < VW; ++e) si[(ri + RI * q) * (TI + 1) + vi * VW + e] = v[e];
      }
#pragma unroll
      for (int q = 0; q < PJ; ++q) {
        VT v = rd.b[q];
        if constexpr (!kUnit) v = v * wt(rd.w[q]);
        v = (kc + rj + RJ * q < kend && cj) ? v : VT(0);
#pragma unroll
        for (int e = 0; e < VW; ++e) sj[(rj + RJ * q) * (TJ + 1) + vj * VW + e] = v[e];
      }
      }
      __syncthreads();
      fetch(kc + DEPTH * KCH, rd);
      if (active) {
#pragma unroll
      for (int kk = 0; kk < KCH; kk += 4) {
        const int kq = kk + (lane >> 4), cl = lane & 15;
        T av[BJ], bv[BI];
#pragma unroll
        for (int a = 0; a < BJ; ++a) av[a] = sj[kq * (TJ + 1) + jb + 16 * a + cl];
#pragma unroll
        for (int b = 0; b < BI; ++b) bv[b] = si[kq * (TI + 1) + ib + 16 * b + cl];
#pragma unroll
        for (int a = 0; a < BJ; ++a)
#pragma unroll
          for (int b = 0; b < BI; ++b) acc[a][b] = Mfma<T>::mac(av[a], bv[b], acc[a][b]);
      }
      }
      __syncthreads();
    };
    if (kbeg >= kend) return;
    wait_mem();
    Round rd[DEPTH];
    // Rounds are consumed in the order they are fetched; the scheduling
    // barriers keep their loads in that order, so a wait for the older one
    // can leave the younger in flight.
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) {
      fetch(kbeg + d * KCH, rd[d]);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int kc = kbeg; kc < kend; kc += DEPTH * KCH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) step(kc + d * KCH, rd[d], kc + d * KCH < kend);
    }
  }

  // Same, with the classic [KC][NB + 1] staging arrays of a 64 x 64 tile.
  template <typename WFun>
  __device__ inline void run(const T* __restrict__ X, int64_t ldx, int kbeg, int kend, int i0,
                             int j0, int ilim, int jlim, WFun wfun, T (*si)[NB + 1],
                             T (*sj)[NB + 1]) {
    static_assert(MI == 1 && MJ == 1, "2-D staging arrays describe a 64 x 64 tile");
    run(X, ldx, kbeg, kend, i0, j0, ilim, jlim, wfun, &si[0][0], &sj[0][0]);
  }

  // Calls f(i, j, value) for every accumulated element of this thread.
  template <typename F>
  __device__ inline void for_each(int i0, int j0, F f) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jb = j0 + 32 * MJ * (wave >> 1), ib = i0 + 32 * MI * (wave & 1);
#pragma unroll
    for (int a = 0; a < BJ; ++a)
#pragma unroll
      for (int b = 0; b < BI; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          f(ib + 16 * b + (lane & 15), jb + 16 * a + Mfma<T>::row(lane, r), acc[a][b][r]);
  }
};

template <typename T>
using RankKTile = RankKTileG<T, 1, 1>;

// Maps a linear index over the lower-triangular tiles (ti >= tj) of an
// nt x nt tile grid to (ti, tj).
__device__ inline void lower_tile(int t, int* ti, int* tj) {
  int r = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  while (r * (r + 1) / 2 > t) --r;
  *ti = r;
  *tj = t - r * (r + 1) / 2;
}

// ---------------------------------------------------------------------------
// LDL' pieces.  The Gram block is positive definite, so its LDL' factor is
// stored in the Cholesky form L D^{1/2} (one triangle, no separate D).
// ---------------------------------------------------------------------------

// Loads the nb x nb diagonal block at (k0, k0) of column-major G into S
// (row-major S[i][j], lower part; padding rows/cols become the identity).
template <typename T>
__device__ inline void load_diag_block(const T* G, int64_t ld, int k0, int nb,
                                       T (*S)[NB + 1]) {
  // 16 loads per thread issued together (clamped in-block addresses, values
  // selected afterwards) so their latencies overlap.
  T v[NB * NB / WG];
#pragma unroll
  for (int u = 0; u < NB * NB / WG; ++u) {
    const int idx = threadIdx.x + u * WG;
    const int i = min(idx & (NB - 1), nb - 1), j = min(idx >> 6, nb - 1);
    v[u] = ldg(G + (int64_t)(k0 + j) * ld + k0 + i);
  }
#pragma unroll
  for (int u = 0; u < NB * NB / WG; ++u) {
    const int idx = threadIdx.x + u * WG;
    const int i = idx & (NB - 1), j = idx >> 6;
    S[i][j] = (i < nb && j < nb) ? (i >= j ? v[u] : T(0)) : T(i == j);
  }
}

// 1 / sqrt(x): the hardware estimate (v_rsq) refined by two Newton steps
// for the reciprocal square root, y <- y + y (1 - x y^2) / 2 (one step in
// FP32).  Within an ulp of the IEEE square root and division, and a much
// shorter dependency chain than the two of them.
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-x * y, y, 1.0);
  y = __builtin_fma(y * e, 0.5, y);
  e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * e, 0.5, y);
}
__device__ inline float rsqrt_nr(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float e = __builtin_fmaf(-x * y, y, 1.0f);
  return __builtin_fmaf(y * e, 0.5f, y);
}

// The value of lane K of each row of 16 lanes, in every lane of that row
// (DPP row_newbcast): independent broadcasts for the four rows of a wave.
template <int K>
__device__ inline float row_bcast(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + K, 0xf, 0xf, false));
}
template <int K>
__device__ inline double row_bcast(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x150 + K, 0xf, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x150 + K, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}
// row_bcast with k known once the caller's loop is unrolled.
template <typename T>
__device__ inline T row_bcast_k(T v, int k) {
  switch (k) {
    case 0: return row_bcast<0>(v);
    case 1: return row_bcast<1>(v);
    case 2: return row_bcast<2>(v);
    case 3: return row_bcast<3>(v);
    case 4: return row_bcast<4>(v);
    case 5: return row_bcast<5>(v);
    case 6: return row_bcast<6>(v);
    case 7: return row_bcast<7>(v);
    case 8: return row_bcast<8>(v);
    case 9: return row_bcast<9>(v);
    case 10: return row_bcast<10>(v);
    case 11: return row_bcast<11>(v);
    case 12: return row_bcast<12>(v);
    case 13: return row_bcast<13>(v);
    case 14: return row_bcast<14>(v);
    default: return row_bcast<15>(v);
  }
}

// In-LDS Cholesky of the 64 x 64 diagonal block fused with the inverse of its
// factor (all workgroup threads).  On entry S holds the block (lower part;
// padding = identity from load_diag_block).  On return S holds L (lower,
// L[j][j] = sqrt(pivot), zeros above), X = L^{-1} (lower, zeros above) and
// sinv[j] = 1 / L[j][j].  Blocked into four 16-column panels, so that it
// needs about 19 barriers instead of one per column:
//   1. every wave factors the 16 x 16 diagonal block D of the panel and
//      inverts its factor in registers (lane = row r, columns 4cg..4cg+3,
//      cg = lane / 16; column broadcasts by cross-lane shuffles);
//   2. the panel below, L_p = A_p L_d^{-T}, one 16-row strip per wave on MFMA
//      16x16x4 (K = 16) with L_d^{-1} read back from X's diagonal block;
//   3. the trailing tiles, S_ij -= L_i L_j^T, on MFMA.
// Then the off-diagonal 16 x 16 blocks of X = L^{-1}, block row by block row:
// X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj.  A non-positive or NaN pivot records
// `kglob + j + 1` in *info (the first such column) and is replaced by 1.
template <typename T>
__device__ inline void potrf_inv_blk(T (*S)[NB + 1], T (*X)[NB + 1], T* sinv, int nb,
                                     int* info, int kglob) {
  using acc_t = typename Mfma<T>::acc_t;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, kq = lane >> 4;
  __shared__ int first_bad;
  if (tid == 0) first_bad = NB;
  // X above its diagonal 16 x 16 blocks is zero (the diagonal blocks are
  // written whole below)
  for (int idx = tid; idx < NB * NB; idx += WG) {
    const int row = idx >> 6, col = idx & (NB - 1);
    if ((col >> 4) > (row >> 4)) X[row][col] = T(0);
  }
  for (int pb = 0; pb < 4; ++pb) {
    const int c0 = 16 * pb;
    // Wave 0 factors the diagonal 16 x 16 block D = L_d L_d': lane r (mod 16)
    // holds row r, and the pivot and each finished entry L[k][j] reach the
    // other lanes of the row by DPP (in vector registers: broadcasts through
    // scalar registers would outnumber them and spill).  The chain of pivots
    // is the critical path (its reciprocal square roots are rsqrt_nr); the
    // other waves do not repeat it.
    if (w == 0) {
      T d[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) d[k] = k <= r ? S[c0 + r][c0 + k] : T(0);
      T drr = T(1);  // L[r][r]
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        T piv = row_bcast_k(d[j], j);
        const bool ok = piv > T(0);
        if (!ok) {
          piv = T(1);
          if (lane == 0 && c0 + j < nb) atomicMin(&first_bad, c0 + j);
        }
        const T rj = rsqrt_nr(piv);
        const T dj = piv * rj;
        const T lrj = r > j ? d[j] * rj : (r == j ? dj : T(0));  // L[r][j]
        // rows below j lose L[r][j] L[k][j] (entries above row r's diagonal
        // are never read back)
#pragma unroll
        for (int k = j + 1; k < 16; ++k) d[k] -= lrj * row_bcast_k(lrj, k);
        d[j] = lrj;
        if (r == j) drr = dj;
      }
      if (lane < 16) {
#pragma unroll
        for (int k = 0; k < 16; ++k) S[c0 + r][c0 + k] = k <= r ? d[k] : T(0);
        sinv[c0 + r] = T(1) / drr;
      }
    }
    __syncthreads();  // L_d in S, 1 / diag(L_d) in sinv
    {
      // X_d = L_d^{-1} by forward substitution, column by column: wave w's
      // lane row kq takes column 4w + kq, lane r its row r; row k's value
      // reaches the other lanes of the row by DPP
      T lr[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) lr[k] = S[c0 + r][c0 + k];
      const int c = 4 * w + kq;
      T xv = T(r == c);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (r == k) xv *= sinv[c0 + k];
        const T xk = row_bcast_k(xv, k);
        if (r > k) xv -= lr[k] * xk;
      }
      X[c0 + r][c0 + c] = xv;
    }
    __syncthreads();  // L_d^{-1} in X
    const int nstrip = 3 - pb;
    if (w < nstrip) {
      const int rs = c0 + 16 + 16 * w;
      acc_t acc;
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = T(0);
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = Mfma<T>::mac(S[rs + r][c0 + k0 + kq], X[c0 + r][c0 + k0 + kq], acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) S[rs + Mfma<T>::row(lane, q)][c0 + r] = acc[q];
    }
    __syncthreads();  // panel strips in S
    const int ntile = nstrip * (nstrip + 1) / 2;
    for (int tix = w; tix < ntile; tix += 4) {
      int si, sj;
      lower_tile(tix, &si, &sj);
      const int ri = c0 + 16 + 16 * si, rj = c0 + 16 + 16 * sj;
      acc_t acc;
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = T(0);
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = Mfma<T>::mac(S[ri + r][c0 + k0 + kq], S[rj + r][c0 + k0 + kq], acc);
#pragma unroll
      for (int q = 0; q < 4; ++q) S[ri + Mfma<T>::row(lane, q)][rj + r] -= acc[q];
    }
    __syncthreads();  // trailing tiles updated
  }
  // off-diagonal blocks of X = L^{-1}, block row bi: wave w takes block column w
  for (int bi = 1; bi < 4; ++bi) {
    acc_t acc;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = T(0);
    if (w < bi) {
      for (int bk = w; bk < bi; ++bk) {
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4)
          acc = Mfma<T>::mac(S[16 * bi + r][16 * bk + k0 + kq], X[16 * bk + k0 + kq][16 * w + r],
                             acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) X[16 * bi + Mfma<T>::row(lane, q)][16 * w + r] = acc[q];
    }
    __syncthreads();  // the sums T_ij are in X
    if (w < bi) {
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = T(0);
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4)
        acc = Mfma<T>::mac(X[16 * bi + r][16 * bi + k0 + kq], X[16 * bi + k0 + kq][16 * w + r], acc);
    }
    __syncthreads();  // every read of T_ij is done
    if (w < bi) {
#pragma unroll
      for (int q = 0; q < 4; ++q) X[16 * bi + Mfma<T>::row(lane, q)][16 * w + r] = -acc[q];
    }
  }
  __syncthreads();
  if (tid == 0 && info && first_bad < nb) atomicCAS(info, 0, kglob + first_bad + 1);
  __syncthreads();
}

// Writes the factored lower diagonal block back into G.
template <typename T>
__device__ inline void store_diag_block(T* G, int64_t ld, int k0, int nb,
                                        const T (*S)[NB + 1]) {
  for (int idx = threadIdx.x; idx < NB * NB; idx += WG) {
    const int i = idx & (NB - 1), j = idx >> 6;
    if (i < nb && j < nb && i >= j) G[(int64_t)(k0 + j) * ld + k0 + i] = S[i][j];
  }
}

// Panel solve on MFMA for rows [rbeg, rbeg + 64*STRIPS) below the diagonal
// block: L21 = A21 L11^{-T}, computed as L21' = Li A21' so that results land
// contiguously in the column-major G.  Wave v handles 16*STRIPS rows as
// 16-row strips; each strip is 4 MFMA 16x16 tiles over the 64 columns.
template <typename T, int STRIPS = 4>
__device__ inline void trsm_mfma_rows(T* G, int64_t ld, int n, int k0, int nb,
                                      int rbeg, const T (*Li)[NB + 1]) {
  using acc_t = typename Mfma<T>::acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane & 15, kq = lane >> 4;
  for (int s = 0; s < STRIPS; ++s) {
    const int r0 = rbeg + 16 * STRIPS * wave + 16 * s;
    if (r0 >= n) break;
    // the strip's whole operand is loaded before its first MFMA (rows
    // clamped below n, columns below nb, masked where used): one memory
    // latency per strip rather than one per group of 4 columns
    T bv[16];
    const int rowc = min(r0 + cl, n - 1);
#pragma unroll
    for (int q = 0; q < 16; ++q)
      bv[q] = ldg(G + (int64_t)(k0 + min(4 * q + kq, nb - 1)) * ld + rowc);
    acc_t acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ct][r] = T(0);
    const int row = r0 + cl;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = 4 * q + kq;
      const T b = (row < n && k < nb) ? bv[q] : T(0);
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) acc[ct] = Mfma<T>::mac(Li[16 * ct + cl][k], b, acc[ct]);
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int c = 16 * ct + Mfma<T>::row(lane, r);
        if (c < nb && row < n) G[(int64_t)(k0 + c) * ld + row] = acc[ct][r];
      }
  }
}

}  // namespace qk
