// Internals of the single-instance dense KKT engine (MI355X, gfx950), shared
// by the LinearSolver backend (kkt.hip) and the device-resident IPM driver
// (ipm.hip), which are compiled into the same library.
//
// K_reg = [ P_off + diag(d_x)   A'       ]     d_x = diag[:n] > 0
//         [ A                   diag(d_y) ]     d_y = diag[n:] < 0
//
// Factorisation: quasi-definite LDL' with the y-block pivoted first.  Its
// pivots are d_y and its trailing update into the x-block is one symmetric
// rank-m MFMA product, giving the positive definite Gram block
//   G = P_off + diag(d_x) + A' W A,   W = diag(-1/d_y)
// (the Schur complement ScipyDenseSolver forms, solvers_dense.py:99-116),
// which is then factored by a blocked right-looking LDL' whose trailing
// updates run on MFMA.  A solve is: g = r_x + A'(W r_y); G x = g by two
// triangular sweeps; y = W(A x - r_y)  (solvers_dense.py:130-142).
//
// Precision modes: FP64 everywhere, or an FP32 factor (FP32 copies of A and
// P_off, FP32 MFMA) whose solves are refined in FP64 against the exact K_reg
// until they are as accurate as an FP64 factor; when the FP32 factor breaks
// down or contracts too slowly the handle refactors in FP64 and stays there.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dense_ops.h"
#include "qtqp_kkt.h"

// Message of the calling thread's last failure (defined in kkt.hip).
extern thread_local std::string g_qtqp_last_error;

using namespace qk;

namespace {


constexpr int RB = 16;  // rows per workgroup in the row/column passes over A

// Switch of a kernel in a fixed launch sequence (one IPM iteration replayed
// from a HIP graph): the kernel returns at once unless `active` is set and,
// for want != 0, the precision `prec` of the current factor is `want`.  The
// protocol entry points pass no gate (always run).
struct Gate {
  int active;
  int prec;  // 32 or 64
};
__device__ inline bool gated_off(const Gate* g, int want) {
  return g != nullptr && (g->active == 0 || (want != 0 && g->prec != want));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <typename T>
__global__ void k_cast(const double* __restrict__ src, T* __restrict__ dst, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x; i < count; i += (int64_t)gridDim.x * WG)
    dst[i] = (T)src[i];
}

// Lower 64 x 64 tile handled by this workgroup.  Workgroups are dispatched to
// the eight XCDs in turn (blockIdx mod 8), each with its own L2; the grid is
// padded to a multiple of 8 and XCD x takes the contiguous tile range
// [x * per, (x+1) * per), whose tiles share rows of the operand.  Returns
// false for padding workgroups.
__device__ inline bool xcd_tile(int ntiles, int* ti, int* tj) {
  const int per = gridDim.x >> 3;
  const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (t >= ntiles) return false;
  lower_tile(t, ti, tj);
  return true;
}
inline unsigned xcd_grid(int ntiles) { return (unsigned)(8 * ((ntiles + 7) / 8)); }

// Rows of the operand staged per LDS round by the tile kernels.
template <typename T> constexpr int kch() { return RankKTile<T>::KCH; }

// G (lower tiles) = P_off + diag(d_x) + A' W A.  VEC: rows of A are 16-byte
// aligned vectors (n a multiple of 16 / sizeof(T)), fetched as such.
template <typename T, bool VEC>
__global__ __launch_bounds__(WG) void k_gram_syrk(const T* __restrict__ A,
                                                  const T* __restrict__ Poff,
                                                  const double* __restrict__ diag,
                                                  int n, int m, T* __restrict__ G,
                                                  int64_t ld, int ntiles, const Gate* gate,
                                                  int want) {
  __shared__ T si[kch<T>()][NB + 1];
  __shared__ T sj[kch<T>()][NB + 1];
  if (gated_off(gate, want)) return;
  int ti, tj;
  if (!xcd_tile(ntiles, &ti, &tj)) return;
  const int i0 = ti * NB, j0 = tj * NB;
  RankKTile<T> t;
  t.zero();
  const double* dy = diag + n;
  if constexpr (VEC) {
    t.run_vec(A, n, 0, m, i0, j0, dy, [](double d) { return (T)(-1.0 / d); }, &si[0][0], &sj[0][0]);
  } else {
    t.run(A, n, 0, m, i0, j0, n, n, [dy](int k) { return (T)(-1.0 / dy[k]); }, si, sj);
  }
  t.for_each(i0, j0, [&](int i, int j, T v) {
    if (i < n && j < n && i >= j) {
      T g = v + Poff[(int64_t)j * n + i];
      if (i == j) g += (T)diag[i];
      G[(int64_t)j * ld + i] = g;
    }
  });
}

// The Gram product split over the rows of A when G has few tiles (a small n
// with many constraints): workgroup (t, c) forms tile t's sum over rows
// [c * chunk, (c + 1) * chunk) into Gp, and k_gram_reduce adds the chunks in
// order, P_off and diag(d_x).  Enough workgroups then stream A at once.
template <typename T, bool VEC>
__global__ __launch_bounds__(WG) void k_gram_part(const T* __restrict__ A,
                                                  const double* __restrict__ diag, int n, int m,
                                                  int nchunk, int chunk, T* __restrict__ Gp,
                                                  const Gate* gate, int want) {
  __shared__ T si[kch<T>()][NB + 1];
  __shared__ T sj[kch<T>()][NB + 1];
  if (gated_off(gate, want)) return;
  const int t = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
  int ti, tj;
  lower_tile(t, &ti, &tj);
  const int i0 = ti * NB, j0 = tj * NB;
  const int k0 = c * chunk, k1 = min(m, k0 + chunk);
  RankKTile<T> acc;
  acc.zero();
  const double* dy = diag + n;
  if constexpr (VEC) {
    acc.run_vec(A, n, k0, k1, i0, j0, dy, [](double d) { return (T)(-1.0 / d); }, &si[0][0],
                &sj[0][0]);
  } else {
    acc.run(A, n, k0, k1, i0, j0, n, n, [dy](int k) { return (T)(-1.0 / dy[k]); }, si, sj);
  }
  T* dst = Gp + ((int64_t)t * nchunk + c) * NB * NB;
  acc.for_each(i0, j0, [&](int i, int j, T v) { dst[(j - j0) * NB + (i - i0)] = v; });
}

//  Workgroup b takes 256 entries of tile b / 16, one per thread.
template <typename T>
__global__ __launch_bounds__(WG) void k_gram_reduce(const T* __restrict__ Gp,
                                                    const T* __restrict__ Poff,
                                                    const double* __restrict__ diag, int n,
                                                    int nchunk, T* __restrict__ G, int64_t ld,
                                                    const Gate* gate, int want) {
  if (gated_off(gate, want)) return;
  constexpr int kParts = NB * NB / WG;
  const int t = blockIdx.x / kParts;
  int ti, tj;
  lower_tile(t, &ti, &tj);
  const int i0 = ti * NB, j0 = tj * NB;
  const int e = (blockIdx.x % kParts) * WG + threadIdx.x;
  const int i = i0 + (e & (NB - 1)), j = j0 + (e >> 6);
  if (i >= n || j >= n || i < j) return;
  const T* src = Gp + (int64_t)t * nchunk * NB * NB + e;
  T v = T(0);
  for (int c = 0; c < nchunk; ++c) v += src[(int64_t)c * NB * NB];
  T g = v + Poff[(int64_t)j * n + i];
  if (i == j) g += (T)diag[i];
  G[(int64_t)j * ld + i] = g;
}

// diag(G) += rel * max(diag(G)); rel = 1e-14 as solvers_dense.py:107-112 (and
// larger on a retry after a breakdown).  One workgroup.
template <typename T>
__global__ __launch_bounds__(WG) void k_jitter(T* G, int64_t ld, int n, double rel,
                                               const Gate* gate, int want) {
  __shared__ double red[4];
  if (gated_off(gate, want)) return;
  double mx = -INFINITY;
  for (int i = threadIdx.x; i < n; i += WG) mx = fmax(mx, (double)G[(int64_t)i * ld + i]);
  mx = block_max(mx, red);
  const double jit = rel * mx;
  for (int i = threadIdx.x; i < n; i += WG) G[(int64_t)i * ld + i] = (T)((double)G[(int64_t)i * ld + i] + jit);
}

// Panel k0, diagonal block: factors it in LDS, stores L11 into G and
// L11^{-1} into linv + (k0/64) * 2 * 64 * 64 twice: column-major (for the
// forward sweep), then row-major (for the panel solve and the backward
// sweep).  One workgroup.
template <typename T>
__global__ __launch_bounds__(WG) void k_potrf_diag(T* G, int64_t ld, int n, int k0, int* info,
                                                  T* linv, const Gate* gate, int want) {
  __shared__ T S[NB][NB + 1];
  __shared__ T Li[NB][NB + 1];
  __shared__ T sinv[NB];
  if (gated_off(gate, want)) return;
  const int nb = min(NB, n - k0);
  load_diag_block(G, ld, k0, nb, S);
  __syncthreads();
  potrf_inv_blk(S, Li, sinv, nb, info, k0);
  store_diag_block(G, ld, k0, nb, S);
  T* dst = linv + (int64_t)(k0 / NB) * 2 * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += WG) {
    dst[idx] = Li[idx & (NB - 1)][idx >> 6];
    dst[NB * NB + idx] = Li[idx >> 6][idx & (NB - 1)];
  }
}

// Panel k0, rows below the diagonal block: L21 = A21 L11^{-T} on MFMA,
// workgroup b handling rows k0+NB+b*64 .. +64 (one 16-row strip per wave).
template <typename T>
__global__ __launch_bounds__(WG) void k_trsm_panel(T* G, int64_t ld, int n, int k0,
                                                  const T* linv, const Gate* gate, int want) {
  __shared__ T Li[NB][NB + 1];
  if (gated_off(gate, want)) return;
  const T* src = linv + (int64_t)(k0 / NB) * 2 * NB * NB + NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += WG) Li[idx >> 6][idx & (NB - 1)] = src[idx];
  __syncthreads();
  trsm_mfma_rows<T, 1>(G, ld, n, k0, min(NB, n - k0), k0 + NB + blockIdx.x * NB, Li);
}

// Trailing update G[t0:, t0:] -= L[t0:, k0:t0] L[t0:, k0:t0]', t0 = k0 + NB
// (VEC as for k_gram_syrk, with ld = n).
template <typename T, bool VEC>
__global__ __launch_bounds__(WG) void k_syrk_trail(T* G, int64_t ld, int n, int k0, int ntiles,
                                                  const Gate* gate, int want) {
  __shared__ T si[kch<T>()][NB + 1];
  __shared__ T sj[kch<T>()][NB + 1];
  if (gated_off(gate, want)) return;
  int ti, tj;
  if (!xcd_tile(ntiles, &ti, &tj)) return;
  const int t0 = k0 + NB;
  const int i0 = t0 + ti * NB, j0 = t0 + tj * NB;
  RankKTile<T> t;
  t.zero();
  // X[k*ldx + i] = L[i][k0+k] (column k0+k of the column-major factor).
  if constexpr (VEC) {
    t.run_vec(G + (int64_t)k0 * ld, ld, 0, NB, i0, j0, nullptr, UnitWeight(), &si[0][0], &sj[0][0]);
  } else {
    t.run(G + (int64_t)k0 * ld, ld, 0, NB, i0, j0, n, n, [](int) { return T(1); }, si, sj);
  }
  t.for_each(i0, j0, [&](int i, int j, T v) {
    if (i < n && j < n && i >= j) G[(int64_t)j * ld + i] -= v;
  });
}

// Row and column passes over a row-major matrix M (rows x ncol) read it once
// in TR x TC tiles, one tile per workgroup, grid (ceil(rows/TR), ceil(ncol/TC)):
//   HX: rpart[blockIdx.y * rows + r] = sum_{j in tile} M[r][j] x[j]
//   HU: cpart[blockIdx.x * ncol + j] = sum_{r in tile} M[r][j] u_r,
//       u_r = u[r] * (w ? -1/w[r] : 1)
// Wave v takes tile rows 16v .. 16v+15 and lane l the four consecutive
// columns 4l .. 4l+3, fetched as one 16-byte vector (two for FP64) when rows
// are 16-byte aligned (VEC).  The row sums are completed across the lanes by
// a transpose-reduce, the column sums across the four waves in LDS.  The
// finish kernels below add the partials in a fixed order, so results are
// deterministic.
constexpr int TR = 64;
constexpr int TC = 256;
// An FP64 pass that forms both products takes tiles of 32 rows (8 per
// wave): with 16 rows of four doubles per lane it would hold more than 256
// registers and run one wave per SIMD.  Its column partials come in
// tiles_rx(rows) row tiles.
constexpr int TRX = 32;
inline int tiles_r(int64_t rows) { return (int)((rows + TR - 1) / TR); }
inline int tiles_rx(int64_t rows) { return (int)((rows + TRX - 1) / TRX); }
inline int tiles_c(int64_t ncol) { return (int)((ncol + TC - 1) / TC); }
__device__ inline int tiles_rx_dev(int rows) { return (rows + TRX - 1) / TRX; }
template <typename T, bool HX, bool HU>
constexpr int tile_rows() { return sizeof(T) == 8 && HX && HU ? TRX : TR; }
// Row tiles of a pass over `rows` rows of an FP64 (dbl) or FP32 matrix
// forming row products (hx) and/or column products (hu).
inline int tiles_r_pass(int64_t rows, bool dbl, bool hx, bool hu) {
  return dbl && hx && hu ? tiles_rx(rows) : tiles_r(rows);
}

template <typename T> struct Vec4;
template <> struct Vec4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  __device__ static void load(const float* p, double* d) {
    const type v = *reinterpret_cast<const type*>(p);
    d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
  }
};
template <> struct Vec4<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  __device__ static void load(const double* p, double* d) {
    const type v0 = reinterpret_cast<const type*>(p)[0];
    const type v1 = reinterpret_cast<const type*>(p)[1];
    d[0] = v0.x; d[1] = v0.y; d[2] = v1.x; d[3] = v1.y;
  }
};

// Tile (bx, by) of the pass; cs: 4 x TC doubles of LDS.
template <typename T, bool HX, bool HU, bool VEC>
__device__ inline void tile_body(const T* __restrict__ M, int rows, int ncol,
                                 const double* __restrict__ x, double* __restrict__ rpart,
                                 const double* __restrict__ u, const double* __restrict__ w,
                                 double* __restrict__ cpart, int bx, int by, double (*cs)[TC]) {
  constexpr int RPW = tile_rows<T, HX, HU>() / 4;  // rows per wave
  // the wave index in a scalar register: the per-row values u[r], w[r] are
  // then scalar loads, which keeps vector registers for the rows
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r0 = bx * (4 * RPW) + wave * RPW;
  const int j0 = by * TC + 4 * lane;  // this lane's first column
  double a[RPW][4];
  if constexpr (VEC) {
    // the wave's rows through one buffer resource: rows past the end read
    // zero without a memory request, and columns past the row end (read
    // from the next row) are masked, so every load is issued without a
    // branch around it
    const int nr = max(0, min(RPW, rows - r0));
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(M + (int64_t)min(r0, rows) * ncol), 0, nr * ncol * (int)sizeof(T),
        0x00020000);
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const uint32_t off = (uint32_t)((q * ncol + j0) * (int)sizeof(T));
      double v[4];
      if constexpr (sizeof(T) == 8) {
        typedef double D2 __attribute__((ext_vector_type(2)));
        const D2 v0 = __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        const D2 v1 =
            __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16, 0, 0));
        v[0] = v0.x;
        v[1] = v0.y;
        v[2] = v1.x;
        v[3] = v1.y;
      } else {
        typedef float F4 __attribute__((ext_vector_type(4)));
        const F4 f = __builtin_bit_cast(F4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        v[0] = f.x;
        v[1] = f.y;
        v[2] = f.z;
        v[3] = f.w;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) a[q][c] = j0 + c < ncol ? v[c] : 0.0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int r = r0 + q;
      const T* row = M + (int64_t)r * ncol + j0;
#pragma unroll
      for (int c = 0; c < 4; ++c) a[q][c] = (r < rows && j0 + c < ncol) ? (double)row[c] : 0.0;
    }
  }
  // the row dots and the column sums in one pass over the rows, so the rows
  // are dead before the row dots are reduced across the lanes (each sum
  // still adds its terms in row / column order)
  double xv[4], rd[RPW], cacc[4] = {0.0, 0.0, 0.0, 0.0};
  if (HX) {
#pragma unroll
    for (int c = 0; c < 4; ++c) xv[c] = j0 + c < ncol ? x[j0 + c] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    if (HX) {
      rd[q] = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) rd[q] += a[q][c] * xv[c];
    }
    if (HU) {
      const int r = r0 + q;
      double ur = 0.0;
      if (r < rows) {
        ur = u[r];
        if (w) ur *= -1.0 / w[r];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) cacc[c] += a[q][c] * ur;
    }
  }
  if (HX) {
    reduce_rows<RPW>(rd);  // lane l: row l / (64 / RPW)
    constexpr int LPR = 64 / RPW;  // lanes per row
    const int q = lane / LPR;
    if (lane % LPR == 0 && r0 + q < rows) rpart[(int64_t)by * rows + r0 + q] = rd[0];
  }
  if (HU) {
#pragma unroll
    for (int c = 0; c < 4; ++c) cs[wave][4 * lane + c] = cacc[c];
    __syncthreads();
    const int j = by * TC + threadIdx.x;
    if (j < ncol)
      cpart[(int64_t)bx * ncol + j] =
          (cs[0][threadIdx.x] + cs[1][threadIdx.x]) + (cs[2][threadIdx.x] + cs[3][threadIdx.x]);
  }
}

template <typename T, bool HX, bool HU, bool VEC>
__global__ __launch_bounds__(WG) void k_tile_pass(const T* __restrict__ M, int rows, int ncol,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ rpart,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ w,
                                                  double* __restrict__ cpart, const Gate* gate,
                                                  int want) {
  __shared__ double cs[4][TC];
  if (gated_off(gate, want)) return;
  tile_body<T, HX, HU, VEC>(M, rows, ncol, x, rpart, u, w, cpart, blockIdx.x, blockIdx.y, cs);
}

// The passes of a K product in one launch: A x_x and A' x_y over the first
// tiles_rx(m) row tiles, P_off x_x over the next tiles_r(n) (P == NULL: none).
template <bool VEC>
__global__ __launch_bounds__(WG) void k_tile_pass_kx(const double* __restrict__ A,
                                                     const double* __restrict__ P, int m, int n,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ rpa,
                                                     double* __restrict__ rpp,
                                                     double* __restrict__ cpart,
                                                     const Gate* gate) {
  __shared__ double cs[4][TC];
  if (gated_off(gate, 0)) return;
  const int ta = tiles_rx_dev(m);
  if ((int)blockIdx.x < ta)
    tile_body<double, true, true, VEC>(A, m, n, x, rpa, x + n, nullptr, cpart, blockIdx.x,
                                       blockIdx.y, cs);
  else
    tile_body<double, true, false, VEC>(P, n, n, x, rpp, nullptr, nullptr, nullptr,
                                        blockIdx.x - ta, blockIdx.y, cs);
}

// out[j] = (base ? base[j] : 0) + sum_{b < nb} cpart[b * ncol + j].  A workgroup
// takes 64 columns; its four waves add every fourth partial, then meet in LDS.
__global__ __launch_bounds__(WG) void k_colfin(const double* __restrict__ cpart, int nb, int ncol,
                                               const double* __restrict__ base,
                                               double* __restrict__ out, const Gate* gate = nullptr,
                                               int want = 0) {
  __shared__ double red[4][64];
  if (gated_off(gate, want)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (j < ncol)
    for (int b = wave; b < nb; b += 4) s += cpart[(int64_t)b * ncol + j];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && j < ncol) {
    const double t = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    out[j] = base ? base[j] + t : t;
  }
}

// sum_{b < nb} rpart[b * rows + r]
__device__ inline double row_total(const double* __restrict__ rpart, int nb, int rows, int r) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += rpart[(int64_t)b * rows + r];
  return s;
}

__global__ void k_rowfin(const double* __restrict__ rpart, int nb, int rows,
                         double* __restrict__ out) {
  const int r = blockIdx.x * WG + threadIdx.x;
  if (r < rows) out[r] = row_total(rpart, nb, rows, r);
}

// A'x_y at column j from the column partials of nbr row tiles (the order of
// k_colfin: four interleaved partial sums, then pairs).
__device__ inline double col_total(const double* __restrict__ cpart, int nbr, int ncol, int j) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int b = 0;
  for (; b + 8 <= nbr; b += 8) {  // eight loads in flight at a time
    double v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = cpart[(int64_t)(b + t) * ncol + j];
#pragma unroll
    for (int t = 0; t < 8; ++t) s[t & 3] += v[t];
  }
  for (; b < nbr; ++b) s[b & 3] += cpart[(int64_t)b * ncol + j];
  return (s[0] + s[1]) + (s[2] + s[3]);
}

// K_reg x (or a residual):  kx = [P_off x_x + d_x*x_x + A'x_y ; A x_x + d_y*x_y]
// from the row partials of P_off x_x (rpp) and A x_x (rpa, ncb tiles each)
// and the column partials of A'x_y (cpart, nbr row tiles; rpp NULL: P_off =
// 0).  If rhs: out = rhs - kx (+ corr*x if corr); with krhs the rhs is taken
// as [rhs_x; -rhs_y] (the quasi-definite form, direct.py:245-248), written to
// krhs, and its inf-norm goes into *rhs_norm.  With norm: norm[0] gets
// max |out| and norm[1] becomes nonzero if x has a NaN.
__device__ inline void kmv_finish(int n, int m, const double* __restrict__ x,
                                  const double* __restrict__ diag, const double* __restrict__ rpp,
                                  const double* __restrict__ rpa, int ncb,
                                  const double* __restrict__ cpart, int nbr,
                                  const double* __restrict__ rhs, double* __restrict__ krhs,
                                  unsigned long long* rhs_norm, const double* __restrict__ corr,
                                  double* __restrict__ out, unsigned long long* norm) {
  __shared__ double red[4];
  const int idx = blockIdx.x * WG + threadIdx.x;
  double v = 0.0, rv = 0.0;
  bool live = idx < n + m;
  if (live) {
    double kx;
    if (idx < n)
      kx = (rpp ? row_total(rpp, ncb, n, idx) : 0.0) + diag[idx] * x[idx] +
           col_total(cpart, nbr, n, idx);
    else
      kx = row_total(rpa, ncb, m, idx - n) + diag[idx] * x[idx];
    if (rhs) {
      rv = rhs[idx];
      if (krhs) {
        if (idx >= n) rv = -rv;
        krhs[idx] = rv;
      }
      v = rv - kx;
      if (corr) v += corr[idx] * x[idx];
    } else {
      v = kx;
    }
    out[idx] = v;
  }
  if (rhs_norm) {
    const double a = block_max(live ? (rv != rv ? INFINITY : fabs(rv)) : 0.0, red);
    if (threadIdx.x == 0) atomic_max_nonneg(rhs_norm, a);
  }
  if (norm) {
    double a = live ? fabs(v) : 0.0;
    if (live && v != v) a = INFINITY;  // fmax drops NaN; report it as inf
    if (live && x[idx] != x[idx]) atomicOr(&norm[1], 1ull);
    a = block_max(a, red);
    if (threadIdx.x == 0) atomic_max_nonneg(norm, a);
  }
}

__global__ __launch_bounds__(WG) void k_kmv_finish(int n, int m, const double* __restrict__ x,
                                                   const double* __restrict__ diag,
                                                   const double* __restrict__ rpp,
                                                   const double* __restrict__ rpa, int ncb,
                                                   const double* __restrict__ cpart, int nbr,
                                                   const double* __restrict__ rhs,
                                                   double* __restrict__ krhs,
                                                   unsigned long long* rhs_norm,
                                                   const double* __restrict__ corr,
                                                   double* __restrict__ out,
                                                   unsigned long long* norm,
                                                   const Gate* gate = nullptr, int want = 0) {
  if (gated_off(gate, want)) return;
  kmv_finish(n, m, x, diag, rpp, rpa, ncb, cpart, nbr, rhs, krhs, rhs_norm, corr, out, norm);
}

// ---------------------------------------------------------------------------
// Triangular sweeps G x = g, G = L L', across many workgroups.
//
// Workgroup "ticket" t (taken with an atomic counter, so every lower ticket
// belongs to a workgroup that is already running) owns row block I = t
// (forward) or I = nt-1-t (backward).  It streams the off-diagonal blocks
// it needs, waiting per block for the published piece of the solution that
// block multiplies, then applies the stored inverse of its diagonal block
// and publishes its own piece.
// The sweep state lives on the device, so a launch needs no host-side
// sequence number and replays from a graph: per direction a ticket counter,
// a done counter and a generation g.  A launch publishes its values with
// the flag value g + 1; the last workgroup to finish resets both counters
// and advances g, so the next launch waits for g + 2 and never mistakes an
// old value for a new one.  Layout of `sync`: [ticket, done, generation]
// forward, the same backward, two words of padding, then the published
// values.  A wait is bounded: past the limit the block writes NaN (which the
// callers' NaN checks report) and still publishes, so every workgroup ends.
// ---------------------------------------------------------------------------
constexpr unsigned kSpinLimit = 1u << 21;
constexpr int kSyncHdr = 8;  // two counter triples, padded to 8-byte words

// Takes this workgroup's ticket and the launch's flag value (generation + 1).
__device__ inline unsigned take_ticket(unsigned* st, unsigned* sh, unsigned* epoch) {
  if (threadIdx.x == 0) {
    sh[1] = __hip_atomic_load(&st[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    sh[0] = atomicAdd(&st[0], 1u);
  }
  __syncthreads();
  *epoch = sh[1];
  return sh[0];
}

// After the workgroup's publish: the last of the nt workgroups resets the
// counters and advances the generation for the next launch.
__device__ inline void end_sweep(unsigned* st, unsigned nt) {
  if (threadIdx.x == 0 && atomicAdd(&st[1], 1u) == nt - 1) {
    __hip_atomic_store(&st[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned g = __hip_atomic_load(&st[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st[2], g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Published solution values.  Each 32-bit half of a value travels in one
// 64-bit word beside the launch's flag value (half << 32 | epoch), written
// and read with single 64-bit atomics: a reader that sees the epoch in a
// word sees the half written with it, so no fence separates data and flag
// and one memory round trip moves both.  FP32 values take one word, FP64
// two.  The words follow the sweep counters in `sync`: 2 n words for the
// forward sweep's v, then 2 n for the backward sweep's x.
__host__ __device__ inline int64_t sync_words(int64_t n) {
  return kSyncHdr + 8 * n;  // unsigned words
}
__device__ inline uint64_t* sync_ll(unsigned* sync) {
  return reinterpret_cast<uint64_t*>(sync + kSyncHdr);
}

__device__ inline void ll_put(uint64_t* w, float v, unsigned ep) {
  __hip_atomic_store(w, ((uint64_t)__float_as_uint(v) << 32) | ep, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void ll_put(uint64_t* w, double v, unsigned ep) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  __hip_atomic_store(w, (b & 0xffffffff00000000ull) | ep, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 1, (b << 32) | ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reading published values in loops that keep loads in flight: ll_poll
// issues the reads of a lane's words, ll_ready tests them (lanes without
// `need` count as ready), ll_value decodes them.  Issuing the first poll of
// a step before that step's prefetch lets the wait for it leave the
// prefetch in flight.
// (ll_poll reads unconditionally: w must be a valid address in every lane.)
template <typename T>
__device__ inline void ll_poll(const uint64_t* w, uint64_t (&a)[sizeof(T) / 4]) {
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k)
    a[k] = __hip_atomic_load(w + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline bool ll_ready(const uint64_t (&a)[sizeof(T) / 4], bool need, unsigned ep) {
  bool rd = true;
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) rd = rd && (unsigned)a[k] == ep;
  return rd || !need;
}
template <typename T>
__device__ inline T ll_value(const uint64_t (&a)[sizeof(T) / 4]) {
  if constexpr (sizeof(T) == 4)
    return __uint_as_float((unsigned)(a[0] >> 32));
  else
    return __longlong_as_double((long long)((a[0] & 0xffffffff00000000ull) | (a[1] >> 32)));
}
// Thread (r = lane, q = wave) part of the 64 x 64 block M: M[c*64 + r] for
// its 16 columns c = 16q .. 16q+15.
template <typename T>
__device__ inline void load_block_part(const T* __restrict__ M, T* mv) {
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) mv[cc] = M[(q * 16 + cc) * NB + r];
}

// out[r] = sum_c M[c*64 + r] w[c] with the block held as load_block_part
// leaves it; the four waves split c and reduce through red[4][64].
template <typename T>
__device__ inline double block_gemv(const T* mv, const double* w, double* red) {
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  double acc = 0.0;
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) acc += (double)mv[cc] * w[q * 16 + cc];
  red[q * NB + r] = acc;
  __syncthreads();
  const double v = (red[r] + red[NB + r]) + (red[2 * NB + r] + red[3 * NB + r]);
  __syncthreads();
  return v;
}

// block_gemv with the vector held by the lanes of every wave (lane c holds
// w[c]) instead of in LDS: wave q reads its 16 columns across lanes; the
// four partial sums meet in red (4 x 64 doubles).  One barrier.
template <typename T>
__device__ inline double block_gemv_lanes(const T* mv, double wv, int qs, double* red) {
  const int r = threadIdx.x & 63;
  double acc = 0.0;
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) acc += (double)mv[cc] * lane_value(wv, qs * 16 + cc);
  red[qs * NB + r] = acc;
  __syncthreads();
  return (red[r] + red[NB + r]) + (red[2 * NB + r] + red[3 * NB + r]);
}

// Forward sweep L v = g with g = gbase + A'u from the column partials of nbr
// row tiles (the right-hand side g = r_x + A'(W r_y) of a K solve, summed as
// k_colfin sums).  Thread (r = lane, q = wave) accumulates
// sum_{J<I} sum_{c in q's 16 columns} L[I*64+r][J*64+c] v[J*64+c].
template <typename T>
__global__ __launch_bounds__(WG) void k_trsv_fwd(const T* __restrict__ L, int64_t ld, int n,
                                                 const T* __restrict__ LiB,
                                                 const double* __restrict__ gbase,
                                                 const double* __restrict__ cpart, int nbr, T* v,
                                                 unsigned* sync, const Gate* gate, int want) {
  __shared__ unsigned sh_ticket[2];
  __shared__ double red[4 * NB], red2[4 * NB];
  if (gated_off(gate, want)) return;
  const int nt = gridDim.x;
  unsigned epoch;
  const int I = (int)take_ticket(sync, sh_ticket, &epoch);
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i0 = I * NB, rows = min(NB, n - i0);
  const bool live = r < rows;
  bool ok = true;
  double acc = 0.0;
  T inv[16];
  // the block inverse and this block's g are needed only after the last
  // wait: fetch them now
  load_block_part(LiB + (int64_t)I * 2 * NB * NB, inv);
  // g for this block's rows: wave q sums the partials b = q, q + 4, ... (the
  // order of k_colfin), wave 0 combines them
  {
    double sq = 0.0;
    if (live)
      for (int b = q; b < nbr; b += 4) sq += cpart[(int64_t)b * n + i0 + r];
    red[q * NB + r] = sq;
  }
  __syncthreads();
  const double gv =
      live ? gbase[i0 + r] + ((red[r] + red[NB + r]) + (red[2 * NB + r] + red[3 * NB + r])) : 0.0;
  __syncthreads();
  // unconditional loads (rows clamped into the block, masked afterwards):
  // a branch around each load would make every wait drain the prefetch
  const int rc = min(r, rows - 1);
  auto load = [&](int J, T* dst) {
    const T* col = L + (int64_t)(J * NB + q * 16) * ld + i0 + rc;
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) dst[cc] = col[(int64_t)cc * ld];
  };
  constexpr int K = sizeof(T) / 4;  // published words per value
  uint64_t* vll = sync_ll(sync);
  // D blocks of this block row in flight while the chain comes: the wait for
  // v_J is issued before block J + D's loads (clamped to the last block, so
  // every step issues the same loads), so it leaves them in flight
  constexpr int D = 3;
  T buf[D + 1][16];
  const int qs = __builtin_amdgcn_readfirstlane(q);
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (I > 0) load(min(d, I - 1), buf[d]);
  for (int J0 = 0; J0 < I; J0 += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      const int J = J0 + u;
      if (J < I) {
        uint64_t a[K];
        ll_poll<T>(vll + K * (J * NB + r), a);  // J < I: a full block
        load(min(J + D, I - 1), buf[(u + D) % (D + 1)]);
        if (!__all(ll_ready<T>(a, true, epoch))) {
          // not published yet: poll again (each retry waits for all loads)
          unsigned spins = 0;
          do {
            if (++spins > kSpinLimit) ok = false;
            __builtin_amdgcn_s_sleep(1);
            ll_poll<T>(vll + K * (J * NB + r), a);
          } while (ok && !__all(ll_ready<T>(a, true, epoch)));
        }
        const T vj = ok ? ll_value<T>(a) : T(NAN);
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          acc += (double)(live ? buf[u][cc] : T(0)) * (double)lane_value(vj, qs * 16 + cc);
      }
    }
  }
  // w = g - (the row sums), formed by every wave; then the diagonal block's
  // inverse applied with w read across lanes (two barriers in all)
  red[q * NB + r] = acc;
  __syncthreads();
  const double sr = (red[r] + red[NB + r]) + (red[2 * NB + r] + red[3 * NB + r]);
  const double vi = block_gemv_lanes(inv, live ? gv - sr : 0.0, qs, red2);
  if (q == 0 && live) {
    const T out = ok ? (T)vi : T(NAN);
    v[i0 + r] = out;
    ll_put(vll + K * (i0 + r), out, epoch);
  }
  end_sweep(sync, (unsigned)nt);
}

// Backward sweep L' x = v.  Thread (rr = lane, q = wave) accumulates, for
// the 16 columns c of q, sum_{J>I} L[J*64+rr][I*64+c] x[J*64+rr].
template <typename T>
__global__ __launch_bounds__(WG) void k_trsv_bwd(const T* __restrict__ L, int64_t ld, int n,
                                                 const T* __restrict__ LiB,
                                                 const T* __restrict__ v, double* x,
                                                 unsigned* sync, const Gate* gate, int want) {
  __shared__ unsigned sh_ticket[2];
  __shared__ double red[4 * NB], red2[4 * NB];
  if (gated_off(gate, want)) return;
  const int nt = gridDim.x;
  unsigned epoch;
  const int I = nt - 1 - (int)take_ticket(sync + 3, sh_ticket, &epoch);
  const int rr = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i0 = I * NB, rows = min(NB, n - i0);
  bool ok = true;
  double acc[16];
#pragma unroll
  for (int cc = 0; cc < 16; ++cc) acc[cc] = 0.0;
  T inv[16];
  load_block_part(LiB + (int64_t)I * 2 * NB * NB + NB * NB, inv);
  // this block's right-hand side, fetched now: it is needed only after the
  // last wait
  const double vin = (double)v[i0 + min(rr, rows - 1)];
  // unconditional loads: rows clamped below n, columns clamped into the
  // block; the terms out of range are masked where they are used
  auto load = [&](int J, T* dst) {
    const T* base_p = L + (int64_t)i0 * ld + min(J * NB + rr, n - 1);
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) dst[cc] = base_p[(int64_t)min(q * 16 + cc, rows - 1) * ld];
  };
  uint64_t* xll = sync_ll(sync) + 2 * (int64_t)n;
  // blocks J = nt-1, nt-2, ..., I+1 (step s: J = nt-1-s), D of them in
  // flight, each step's wait issued before its prefetch as in k_trsv_fwd
  constexpr int D = 3;
  T buf[D + 1][16];
  const int nJ = nt - 1 - I;
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (nJ > 0) load(nt - 1 - min(d, nJ - 1), buf[d]);
  for (int s0 = 0; s0 < nJ; s0 += D + 1) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      const int st = s0 + u;
      if (st < nJ) {
        const int row = (nt - 1 - st) * NB + rr;
        const bool need = row < n;
        const uint64_t* wp = xll + 2 * (int64_t)min(row, n - 1);
        uint64_t a[2];
        ll_poll<double>(wp, a);
        load(nt - 1 - min(st + D, nJ - 1), buf[(u + D) % (D + 1)]);
        if (!__all(ll_ready<double>(a, need, epoch))) {
          unsigned spins = 0;
          do {
            if (++spins > kSpinLimit) ok = false;
            __builtin_amdgcn_s_sleep(1);
            ll_poll<double>(wp, a);
          } while (ok && !__all(ll_ready<double>(a, need, epoch)));
        }
        const double xj = !need ? 0.0 : ok ? ll_value<double>(a) : (double)NAN;
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
          acc[cc] += (double)(q * 16 + cc < rows ? buf[u][cc] : T(0)) * xj;
      }
    }
  }
  // column sums over the wave's 64 rows by lane exchanges (no LDS): lane l
  // ends with column q * 16 + (l >> 2)
  reduce_rows<16>(acc);
  if ((rr & 3) == 0) red[q * 16 + (rr >> 2)] = acc[0];
  __syncthreads();
  const int qs = __builtin_amdgcn_readfirstlane(q);
  const double xi = block_gemv_lanes(inv, rr < rows ? vin - red[rr] : 0.0, qs, red2);
  if (q == 0 && rr < rows) {
    const double out = ok ? xi : (double)NAN;
    x[i0 + rr] = out;
    ll_put(xll + 2 * (int64_t)(i0 + rr), out, epoch);
  }
  end_sweep(sync + 3, (unsigned)nt);
}

// y = W (Ax - r_y) with W = -1/d_y, A x from its row partials (rpa, ncb
// tiles); writes [x; y] into out (or adds it).
__global__ __launch_bounds__(WG) void k_solve_finish(int n, int m, const double* __restrict__ x,
                                                     const double* __restrict__ rpa, int ncb,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ diag,
                                                     double* __restrict__ out, int accumulate,
                                                     const Gate* gate, int want) {
  if (gated_off(gate, want)) return;
  const int idx = blockIdx.x * WG + threadIdx.x;
  if (idx >= n + m) return;
  double v;
  if (idx < n) v = x[idx];
  else v = (-1.0 / diag[idx]) * (row_total(rpa, ncb, m, idx - n) - r[idx]);
  out[idx] = accumulate ? out[idx] + v : v;
}

// sol += [x; W (A x - r_y)] (the solve's finish added to the iterate).
__global__ __launch_bounds__(WG) void k_solve_add(int n, int m, const double* __restrict__ x,
                                                  const double* __restrict__ rpa, int ncb,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ diag,
                                                  double* __restrict__ sol, const Gate* gate,
                                                  int want) {
  if (gated_off(gate, want)) return;
  const int idx = blockIdx.x * WG + threadIdx.x;
  if (idx >= n + m) return;
  double v;
  if (idx < n) v = x[idx];
  else v = (-1.0 / diag[idx]) * (row_total(rpa, ncb, m, idx - n) - r[idx]);
  sol[idx] += v;
}

__global__ void k_axpy(int64_t count, double alpha, const double* __restrict__ x,
                       double* __restrict__ y, const Gate* gate = nullptr) {
  if (gated_off(gate, 0)) return;
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  if (i < count) y[i] += alpha * x[i];
}

// krhs = [r_x; -r_y] (direct.py:245-248) and ||krhs||_inf into *norm (a NaN
// entry is reported as inf, since fmax drops NaN).
__global__ __launch_bounds__(WG) void k_negate_y(int64_t count, int64_t n,
                                                 const double* __restrict__ r,
                                                 double* __restrict__ krhs,
                                                 unsigned long long* norm) {
  __shared__ double red[4];
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  double a = 0.0;
  if (i < count) {
    const double v = i < n ? r[i] : -r[i];
    krhs[i] = v;
    a = v != v ? INFINITY : fabs(v);
  }
  a = block_max(a, red);
  if (threadIdx.x == 0) atomic_max_nonneg(norm, a);
}

// dst (rows x cols, row-major) = src', where src is cols x rows row-major
// (i.e. the column-major rows x cols matrix).  64 x 64 tiles through LDS;
// blockIdx.x walks row tiles, blockIdx.y column tiles.
__global__ __launch_bounds__(WG) void k_transpose(const double* __restrict__ src,
                                                  double* __restrict__ dst, int64_t rows,
                                                  int64_t cols) {
  __shared__ double t[NB][NB + 1];
  const int64_t r0 = (int64_t)blockIdx.x * NB, c0 = (int64_t)blockIdx.y * NB;
  for (int e = threadIdx.x; e < NB * NB; e += WG) {
    const int a = e >> 6, b = e & (NB - 1);  // src row c0 + a, src column r0 + b
    if (c0 + a < cols && r0 + b < rows) t[a][b] = src[(c0 + a) * rows + r0 + b];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NB * NB; e += WG) {
    const int a = e >> 6, b = e & (NB - 1);  // dst row r0 + a, dst column c0 + b
    if (r0 + a < rows && c0 + b < cols) dst[(r0 + a) * cols + c0 + b] = t[b][a];
  }
}

// dst = src and ||src||_inf into *norm (NaN reported as inf).
__global__ __launch_bounds__(WG) void k_copy_norm(int64_t count, const double* __restrict__ src,
                                                  double* __restrict__ dst,
                                                  unsigned long long* norm) {
  __shared__ double red[4];
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  double a = 0.0;
  if (i < count) {
    const double v = src[i];
    dst[i] = v;
    a = v != v ? INFINITY : fabs(v);
  }
  a = block_max(a, red);
  if (threadIdx.x == 0) atomic_max_nonneg(norm, a);
}

// Zeroes the diagonal of the row-major n x n matrix M.
__global__ void k_zero_diag(double* M, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)WG + threadIdx.x;
  if (i < n) M[i * n + i] = 0.0;
}

inline unsigned blocks_for(int64_t count) { return (unsigned)((count + WG - 1) / WG); }

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct qtqp_kkt {
  int64_t n = 0, m = 0;
  int precision = QTQP_PREC_FP64;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // problem data
  double *A64 = nullptr, *Poff64 = nullptr;
  float *A32 = nullptr, *Poff32 = nullptr;
  double *diag = nullptr, *corr = nullptr;
  // factor storage (column-major, ld = n)
  double* G64 = nullptr;
  float* G32 = nullptr;
  int* info = nullptr;
  // vectors
  double *vin = nullptr, *vout = nullptr, *vsol = nullptr, *vres = nullptr;
  double *rin = nullptr, *racc = nullptr, *rres = nullptr, *rtmp = nullptr;
  double *tx = nullptr, *tg = nullptr, *part = nullptr;
  double *rpa = nullptr, *rpp = nullptr, *aty = nullptr;  // row partials of A, P_off; A'v
  double* tv = nullptr;    // forward-sweep result of the triangular solves
  double* linv = nullptr;  // inverses of the 64x64 diagonal blocks of the factor
  // row split of the Gram product (gsplit chunks of gchunk rows; 1: none)
  int gsplit = 1, gchunk = 0;
  double* gpart = nullptr;
  unsigned* sync = nullptr;  // counters and published values of the triangular sweeps
  unsigned long long* dnorm = nullptr;
  // state
  bool have_diag = false, have_factor = false, force64 = false;
  int fprec = 64;
  // backend solves and their FP32 refinement steps on the current FP32 factor
  int64_t f32_solves = 0, f32_steps = 0;
  qtqp_kkt_info info_acc{};
};

namespace {

int fail(int code, const std::string& msg) {
  g_qtqp_last_error = msg;
  return code;
}

#define HIPC(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(QTQP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

int nblk_rows(int64_t rows) { return (int)((rows + RB - 1) / RB); }

// One tile pass over M (k_tile_pass) on stream st: row partials into rpart
// (if x) and column partials into cpart (if u).
template <typename T, bool VEC>
void launch_tile_pass(hipStream_t st, const T* M, int64_t rows, int64_t ncol, const double* x,
                      double* rpart, const double* u, const double* w, double* cpart,
                      const Gate* g, int want) {
  const dim3 grid(tiles_r_pass(rows, sizeof(T) == 8, x != nullptr, u != nullptr), tiles_c(ncol));
  if (x && u)
    hipLaunchKernelGGL((k_tile_pass<T, true, true, VEC>), grid, dim3(WG), 0, st, M,
                       (int)rows, (int)ncol, x, rpart, u, w, cpart, g, want);
  else if (x)
    hipLaunchKernelGGL((k_tile_pass<T, true, false, VEC>), grid, dim3(WG), 0, st, M,
                       (int)rows, (int)ncol, x, rpart, u, w, cpart, g, want);
  else
    hipLaunchKernelGGL((k_tile_pass<T, false, true, VEC>), grid, dim3(WG), 0, st, M,
                       (int)rows, (int)ncol, x, rpart, u, w, cpart, g, want);
}

template <typename T>
int tile_pass_on(hipStream_t st, double* cpart, const T* M, int64_t rows, int64_t ncol,
                 const double* x, double* rpart, const double* u, const double* w,
                 const Gate* g = nullptr, int want = 0) {
  if (rows == 0) return QTQP_OK;
  // rows start on 16-byte boundaries when ncol * sizeof(T) is a multiple of 16
  // and M itself is 16-byte aligned
  const bool vec = (ncol * (int64_t)sizeof(T)) % 16 == 0 &&
                   (reinterpret_cast<uintptr_t>(M) & 15) == 0;
  if (vec) launch_tile_pass<T, true>(st, M, rows, ncol, x, rpart, u, w, cpart, g, want);
  else launch_tile_pass<T, false>(st, M, rows, ncol, x, rpart, u, w, cpart, g, want);
  HIPC(hipGetLastError());
  return QTQP_OK;
}

template <typename T>
int tile_pass(qtqp_kkt* h, const T* M, int64_t rows, int64_t ncol, const double* x,
              double* rpart, const double* u, const double* w, const Gate* g = nullptr,
              int want = 0) {
  return tile_pass_on<T>(h->stream, h->part, M, rows, ncol, x, rpart, u, w, g, want);
}

inline unsigned colfin_blocks(int64_t ncol) { return (unsigned)((ncol + 63) / 64); }

// Launches out = K_reg x, or out = rhs - K_reg x + corr*x.  With norm (two
// device words, zero on entry): norm[0] gets ||out||_inf and norm[1] becomes
// nonzero if x has a NaN.  Nothing is read back.
//  With krhs, rhs is taken as [rhs_x; -rhs_y], written to krhs, and its
//  inf-norm goes into *rhs_norm (zero on entry).
int kmatvec_dev(qtqp_kkt* h, const double* x, const double* rhs, const double* corr, double* out,
                unsigned long long* norm, const Gate* g = nullptr, double* krhs = nullptr,
                unsigned long long* rhs_norm = nullptr) {
  const int64_t n = h->n, m = h->m, N = n + m;
  // A x_x, A' x_y and P_off x_x in one launch: rows of A and P_off start on
  // 16-byte boundaries when n is even (allocations are 256-byte aligned)
  const dim3 grid(tiles_rx(m) + tiles_r(n), tiles_c(n));
  if (n % 2 == 0)
    hipLaunchKernelGGL(k_tile_pass_kx<true>, grid, dim3(WG), 0, h->stream, h->A64, h->Poff64,
                       (int)m, (int)n, x, h->rpa, h->rpp, h->part, g);
  else
    hipLaunchKernelGGL(k_tile_pass_kx<false>, grid, dim3(WG), 0, h->stream, h->A64, h->Poff64,
                       (int)m, (int)n, x, h->rpa, h->rpp, h->part, g);
  hipLaunchKernelGGL(k_kmv_finish, dim3(blocks_for(N)), dim3(WG), 0, h->stream, (int)n, (int)m, x,
                     h->diag, h->rpp, h->rpa, tiles_c(n), h->part, tiles_rx(m), rhs, krhs, rhs_norm,
                     corr, out, norm, g, 0);
  HIPC(hipGetLastError());
  return QTQP_OK;
}

// out = K_reg x, or out = rhs - K_reg x + corr*x with ||out||_inf into *norm_host.
//  With norm_host the device words dnorm = {||out||_inf, NaN-in-x flag,
//  ||rhs||_inf (set by k_negate_y)} are read back: the flag into *nan_host and
//  the rhs norm into *rhs_norm_host when given.
int kmatvec(qtqp_kkt* h, const double* x, const double* rhs, const double* corr, double* out,
            double* norm_host, double* rhs_norm_host = nullptr, int* nan_host = nullptr) {
  if (norm_host) HIPC(hipMemsetAsync(h->dnorm, 0, 2 * sizeof(unsigned long long), h->stream));
  int rc = kmatvec_dev(h, x, rhs, corr, out, norm_host ? h->dnorm : nullptr);
  if (rc) return rc;
  if (norm_host) {
    unsigned long long bits[3] = {0, 0, 0};
    const int cnt = rhs_norm_host ? 3 : 2;
    HIPC(hipMemcpyAsync(bits, h->dnorm, sizeof(bits[0]) * cnt, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    std::memcpy(norm_host, &bits[0], sizeof(double));
    if (nan_host) *nan_host = bits[1] != 0;
    if (rhs_norm_host) std::memcpy(rhs_norm_host, &bits[2], sizeof(double));
  }
  return QTQP_OK;
}

// Launches the factorisation of G in precision T: Gram SYRK, diagonal shift
// jit_rel * max(diag(G)), blocked LDL'.  A breakdown stores its column
// (1-based) in *info, which must be zero on entry.  Nothing is read back.
template <typename T>
int factor_launch(qtqp_kkt* h, const T* A, const T* Poff, T* G, int* info, double jit_rel,
                  const Gate* g = nullptr, int want = 0) {
  const int n = (int)h->n, m = (int)h->m;
  const int nt = (n + NB - 1) / NB;
  // 16-byte vector fetches need rows of n elements to stay 16-byte aligned
  const bool vec = (n * (int)sizeof(T)) % 16 == 0;
  const int ng = nt * (nt + 1) / 2;
  if (h->gsplit > 1) {
    T* gp = reinterpret_cast<T*>(h->gpart);
    const unsigned parts = (unsigned)(ng * h->gsplit);
    if (vec)
      hipLaunchKernelGGL((k_gram_part<T, true>), dim3(parts), dim3(WG), 0, h->stream, A, h->diag, n,
                         m, h->gsplit, h->gchunk, gp, g, want);
    else
      hipLaunchKernelGGL((k_gram_part<T, false>), dim3(parts), dim3(WG), 0, h->stream, A, h->diag,
                         n, m, h->gsplit, h->gchunk, gp, g, want);
    hipLaunchKernelGGL(k_gram_reduce<T>, dim3(ng * (NB * NB / WG)), dim3(WG), 0, h->stream,
                       (const T*)gp, Poff,
                       h->diag, n, h->gsplit, G, (int64_t)n, g, want);
  } else if (vec) {
    hipLaunchKernelGGL((k_gram_syrk<T, true>), dim3(xcd_grid(ng)), dim3(WG), 0, h->stream, A, Poff,
                       h->diag, n, m, G, (int64_t)n, ng, g, want);
  } else {
    hipLaunchKernelGGL((k_gram_syrk<T, false>), dim3(xcd_grid(ng)), dim3(WG), 0, h->stream, A,
                       Poff, h->diag, n, m, G, (int64_t)n, ng, g, want);
  }
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_jitter<T>, dim3(1), dim3(WG), 0, h->stream, G, (int64_t)n, n, jit_rel, g,
                     want);
  HIPC(hipGetLastError());
  T* linv = reinterpret_cast<T*>(h->linv);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int below = std::max(0, n - k0 - NB);
    hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(WG), 0, h->stream, G, (int64_t)n, n, k0,
                       info, linv, g, want);
    HIPC(hipGetLastError());
    if (below > 0) {
      hipLaunchKernelGGL(k_trsm_panel<T>, dim3((below + NB - 1) / NB), dim3(WG), 0, h->stream, G,
                         (int64_t)n, n, k0, (const T*)linv, g, want);
      HIPC(hipGetLastError());
      const int tt = (below + NB - 1) / NB;
      const int nst = tt * (tt + 1) / 2;
      if (vec)
        hipLaunchKernelGGL((k_syrk_trail<T, true>), dim3(xcd_grid(nst)), dim3(WG), 0, h->stream,
                           G, (int64_t)n, n, k0, nst, g, want);
      else
        hipLaunchKernelGGL((k_syrk_trail<T, false>), dim3(xcd_grid(nst)), dim3(WG), 0, h->stream,
                           G, (int64_t)n, n, k0, nst, g, want);
      HIPC(hipGetLastError());
    }
  }
  return QTQP_OK;
}

// Factor G in precision T and read the breakdown column (1-based, or 0)
// back into *brk.
template <typename T>
int factor_prec(qtqp_kkt* h, const T* A, const T* Poff, T* G, int* brk, double jit_rel = 1e-14) {
  HIPC(hipMemsetAsync(h->info, 0, sizeof(int), h->stream));
  int rc = factor_launch<T>(h, A, Poff, G, h->info, jit_rel);
  if (rc) return rc;
  int hinfo = 0;
  HIPC(hipMemcpyAsync(&hinfo, h->info, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  *brk = hinfo;
  return QTQP_OK;
}

// Diagonal shifts of the FP64 factor, relative to max(diag(G)): the
// reference's 1e-14 (solvers_dense.py:107-112), then two retries after a
// breakdown.  Near the end of an IPM on an unbounded or infeasible problem
// G's condition number reaches 1e18 and whether its last pivots come out
// positive depends on rounding (the reference's LAPACK Gram and Cholesky can
// land on either side too); a larger shift keeps the factor usable, and the
// refinement against the exact K, which never sees the shift, removes its
// effect, as the reference relies on for its own shift.
constexpr double kJitterRel[3] = {1e-14, 1e-12, 1e-10};

int factor64(qtqp_kkt* h) {
  int brk = 0, rc;
  for (int t = 0; t < 3; ++t) {
    if ((rc = factor_prec<double>(h, h->A64, h->Poff64, h->G64, &brk, kJitterRel[t]))) return rc;
    if (brk == 0) break;
    if (t < 2) h->info_acc.jitter_retries++;
  }
  if (brk != 0) {
    h->have_factor = false;
    char buf[128];
    snprintf(buf, sizeof(buf), "LDL' breakdown of the FP64 Gram block at column %d", brk);
    return fail(QTQP_ERR_FACTOR, buf);
  }
  h->fprec = 64;
  h->have_factor = true;
  return QTQP_OK;
}

// K_reg solve of r with the factor in precision T: [x; y] into out (or added
// to it): g = r_x + A'(W r_y), two triangular sweeps for x, y = W(A x - r_y).
//  With add_to, [x; y] is added to add_to instead (out is not used).
template <typename T>
int solve_with(qtqp_kkt* h, const T* A, const T* L, const double* r, double* out, int accumulate,
               const Gate* g = nullptr, int want = 0, double* add_to = nullptr) {
  const int64_t n = h->n, m = h->m, N = n + m;
  int rc;
  if ((rc = tile_pass<T>(h, A, m, n, nullptr, nullptr, r + n, h->diag + n, g, want))) return rc;
  const int nt = (int)((n + NB - 1) / NB);
  const T* LiB = reinterpret_cast<const T*>(h->linv);
  T* tv = reinterpret_cast<T*>(h->tv);
  hipLaunchKernelGGL(k_trsv_fwd<T>, dim3(nt), dim3(WG), 0, h->stream, L, (int64_t)n, (int)n, LiB,
                     r, h->part, tiles_r(m), tv, h->sync, g, want);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_trsv_bwd<T>, dim3(nt), dim3(WG), 0, h->stream, L, (int64_t)n, (int)n, LiB,
                     tv, h->tx, h->sync, g, want);
  HIPC(hipGetLastError());
  if ((rc = tile_pass<T>(h, A, m, n, h->tx, h->rpa, nullptr, nullptr, g, want))) return rc;
  if (add_to)
    hipLaunchKernelGGL(k_solve_add, dim3(blocks_for(N)), dim3(WG), 0, h->stream, (int)n, (int)m,
                       h->tx, h->rpa, tiles_c(n), r, h->diag, add_to, g, want);
  else
    hipLaunchKernelGGL(k_solve_finish, dim3(blocks_for(N)), dim3(WG), 0, h->stream, (int)n,
                       (int)m, h->tx, h->rpa, tiles_c(n), r, h->diag, out, accumulate, g, want);
  HIPC(hipGetLastError());
  if (!g) h->info_acc.backend_solves++;
  return QTQP_OK;
}

int solve_current(qtqp_kkt* h, const double* r, double* out, int accumulate) {
  if (h->fprec == 64) return solve_with<double>(h, h->A64, h->G64, r, out, accumulate);
  return solve_with<float>(h, h->A32, h->G32, r, out, accumulate);
}

// Backend solve: out = K_reg^{-1} r to FP64 accuracy (device buffers).  With
// abs_target > 0 the FP32 refinement may stop once ||r - K_reg out||_inf is
// below it: a caller refining against a looser tolerance gains nothing from
// the last digits.
//  r0_known >= 0 is ||r||_inf when the caller has it (saves a readback).
int backend_solve(qtqp_kkt* h, const double* r, double* out, double abs_target = 0.0,
                  double r0_known = -1.0) {
  const int64_t N = h->n + h->m;
  if (h->fprec == 64) return solve_current(h, r, out, 0);
  // FP32 factor: refine in FP64 against the exact K_reg.
  // An FP32 factor of a Gram block with condition number k contracts the
  // residual by about n * 6e-8 * k per step; past kPoorContraction it would
  // need too many steps, and an FP64 refactor is cheaper.
  constexpr int kMaxInner = 40;
  constexpr double kTarget = 1e-15;  // relative residual reached by an FP64 factor
  constexpr double kPoorContraction = 0.25;
  constexpr double kFloor = 1e-13;
  int rc;
  double r0 = 0.0, prev = 0.0, cur = 0.0;
  // the residual of x = 0 is r itself: rres = r, r0 = ||r||_inf, racc = 0
  if (r0_known >= 0.0) {
    r0 = r0_known;
    HIPC(hipMemsetAsync(h->racc, 0, sizeof(double) * N, h->stream));
    HIPC(hipMemcpyAsync(h->rres, r, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
  } else {
    HIPC(hipMemsetAsync(h->dnorm, 0, sizeof(unsigned long long), h->stream));
    HIPC(hipMemsetAsync(h->racc, 0, sizeof(double) * N, h->stream));
    hipLaunchKernelGGL(k_copy_norm, dim3(blocks_for(N)), dim3(WG), 0, h->stream, N, r, h->rres,
                       h->dnorm);
    HIPC(hipGetLastError());
    unsigned long long bits = 0;
    HIPC(hipMemcpyAsync(&bits, h->dnorm, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    std::memcpy(&r0, &bits, sizeof(double));
  }
  if (r0 == 0.0) {
    HIPC(hipMemsetAsync(out, 0, sizeof(double) * N, h->stream));
    return QTQP_OK;
  }
  prev = r0;
  h->f32_solves++;
  for (int step = 0; step < kMaxInner; ++step) {
    if ((rc = solve_current(h, h->rres, h->racc, 1))) return rc;
    if (h->fprec == 32) {
      h->info_acc.inner_steps++;
      h->f32_steps++;
    }
    if ((rc = kmatvec(h, h->racc, r, nullptr, h->rres, &cur))) return rc;
    if (!(cur == cur)) break;
    if (cur <= kTarget * r0 || cur <= abs_target) break;
    if (cur > kPoorContraction * prev) {
      // Stalled.  Below kFloor the FP64 residual itself is at its rounding
      // floor (an FP64 factor would stop there too); above it the FP32
      // factor is too inaccurate at this conditioning: refactor in FP64.
      if (cur <= kFloor * r0 || h->fprec == 64) break;
      h->info_acc.fp64_fallbacks++;
      h->force64 = true;
      if ((rc = factor64(h))) return rc;
      continue;  // keep the accumulated correction; refine with the FP64 factor
    }
    prev = cur;
  }
  HIPC(hipMemcpyAsync(out, h->racc, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
  return QTQP_OK;
}

template <typename T>
T* dalloc(int64_t count, hipError_t* err) {
  void* p = nullptr;
  *err = hipMalloc(&p, sizeof(T) * std::max<int64_t>(count, 1));
  return static_cast<T*>(p);
}

void release(qtqp_kkt* h) {
  void* bufs[] = {h->A64, h->Poff64, h->A32, h->Poff32, h->diag, h->corr, h->G64, h->G32,
                  h->info, h->vin, h->vout, h->vsol, h->vres, h->rin, h->racc, h->rres,
                  h->rtmp, h->tx, h->tg, h->part, h->dnorm,
                  h->tv, h->linv, h->sync, h->rpa, h->rpp, h->aty, h->gpart};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

double elapsed_ms(qtqp_kkt* h) {
  float ms = 0.f;
  if (hipEventRecord(h->ev1, h->stream) != hipSuccess) return 0.0;
  if (hipEventSynchronize(h->ev1) != hipSuccess) return 0.0;
  if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return 0.0;
  return ms;
}

// Allocates a handle and uploads A (m x n row-major) and P (n x n row-major,
// its diagonal dropped; NULL for P = 0).  With cast, the FP32 copies of a
// mixed-precision handle are made now; otherwise the caller makes them
// (after rescaling the FP64 data) with cast_fp32.
int cast_fp32(qtqp_kkt* h) {
  if (h->precision != QTQP_PREC_MIXED) return QTQP_OK;
  hipLaunchKernelGGL(k_cast<float>, dim3(1024), dim3(WG), 0, h->stream, h->A64, h->A32,
                     h->m * h->n);
  hipLaunchKernelGGL(k_cast<float>, dim3(256), dim3(WG), 0, h->stream, h->Poff64, h->Poff32,
                     h->n * h->n);
  HIPC(hipGetLastError());
  return QTQP_OK;
}

//  With col_major, a and p are given column-major (transposed on the device).
int engine_create(int64_t n, int64_t m, const double* a, const double* p, int32_t precision,
                  bool cast, qtqp_kkt** out, bool col_major = false) {
  if (!out) return fail(QTQP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (n < 1 || m < 1) return fail(QTQP_ERR_ARG, "n and m must be >= 1");
  if (n > 46340 || n * m > (int64_t)1 << 40) return fail(QTQP_ERR_ARG, "problem too large");
  if (!a) return fail(QTQP_ERR_ARG, "a is NULL");
  if (precision != QTQP_PREC_FP64 && precision != QTQP_PREC_MIXED)
    return fail(QTQP_ERR_ARG, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(QTQP_ERR_HIP, "no HIP device available");
  auto* h = new qtqp_kkt();
  h->n = n;
  h->m = m;
  h->precision = precision;
  h->info_acc.precision = precision;
  const int64_t N = n + m;
  hipError_t e = hipSuccess;
  auto bad = [&](const char* what) {
    std::string msg = std::string("allocation/copy failed: ") + what + ": " + hipGetErrorString(e);
    release(h);
    delete h;
    return fail(QTQP_ERR_HIP, msg);
  };
  if ((e = hipStreamCreate(&h->stream)) != hipSuccess) return bad("stream");
  if ((e = hipEventCreate(&h->ev0)) != hipSuccess) return bad("event");
  if ((e = hipEventCreate(&h->ev1)) != hipSuccess) return bad("event");
#define ALLOC(field, T, count) \
  if ((h->field = dalloc<T>((count), &e)), e != hipSuccess) return bad(#field)
  ALLOC(A64, double, m * n);
  ALLOC(Poff64, double, n * n);
  ALLOC(diag, double, N);
  ALLOC(corr, double, N);
  ALLOC(G64, double, n * n);
  ALLOC(info, int, 1);
  ALLOC(vin, double, N);
  ALLOC(vout, double, N);
  ALLOC(vsol, double, N);
  ALLOC(vres, double, N);
  ALLOC(rin, double, N);
  ALLOC(racc, double, N);
  ALLOC(rres, double, N);
  ALLOC(rtmp, double, N);
  ALLOC(tx, double, n);
  ALLOC(tg, double, n);
  ALLOC(tv, double, n);
  const int64_t nt = (n + NB - 1) / NB;
  ALLOC(linv, double, nt * 2 * NB * NB);
  ALLOC(sync, unsigned, sync_words(n));
  {
    // Split the Gram product over the rows of A when its lower tiles alone
    // would leave most of the chip idle: about 512 workgroups, chunks of at
    // least 128 rows.
    const int64_t ng = nt * (nt + 1) / 2;
    const int64_t want_chunks = std::min<int64_t>((512 + ng - 1) / ng, (m + 127) / 128);
    if (want_chunks > 1) {
      h->gchunk = (int)((m + want_chunks - 1) / want_chunks);
      h->gsplit = (int)((m + h->gchunk - 1) / h->gchunk);
    }
  }
  if (h->gsplit > 1) ALLOC(gpart, double, (nt * (nt + 1) / 2) * h->gsplit * NB * NB);
  ALLOC(part, double, (int64_t)nblk_rows(std::max(m, n)) * n);
  ALLOC(rpa, double, m * tiles_c(n));
  ALLOC(rpp, double, n * tiles_c(n));
  ALLOC(aty, double, n);
  ALLOC(dnorm, unsigned long long, 3);
  if (precision == QTQP_PREC_MIXED) {
    ALLOC(A32, float, m * n);
    ALLOC(Poff32, float, n * n);
    ALLOC(G32, float, n * n);
  }
#undef ALLOC
  // Column-major inputs are staged in G64 (P) and a temporary buffer (A).
  double* stage = nullptr;
  if (col_major) {
    if ((e = hipMalloc(&stage, sizeof(double) * m * n)) != hipSuccess) return bad("staging");
    if ((e = hipMemcpyAsync(stage, a, sizeof(double) * m * n, hipMemcpyHostToDevice, h->stream)) ==
        hipSuccess) {
      hipLaunchKernelGGL(k_transpose, dim3((m + NB - 1) / NB, (n + NB - 1) / NB), dim3(WG), 0,
                         h->stream, stage, h->A64, m, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return bad("A");
  } else if ((e = hipMemcpyAsync(h->A64, a, sizeof(double) * m * n, hipMemcpyHostToDevice,
                                 h->stream)) != hipSuccess) {
    return bad("A");
  }
  if (p) {
    double* pdst = col_major ? h->G64 : h->Poff64;
    if ((e = hipMemcpyAsync(pdst, p, sizeof(double) * n * n, hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
      return bad("P");
    if (col_major) {
      hipLaunchKernelGGL(k_transpose, dim3((n + NB - 1) / NB, (n + NB - 1) / NB), dim3(WG), 0,
                         h->stream, h->G64, h->Poff64, n, n);
      if ((e = hipGetLastError()) != hipSuccess) return bad("P transpose");
    }
    hipLaunchKernelGGL(k_zero_diag, dim3(blocks_for(n)), dim3(WG), 0, h->stream, h->Poff64, n);
    if ((e = hipGetLastError()) != hipSuccess) return bad("P diagonal");
  } else if ((e = hipMemsetAsync(h->Poff64, 0, sizeof(double) * n * n, h->stream)) != hipSuccess) {
    return bad("P");
  }
  if ((e = hipMemsetAsync(h->corr, 0, sizeof(double) * N, h->stream)) != hipSuccess) return bad("corr");
  if ((e = hipMemsetAsync(h->sync, 0, sizeof(unsigned) * sync_words(n), h->stream)) !=
      hipSuccess)
    return bad("sync");
  if (cast && cast_fp32(h) != QTQP_OK) {
    e = hipGetLastError();
    return bad("cast");
  }
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bad("sync");
  *out = h;
  return QTQP_OK;
}

// Mean FP32 refinement steps per backend solve above which a mixed handle
// factors in FP64 from the next factorisation on: an FP64 factor solves in
// one pair of triangular sweeps, and the extra sweeps and FP64 residuals of
// the FP32 refinement then cost more than the FP32 factor saves.
constexpr int64_t kF32StepsPerSolve = 2;

// LinearSolver.factorize (solvers_dense.py:99-116) on the diagonal in
// h->diag: FP32 factor for a mixed handle (FP64 refactor on breakdown, and
// FP64 from then on, as after a factor whose solves refined too slowly),
// FP64 otherwise.
int engine_factorize(qtqp_kkt* h) {
  HIPC(hipEventRecord(h->ev0, h->stream));
  int rc = QTQP_OK;
  h->info_acc.factorizations++;
  if (h->precision == QTQP_PREC_MIXED && !h->force64 && h->fprec == 32 &&
      h->f32_steps > kF32StepsPerSolve * h->f32_solves) {
    h->info_acc.fp64_fallbacks++;
    h->force64 = true;
  }
  h->f32_solves = h->f32_steps = 0;
  if (h->precision == QTQP_PREC_MIXED && !h->force64) {
    int brk = 0;
    rc = factor_prec<float>(h, h->A32, h->Poff32, h->G32, &brk);
    if (rc == QTQP_OK && brk == 0) {
      h->fprec = 32;
      h->have_factor = true;
    } else if (rc == QTQP_OK) {
      h->info_acc.fp64_fallbacks++;
      h->force64 = true;
      rc = factor64(h);
    }
  } else {
    rc = factor64(h);
  }
  h->info_acc.factor_ms += elapsed_ms(h);
  return rc;
}

// DirectKktSolver.solve (direct.py:218-307) on device buffers: refines sol
// (warm start in, solution out) against K_true = K_reg - diag(corr) for the
// right-hand side [rhs_x; -rhs_y] until ||res||_inf < atol + rtol ||rhs||_inf,
// a stall (the residual stops decreasing), or max_steps backend solves.
// Returns QTQP_ERR_NAN when sol has NaNs (direct.py:293-294).
int ir_refine(qtqp_kkt* h, const double* rhs, double* sol, int max_steps, double atol,
              double rtol, qtqp_ir_stats* st) {
  const int64_t n = h->n, N = h->n + h->m;
  HIPC(hipMemsetAsync(h->dnorm + 2, 0, sizeof(unsigned long long), h->stream));
  hipLaunchKernelGGL(k_negate_y, dim3(blocks_for(N)), dim3(WG), 0, h->stream, N, n, rhs, h->vin,
                     h->dnorm + 2);
  HIPC(hipGetLastError());
  double res = 0.0, old = 0.0, rhs_norm = 0.0;
  int nanflag = 0;
  int rc = kmatvec(h, sol, h->vin, h->corr, h->vres, &res, &rhs_norm, &nanflag);
  if (rc) return rc;
  // one rounding, as the device-side control computes it
  const double tol = std::fma(rtol, rhs_norm, atol);
  int status = 1, solves = 0;
  for (solves = 1; solves <= max_steps; ++solves) {
    old = res;
    // An inner residual a tenth of the tolerance leaves the next outer
    // residual below it (up to the correction term).
    if ((rc = backend_solve(h, h->vres, h->vout, 0.1 * tol, res))) return rc;
    hipLaunchKernelGGL(k_axpy, dim3(blocks_for(N)), dim3(WG), 0, h->stream, N, 1.0, h->vout, sol);
    HIPC(hipGetLastError());
    if ((rc = kmatvec(h, sol, h->vin, h->corr, h->vres, &res, nullptr, &nanflag))) return rc;
    if (res < tol) {
      status = 0;
      break;
    }
    if (res >= old) {
      status = 2;
      break;
    }
  }
  if (solves > max_steps) solves = max_steps;
  // nanflag comes from the last residual, which was formed from the final sol
  if (st) {
    st->solves = solves;
    st->status = status;
    st->final_residual_norm = res;
    st->rhs_norm = rhs_norm;
    st->tolerance = tol;
  }
  if (nanflag) return fail(QTQP_ERR_NAN, "Linear solver returned NaNs.");
  return QTQP_OK;
}

}  // namespace
