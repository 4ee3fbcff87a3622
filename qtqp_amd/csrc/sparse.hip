// Sparse KKT engine: supernodal LDL' of the quasi-definite KKT matrix
//
//     K = [ P + R_x    A'   ]
//         [   A      -R_y   ]     (R_x, R_y > 0 diagonal)
//
// on one MI355X.  A quasi-definite matrix has an LDL' factorisation for every
// symmetric permutation (D holds n positive and m negative pivots), so the
// ordering is chosen for fill alone and the pivots are static, as in the
// reference's sparse backends (QdldlSolver, solvers_sparse.py:74-93;
// CuDssSolver, solvers_gpu.py:27-115).
//
// Host (once per pattern, qtqp_sparse_analyze / create):
//   1. minimum-degree ordering on a bitset elimination graph; once the
//      remaining graph is a clique the rest is taken in index order;
//   2. elimination tree (path-compressed ancestor walk) and a postorder, so
//      every subtree is a contiguous column range;
//   3. column structures of L by merging each child's structure into its
//      parent (each structure is freed once its parent has consumed it);
//   4. supernodes: column j joins j-1's supernode when parent(j-1) = j and
//      |struct(j-1)| = |struct(j)| + 1 (then struct(j-1) = {j} u struct(j)),
//      at most kMaxCols columns each;
//   5. levels of the supernodal tree (a supernode's level is one above its
//      highest child) and, per level, the list of 64x64 Schur-update tiles.
//
// Device (per factorize / solve), one launch per level and kernel:
//   * k_panel: one workgroup per supernode factors its dense r x c column
//     panel (rows = the supernode's row structure) in place: right-looking
//     LDL' over the c pivot columns, L scaled by D, D on the diagonal;
//   * k_update: one workgroup per 64x64 tile of the supernode's Schur update
//     U = L21 D L21' (L21 the panel's rows below the pivot block), scattered
//     with FP64 atomics into the ancestor supernodes that own those entries
//     (several supernodes of one level may update one ancestor);
//   * k_forward / k_backward: level-ordered triangular solves, one
//     workgroup per supernode; the forward pass scatters its below-block
//     products into the right-hand side with atomics, the backward pass
//     gathers the solved ancestor values.
// Storage in HBM: each supernode's panel column-major (ld = its row count),
// panels back to back; an untouched copy holds K's off-diagonal values in
// the same layout, copied over the panel storage before each factorisation.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "qtqp_kkt.h"
#include "qtqp_sparse.h"

namespace {

constexpr int kMaxCols = 64;      // pivot columns per supernode
constexpr int kTile = 64;         // Schur-update tile edge
constexpr int kThreads = 256;     // workgroup size of every kernel
constexpr int64_t kMaxOrderN = 40000;  // larger systems keep their given order

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return fail(QTQP_ERR_HIP, std::string(#expr ": ") +      \
                                                        hipGetErrorString(e_));    \
  } while (0)

// ---------------------------------------------------------------------------
// Symbolic analysis (host)
// ---------------------------------------------------------------------------

struct Symbolic {
  int64_t n = 0, nnz_kkt = 0, nnz_l = 0, max_rows = 0;
  double flops = 0.0;
  std::vector<int32_t> perm, iperm;       // perm[k] = original index of column k
  std::vector<int32_t> sn_first;          // S+1: first column of each supernode
  std::vector<int64_t> sn_rows;           // S+1: offsets into rows
  std::vector<int64_t> sn_lx;             // S+1: offsets into the panel storage
  std::vector<int32_t> rows;              // row structure of each supernode (sorted)
  std::vector<int32_t> col_sn;            // supernode of each column
  std::vector<int32_t> level_sn;          // supernodes ordered by level
  std::vector<int64_t> level_off;         // L+1
  std::vector<int32_t> tiles;             // (supernode, a0, b0) per update tile
  std::vector<int64_t> tile_off;          // L+1
};

// Symmetric adjacency (no diagonal) of the upper triangle given as
// compressed major/minor arrays; rejects entries outside the upper triangle.
int adjacency(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
              std::vector<std::vector<int32_t>>& adj, int64_t& nnz) {
  if (format != QTQP_SPARSE_CSC && format != QTQP_SPARSE_CSR)
    return fail(QTQP_ERR_ARG, "format must be QTQP_SPARSE_CSC or QTQP_SPARSE_CSR");
  if (ptr[0] != 0) return fail(QTQP_ERR_ARG, "ptr[0] must be 0");
  adj.assign(n, {});
  for (int64_t j = 0; j < n; ++j) {
    if (ptr[j + 1] < ptr[j]) return fail(QTQP_ERR_ARG, "ptr must be non-decreasing");
    for (int64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
      const int64_t i = ind[p];
      if (i < 0 || i >= n) return fail(QTQP_ERR_ARG, "index out of range");
      const bool upper = format == QTQP_SPARSE_CSC ? i <= j : i >= j;
      if (!upper) return fail(QTQP_ERR_ARG, "entry outside the upper triangle");
      if (i == j) continue;
      adj[j].push_back((int32_t)i);
      adj[i].push_back((int32_t)j);
    }
  }
  nnz = ptr[n];
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  return QTQP_OK;
}

// Minimum degree on the explicit elimination graph, held as one bitset row
// per vertex: eliminating v joins its neighbours into a clique.  Ties go to
// the lowest index, so the order is deterministic.
std::vector<int32_t> min_degree(int64_t n, const std::vector<std::vector<int32_t>>& adj) {
  std::vector<int32_t> order;
  order.reserve(n);
  if (n > kMaxOrderN) {
    for (int64_t i = 0; i < n; ++i) order.push_back((int32_t)i);
    return order;
  }
  const int64_t w = (n + 63) / 64;
  std::vector<uint64_t> bits((size_t)(n * w), 0);
  std::vector<int64_t> deg(n);
  std::vector<char> done(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    for (int32_t j : adj[i]) bits[i * w + j / 64] |= 1ull << (j % 64);
    deg[i] = (int64_t)adj[i].size();
  }
  int64_t left = n;
  while (left > 0) {
    int64_t v = -1, best = INT64_MAX;
    for (int64_t i = 0; i < n; ++i)
      if (!done[i] && deg[i] < best) best = deg[i], v = i;
    if (best == left - 1) {  // a clique remains: any order adds no fill
      for (int64_t i = 0; i < n; ++i)
        if (!done[i]) order.push_back((int32_t)i);
      break;
    }
    done[v] = 1;
    order.push_back((int32_t)v);
    --left;
    const uint64_t* bv = &bits[v * w];
    for (int64_t q = 0; q < w; ++q) {
      uint64_t word = bv[q];
      while (word) {
        const int64_t u = q * 64 + __builtin_ctzll(word);
        word &= word - 1;
        uint64_t* bu = &bits[u * w];
        int64_t d = 0;
        for (int64_t k = 0; k < w; ++k) {
          bu[k] |= bv[k];
          if (k == u / 64) bu[k] &= ~(1ull << (u % 64));
          if (k == v / 64) bu[k] &= ~(1ull << (v % 64));
          d += __builtin_popcountll(bu[k]);
        }
        deg[u] = d;
      }
    }
  }
  return order;
}

// Elimination tree of the pattern in the order `perm` (padj: neighbours in
// permuted indices).
std::vector<int32_t> etree(int64_t n, const std::vector<std::vector<int32_t>>& padj) {
  std::vector<int32_t> parent(n, -1), anc(n, -1);
  for (int64_t k = 0; k < n; ++k) {
    for (int32_t i : padj[k]) {
      while (i != -1 && i < k) {
        const int32_t next = anc[i];
        anc[i] = (int32_t)k;
        if (next == -1) parent[i] = (int32_t)k;
        i = next;
      }
    }
  }
  return parent;
}

std::vector<std::vector<int32_t>> permuted_adjacency(const std::vector<std::vector<int32_t>>& adj,
                                                     const std::vector<int32_t>& perm,
                                                     const std::vector<int32_t>& iperm) {
  const int64_t n = (int64_t)perm.size();
  std::vector<std::vector<int32_t>> padj(n);
  for (int64_t k = 0; k < n; ++k) {
    padj[k].reserve(adj[perm[k]].size());
    for (int32_t j : adj[perm[k]]) padj[k].push_back(iperm[j]);
    std::sort(padj[k].begin(), padj[k].end());
  }
  return padj;
}

int analyze(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind, Symbolic& S) {
  if (n < 1) return fail(QTQP_ERR_ARG, "n must be positive");
  if (n >= INT32_MAX) return fail(QTQP_ERR_ARG, "n too large");
  if (!ptr || !ind) return fail(QTQP_ERR_ARG, "null pattern");
  std::vector<std::vector<int32_t>> adj;
  if (int rc = adjacency(n, format, ptr, ind, adj, S.nnz_kkt)) return rc;
  S.n = n;

  // 1-2. ordering, then a postorder of its elimination tree.
  std::vector<int32_t> perm = min_degree(n, adj), iperm(n);
  for (int64_t k = 0; k < n; ++k) iperm[perm[k]] = (int32_t)k;
  {
    const auto padj = permuted_adjacency(adj, perm, iperm);
    const auto parent = etree(n, padj);
    std::vector<int32_t> head(n, -1), next(n, -1), post;
    post.reserve(n);
    for (int64_t k = n - 1; k >= 0; --k)  // children lists in increasing order
      if (parent[k] != -1) next[k] = head[parent[k]], head[parent[k]] = (int32_t)k;
    std::vector<int32_t> stack;
    for (int64_t r = 0; r < n; ++r) {
      if (parent[r] != -1) continue;
      stack.push_back((int32_t)r);
      while (!stack.empty()) {
        const int32_t v = stack.back();
        if (head[v] != -1) {  // descend into the next unvisited child
          const int32_t c = head[v];
          head[v] = next[c];
          stack.push_back(c);
        } else {
          post.push_back(v);
          stack.pop_back();
        }
      }
    }
    std::vector<int32_t> p2(n);
    for (int64_t k = 0; k < n; ++k) p2[k] = perm[post[k]];
    perm.swap(p2);
    for (int64_t k = 0; k < n; ++k) iperm[perm[k]] = (int32_t)k;
  }
  const auto padj = permuted_adjacency(adj, perm, iperm);
  const auto parent = etree(n, padj);

  // 3-4. column structures and supernodes.
  std::vector<int32_t> nchild(n, 0), head(n, -1), next(n, -1);
  for (int64_t k = n - 1; k >= 0; --k)
    if (parent[k] != -1) next[k] = head[parent[k]], head[parent[k]] = (int32_t)k, ++nchild[parent[k]];
  std::vector<std::vector<int32_t>> live(n);
  std::vector<int32_t> mark(n, -1);
  std::vector<int64_t> count(n);
  S.sn_first.clear();
  S.sn_rows.assign(1, 0);
  S.rows.clear();
  S.col_sn.assign(n, 0);
  int64_t start = 0;
  for (int64_t k = 0; k < n; ++k) {
    std::vector<int32_t> s;
    mark[k] = (int32_t)k;
    for (int32_t i : padj[k])
      if (i > k && mark[i] != k) mark[i] = (int32_t)k, s.push_back(i);
    for (int32_t c = head[k]; c != -1; c = next[c]) {
      for (int32_t i : live[c])
        if (mark[i] != k) mark[i] = (int32_t)k, s.push_back(i);
      std::vector<int32_t>().swap(live[c]);
    }
    std::sort(s.begin(), s.end());
    count[k] = (int64_t)s.size() + 1;
    const bool cont = k > 0 && parent[k - 1] == k && count[k - 1] == count[k] + 1 &&
                      k - start < kMaxCols;
    if (!cont) {
      start = k;
      if (!S.sn_first.empty()) S.sn_rows.push_back((int64_t)S.rows.size());
      S.sn_first.push_back((int32_t)k);
      S.rows.push_back((int32_t)k);
      S.rows.insert(S.rows.end(), s.begin(), s.end());
    }
    S.col_sn[k] = (int32_t)S.sn_first.size() - 1;
    live[k].swap(s);
    if (parent[k] == -1) std::vector<int32_t>().swap(live[k]);
  }
  S.sn_rows.push_back((int64_t)S.rows.size());
  const int64_t ns = (int64_t)S.sn_first.size();
  S.sn_first.push_back((int32_t)n);

  // Panel offsets, fill and flop counts.
  S.sn_lx.assign(ns + 1, 0);
  S.nnz_l = 0;
  S.flops = 0.0;
  S.max_rows = 0;
  for (int64_t s = 0; s < ns; ++s) {
    const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
    const int64_t r = S.sn_rows[s + 1] - S.sn_rows[s];
    S.sn_lx[s + 1] = S.sn_lx[s] + r * c;
    S.max_rows = std::max(S.max_rows, r);
    for (int64_t j = 0; j < c; ++j) {
      const double below = (double)(r - j - 1);
      S.nnz_l += r - j;
      S.flops += below + below * (below + 1.0);  // scale + symmetric rank-1 update
    }
  }

  // 5. levels and update tiles.
  std::vector<int32_t> level(ns, 0);
  int32_t nlev = 0;
  for (int64_t s = 0; s < ns; ++s) {
    const int32_t last = S.sn_first[s + 1] - 1;
    if (parent[last] != -1) {
      const int32_t p = S.col_sn[parent[last]];
      level[p] = std::max(level[p], level[s] + 1);
    }
    nlev = std::max(nlev, level[s] + 1);
  }
  std::vector<std::vector<int32_t>> by_level(nlev);
  for (int64_t s = 0; s < ns; ++s) by_level[level[s]].push_back((int32_t)s);
  S.level_sn.clear();
  S.level_off.assign(1, 0);
  S.tiles.clear();
  S.tile_off.assign(1, 0);
  for (int32_t l = 0; l < nlev; ++l) {
    for (int32_t s : by_level[l]) {
      S.level_sn.push_back(s);
      const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
      const int64_t nb = (S.sn_rows[s + 1] - S.sn_rows[s] - c + kTile - 1) / kTile;
      for (int64_t bj = 0; bj < nb; ++bj)
        for (int64_t bi = bj; bi < nb; ++bi)
          S.tiles.insert(S.tiles.end(), {s, (int32_t)(bi * kTile), (int32_t)(bj * kTile)});
    }
    S.level_off.push_back((int64_t)S.level_sn.size());
    S.tile_off.push_back((int64_t)S.tiles.size() / 3);
  }
  S.perm = perm;
  S.iperm = iperm;
  return QTQP_OK;
}

// Position in the panel storage of the lower entry (row, col), both permuted,
// row >= col; -1 when the pattern does not hold it.
int64_t panel_pos(const Symbolic& S, int32_t row, int32_t col) {
  const int32_t t = S.col_sn[col];
  const int32_t* rb = S.rows.data() + S.sn_rows[t];
  const int64_t r = S.sn_rows[t + 1] - S.sn_rows[t];
  const int32_t lc = col - S.sn_first[t];
  const int32_t* it = std::lower_bound(rb + lc, rb + r, row);
  if (it == rb + r || *it != row) return -1;
  return S.sn_lx[t] + (int64_t)lc * r + (it - rb);
}

void fill_info(const Symbolic& S, qtqp_sparse_info* info) {
  info->n = S.n;
  info->nnz_kkt = S.nnz_kkt;
  info->nnz_l = S.nnz_l;
  info->supernodes = (int64_t)S.sn_first.size() - 1;
  info->levels = (int64_t)S.level_off.size() - 1;
  info->max_rows = S.max_rows;
  info->update_tiles = (int64_t)S.tiles.size() / 3;
  info->factor_flops = S.flops;
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// Supernode tables, device pointers.
struct Tables {
  const int32_t* first;    // S+1
  const int64_t* rows_off; // S+1
  const int64_t* lx_off;   // S+1
  const int32_t* rows;
  const int32_t* col_sn;
  const int32_t* level_sn;
  const int32_t* tiles;
};

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_load_diag(double* __restrict__ lx, const int64_t* __restrict__ dpos,
                            const double* __restrict__ diag, const int32_t* __restrict__ perm,
                            int64_t n) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) lx[dpos[k]] = diag[perm[k]];
}

// Right-looking LDL' of one supernode's r x c panel (column-major, ld = r).
// On exit column j holds d_j on the diagonal and l_ij = a_ij / d_j below it.
// A zero or non-finite pivot records (original column + 1) in *bad.
__global__ void __launch_bounds__(kThreads) k_panel(Tables T, double* __restrict__ lx,
                                                    int64_t lvl_begin, int* __restrict__ bad,
                                                    const int32_t* __restrict__ perm) {
  __shared__ double w[kMaxCols];
  const int32_t s = T.level_sn[lvl_begin + blockIdx.x];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int r = (int)(T.rows_off[s + 1] - T.rows_off[s]);
  double* L = lx + T.lx_off[s];
  const int tid = threadIdx.x;
  for (int k = 0; k < c; ++k) {
    double* Lk = L + (int64_t)k * r;
    const double d = Lk[k];
    if (!(d != 0.0) || !isfinite(d)) {  // uniform across the workgroup
      if (tid == 0) atomicCAS(bad, 0, perm[f + k] + 1);
      return;
    }
    const double rd = 1.0 / d;
    for (int j = k + 1 + tid; j < c; j += kThreads) w[j] = Lk[j] * rd;
    __syncthreads();
    for (int i = k + 1 + tid; i < r; i += kThreads) {
      const double v = Lk[i];
      Lk[i] = v * rd;
      const int jend = min(i, c - 1);
      for (int j = k + 1; j <= jend; ++j) L[(int64_t)j * r + i] -= v * w[j];
    }
    __syncthreads();
  }
}

// Binary search of `row` in the sorted list rows[lo, hi).
__device__ __forceinline__ int find_row(const int32_t* __restrict__ rows, int lo, int hi, int row) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rows[mid] < row) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One 64x64 tile of a supernode's Schur update U = L21 D L21' (rows a, b of
// the panel below its pivot block, a >= b), subtracted from the ancestor
// entries (rows[c+a], rows[c+b]).  Each thread owns a 4x4 block of the tile.
__global__ void __launch_bounds__(kThreads) k_update(Tables T, double* __restrict__ lx,
                                                     int64_t tile_begin, int* __restrict__ bad) {
  constexpr int kChunk = 32;  // pivot columns staged at a time (33 KB of LDS)
  __shared__ double As[kChunk][kTile + 1];
  __shared__ double Bs[kChunk][kTile + 1];
  const int32_t* tl = T.tiles + 3 * (tile_begin + blockIdx.x);
  const int32_t s = tl[0];
  const int a0 = tl[1], b0 = tl[2];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const int nb = r - c;  // rows below the pivot block
  const double* L = lx + T.lx_off[s];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < c; k0 += kChunk) {
    const int kc = min(kChunk, c - k0);
    __syncthreads();  // the previous chunk has been consumed
    // Stage L21 rows a0.. and (scaled by D) rows b0.. of these pivot columns.
    for (int e = tid; e < kc * kTile; e += kThreads) {
      const int kk = e / kTile, t = e % kTile;
      const double* Lk = L + (int64_t)(k0 + kk) * r;
      const int a = a0 + t, b = b0 + t;
      As[kk][t] = a < nb ? Lk[c + a] : 0.0;
      Bs[kk][t] = b < nb ? Lk[c + b] * Lk[k0 + kk] : 0.0;
    }
    __syncthreads();
    for (int kk = 0; kk < kc; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) av[u] = As[kk][ty + 16 * u], bv[u] = Bs[kk][tx + 16 * u];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
  }
  const int32_t* R = T.rows + roff;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int b = b0 + tx + 16 * v;
    if (b >= nb) continue;
    const int32_t col = R[c + b];
    const int32_t t = T.col_sn[col];
    const int64_t troff = T.rows_off[t];
    const int tr = (int)(T.rows_off[t + 1] - troff);
    const int lc = col - T.first[t];
    const int32_t* tR = T.rows + troff;
    double* dst = lx + T.lx_off[t] + (int64_t)lc * tr;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int a = a0 + ty + 16 * u;
      if (a >= nb || a < b) continue;
      const int32_t row = R[c + a];
      const int pos = find_row(tR, lc, tr, row);
      if (pos < tr && tR[pos] == row) unsafeAtomicAdd(dst + pos, -acc[u][v]);
      else atomicCAS(bad, 0, -1);  // the symbolic structure misses a fill entry
    }
  }
}

// Forward substitution L y = b for the supernodes of one level, then
// z = D^{-1} y; x holds b on entry (with earlier levels' updates) and z on
// exit for these columns.
__global__ void __launch_bounds__(kThreads) k_forward(Tables T, const double* __restrict__ lx,
                                                      int64_t lvl_begin, double* __restrict__ x) {
  __shared__ double y[kMaxCols];
  const int32_t s = T.level_sn[lvl_begin + blockIdx.x];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const double* L = lx + T.lx_off[s];
  const int tid = threadIdx.x;
  if (tid < c) y[tid] = x[f + tid];
  __syncthreads();
  for (int k = 0; k < c - 1; ++k) {
    const double yk = y[k];
    const double* Lk = L + (int64_t)k * r;
    if (tid > k && tid < c) y[tid] -= Lk[tid] * yk;
    __syncthreads();
  }
  const int32_t* R = T.rows + roff;
  for (int a = c + tid; a < r; a += kThreads) {
    double t = 0.0;
    for (int k = 0; k < c; ++k) t += L[(int64_t)k * r + a] * y[k];
    unsafeAtomicAdd(x + R[a], -t);
  }
  if (tid < c) x[f + tid] = y[tid] / L[(int64_t)tid * r + tid];
}

// Backward substitution L' x = z for the supernodes of one level; the
// ancestors' entries of x are final.
__global__ void __launch_bounds__(kThreads) k_backward(Tables T, const double* __restrict__ lx,
                                                       int64_t lvl_begin, double* __restrict__ x) {
  __shared__ double xs[kMaxCols];
  const int32_t s = T.level_sn[lvl_begin + blockIdx.x];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const double* L = lx + T.lx_off[s];
  const int32_t* R = T.rows + roff;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  for (int k = wave; k < c; k += kThreads / 64) {
    const double* Lk = L + (int64_t)k * r;
    double t = 0.0;
    for (int a = c + lane; a < r; a += 64) t += Lk[a] * x[R[a]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0) xs[k] = x[f + k] - t;
  }
  __syncthreads();
  for (int i = c - 1; i > 0; --i) {
    const double xi = xs[i];
    if (tid < i) xs[tid] -= L[(int64_t)tid * r + i] * xi;
    __syncthreads();
  }
  if (tid < c) x[f + tid] = xs[tid];
}

__global__ void k_gather(double* __restrict__ out, const double* __restrict__ in,
                         const int32_t* __restrict__ perm, int64_t n) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[perm[k]];
}

__global__ void k_scatter(double* __restrict__ out, const double* __restrict__ in,
                          const int32_t* __restrict__ perm, int64_t n) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) out[perm[k]] = in[k];
}

// out = K x from the full symmetric pattern (off-diagonal entries, CSR) plus
// the current diagonal; one wavefront per row.
__global__ void __launch_bounds__(kThreads) k_spmv(const int64_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ diag,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)(kThreads / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= n) return;
  double t = 0.0;
  for (int64_t p = ptr[i] + lane; p < ptr[i + 1]; p += 64) t += val[p] * x[col[p]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  if (lane == 0) out[i] = t + diag[i] * x[i];
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }

}  // namespace

// ---------------------------------------------------------------------------
// Handle and C ABI
// ---------------------------------------------------------------------------

struct qtqp_sparse {
  Symbolic S;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> allocs;
  Tables T{};
  double* lx = nullptr;        // panels (factor)
  double* lx0 = nullptr;       // off-diagonal values of K in panel layout
  int64_t* dpos = nullptr;     // panel position of each (permuted) diagonal
  int32_t* perm = nullptr;
  double* diag = nullptr;      // current diagonal of K, original order
  int64_t* sym_ptr = nullptr;  // full symmetric off-diagonal pattern, CSR
  int32_t* sym_col = nullptr;
  double* sym_val = nullptr;
  double* xin = nullptr;       // N-vector staging (original order)
  double* xp = nullptr;        // N-vector work (permuted order)
  int* bad = nullptr;
  bool factored = false;
  int64_t factorizations = 0, solves = 0;
  double analyze_ms = 0, factor_ms = 0, solve_ms = 0;

  template <class T>
  int alloc(T** p, int64_t count) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    allocs.push_back(q);
    *p = (T*)q;
    return QTQP_OK;
  }
  template <class T>
  int upload(T** p, const std::vector<T>& v) {
    if (int rc = alloc(p, (int64_t)v.size())) return rc;
    if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return QTQP_OK;
  }
  ~qtqp_sparse() {
    for (void* p : allocs) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int build_handle(qtqp_sparse* h, int32_t format, const int64_t* ptr, const int64_t* ind,
                 const double* val) {
  const Symbolic& S = h->S;
  const int64_t n = S.n;
  // Off-diagonal values in panel layout, the diagonal positions, the stored
  // diagonal, and the full symmetric pattern for the product.
  std::vector<double> lx0((size_t)S.sn_lx.back(), 0.0), diag(n, 0.0);
  std::vector<int64_t> dpos(n), sptr(n + 1, 0);
  for (int64_t k = 0; k < n; ++k) dpos[k] = panel_pos(S, (int32_t)k, (int32_t)k);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t p = ptr[j]; p < ptr[j + 1]; ++p)
      if (ind[p] != j) ++sptr[j + 1], ++sptr[ind[p] + 1];
  for (int64_t i = 0; i < n; ++i) sptr[i + 1] += sptr[i];
  std::vector<int32_t> scol((size_t)sptr[n]);
  std::vector<double> sval((size_t)sptr[n]);
  std::vector<int64_t> fillp(sptr.begin(), sptr.end() - 1);
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t p = ptr[j]; p < ptr[j + 1]; ++p) {
      const int64_t i = ind[p];
      const double v = val[p];
      if (!std::isfinite(v)) return fail(QTQP_ERR_ARG, "non-finite value in K");
      if (i == j) {
        diag[j] += v;
        continue;
      }
      scol[fillp[j]] = (int32_t)i, sval[fillp[j]++] = v;
      scol[fillp[i]] = (int32_t)j, sval[fillp[i]++] = v;
      const int32_t pi = S.iperm[i], pj = S.iperm[j];
      const int64_t pos = panel_pos(S, std::max(pi, pj), std::min(pi, pj));
      if (pos < 0) return fail(QTQP_ERR_STATE, "symbolic structure misses an entry of K");
      lx0[pos] += v;
    }
  }
  (void)format;
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&h->ev0));
  HIP_TRY(hipEventCreate(&h->ev1));
  int32_t *first, *rows, *col_sn, *level_sn, *tiles;
  int64_t *rows_off, *lx_off;
  if (int rc = h->upload(&first, S.sn_first)) return rc;
  if (int rc = h->upload(&rows_off, S.sn_rows)) return rc;
  if (int rc = h->upload(&lx_off, S.sn_lx)) return rc;
  if (int rc = h->upload(&rows, S.rows)) return rc;
  if (int rc = h->upload(&col_sn, S.col_sn)) return rc;
  if (int rc = h->upload(&level_sn, S.level_sn)) return rc;
  if (int rc = h->upload(&tiles, S.tiles)) return rc;
  h->T = Tables{first, rows_off, lx_off, rows, col_sn, level_sn, tiles};
  if (int rc = h->upload(&h->lx0, lx0)) return rc;
  if (int rc = h->alloc(&h->lx, (int64_t)lx0.size())) return rc;
  if (int rc = h->upload(&h->dpos, dpos)) return rc;
  if (int rc = h->upload(&h->perm, S.perm)) return rc;
  if (int rc = h->upload(&h->diag, diag)) return rc;
  if (int rc = h->upload(&h->sym_ptr, sptr)) return rc;
  if (int rc = h->upload(&h->sym_col, scol)) return rc;
  if (int rc = h->upload(&h->sym_val, sval)) return rc;
  if (int rc = h->alloc(&h->xin, n)) return rc;
  if (int rc = h->alloc(&h->xp, n)) return rc;
  if (int rc = h->alloc(&h->bad, 1)) return rc;
  return QTQP_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  return hipEventElapsedTime(&ms, a, b) == hipSuccess ? (double)ms : 0.0;
}

}  // namespace

extern "C" {

int qtqp_sparse_analyze(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                        int64_t* perm_out, qtqp_sparse_info* info) {
  Symbolic S;
  const auto t0 = std::chrono::steady_clock::now();
  if (int rc = analyze(n, format, ptr, ind, S)) return rc;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (perm_out)
    for (int64_t k = 0; k < n; ++k) perm_out[k] = S.perm[k];
  if (info) {
    std::memset(info, 0, sizeof(*info));
    fill_info(S, info);
    info->analyze_ms = ms;
  }
  return QTQP_OK;
}

int qtqp_sparse_create(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                       const double* val, qtqp_sparse** out) {
  if (!out) return fail(QTQP_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!val) return fail(QTQP_ERR_ARG, "null values");
  auto* h = new qtqp_sparse();
  const auto t0 = std::chrono::steady_clock::now();
  int rc = analyze(n, format, ptr, ind, h->S);
  h->analyze_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc == QTQP_OK) rc = build_handle(h, format, ptr, ind, val);
  if (rc != QTQP_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return QTQP_OK;
}

int qtqp_sparse_update_diag(qtqp_sparse* h, const double* diag) {
  if (!h || !diag) return fail(QTQP_ERR_ARG, "null argument");
  for (int64_t i = 0; i < h->S.n; ++i)
    if (!std::isfinite(diag[i])) return fail(QTQP_ERR_ARG, "non-finite diagonal entry");
  HIP_TRY(hipMemcpyAsync(h->diag, diag, sizeof(double) * h->S.n, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->factored = false;
  return QTQP_OK;
}

int qtqp_sparse_factorize(qtqp_sparse* h) {
  if (!h) return fail(QTQP_ERR_ARG, "null handle");
  const Symbolic& S = h->S;
  h->factored = false;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  HIP_TRY(hipMemsetAsync(h->bad, 0, sizeof(int), h->stream));
  const int64_t nlx = S.sn_lx.back();
  k_copy<<<(unsigned)std::min<int64_t>(4096, blocks_for(nlx)), kThreads, 0, h->stream>>>(
      h->lx, h->lx0, nlx);
  k_load_diag<<<blocks_for(S.n), kThreads, 0, h->stream>>>(h->lx, h->dpos, h->diag, h->perm, S.n);
  const int64_t nlev = (int64_t)S.level_off.size() - 1;
  for (int64_t l = 0; l < nlev; ++l) {
    const int64_t ns = S.level_off[l + 1] - S.level_off[l];
    k_panel<<<(unsigned)ns, kThreads, 0, h->stream>>>(h->T, h->lx, S.level_off[l], h->bad,
                                                      h->perm);
    const int64_t nt = S.tile_off[l + 1] - S.tile_off[l];
    if (nt > 0)
      k_update<<<(unsigned)nt, kThreads, 0, h->stream>>>(h->T, h->lx, S.tile_off[l], h->bad);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  int bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, h->bad, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->factor_ms += elapsed_ms(h->ev0, h->ev1);
  ++h->factorizations;
  if (bad < 0) return fail(QTQP_ERR_STATE, "symbolic structure misses a fill entry");
  if (bad != 0)
    return fail(QTQP_ERR_FACTOR, "zero or non-finite pivot at KKT row " + std::to_string(bad - 1));
  h->factored = true;
  return QTQP_OK;
}

int qtqp_sparse_solve(qtqp_sparse* h, const double* rhs, double* out) {
  if (!h || !rhs || !out) return fail(QTQP_ERR_ARG, "null argument");
  if (!h->factored) return fail(QTQP_ERR_STATE, "solve before a successful factorize");
  const Symbolic& S = h->S;
  const int64_t n = S.n;
  HIP_TRY(hipMemcpyAsync(h->xin, rhs, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  k_gather<<<blocks_for(n), kThreads, 0, h->stream>>>(h->xp, h->xin, h->perm, n);
  const int64_t nlev = (int64_t)S.level_off.size() - 1;
  for (int64_t l = 0; l < nlev; ++l)
    k_forward<<<(unsigned)(S.level_off[l + 1] - S.level_off[l]), kThreads, 0, h->stream>>>(
        h->T, h->lx, S.level_off[l], h->xp);
  for (int64_t l = nlev - 1; l >= 0; --l)
    k_backward<<<(unsigned)(S.level_off[l + 1] - S.level_off[l]), kThreads, 0, h->stream>>>(
        h->T, h->lx, S.level_off[l], h->xp);
  k_scatter<<<blocks_for(n), kThreads, 0, h->stream>>>(h->xin, h->xp, h->perm, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipMemcpyAsync(out, h->xin, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->solve_ms += elapsed_ms(h->ev0, h->ev1);
  ++h->solves;
  return QTQP_OK;
}

int qtqp_sparse_matvec(qtqp_sparse* h, const double* x, double* out) {
  if (!h || !x || !out) return fail(QTQP_ERR_ARG, "null argument");
  const int64_t n = h->S.n;
  HIP_TRY(hipMemcpyAsync(h->xin, x, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  k_spmv<<<(unsigned)((n + kThreads / 64 - 1) / (kThreads / 64)), kThreads, 0, h->stream>>>(
      h->sym_ptr, h->sym_col, h->sym_val, h->diag, h->xin, h->xp, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipMemcpyAsync(out, h->xp, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->solve_ms += elapsed_ms(h->ev0, h->ev1);
  return QTQP_OK;
}

int qtqp_sparse_get_info(qtqp_sparse* h, qtqp_sparse_info* info) {
  if (!h || !info) return fail(QTQP_ERR_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  fill_info(h->S, info);
  info->factorizations = h->factorizations;
  info->solves = h->solves;
  info->analyze_ms = h->analyze_ms;
  info->factor_ms = h->factor_ms;
  info->solve_ms = h->solve_ms;
  return QTQP_OK;
}

void qtqp_sparse_free(qtqp_sparse* h) { delete h; }

const char* qtqp_sparse_last_error(void) { return g_err.c_str(); }

int qtqp_sparse_device_count(void) {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

}  // extern "C"
