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
//   1. minimum-degree ordering on a bitset elimination graph; neighbours left
//      simplicial by an elimination follow it at once, and once the remaining
//      graph is a clique the rest is taken in index order;
//   2. elimination tree (path-compressed ancestor walk), relabelled by a
//      postorder so that every subtree is a contiguous column range;
//   3. column structures of L by merging each child's structure into its
//      parent (each structure is freed once its parent has consumed it);
//   4. supernodes: column j joins j-1's supernode when parent(j-1) = j and
//      |struct(j-1)| = |struct(j)| + 1 (then struct(j-1) = {j} u struct(j)),
//      at most kMaxCols columns each;
//   5. levels of the supernodal tree (a supernode's level is one above its
//      highest child) and, per level, the job lists of the kernels below.
//
// Device, one launch per level and kernel (levels bottom-up to factor and
// solve forward, top-down to solve backward):
//   * k_panel: jobs of 256 rows below a pivot block; each factors the
//     supernode's pivot block in LDS, then its rows left-looking in registers;
//   * k_update: 64x64 tiles of each supernode's Schur update U = L21 D L21',
//     subtracted from the ancestor panels that own the entries (FP64 atomics
//     when several supernodes share the level);
//   * k_fwd: pivot-block solve by one wavefront, then 64-row jobs that
//     update the right-hand side of the ancestor rows (atomics);
//   * k_bwd_off / k_bwd_diag: 64-row jobs of partial sums over the solved
//     ancestor values, then the pivot-block solve by one wavefront.
// Storage in HBM: each supernode's panel column-major (ld = its row count),
// panels back to back; an untouched copy holds K's off-diagonal values in
// the same layout, copied over the panel storage before each factorisation;
// the factored pivot blocks in a separate array.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "dense_ops.h"
#include "qtqp_kkt.h"
#include "qtqp_sparse.h"

namespace {

constexpr int kMaxCols = 64;      // pivot columns per supernode
constexpr int kTile = 64;         // Schur-update tile edge
constexpr int kThreads = 256;     // workgroup size of the row kernels
constexpr int kPanelRows = 256;   // rows below the pivot block per panel job
constexpr int kSolveRows = 64;    // rows below the pivot block per solve job
constexpr int64_t kMaxOrderN = 40000;  // exact minimum degree up to here, AMD above
constexpr int64_t kMaxPanelEntries = int64_t(8) << 30;  // 64 GiB of FP64 panels

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
  std::vector<int64_t> dblk_off;          // S+1: pivot blocks, c*c each
  std::vector<int32_t> pjobs, sjobs;      // (supernode, a0) per panel / solve job
  std::vector<int32_t> contig;            // S: rows are consecutive indices
  std::vector<int64_t> pjob_off, sjob_off;  // L+1
  // Phases: the supernodes of a level in colour classes that share no row
  // below their pivot blocks, so that each class writes disjoint entries of
  // the ancestors; classes run in a fixed order.  tiles and sjobs are in
  // phase order (phases of a level contiguous, levels in order).
  std::vector<int64_t> phase_tile_off, phase_sjob_off;  // P+1
  std::vector<int64_t> level_phase_off;                  // L+1: phases of each level
  std::vector<int64_t> sn_sjob;                          // S+1 by supernode: its solve jobs
  std::vector<int32_t> sjob_sorted;                      // solve-job indices by supernode
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

// Approximate minimum degree on a quotient graph (the method of Amestoy,
// Davis and Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996), for systems too
// large for the explicit elimination graph below.  Eliminated vertices
// become elements: an element e keeps the list L_e of live variables it
// connects, and a variable keeps its remaining variable neighbours A_i and
// its adjacent elements E_i, so memory stays near the size of the pattern.
// Eliminating the pivot p forms L_p from A_p and the lists of p's elements,
// which are absorbed into p.  Each variable i in L_p then gets the degree
// bound
//     d_i = min(left - 1, d_i + |L_p| - 1,
//               |A_i| + |L_p| - 1 + sum over e in E_i, e != p, of |L_e \ L_p|),
// with |L_e \ L_p| counted by decrementing |L_e| once per variable of
// L_p in L_e; an element with |L_e \ L_p| = 0 is absorbed into p as well.
// Variables with more than max(16, 10 sqrt(n)) neighbours (dense rows of A)
// are ordered last.  Pivots are taken from degree buckets, lowest degree
// first, most recently inserted first, so the order is deterministic.
std::vector<int32_t> amd_order(int64_t n, const std::vector<std::vector<int32_t>>& adj) {
  std::vector<int32_t> order;
  order.reserve(n);
  const int64_t dense = std::max<int64_t>(16, (int64_t)(10.0 * std::sqrt((double)n)));
  std::vector<std::vector<int32_t>> A(n), E(n), L(n);
  std::vector<int64_t> deg(n, 0), w(n, 0);
  std::vector<int64_t> stamp(n, -1), wstamp(n, -1);
  // 0 live variable, 1 eliminated (live element), 2 absorbed element, 3 dense
  std::vector<char> state(n, 0);
  std::vector<int32_t> dense_list;
  for (int64_t i = 0; i < n; ++i) {
    if ((int64_t)adj[i].size() > dense) {
      state[i] = 3;
      dense_list.push_back((int32_t)i);
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    if (state[i] == 3) continue;
    for (int32_t j : adj[i])
      if (j != i && state[j] != 3) A[i].push_back(j);
    deg[i] = (int64_t)A[i].size();
  }
  // degree buckets: doubly linked lists
  std::vector<int32_t> head(n + 1, -1), nxt(n, -1), prv(n, -1);
  auto insert = [&](int32_t i) {
    const int64_t d = deg[i];
    nxt[i] = head[d];
    prv[i] = -1;
    if (head[d] != -1) prv[head[d]] = i;
    head[d] = i;
  };
  auto remove = [&](int32_t i) {
    if (prv[i] != -1) nxt[prv[i]] = nxt[i];
    else head[deg[i]] = nxt[i];
    if (nxt[i] != -1) prv[nxt[i]] = prv[i];
    nxt[i] = prv[i] = -1;
  };
  int64_t left = 0;
  for (int64_t i = n - 1; i >= 0; --i)
    if (state[i] == 0) insert((int32_t)i), ++left;
  int64_t mindeg = 0, tag = 0;
  std::vector<int32_t> lp;
  while (left > 0) {
    while (head[mindeg] == -1) ++mindeg;
    const int32_t p = head[mindeg];
    remove(p);
    state[p] = 1;
    order.push_back(p);
    --left;
    // L_p: live variables of A_p and of p's elements (which are absorbed)
    ++tag;
    stamp[p] = tag;
    lp.clear();
    for (int32_t j : A[p])
      if (state[j] == 0 && stamp[j] != tag) stamp[j] = tag, lp.push_back(j);
    for (int32_t e : E[p]) {
      if (state[e] != 1) continue;
      for (int32_t j : L[e])
        if (state[j] == 0 && stamp[j] != tag) stamp[j] = tag, lp.push_back(j);
      state[e] = 2;
      std::vector<int32_t>().swap(L[e]);
    }
    std::vector<int32_t>().swap(A[p]);
    std::vector<int32_t>().swap(E[p]);
    const int64_t np = (int64_t)lp.size();
    // |L_e \ L_p| for the elements next to L_p
    for (int32_t i : lp) {
      remove(i);
      for (int32_t e : E[i]) {
        if (state[e] != 1) continue;
        if (wstamp[e] != tag) wstamp[e] = tag, w[e] = (int64_t)L[e].size();
        --w[e];
      }
    }
    for (int32_t i : lp) {
      // elements: drop absorbed ones and those inside L_p, then add p
      int64_t ext = 0;
      size_t k = 0;
      for (int32_t e : E[i]) {
        if (state[e] != 1) continue;
        if (w[e] <= 0) {  // L_e within L_p: absorbed into p
          state[e] = 2;
          std::vector<int32_t>().swap(L[e]);
          continue;
        }
        ext += w[e];
        E[i][k++] = e;
      }
      E[i].resize(k);
      E[i].push_back(p);
      // variables: those in L_p are now reached through p
      k = 0;
      for (int32_t j : A[i])
        if (state[j] == 0 && stamp[j] != tag) A[i][k++] = j;
      A[i].resize(k);
      int64_t d = std::min<int64_t>(left - 1, deg[i] + np - 1);
      d = std::min<int64_t>(d, (int64_t)A[i].size() + np - 1 + ext);
      deg[i] = std::max<int64_t>(d, 0);
      insert(i);
      mindeg = std::min(mindeg, deg[i]);
    }
    L[p].swap(lp);
    lp.clear();
  }
  for (int32_t i : dense_list) order.push_back(i);
  return order;
}

// Minimum degree on the explicit elimination graph, held as one bitset row
// per vertex: eliminating v joins its neighbours into a clique.  Ties go to
// the lowest index, so the order is deterministic.  ordering: 0 this one up
// to kMaxOrderN and amd_order above, 1 this one, 2 amd_order.
std::vector<int32_t> min_degree(int64_t n, const std::vector<std::vector<int32_t>>& adj,
                                int ordering = 0) {
  if (ordering == 2 || (ordering == 0 && n > kMaxOrderN) || (ordering == 1 && n > kMaxOrderN))
    return amd_order(n, adj);
  std::vector<int32_t> order;
  order.reserve(n);
  const int64_t w = (n + 63) / 64;
  std::vector<uint64_t> bits((size_t)(n * w), 0);
  std::vector<int64_t> deg(n);
  std::vector<char> done(n, 0);
  std::vector<int64_t> clique;
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
    // v's neighbours now form a clique C.  A neighbour whose adjacency is
    // exactly C minus itself is simplicial: eliminating it adds no fill, so
    // it goes next (in index order), at the cost of clearing one bit in the
    // rest of C instead of a bitset merge.
    clique.clear();
    for (int64_t q = 0; q < w; ++q)
      for (uint64_t word = bv[q]; word; word &= word - 1) clique.push_back(q * 64 + __builtin_ctzll(word));
    int64_t csize = (int64_t)clique.size();
    for (int64_t u : clique) {
      if (deg[u] != csize - 1) continue;
      done[u] = 1;
      order.push_back((int32_t)u);
      --left;
      --csize;
      for (int64_t x : clique) {
        if (done[x]) continue;
        bits[x * w + u / 64] &= ~(1ull << (u % 64));
        --deg[x];
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
  }
  return padj;
}

int analyze(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind, Symbolic& S,
            int32_t ordering = 0) {
  if (n < 1) return fail(QTQP_ERR_ARG, "n must be positive");
  if (n >= INT32_MAX) return fail(QTQP_ERR_ARG, "n too large");
  if (!ptr || !ind) return fail(QTQP_ERR_ARG, "null pattern");
  std::vector<std::vector<int32_t>> adj;
  if (int rc = adjacency(n, format, ptr, ind, adj, S.nnz_kkt)) return rc;
  S.n = n;

  // 1-2. ordering, then a postorder of its elimination tree.  A postorder is
  // an equivalent elimination order: same fill, same tree, relabelled.
  if (ordering < 0 || ordering > 2) return fail(QTQP_ERR_ARG, "ordering must be 0, 1 or 2");
  std::vector<int32_t> perm = min_degree(n, adj, ordering), iperm(n);
  for (int64_t k = 0; k < n; ++k) iperm[perm[k]] = (int32_t)k;
  std::vector<std::vector<int32_t>> padj = permuted_adjacency(adj, perm, iperm);
  std::vector<int32_t> parent = etree(n, padj);
  {
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
    std::vector<int32_t> ipost(n), p2(n), parent2(n);
    for (int64_t k = 0; k < n; ++k) ipost[post[k]] = (int32_t)k;
    std::vector<std::vector<int32_t>> padj2(n);
    for (int64_t k = 0; k < n; ++k) {
      const int32_t old = post[k];
      p2[k] = perm[old];
      parent2[k] = parent[old] == -1 ? -1 : ipost[parent[old]];
      padj2[k].swap(padj[old]);
      for (int32_t& i : padj2[k]) i = ipost[i];
    }
    perm.swap(p2);
    parent.swap(parent2);
    padj.swap(padj2);
    for (int64_t k = 0; k < n; ++k) iperm[perm[k]] = (int32_t)k;
  }

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
    count[k] = (int64_t)s.size() + 1;
    const bool cont = k > 0 && parent[k - 1] == k && count[k - 1] == count[k] + 1 &&
                      k - start < kMaxCols;
    if (!cont) {  // only a supernode's first structure needs its rows in order
      std::sort(s.begin(), s.end());
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
  if (S.sn_lx[ns] > kMaxPanelEntries)
    return fail(QTQP_ERR_ARG, "the factor's panels would hold " + std::to_string(S.sn_lx[ns]) +
                                  " values (" + std::to_string(S.sn_lx[ns] * 8 >> 30) +
                                  " GiB), over the limit of " +
                                  std::to_string(kMaxPanelEntries * 8 >> 30) + " GiB");

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
  S.pjobs.clear();
  S.pjob_off.assign(1, 0);
  S.sjobs.clear();
  S.sjob_off.assign(1, 0);
  S.phase_tile_off.assign(1, 0);
  S.phase_sjob_off.assign(1, 0);
  S.level_phase_off.assign(1, 0);
  // colours of the current level's rows below pivot blocks (greedy colouring:
  // a supernode takes the lowest colour none of its rows carries yet)
  std::vector<std::vector<int32_t>> row_colours(n);
  std::vector<int32_t> touched;
  std::vector<char> forbid;
  for (int32_t l = 0; l < nlev; ++l) {
    std::vector<int32_t> colour(by_level[l].size());
    int32_t ncol = 0;
    for (size_t k = 0; k < by_level[l].size(); ++k) {
      const int32_t s = by_level[l][k];
      const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
      forbid.assign((size_t)ncol + 1, 0);
      for (int64_t q = S.sn_rows[s] + c; q < S.sn_rows[s + 1]; ++q)
        for (int32_t col : row_colours[S.rows[q]]) forbid[col] = 1;
      int32_t col = 0;
      while (forbid[col]) ++col;
      colour[k] = col;
      ncol = std::max(ncol, col + 1);
      for (int64_t q = S.sn_rows[s] + c; q < S.sn_rows[s + 1]; ++q) {
        if (row_colours[S.rows[q]].empty()) touched.push_back(S.rows[q]);
        row_colours[S.rows[q]].push_back(col);
      }
    }
    for (int32_t r : touched) row_colours[r].clear();
    touched.clear();
    for (int32_t s : by_level[l]) {
      S.level_sn.push_back(s);
      const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
      const int64_t below = S.sn_rows[s + 1] - S.sn_rows[s] - c;
      for (int64_t a0 = 0; a0 == 0 || a0 < below; a0 += kPanelRows)
        S.pjobs.insert(S.pjobs.end(), {s, (int32_t)a0});
    }
    for (int32_t col = 0; col < ncol; ++col) {
      for (size_t k = 0; k < by_level[l].size(); ++k) {
        if (colour[k] != col) continue;
        const int32_t s = by_level[l][k];
        const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
        const int64_t below = S.sn_rows[s + 1] - S.sn_rows[s] - c;
        const int64_t nb = (below + kTile - 1) / kTile;
        for (int64_t bj = 0; bj < nb; ++bj)
          for (int64_t bi = bj; bi < nb; ++bi)
            S.tiles.insert(S.tiles.end(), {s, (int32_t)(bi * kTile), (int32_t)(bj * kTile)});
        for (int64_t a0 = 0; a0 == 0 || a0 < below; a0 += kSolveRows)
          S.sjobs.insert(S.sjobs.end(), {s, (int32_t)a0});
      }
      S.phase_tile_off.push_back((int64_t)S.tiles.size() / 3);
      S.phase_sjob_off.push_back((int64_t)S.sjobs.size() / 2);
    }
    S.level_phase_off.push_back((int64_t)S.phase_tile_off.size() - 1);
    S.level_off.push_back((int64_t)S.level_sn.size());
    S.tile_off.push_back((int64_t)S.tiles.size() / 3);
    S.pjob_off.push_back((int64_t)S.pjobs.size() / 2);
    S.sjob_off.push_back((int64_t)S.sjobs.size() / 2);
  }
  // each supernode's solve jobs, in increasing a0 (they are contiguous)
  S.sn_sjob.assign(ns + 1, 0);
  {
    std::vector<int64_t> cnt(ns, 0), first(ns, -1);
    const int64_t nj = (int64_t)S.sjobs.size() / 2;
    for (int64_t j = 0; j < nj; ++j) {
      const int32_t s = S.sjobs[2 * j];
      if (first[s] < 0) first[s] = j;
      ++cnt[s];
    }
    S.sjob_sorted.assign(nj, 0);
    for (int64_t s = 0; s < ns; ++s) S.sn_sjob[s + 1] = S.sn_sjob[s] + cnt[s];
    for (int64_t s = 0; s < ns; ++s)
      for (int64_t k = 0; k < cnt[s]; ++k) S.sjob_sorted[S.sn_sjob[s] + k] = (int32_t)(first[s] + k);
  }
  S.contig.assign(ns, 0);
  for (int64_t s = 0; s < ns; ++s)
    S.contig[s] = S.rows[S.sn_rows[s + 1] - 1] - S.rows[S.sn_rows[s]] ==
                  S.sn_rows[s + 1] - S.sn_rows[s] - 1;
  S.dblk_off.assign(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s) {
    const int64_t c = S.sn_first[s + 1] - S.sn_first[s];
    S.dblk_off[s + 1] = S.dblk_off[s] + c * c;
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
  info->phases = (int64_t)S.phase_tile_off.size() - 1;
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// Records the launches `enqueue` makes on `stream` into a graph and
// instantiates it (no memset nodes: the captured kernels clear what they need).
template <class F>
int capture(hipStream_t stream, F enqueue, hipGraphExec_t* exec) {
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  enqueue();
  const hipError_t launch = hipGetLastError();
  const hipError_t end = hipStreamEndCapture(stream, &graph);
  if (launch != hipSuccess || end != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(QTQP_ERR_HIP, std::string("graph capture: ") +
                                  hipGetErrorString(launch != hipSuccess ? launch : end));
  }
  const hipError_t inst = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (inst != hipSuccess) {
    *exec = nullptr;
    return fail(QTQP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
  }
  return QTQP_OK;
}

// Supernode tables, device pointers.
struct Tables {
  const int32_t* first;    // S+1
  const int64_t* rows_off; // S+1
  const int64_t* lx_off;   // S+1
  const int64_t* dblk_off; // S+1: factored diagonal blocks (c x c, column-major)
  const int32_t* rows;
  const int32_t* col_sn;
  const int32_t* level_sn;
  const int32_t* tiles;    // (supernode, a0, b0) per Schur-update tile
  const int32_t* pjobs;    // (supernode, a0) per panel job: 256 rows below the block
  const int32_t* sjobs;    // (supernode, a0) per solve job: 64 rows below the block
  const int32_t* contig;   // S: 1 when a supernode's rows are consecutive indices
  const int64_t* sn_sjob;  // S+1: each supernode's solve jobs in sjob_sorted
  const int32_t* sjob_sorted;
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

__device__ __forceinline__ double bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Right-looking LDL' of a pivot block by one wavefront, lane i holding row
// i in registers, padded to CM x CM with identity rows (pivot 1, zero
// coupling), so every step runs without a branch: step k reads the pivot
// and column k of the other rows with readlane, and no step touches LDS or
// waits at a barrier.  On return row[j] = l_ij for j < i and row[i] = d_i.
// Returns false on a zero or non-finite pivot (the same in every lane).
template <int CM>
__device__ __forceinline__ bool wave_ldl(double (&row)[kMaxCols], int lane, int& bad_k) {
#pragma unroll
  for (int k = 0; k < CM; ++k) {
    const double d = bcast(row[k], k);
    if (!(d != 0.0) || !isfinite(d)) {  // uniform
      bad_k = k;
      return false;
    }
    const double vik = row[k] * (1.0 / d);
#pragma unroll
    for (int j = k + 1; j < CM; ++j) {
      const double ajk = bcast(row[k], j);
      if (lane >= j) row[j] -= vik * ajk;
    }
    if (lane > k) row[k] = vik;
  }
  return true;
}

// One row of L below a supernode's pivot block, left-looking along the row:
// w_k = a_ik - sum_{j<k} w_j l_kj, l_ik = w_k / d_k (w_j = l_ij d_j stays in
// registers).  The row is loaded 16 columns at a time with unconditional,
// clamped loads, so one memory round trip serves 16 steps.  Dl holds the
// factored pivot block (l_kj below its diagonal) and rd the reciprocal pivots.
template <int CM>
__device__ __forceinline__ void panel_row(const double (*Dl)[kMaxCols + 1], const double* rd,
                                          double* __restrict__ L, int r, int c, int i) {
  constexpr int B = CM < 16 ? CM : 16;
  double w[CM];
  double* p = L + i;
#pragma unroll
  for (int kb = 0; kb < CM; kb += B) {
    double a[B];
#pragma unroll
    for (int q = 0; q < B; ++q) a[q] = p[(int64_t)min(kb + q, c - 1) * r];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int k = kb + q;
      if (k < c) {  // uniform across the workgroup
        double v = a[q], v2 = 0.0;  // two partial sums halve the FMA chain
#pragma unroll
        for (int j = 0; j + 1 < k; j += 2) v -= w[j] * Dl[k][j], v2 += w[j + 1] * Dl[k][j + 1];
        if (k % 2) v -= w[k - 1] * Dl[k][k - 1];
        v -= v2;
        w[k] = v;
        p[(int64_t)k * r] = v * rd[k];
      }
    }
  }
}

// One panel job: factor the supernode's c x c pivot block in LDS (every job
// of the supernode does, from the same unmodified input), then the job's
// rows below it.  The job with a0 == 0 stores the factored block (d_k on the
// diagonal, l_jk below) in dblk; the panel's own copy of the block is left
// as it was, so no job reads what another writes.  A zero or non-finite
// pivot records (original column + 1) in *bad.
__global__ void __launch_bounds__(kThreads) k_panel(Tables T, double* __restrict__ lx,
                                                    double* __restrict__ dblk, int64_t job_begin,
                                                    int* __restrict__ bad,
                                                    const int32_t* __restrict__ perm) {
  __shared__ double Dl[kMaxCols][kMaxCols + 1];
  __shared__ double rd[kMaxCols];
  const int32_t* job = T.pjobs + 2 * (job_begin + blockIdx.x);
  const int32_t s = job[0];
  const int a0 = job[1];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int r = (int)(T.rows_off[s + 1] - T.rows_off[s]);
  double* L = lx + T.lx_off[s];
  const int tid = threadIdx.x;
  __shared__ int failed;
  // The block's lower triangle, staged by the whole workgroup (coalesced).
  for (int e = tid; e < kMaxCols * kMaxCols; e += kThreads) {
    const int j = e / kMaxCols, i = e % kMaxCols;
    Dl[i][j] = (i < c && j < c && i >= j) ? L[(int64_t)j * r + i] : 0.0;
  }
  __syncthreads();
  if (tid < kMaxCols) {  // wave 0 factors the pivot block in registers
    double row[kMaxCols];
#pragma unroll
    for (int j = 0; j < kMaxCols; ++j) row[j] = tid < c ? Dl[tid][j] : (j == tid ? 1.0 : 0.0);
    int bad_k = -1;
    const bool ok = c <= 8    ? wave_ldl<8>(row, tid, bad_k)
                    : c <= 16 ? wave_ldl<16>(row, tid, bad_k)
                    : c <= 32 ? wave_ldl<32>(row, tid, bad_k)
                              : wave_ldl<64>(row, tid, bad_k);
    if (!ok) {
      if (tid == 0) {
        failed = 1;
        atomicCAS(bad, 0, perm[f + bad_k] + 1);
      }
    } else {
      if (tid == 0) failed = 0;
#pragma unroll
      for (int j = 0; j < kMaxCols; ++j) Dl[tid][j] = (j <= tid && tid < c) ? row[j] : 0.0;
      // (row[] is indexed by constants only, so it stays in registers; the
      // pivot comes back from LDS)
      rd[tid] = tid < c ? 1.0 / Dl[tid][tid] : 0.0;
    }
  }
  __syncthreads();
  if (failed) return;  // uniform: one value in LDS
  if (a0 == 0) {
    double* D = dblk + T.dblk_off[s];
    for (int e = tid; e < c * c; e += kThreads) {
      const int j = e / c, i = e % c;
      D[e] = i >= j ? Dl[i][j] : 0.0;  // l_ij below the diagonal, d_i on it
    }
  }
  const int a = a0 + tid;
  if (a < r - c) {
    if (c <= 8) panel_row<8>(Dl, rd, L, r, c, c + a);
    else if (c <= 16) panel_row<16>(Dl, rd, L, r, c, c + a);
    else if (c <= 32) panel_row<32>(Dl, rd, L, r, c, c + a);
    else panel_row<64>(Dl, rd, L, r, c, c + a);
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
// entries (rows[c+a], rows[c+b]).  The panel read as a row-major matrix with
// ld = r has the pivot columns as its rows, so U is the weighted rank-c
// product of the batched Gram engine: qk::RankKTile on MFMA 16x16x4 f64 with
// the pivots as weights.
__global__ void __launch_bounds__(kThreads) k_update(Tables T, double* __restrict__ lx,
                                                     const double* __restrict__ dblk,
                                                     int64_t tile_begin, int* __restrict__ bad) {
  __shared__ double si[qk::KC][qk::NB + 1];
  __shared__ double sj[qk::KC][qk::NB + 1];
  const int32_t* tl = T.tiles + 3 * (tile_begin + blockIdx.x);
  const int32_t s = tl[0];
  const int a0 = tl[1], b0 = tl[2];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const double* L = lx + T.lx_off[s];
  const double* D = dblk + T.dblk_off[s];
  qk::RankKTile<double> tile;
  tile.zero();
  tile.run(L, r, 0, c, c + a0, c + b0, r, r, [=](int k) { return D[(int64_t)k * (c + 1)]; }, si,
           sj);
  const int32_t* R = T.rows + roff;
  tile.for_each(c + a0, c + b0, [&](int i, int j, double v) {
    if (i >= r || j >= r || i < j) return;  // outside the panel or above the diagonal
    const int32_t row = R[i], col = R[j];
    const int32_t t = T.col_sn[col];
    const int64_t troff = T.rows_off[t];
    const int tr = (int)(T.rows_off[t + 1] - troff);
    const int lc = col - T.first[t];
    const int32_t* tR = T.rows + troff;
    const int pos = T.contig[t] ? row - tR[0] : find_row(tR, lc, tr, row);
    if (pos >= lc && pos < tr && tR[pos] == row) {
      // The supernodes of one phase share no row below their pivot blocks,
      // so each target entry has one writer in the launch.
      lx[T.lx_off[t] + (int64_t)lc * tr + pos] -= v;
    } else {
      atomicCAS(bad, 0, -1);  // the symbolic structure misses a fill entry
    }
  });
}

// L11 y = b by one wavefront: lane i holds b_i and row i of the unit lower
// pivot block (D column-major, c x c); each solved y_k is broadcast with
// readlane.  Returns y_i (0 for lanes >= c).
__device__ __forceinline__ double wave_lower_solve(const double* __restrict__ D, int c, int lane,
                                                   double b) {
  double row[kMaxCols];
#pragma unroll
  for (int k = 0; k < kMaxCols; ++k)
    row[k] = (k < c && lane < c) ? D[k * c + lane] : 0.0;
  double y = lane < c ? b : 0.0;
#pragma unroll
  for (int k = 0; k < kMaxCols - 1; ++k) {
    if (k + 1 < c) {  // uniform
      const double yk = bcast(y, k);
      if (lane > k) y -= row[k] * yk;
    }
  }
  return y;
}

// L11' x = v by one wavefront: lane j holds v_j and column j of the pivot
// block; x_i is broadcast from the last row up.  Returns x_j.
__device__ __forceinline__ double wave_upper_solve(const double* __restrict__ D, int c, int lane,
                                                   double v) {
  double col[kMaxCols];
#pragma unroll
  for (int k = 0; k < kMaxCols; ++k)
    col[k] = (k < c && lane < c) ? D[lane * c + k] : 0.0;
  double x = lane < c ? v : 0.0;
#pragma unroll
  for (int i = kMaxCols - 1; i > 0; --i) {
    if (i < c) {  // uniform
      const double xi = bcast(x, i);
      if (lane < i) x -= col[i] * xi;
    }
  }
  return x;
}

// Forward substitution for one phase of a level, one job per 64 rows below
// a pivot block (a single job when there are none): every job solves
// L11 y = b for its supernode in LDS (b unchanged in x during the launch),
// the job with a0 == 0 stores z = D^{-1} y in xo, and the job's rows update
// b[rows] -= L21 y (the supernodes of a phase share no ancestor row, and the
// jobs of one supernode take disjoint rows, so each row has one writer).
__global__ void __launch_bounds__(kThreads) k_fwd(Tables T, const double* __restrict__ lx,
                                                  const double* __restrict__ dblk,
                                                  int64_t job_begin, double* __restrict__ x,
                                                  double* __restrict__ xo) {
  __shared__ double y[kMaxCols];
  __shared__ double part[4][kSolveRows];
  const int32_t* job = T.sjobs + 2 * (job_begin + blockIdx.x);
  const int32_t s = job[0];
  const int a0 = job[1];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const double* L = lx + T.lx_off[s];
  const double* D = dblk + T.dblk_off[s];
  const int tid = threadIdx.x, lane = tid % kSolveRows, g = tid / kSolveRows;
  if (g == 0) {
    const double yi = wave_lower_solve(D, c, lane, lane < c ? x[f + lane] : 0.0);
    if (lane < c) {
      y[lane] = yi;
      if (a0 == 0) xo[f + lane] = yi / D[lane * (c + 1)];
    }
  }
  __syncthreads();
  const int a = a0 + lane;
  double t = 0.0;
  if (a < r - c)
    for (int k = g; k < c; k += 4) t += L[(int64_t)k * r + c + a] * y[k];
  part[g][lane] = t;
  __syncthreads();
  if (g == 0 && a < r - c)
    x[T.rows[roff + c + a]] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
}

// Backward substitution, rows below the pivot blocks: each job's partial sums
// sum_a L21[a][k] x[rows[a]] over its 64 rows into its own slot of tb
// (kMaxCols per job), which k_bwd_diag adds up in job order; staged through
// LDS so the loads stay coalesced.
__global__ void __launch_bounds__(kThreads) k_bwd_off(Tables T, const double* __restrict__ lx,
                                                      int64_t job_begin,
                                                      const double* __restrict__ xo,
                                                      double* __restrict__ tb) {
  __shared__ double Ls[kMaxCols][kSolveRows + 1];
  __shared__ double xa[kSolveRows];
  __shared__ double part[4][kMaxCols];
  const int32_t* job = T.sjobs + 2 * (job_begin + blockIdx.x);
  const int32_t s = job[0];
  const int a0 = job[1];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int64_t roff = T.rows_off[s];
  const int r = (int)(T.rows_off[s + 1] - roff);
  const int nb = r - c;
  const double* L = lx + T.lx_off[s];
  const int tid = threadIdx.x;
  for (int e = tid; e < c * kSolveRows; e += kThreads) {
    const int k = e / kSolveRows, t = e % kSolveRows;
    Ls[k][t] = a0 + t < nb ? L[(int64_t)k * r + c + a0 + t] : 0.0;
  }
  if (tid < kSolveRows) xa[tid] = a0 + tid < nb ? xo[T.rows[roff + c + a0 + tid]] : 0.0;
  __syncthreads();
  const int k = tid % kMaxCols, g = tid / kMaxCols;
  double t = 0.0;
  if (k < c)
    for (int q = 16 * g; q < 16 * g + 16; ++q) t += Ls[k][q] * xa[q];
  part[g][k] = t;
  __syncthreads();
  if (tid < c)
    tb[(job_begin + blockIdx.x) * kMaxCols + tid] = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
}

// Backward substitution, pivot blocks of one level: x = L11^{-T} (z - t),
// one wavefront per supernode.
__global__ void __launch_bounds__(64) k_bwd_diag(Tables T, const double* __restrict__ dblk,
                                                 int64_t lvl_begin, const double* __restrict__ tb,
                                                 double* __restrict__ xo) {
  const int32_t s = T.level_sn[lvl_begin + blockIdx.x];
  const int32_t f = T.first[s];
  const int c = T.first[s + 1] - f;
  const int lane = threadIdx.x;
  // the jobs' partial sums, four running sums over the jobs in increasing row
  // order, combined in a fixed order
  double t4[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t qb = T.sn_sjob[s], qe = T.sn_sjob[s + 1];
  if (lane < c) {
    int64_t q = qb;
    for (; q + 4 <= qe; q += 4)
#pragma unroll
      for (int u = 0; u < 4; ++u) t4[u] += tb[(int64_t)T.sjob_sorted[q + u] * kMaxCols + lane];
    for (; q < qe; ++q) t4[0] += tb[(int64_t)T.sjob_sorted[q] * kMaxCols + lane];
  }
  const double t = (t4[0] + t4[1]) + (t4[2] + t4[3]);
  const double v = lane < c ? xo[f + lane] - t : 0.0;
  const double xj = wave_upper_solve(dblk + T.dblk_off[s], c, lane, v);
  if (lane < c) xo[f + lane] = xj;
}

__global__ void k_zero_int(int* __restrict__ x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *x = 0;
}

__global__ void k_zero(double* __restrict__ x, int64_t n) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) x[k] = 0.0;
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
  double* dblk = nullptr;      // factored pivot blocks
  double* tb = nullptr;        // N-vector of backward partial sums
  int64_t* dpos = nullptr;     // panel position of each (permuted) diagonal
  int32_t* perm = nullptr;
  double* diag = nullptr;      // current diagonal of K, original order
  int64_t* sym_ptr = nullptr;  // full symmetric off-diagonal pattern, CSR
  int32_t* sym_col = nullptr;
  double* sym_val = nullptr;
  double* xin = nullptr;       // N-vector staging (original order)
  double* xp = nullptr;        // N-vector work (permuted order): b, then y
  double* xo = nullptr;        // N-vector work (permuted order): z, then x
  int* bad = nullptr;
  // the factorisation's and the solve's launches, captured once per handle
  // (the pattern and the buffers are fixed) and replayed
  hipGraphExec_t gfac = nullptr, gsol = nullptr;
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
    if (gfac) (void)hipGraphExecDestroy(gfac);
    if (gsol) (void)hipGraphExecDestroy(gsol);
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
  int32_t *first, *rows, *col_sn, *level_sn, *tiles, *pjobs, *sjobs, *contig, *sjob_sorted;
  int64_t *rows_off, *lx_off, *dblk_off, *sn_sjob;
  if (int rc = h->upload(&first, S.sn_first)) return rc;
  if (int rc = h->upload(&rows_off, S.sn_rows)) return rc;
  if (int rc = h->upload(&lx_off, S.sn_lx)) return rc;
  if (int rc = h->upload(&rows, S.rows)) return rc;
  if (int rc = h->upload(&col_sn, S.col_sn)) return rc;
  if (int rc = h->upload(&level_sn, S.level_sn)) return rc;
  if (int rc = h->upload(&tiles, S.tiles)) return rc;
  if (int rc = h->upload(&dblk_off, S.dblk_off)) return rc;
  if (int rc = h->upload(&pjobs, S.pjobs)) return rc;
  if (int rc = h->upload(&sjobs, S.sjobs)) return rc;
  if (int rc = h->upload(&contig, S.contig)) return rc;
  if (int rc = h->upload(&sn_sjob, S.sn_sjob)) return rc;
  if (int rc = h->upload(&sjob_sorted, S.sjob_sorted)) return rc;
  h->T = Tables{first,    rows_off, lx_off, dblk_off, rows,   col_sn,  level_sn,
                tiles,    pjobs,    sjobs,  contig,   sn_sjob, sjob_sorted};
  if (int rc = h->alloc(&h->dblk, S.dblk_off.back())) return rc;
  if (int rc = h->alloc(&h->tb, std::max<int64_t>(1, (int64_t)S.sjobs.size() / 2 * kMaxCols)))
    return rc;
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
  if (int rc = h->alloc(&h->xo, n)) return rc;
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
  return qtqp_sparse_analyze_ex(n, format, ptr, ind, 0, perm_out, info);
}

int qtqp_sparse_analyze_ex(int64_t n, int32_t format, const int64_t* ptr, const int64_t* ind,
                           int32_t ordering, int64_t* perm_out, qtqp_sparse_info* info) {
  Symbolic S;
  const auto t0 = std::chrono::steady_clock::now();
  if (int rc = analyze(n, format, ptr, ind, S, ordering)) return rc;
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
  if (!h->gfac) {
    auto enqueue = [&]() {
      k_zero_int<<<1, 64, 0, h->stream>>>(h->bad);
      const int64_t nlx = S.sn_lx.back();
      k_copy<<<(unsigned)std::min<int64_t>(4096, blocks_for(nlx)), kThreads, 0, h->stream>>>(
          h->lx, h->lx0, nlx);
      k_load_diag<<<blocks_for(S.n), kThreads, 0, h->stream>>>(h->lx, h->dpos, h->diag, h->perm,
                                                                S.n);
      const int64_t nlev = (int64_t)S.level_off.size() - 1;
      for (int64_t l = 0; l < nlev; ++l) {
        const int64_t np = S.pjob_off[l + 1] - S.pjob_off[l];
        k_panel<<<(unsigned)np, kThreads, 0, h->stream>>>(h->T, h->lx, h->dblk, S.pjob_off[l],
                                                          h->bad, h->perm);
        for (int64_t ph = S.level_phase_off[l]; ph < S.level_phase_off[l + 1]; ++ph) {
          const int64_t nt = S.phase_tile_off[ph + 1] - S.phase_tile_off[ph];
          if (nt > 0)
            k_update<<<(unsigned)nt, kThreads, 0, h->stream>>>(h->T, h->lx, h->dblk,
                                                               S.phase_tile_off[ph], h->bad);
        }
      }
    };
    if (int rc = capture(h->stream, enqueue, &h->gfac)) return rc;
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  HIP_TRY(hipGraphLaunch(h->gfac, h->stream));
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
  if (!h->gsol) {
    auto enqueue = [&]() {
      k_gather<<<blocks_for(n), kThreads, 0, h->stream>>>(h->xp, h->xin, h->perm, n);
      const int64_t nph = (int64_t)S.phase_sjob_off.size() - 1;
      for (int64_t ph = 0; ph < nph; ++ph)
        k_fwd<<<(unsigned)(S.phase_sjob_off[ph + 1] - S.phase_sjob_off[ph]), kThreads, 0,
                h->stream>>>(h->T, h->lx, h->dblk, S.phase_sjob_off[ph], h->xp, h->xo);
      const int64_t nlev = (int64_t)S.level_off.size() - 1;
      for (int64_t l = nlev - 1; l >= 0; --l) {
        k_bwd_off<<<(unsigned)(S.sjob_off[l + 1] - S.sjob_off[l]), kThreads, 0, h->stream>>>(
            h->T, h->lx, S.sjob_off[l], h->xo, h->tb);
        k_bwd_diag<<<(unsigned)(S.level_off[l + 1] - S.level_off[l]), 64, 0, h->stream>>>(
            h->T, h->dblk, S.level_off[l], h->tb, h->xo);
      }
      k_scatter<<<blocks_for(n), kThreads, 0, h->stream>>>(h->xin, h->xo, h->perm, n);
    };
    if (int rc = capture(h->stream, enqueue, &h->gsol)) return rc;
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  HIP_TRY(hipGraphLaunch(h->gsol, h->stream));
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
