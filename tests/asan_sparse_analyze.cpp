// Host-only AddressSanitizer check of the sparse engine's symbolic analysis:
// compiled with the sanitizer on the host side only, it runs
// qtqp_sparse_analyze_ex (which touches no device) on random KKT patterns
// [P + I, A'; 0, I] of n variables and 2n constraints with three entries each,
// with each ordering, and checks that every order is a permutation.  Built and
// run by tests/test_sparse_asan.py; exits with the number of failed cases.
#include "sparse.hip"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 g(3);
  int fails = 0;
  for (int trial = 0; trial < 6; ++trial) {
    const int64_t n = 40 + 400 * trial, m = 2 * n, N = n + m;
    std::vector<std::vector<int64_t>> cols(N);
    for (int64_t j = 0; j < N; ++j) cols[j].push_back(j);
    for (int64_t i = 0; i < m; ++i)
      for (int k = 0; k < 3; ++k) cols[n + i].push_back((int64_t)(g() % n));
    std::vector<int64_t> ptr(1, 0), ind;
    for (auto& c : cols) { std::sort(c.begin(), c.end()); c.erase(std::unique(c.begin(), c.end()), c.end());
      ind.insert(ind.end(), c.begin(), c.end()); ptr.push_back((int64_t)ind.size()); }
    for (int ord = 0; ord <= 2; ++ord) {
      std::vector<int64_t> perm(N);
      qtqp_sparse_info info;
      const int rc = qtqp_sparse_analyze_ex(N, 0, ptr.data(), ind.data(), ord, perm.data(), &info);
      std::vector<char> seen(N, 0);
      bool ok = rc == 0;
      for (int64_t k = 0; k < N && ok; ++k) ok = perm[k] >= 0 && perm[k] < N && !seen[perm[k]]++;
      printf("N=%lld ordering=%d rc=%d nnz_l=%lld ok=%d\n", (long long)N, ord, rc, (long long)info.nnz_l, ok);
      fails += !ok;
    }
  }
  return fails;
}
