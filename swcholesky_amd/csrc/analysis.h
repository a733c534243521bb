// Host symbolic analysis (restatement of analyze_p2 for a given ordering).
#pragma once
#include <cstdint>
#include <vector>

namespace swc {

struct Symbolic {
  int n = 0;
  int nsuper = 0;
  std::vector<int> perm;       // perm[new] = old, after the weighted postorder
  std::vector<int> colcount;   // nnz per column of L, diagonal included
  std::vector<int> parent;     // column elimination tree
  std::vector<int> super;      // nsuper+1, first column of each supernode
  std::vector<int> sparent;    // supernodal elimination tree
  std::vector<int> col2sup;    // n
  std::vector<int64_t> pi;     // nsuper+1, row-list pointer per supernode
  std::vector<int> s;          // ssize row indices, sorted within each supernode
  std::vector<int64_t> i_ptr;  // n+1, row-list pointer per column (Li_ptr)
  std::vector<int64_t> p;      // n+1, value pointer per column (lC)
  std::vector<int> pruneptr;   // nsuper+1
  std::vector<int> pruneset;   // descendants whose rows meet each supernode
  std::vector<int> a2p, a2i;   // tril(P A P^T), rows sorted
  std::vector<int64_t> a2src;  // a2 entry -> index of the input entry of A
  int64_t ssize = 0, xsize = 0;
  int max_sup_wid = 0;
  int64_t max_col = 0;
  double flops = 0;            // sum of colcount^2
};

// A: lower triangle, CSC, n columns; perm[new] = old; nrelax/zrelax as CHOLMOD.
void analyze(int n, const int* ap, const int* ai, const int64_t* perm, const int nrelax[3],
             const double zrelax[3], Symbolic& S);

}  // namespace swc
