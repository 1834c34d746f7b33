// Bandwidth-reducing row order for the batched graphs (host code).
//
// Reverse Cuthill-McKee: a breadth-first numbering that visits each level's
// neighbours in increasing degree, started from a pseudo-peripheral node
// (George & Liu, "An implementation of a pseudoperipheral node finder", ACM
// TOMS 5 (1979); Cuthill & McKee, Proc. 24th ACM National Conference, 1969),
// then reversed.  Disconnected parts are numbered one after another, each
// from its own pseudo-peripheral node.  The pattern is taken as given (rows'
// column lists); for a symmetric pattern this is the textbook algorithm.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/psfun_hip.h"

namespace {

struct Graph {
  int n;
  const int32_t *rp;
  const int32_t *ci;
  int base;  // column index of local vertex 0
  int deg(int v) const { return rp[v + 1] - rp[v]; }
};

// Breadth-first levels from `root` over the vertices not yet numbered
// (done == 0); returns the number of levels, fills `order` with the visit
// order and `lastlevel_start` with the index in `order` where the last level
// begins.
int bfs_levels(const Graph &g, int root, const std::vector<int> &done, std::vector<int> &mark, int stamp,
               std::vector<int> &order, size_t &lastlevel_start) {
  order.clear();
  order.push_back(root);
  mark[root] = stamp;
  size_t lo = 0;
  int levels = 0;
  lastlevel_start = 0;
  while (lo < order.size()) {
    const size_t hi = order.size();
    lastlevel_start = lo;
    ++levels;
    for (size_t i = lo; i < hi; ++i) {
      const int v = order[i];
      for (int q = g.rp[v]; q < g.rp[v + 1]; ++q) {
        const int w = g.ci[q] - g.base;
        if (w < 0 || w >= g.n || done[w] || mark[w] == stamp) continue;
        mark[w] = stamp;
        order.push_back(w);
      }
    }
    lo = hi;
  }
  return levels;
}

}  // namespace

// perm[p] = old local index of new local index p, for one graph with n rows
// whose column indices are base + local index.  Returns PSFUN_SUCCESS.
extern "C" int psfun_rcm_order(int n, const int32_t *rowptr, const int32_t *colind, int base, int32_t *perm) {
  if (n < 0 || (n && (!rowptr || !colind || !perm))) return PSFUN_ERR_INVALID_ARG;
  if (n == 0) return PSFUN_SUCCESS;
  Graph g{n, rowptr, colind, base};
  std::vector<int> done(n, 0), mark(n, -1), level, next_order;
  std::vector<int> out;
  out.reserve(n);
  int stamp = 0;
  // vertices by increasing degree: start candidates for each part
  std::vector<int> by_deg(n);
  for (int v = 0; v < n; ++v) by_deg[v] = v;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return g.deg(a) < g.deg(b); });
  std::vector<int> nb;
  for (int cand : by_deg) {
    if (done[cand]) continue;
    // pseudo-peripheral node: repeat BFS from a minimum-degree vertex of the
    // last level while the number of levels grows
    int root = cand;
    size_t last = 0;
    int nlev = bfs_levels(g, root, done, mark, stamp++, level, last);
    for (int it = 0; it < 8; ++it) {
      int best = level[last];
      for (size_t i = last; i < level.size(); ++i)
        if (g.deg(level[i]) < g.deg(best)) best = level[i];
      size_t last2 = 0;
      const int nlev2 = bfs_levels(g, best, done, mark, stamp++, next_order, last2);
      if (nlev2 <= nlev) break;
      root = best;
      nlev = nlev2;
      level.swap(next_order);
      last = last2;
    }
    // Cuthill-McKee from root: neighbours in increasing degree
    const size_t start = out.size();
    out.push_back(root);
    done[root] = 1;
    for (size_t i = start; i < out.size(); ++i) {
      const int v = out[i];
      nb.clear();
      for (int q = g.rp[v]; q < g.rp[v + 1]; ++q) {
        const int w = g.ci[q] - g.base;
        if (w < 0 || w >= n || done[w]) continue;
        done[w] = 1;
        nb.push_back(w);
      }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return g.deg(a) < g.deg(b); });
      out.insert(out.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(out.begin(), out.end());
  std::copy(out.begin(), out.end(), perm);
  return PSFUN_SUCCESS;
}
