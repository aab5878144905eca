// Host compile step for stdBrick batches: the data the device kernels read.
//
// Produces what the reference's compile::compile_model produces for one stdBrick batch, with
// the same bits (connectivity, colouring and DOF numbering exact):
//   node order       reverse Cuthill-McKee from a George-Liu pseudo-peripheral root
//                    (src/compile/renumber.cpp:30-143), components in file order
//   element order    by smallest internal node index, ties by file position (:164-181)
//   DOF numbering    union-find over DOF ids, smallest id as root; groups holding a fixed
//                    DOF fixed; equations in root-id order (dof_numbering.cpp:21-40, 452-463)
//   batch maps       value[k * padded + e], padding -1 (batches.cpp:153-203)
//   param sets       per (group, region) in first-appearance order (batches.cpp:161-190)
//   colouring        greedy in batch order over shared equations (batches.cpp:23-50)
// Linear constraints and the OpenSees handler cross-check are outside this path.
#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <utility>
#include <vector>

#include "fx_internal.hpp"

namespace fx {

namespace {

struct Graph {
  std::vector<int64_t> start;
  std::vector<int32_t> adj;
  int32_t degree(int32_t v) const { return static_cast<int32_t>(start[v + 1] - start[v]); }
};

Graph build_node_graph(const fx_model_desc& m) {
  const int32_t N = m.nnode;
  std::vector<std::vector<int32_t>> nbr(N);
  for (int32_t e = 0; e < m.nelem; ++e) {
    const int32_t* c = m.conn + 8 * static_cast<int64_t>(e);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        if (a != b) nbr[c[a]].push_back(c[b]);
  }
  for (int32_t r = 0; r < m.ntie; ++r) {
    nbr[m.tie_retained[r]].push_back(m.tie_constrained[r]);
    nbr[m.tie_constrained[r]].push_back(m.tie_retained[r]);
  }
  Graph g;
  g.start.assign(static_cast<size_t>(N) + 1, 0);
  size_t total = 0;
  for (int32_t v = 0; v < N; ++v) {
    auto& l = nbr[v];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    // a node listed at two corners of one element keeps itself as a neighbour (the pairs are
    // distinct corners, not distinct nodes: renumber.cpp:38-41), which counts in its degree
    total += l.size();
    g.start[v + 1] = static_cast<int64_t>(total);
  }
  g.adj.reserve(total);
  for (auto& l : nbr) {
    g.adj.insert(g.adj.end(), l.begin(), l.end());
    std::vector<int32_t>().swap(l);
  }
  return g;
}

// Reverse Cuthill-McKee over all components.
std::vector<int32_t> reverse_cuthill_mckee(const Graph& g, int32_t N) {
  auto before = [&](int32_t a, int32_t b) {
    const int32_t da = g.degree(a), db = g.degree(b);
    return da != db ? da < db : a < b;
  };
  std::vector<char> ordered(N, 0), mark(N, 0);
  std::vector<int32_t> order, queue;
  order.reserve(N);
  queue.reserve(N);
  // BFS from root over unordered nodes; returns depth and leaves the visit in `queue`,
  // with [last_begin, queue.size()) the deepest level.
  auto levels = [&](int32_t root, size_t& last_begin) {
    queue.clear();
    queue.push_back(root);
    mark[root] = 1;
    size_t lb = 0, le = 1;
    int depth = 0;
    last_begin = 0;
    while (lb < le) {
      last_begin = lb;
      for (size_t h = lb; h < le; ++h) {
        const int32_t v = queue[h];
        for (int64_t k = g.start[v]; k < g.start[v + 1]; ++k) {
          const int32_t w = g.adj[k];
          if (!mark[w]) {
            mark[w] = 1;
            queue.push_back(w);
          }
        }
      }
      if (queue.size() > le) ++depth;
      lb = le;
      le = queue.size();
    }
    for (int32_t v : queue) mark[v] = ordered[v];
    return depth;
  };
  std::vector<int32_t> next;
  for (int32_t s = 0; s < N; ++s) {
    if (ordered[s]) continue;
    size_t lb;
    levels(s, lb);
    int32_t start = s;
    for (int32_t v : queue)
      if (before(v, start)) start = v;
    int32_t root = start;
    int depth = -1;
    for (int sweep = 0; sweep < 32; ++sweep) {
      const int d = levels(root, lb);
      if (d <= depth) break;
      depth = d;
      root = *std::min_element(queue.begin() + static_cast<std::ptrdiff_t>(lb), queue.end(), before);
    }
    const size_t first = order.size();
    order.push_back(root);
    ordered[root] = 1;
    for (size_t h = first; h < order.size(); ++h) {
      const int32_t v = order[h];
      next.clear();
      for (int64_t k = g.start[v]; k < g.start[v + 1]; ++k) {
        const int32_t w = g.adj[k];
        if (!ordered[w]) {
          ordered[w] = 1;
          next.push_back(w);
        }
      }
      std::sort(next.begin(), next.end(), before);
      order.insert(order.end(), next.begin(), next.end());
    }
    std::reverse(order.begin() + static_cast<std::ptrdiff_t>(first), order.end());
    for (size_t i = first; i < order.size(); ++i) mark[order[i]] = 1;
  }
  return order;
}

struct DisjointSets {
  std::vector<int32_t> up;
  explicit DisjointSets(int32_t n) : up(n) { std::iota(up.begin(), up.end(), 0); }
  int32_t root(int32_t a) {
    while (up[a] != a) {
      up[a] = up[up[a]];
      a = up[a];
    }
    return a;
  }
  void unite(int32_t a, int32_t b) {
    a = root(a);
    b = root(b);
    if (a == b) return;
    if (b < a) std::swap(a, b);
    up[b] = a;  // the smaller id stays the root
  }
};

}  // namespace

int compile_model(const fx_model_desc& m, int renumber_rcm, int simd_width, int build_coloring, Mesh& out,
                  std::string& err) {
  const int32_t N = m.nnode, E = m.nelem;
  for (int32_t e = 0; e < E; ++e)
    for (int a = 0; a < 8; ++a) {
      const int32_t n = m.conn[8 * static_cast<int64_t>(e) + a];
      if (n < 0 || n >= N) {
        err = "fx_compile_model: element " + std::to_string(e) + " names node " + std::to_string(n) +
              " outside [0, " + std::to_string(N) + ")";
        return FX_EINVAL;
      }
    }
  for (int32_t r = 0; r < m.nfix; ++r)
    if (m.fix_node[r] < 0 || m.fix_node[r] >= N || m.fix_dof[r] < 0 || m.fix_dof[r] > 2) {
      err = "fx_compile_model: fix row " + std::to_string(r) + " is out of range";
      return FX_EINVAL;
    }
  for (int32_t r = 0; r < m.ntie; ++r)
    if (m.tie_retained[r] < 0 || m.tie_retained[r] >= N || m.tie_constrained[r] < 0 ||
        m.tie_constrained[r] >= N || m.tie_dof[r] < 0 || m.tie_dof[r] > 2) {
      err = "fx_compile_model: equalDOF row " + std::to_string(r) + " is out of range";
      return FX_EINVAL;
    }
  if (int rc = check_equal_dof_commands(m, err)) return rc;
  out.nnode = N;
  out.nelem = E;
  // --- node and element order -------------------------------------------------------------
  if (renumber_rcm) {
    const Graph g = build_node_graph(m);
    out.node_order = reverse_cuthill_mckee(g, N);
  } else {
    out.node_order.resize(N);
    std::iota(out.node_order.begin(), out.node_order.end(), 0);
  }
  out.node_rank.assign(N, 0);
  for (int32_t i = 0; i < N; ++i) out.node_rank[out.node_order[i]] = i;
  out.elem_order.resize(E);
  std::iota(out.elem_order.begin(), out.elem_order.end(), 0);
  if (renumber_rcm) {
    std::vector<int32_t> key(E);
    for (int32_t e = 0; e < E; ++e) {
      int32_t k = INT32_MAX;
      for (int a = 0; a < 8; ++a) k = std::min(k, out.node_rank[m.conn[8 * static_cast<int64_t>(e) + a]]);
      key[e] = k;
    }
    std::stable_sort(out.elem_order.begin(), out.elem_order.end(),
                     [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  }
  // --- DOF numbering ----------------------------------------------------------------------
  const int32_t ndof = 3 * N;
  DisjointSets sets(ndof);
  for (int32_t r = 0; r < m.ntie; ++r)
    sets.unite(3 * out.node_rank[m.tie_retained[r]] + m.tie_dof[r],
               3 * out.node_rank[m.tie_constrained[r]] + m.tie_dof[r]);
  std::vector<char> fixed(ndof, 0);
  for (int32_t r = 0; r < m.nfix; ++r) fixed[sets.root(3 * out.node_rank[m.fix_node[r]] + m.fix_dof[r])] = 1;
  out.node_eq.assign(ndof, -1);
  std::vector<int32_t> eq_of_root(ndof, -1);
  int32_t neq = 0;
  for (int32_t a = 0; a < ndof; ++a) {
    const int32_t r = sets.root(a);
    if (fixed[r]) continue;
    if (eq_of_root[r] < 0) eq_of_root[r] = neq++;
    out.node_eq[a] = eq_of_root[r];
  }
  out.neq = neq;
  if (int rc = check_transformation_handler(m, out.node_rank, out.node_eq, err)) return rc;
  // --- coordinates SoA ---------------------------------------------------------------------
  out.coords.resize(static_cast<size_t>(3) * N);
  for (int32_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d)
      out.coords[static_cast<size_t>(d) * N + i] = m.xyz[3 * static_cast<int64_t>(out.node_order[i]) + d];
  // --- maps --------------------------------------------------------------------------------
  const int w = std::max(simd_width, 1);
  const int32_t P = (E + w - 1) / w * w;
  out.padded = P;
  out.node.assign(static_cast<size_t>(8) * P, -1);
  out.eq.assign(static_cast<size_t>(24) * P, -1);
  out.param.assign(P, 0);
  std::map<std::pair<int32_t, int32_t>, int32_t> param_of;
  out.param_group.clear();
  out.param_region.clear();
  for (int32_t e = 0; e < E; ++e) {
    const int32_t ge = out.elem_order[e];
    const int32_t grp = m.group ? m.group[ge] : 0;
    const int32_t reg = m.region ? m.region[ge] : -1;
    auto it = param_of.emplace(std::make_pair(grp, reg), static_cast<int32_t>(out.param_group.size()));
    if (it.second) {
      out.param_group.push_back(grp);
      out.param_region.push_back(reg);
    }
    out.param[e] = it.first->second;
    for (int a = 0; a < 8; ++a) {
      const int32_t n = out.node_rank[m.conn[8 * static_cast<int64_t>(ge) + a]];
      out.node[static_cast<size_t>(a) * P + e] = n;
      for (int d = 0; d < 3; ++d) out.eq[static_cast<size_t>(3 * a + d) * P + e] = out.node_eq[3 * n + d];
    }
  }
  out.nparam = static_cast<int32_t>(out.param_group.size());
  // --- greedy colouring ----------------------------------------------------------------------
  out.color_start.clear();
  out.color_elements.clear();
  out.ncolor = 0;
  if (build_coloring) {
    std::vector<std::vector<int32_t>> colours_at(neq);
    std::vector<int32_t> colour(E);
    std::vector<char> busy;
    int32_t ncol = 0;
    for (int32_t e = 0; e < E; ++e) {
      busy.assign(static_cast<size_t>(ncol) + 1, 0);
      for (int k = 0; k < 24; ++k) {
        const int32_t q = out.eq[static_cast<size_t>(k) * P + e];
        if (q >= 0)
          for (int32_t c : colours_at[q]) busy[c] = 1;
      }
      int32_t c = 0;
      while (busy[c]) ++c;
      colour[e] = c;
      ncol = std::max(ncol, c + 1);
      for (int k = 0; k < 24; ++k) {
        const int32_t q = out.eq[static_cast<size_t>(k) * P + e];
        if (q >= 0 && std::find(colours_at[q].begin(), colours_at[q].end(), c) == colours_at[q].end())
          colours_at[q].push_back(c);
      }
    }
    out.ncolor = ncol;
    out.color_start.assign(static_cast<size_t>(ncol) + 1, 0);
    for (int32_t e = 0; e < E; ++e) ++out.color_start[colour[e] + 1];
    for (int32_t c = 0; c < ncol; ++c) out.color_start[c + 1] += out.color_start[c];
    out.color_elements.resize(E);
    std::vector<int32_t> fill(out.color_start.begin(), out.color_start.end() - 1);
    for (int32_t e = 0; e < E; ++e) out.color_elements[fill[colour[e]]++] = e;
  }
  return FX_OK;
}

}  // namespace fx
