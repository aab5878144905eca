/*
 * fxo compile step — TEST INFRASTRUCTURE (the oracle; see fxo.h).
 *
 * Restates the reference's compile-step data that the hot path reads:
 *   node graph + RCM          src/compile/renumber.cpp:30-143 (node_graph, bfs_last_level,
 *                             pseudo_peripheral, rcm), compute_ordering :153-183
 *   DOF numbering             src/compile/dof_numbering.cpp:21-40 (union-find, smallest id as
 *                             root), :269-463 (fix/equalDOF groups, equations by root id)
 *   batch SoA maps            src/compile/batches.cpp:153-227 (padding, node/eq/param maps)
 *   greedy colouring          src/compile/batches.cpp:23-50
 * Linear constraints (rigid links, diaphragms, embedded nodes), the OpenSees handler check and
 * multi-pair batches are not part of this path and are not restated.
 */
#include <stdlib.h>
#include <string.h>

#include "fxo.h"

double fxo_lcg_symmetric(uint64_t* s) {
  /* driver_common.hpp:21-32: Knuth MMIX LCG, top 53 bits mapped to [-1, 1). */
  *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(*s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

/* ---- small growable int32 arrays ----------------------------------------------------------- */
typedef struct { int32_t* a; int64_t n, cap; } ivec;
static void iv_push(ivec* v, int32_t x) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 16;
    v->a = (int32_t*)realloc(v->a, (size_t)v->cap * sizeof(int32_t));
  }
  v->a[v->n++] = x;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* CSR node graph, neighbour lists sorted and unique (renumber.cpp:30-57). Pairs are of distinct
 * corners (:38-41), so a node listed twice in one element keeps itself as a neighbour and in its
 * degree. */
typedef struct { int32_t n; int64_t* start; int32_t* adj; } graph;

static graph node_graph(const fxo_model* m) {
  graph g;
  g.n = m->nnode;
  /* count candidate entries per node, fill, then sort + unique per row */
  int64_t* cnt = (int64_t*)calloc((size_t)m->nnode + 1, sizeof(int64_t));
  for (int32_t e = 0; e < m->nelem; ++e)
    for (int a = 0; a < 8; ++a) cnt[m->conn[8 * (int64_t)e + a] + 1] += 7;
  for (int32_t r = 0; r < m->ntie; ++r) {
    cnt[m->tie_retained[r] + 1] += 1;
    cnt[m->tie_constrained[r] + 1] += 1;
  }
  for (int32_t i = 0; i < m->nnode; ++i) cnt[i + 1] += cnt[i];
  int32_t* tmp = (int32_t*)malloc((size_t)(cnt[m->nnode] > 0 ? cnt[m->nnode] : 1) * sizeof(int32_t));
  int64_t* fill = (int64_t*)malloc((size_t)m->nnode * sizeof(int64_t) + 8);
  for (int32_t i = 0; i < m->nnode; ++i) fill[i] = cnt[i];
  for (int32_t e = 0; e < m->nelem; ++e)
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        if (a != b) {
          int32_t na = m->conn[8 * (int64_t)e + a];
          tmp[fill[na]++] = m->conn[8 * (int64_t)e + b];
        }
  for (int32_t r = 0; r < m->ntie; ++r) {
    int32_t a = m->tie_retained[r], b = m->tie_constrained[r];
    tmp[fill[a]++] = b;
    tmp[fill[b]++] = a;
  }
  g.start = (int64_t*)malloc(((size_t)m->nnode + 1) * sizeof(int64_t));
  g.start[0] = 0;
  int64_t w = 0;
  for (int32_t i = 0; i < m->nnode; ++i) {
    int64_t b = cnt[i], e = cnt[i + 1];
    qsort(tmp + b, (size_t)(e - b), sizeof(int32_t), cmp_i32);
    int64_t r0 = w;
    for (int64_t k = b; k < e; ++k) {
      int32_t v = tmp[k];
      if (w > r0 && tmp[w - 1] == v) continue;
      tmp[w++] = v;
    }
    g.start[i + 1] = w;
  }
  g.adj = (int32_t*)realloc(tmp, (size_t)(w > 0 ? w : 1) * sizeof(int32_t));
  free(cnt);
  free(fill);
  return g;
}

static inline int32_t deg(const graph* g, int32_t v) { return (int32_t)(g->start[v + 1] - g->start[v]); }
/* (degree, index) order used for roots and neighbour sorting (renumber.cpp:91-93, 135-137) */
static inline int deg_less(const graph* g, int32_t a, int32_t b) {
  int32_t da = deg(g, a), db = deg(g, b);
  return da != db ? da < db : a < b;
}

static const graph* g_sort_graph;
static int cmp_deg(const void* pa, const void* pb) {
  int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
  if (a == b) return 0;
  return deg_less(g_sort_graph, a, b) ? -1 : 1;
}

/* Breadth-first levels from root over nodes not yet ordered: returns the last level in
 * `last` (count *nlast) and the depth (renumber.cpp:59-79). */
static int bfs_depth(const graph* g, int32_t root, const char* done, char* seen, int32_t* q,
                     int32_t** last, int32_t* nlast) {
  /* seen: scratch of size n, must equal `done` on entry for the reachable set; restored */
  int64_t head = 0, tail = 0;
  q[tail++] = root;
  seen[root] = 1;
  int depth = 0;
  int64_t lvl_begin = 0, lvl_end = 1;
  int64_t last_b = 0, last_e = 1;
  while (lvl_begin < lvl_end) {
    last_b = lvl_begin;
    last_e = lvl_end;
    for (head = lvl_begin; head < lvl_end; ++head) {
      int32_t v = q[head];
      for (int64_t k = g->start[v]; k < g->start[v + 1]; ++k) {
        int32_t w = g->adj[k];
        if (!seen[w]) {
          seen[w] = 1;
          q[tail++] = w;
        }
      }
    }
    if (tail > lvl_end) ++depth;
    lvl_begin = lvl_end;
    lvl_end = tail;
  }
  for (int64_t i = 0; i < tail; ++i) seen[q[i]] = done[q[i]];
  *last = q + last_b;
  *nlast = (int32_t)(last_e - last_b);
  return depth;
}

static int32_t* rcm(const graph* g) {
  const int32_t n = g->n;
  char* done = (char*)calloc((size_t)n + 1, 1);
  char* seen = (char*)calloc((size_t)n + 1, 1);
  int32_t* q = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
  int32_t* order = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
  int32_t* next = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
  int64_t norder = 0;
  for (int32_t s = 0; s < n; ++s) {
    if (done[s]) continue;
    /* smallest-degree node of the component (ties: lowest index), renumber.cpp:105-121 */
    int32_t best = s;
    {
      int64_t t = 0;
      q[t++] = s;
      seen[s] = 1;
      for (int64_t h = 0; h < t; ++h) {
        int32_t v = q[h];
        if (deg_less(g, v, best)) best = v;
        for (int64_t k = g->start[v]; k < g->start[v + 1]; ++k)
          if (!seen[g->adj[k]]) {
            seen[g->adj[k]] = 1;
            q[t++] = g->adj[k];
          }
      }
      for (int64_t h = 0; h < t; ++h) seen[q[h]] = done[q[h]];
    }
    /* George-Liu pseudo-peripheral root, at most 32 sweeps (renumber.cpp:83-96) */
    int32_t root = best;
    int depth = -1;
    for (int it = 0; it < 32; ++it) {
      int32_t* last;
      int32_t nlast;
      int d = bfs_depth(g, root, done, seen, q, &last, &nlast);
      if (d <= depth) break;
      depth = d;
      int32_t r = last[0];
      for (int32_t i = 1; i < nlast; ++i)
        if (deg_less(g, last[i], r)) r = last[i];
      root = r;
    }
    /* Cuthill-McKee from root, neighbours by (degree, index); reversed per component */
    int64_t first = norder;
    order[norder++] = root;
    done[root] = 1;
    for (int64_t head = first; head < norder; ++head) {
      int32_t v = order[head];
      int32_t nn = 0;
      for (int64_t k = g->start[v]; k < g->start[v + 1]; ++k)
        if (!done[g->adj[k]]) {
          done[g->adj[k]] = 1;
          next[nn++] = g->adj[k];
        }
      g_sort_graph = g;
      qsort(next, (size_t)nn, sizeof(int32_t), cmp_deg);
      for (int32_t i = 0; i < nn; ++i) order[norder++] = next[i];
    }
    for (int64_t i = first, j = norder - 1; i < j; ++i, --j) {
      int32_t t = order[i];
      order[i] = order[j];
      order[j] = t;
    }
    for (int64_t i = first; i < norder; ++i) seen[order[i]] = 1;
  }
  free(done);
  free(seen);
  free(q);
  free(next);
  return order;
}

/* element order: by smallest internal node index, ties by file position (stable) */
static const int32_t* g_key;
static int cmp_key(const void* pa, const void* pb) {
  int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
  if (g_key[a] != g_key[b]) return g_key[a] < g_key[b] ? -1 : 1;
  return (a > b) - (a < b); /* stable: file position */
}

/* union-find with the smallest id as root (dof_numbering.cpp:21-40) */
static int32_t uf_find(int32_t* parent, int32_t a) {
  while (parent[a] != a) {
    parent[a] = parent[parent[a]];
    a = parent[a];
  }
  return a;
}
static void uf_join(int32_t* parent, int32_t a, int32_t b) {
  a = uf_find(parent, a);
  b = uf_find(parent, b);
  if (a == b) return;
  if (b < a) {
    int32_t t = a;
    a = b;
    b = t;
  }
  parent[b] = a;
}

/* ---- Constraint refusals (dof_numbering.cpp:44-262, 303-339) ------------------------------
 * Rule 1 (:303-339): one equalDOF command per constrained node; a command is the row's
 * command id, or without ids the (retained, constrained) pair.
 * Rule 2 (:61-150, 241-262): emulate OpenSees' TransformationConstraintHandler (nodes in
 * ascending tag order, tag = model index + 1) and compare with the union-find numbering: the
 * fixed sets must agree and the handler's values and the equations be one to one. On failure
 * the pattern is classified as refusal() does (:162-237). */
#define FXO_MAXID 8
static int classify_refusal(const fxo_model* m) {
  /* a fix on a tied DOF whose retained DOF is free (:189-206) */
  for (int32_t r = 0; r < m->nfix; ++r) {
    const int32_t C = m->fix_node[r], d = m->fix_dof[r];
    int32_t R = -1, tied = 0;
    for (int32_t t = 0; t < m->ntie; ++t)
      if (m->tie_constrained[t] == C) {
        R = m->tie_retained[t];
        if (m->tie_dof[t] == d) tied = 1;
      }
    if (!tied) continue;
    int rfixed = 0;
    for (int32_t q = 0; q < m->nfix; ++q) rfixed |= m->fix_node[q] == R && m->fix_dof[q] == d;
    if (!rfixed) return FXO_REFUSE_FIX_LOST;
  }
  /* chains C -> R -> M (:207-233): partial if the DOF lists differ, else refused if R > C */
  for (int32_t t = 0; t < m->ntie; ++t) {
    const int32_t C = m->tie_constrained[t];
    int32_t R = -1;
    for (int32_t u = 0; u < m->ntie; ++u)
      if (m->tie_constrained[u] == C) R = m->tie_retained[u];
    int chained = 0;
    for (int32_t u = 0; u < m->ntie; ++u) chained |= m->tie_constrained[u] == R;
    if (!chained) continue;
    /* DOF lists in row order */
    int32_t dc[FXO_MAXID], dr[FXO_MAXID], nc = 0, nr = 0;
    for (int32_t u = 0; u < m->ntie; ++u) {
      if (m->tie_constrained[u] == C && nc < FXO_MAXID) dc[nc++] = m->tie_dof[u];
      if (m->tie_constrained[u] == R && nr < FXO_MAXID) dr[nr++] = m->tie_dof[u];
    }
    int same = nc == nr;
    for (int32_t k = 0; same && k < nc; ++k) same = dc[k] == dr[k];
    if (!same) return FXO_REFUSE_PARTIAL_CHAIN;
    if (R > C) return FXO_REFUSE_LARGER_TAG;
  }
  return FXO_REFUSE_OTHER;
}

static int check_commands(const fxo_model* m) {
  for (int32_t r = 0; r < m->ntie; ++r) {
    const int64_t id = m->tie_command ? (int64_t)m->tie_command[r] : -1 - (int64_t)m->tie_retained[r];
    for (int32_t q = 0; q < r; ++q) {
      if (m->tie_constrained[q] != m->tie_constrained[r]) continue;
      const int64_t iq = m->tie_command ? (int64_t)m->tie_command[q] : -1 - (int64_t)m->tie_retained[q];
      if (iq != id) return FXO_REFUSE_COMMANDS;
    }
  }
  return 0;
}

static int check_handler(const fxo_model* m, const int32_t* rank, const int32_t* node_eq) {
  const int32_t N = m->nnode;
  int32_t* retained = (int32_t*)malloc(((size_t)N + 1) * sizeof(int32_t));
  int32_t* nd = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));        /* MP DOF count */
  int32_t* dofs = (int32_t*)malloc(((size_t)N * FXO_MAXID + 1) * sizeof(int32_t));
  int32_t* nid = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  int64_t* id = (int64_t*)malloc(((size_t)N * FXO_MAXID + 1) * sizeof(int64_t));
  int32_t* nunc = (int32_t*)calloc((size_t)N + 1, sizeof(int32_t));
  int32_t* unc = (int32_t*)malloc(((size_t)N * 3 + 1) * sizeof(int32_t));
  for (int32_t i = 0; i < N; ++i) retained[i] = -1;
  for (int32_t r = 0; r < m->ntie; ++r) {
    const int32_t c = m->tie_constrained[r];
    retained[c] = m->tie_retained[r];
    if (nd[c] < FXO_MAXID - 3) dofs[(size_t)c * FXO_MAXID + nd[c]++] = m->tie_dof[r];
  }
  for (int32_t i = 0; i < N; ++i) {
    int64_t* I = id + (size_t)i * FXO_MAXID;
    const int32_t* D = dofs + (size_t)i * FXO_MAXID;
    if (retained[i] < 0) {
      nid[i] = 3;
      for (int d = 0; d < 3; ++d) I[d] = 3 * (int64_t)i + d;
      for (int32_t r = 0; r < m->nfix; ++r)
        if (m->fix_node[r] == i) I[m->fix_dof[r]] = -1;
      continue;
    }
    for (int d = 0; d < 3; ++d) {
      int in = 0;
      for (int32_t k = 0; k < nd[i]; ++k) in |= D[k] == d;
      if (!in) unc[3 * (size_t)i + nunc[i]++] = d;
    }
    nid[i] = nunc[i] + nd[i];
    for (int32_t p = 0; p < nid[i]; ++p) I[p] = -1;
    for (int32_t p = 0; p < nunc[i]; ++p) I[p] = 3 * (int64_t)i + unc[3 * (size_t)i + p];
    for (int32_t r = 0; r < m->nfix; ++r) {
      if (m->fix_node[r] != i) continue;
      int32_t loc = 0;
      for (int32_t d = 0; d < m->fix_dof[r]; ++d) {
        int in = 0;
        for (int32_t k = 0; k < nd[i]; ++k) in |= D[k] == d;
        loc += !in;
      }
      if (loc < nid[i]) I[loc] = -1;
    }
  }
  for (int32_t i = 0; i < N; ++i) { /* doneID in ascending tag order */
    if (retained[i] < 0) continue;
    const int64_t* O = id + (size_t)retained[i] * FXO_MAXID;
    for (int32_t k = 0; k < nd[i]; ++k) {
      const int32_t pos = dofs[(size_t)i * FXO_MAXID + k];
      id[(size_t)i * FXO_MAXID + nunc[i] + k] = pos < nid[retained[i]] ? O[pos] : -1;
    }
  }
  /* one-to-one check through dense maps: value (a DOF id or -1) <-> equation */
  const int64_t ndof = 3 * (int64_t)N;
  int32_t* eq_of_value = (int32_t*)malloc(((size_t)ndof + 1) * sizeof(int32_t));
  int64_t* value_of_eq = (int64_t*)malloc(((size_t)ndof + 1) * sizeof(int64_t));
  for (int64_t a = 0; a < ndof; ++a) {
    eq_of_value[a] = -1;
    value_of_eq[a] = -2;
  }
  int bad = 0;
  for (int32_t i = 0; i < N && !bad; ++i)
    for (int d = 0; d < 3 && !bad; ++d) {
      int64_t v;
      const int32_t* D = dofs + (size_t)i * FXO_MAXID;
      const int64_t* I = id + (size_t)i * FXO_MAXID;
      if (retained[i] < 0) {
        v = I[d];
      } else {
        int32_t k = 0;
        while (k < nd[i] && D[k] != d) ++k;
        if (k == nd[i]) {
          int32_t p = 0;
          while (unc[3 * (size_t)i + p] != d) ++p;
          v = I[p];
        } else {
          v = I[nunc[i] + k];
        }
      }
      const int32_t e = node_eq[3 * (size_t)rank[i] + d];
      if ((v < 0) != (e < 0)) {
        bad = 1;
      } else if (e >= 0) {
        if (eq_of_value[v] < 0) eq_of_value[v] = e;
        if (value_of_eq[e] == -2) value_of_eq[e] = v;
        bad = eq_of_value[v] != e || value_of_eq[e] != v;
      }
    }
  free(retained);
  free(nd);
  free(dofs);
  free(nid);
  free(id);
  free(nunc);
  free(unc);
  free(eq_of_value);
  free(value_of_eq);
  return bad ? classify_refusal(m) : 0;
}

int fxo_compile(const fxo_model* m, int renumber_rcm, int simd_width, int build_coloring,
                fxo_compiled* c) {
  memset(c, 0, sizeof(*c));
  const int32_t N = m->nnode, E = m->nelem;
  if (check_commands(m)) return FXO_REFUSE_COMMANDS;
  c->nnode = N;
  c->nelem = E;
  /* --- ordering ------------------------------------------------------------------------ */
  if (renumber_rcm) {
    graph g = node_graph(m);
    c->node_order = rcm(&g);
    free(g.start);
    free(g.adj);
  } else {
    c->node_order = (int32_t*)malloc(((size_t)N + 1) * sizeof(int32_t));
    for (int32_t i = 0; i < N; ++i) c->node_order[i] = i;
  }
  c->node_rank = (int32_t*)malloc(((size_t)N + 1) * sizeof(int32_t));
  for (int32_t i = 0; i < N; ++i) c->node_rank[c->node_order[i]] = i;
  c->elem_order = (int32_t*)malloc(((size_t)E + 1) * sizeof(int32_t));
  for (int32_t e = 0; e < E; ++e) c->elem_order[e] = e;
  if (renumber_rcm) {
    int32_t* key = (int32_t*)malloc(((size_t)E + 1) * sizeof(int32_t));
    for (int32_t e = 0; e < E; ++e) {
      int32_t k = INT32_MAX;
      for (int a = 0; a < 8; ++a) {
        int32_t r = c->node_rank[m->conn[8 * (int64_t)e + a]];
        if (r < k) k = r;
      }
      key[e] = k;
    }
    g_key = key;
    qsort(c->elem_order, (size_t)E, sizeof(int32_t), cmp_key);
    free(key);
  }
  /* --- DOF numbering (3 DOFs per node: dof id = 3 * internal + d) ------------------------ */
  const int32_t ndof = 3 * N;
  int32_t* parent = (int32_t*)malloc(((size_t)ndof + 1) * sizeof(int32_t));
  for (int32_t a = 0; a < ndof; ++a) parent[a] = a;
  for (int32_t r = 0; r < m->ntie; ++r)
    uf_join(parent, 3 * c->node_rank[m->tie_retained[r]] + m->tie_dof[r],
            3 * c->node_rank[m->tie_constrained[r]] + m->tie_dof[r]);
  char* fixed_root = (char*)calloc((size_t)ndof + 1, 1);
  for (int32_t r = 0; r < m->nfix; ++r)
    fixed_root[uf_find(parent, 3 * c->node_rank[m->fix_node[r]] + m->fix_dof[r])] = 1;
  c->node_eq = (int32_t*)malloc(((size_t)ndof + 1) * sizeof(int32_t));
  int32_t* eq_of_root = (int32_t*)malloc(((size_t)ndof + 1) * sizeof(int32_t));
  for (int32_t a = 0; a < ndof; ++a) eq_of_root[a] = -1;
  int32_t next = 0;
  for (int32_t a = 0; a < ndof; ++a) {
    int32_t root = uf_find(parent, a);
    c->node_eq[a] = -1;
    if (fixed_root[root]) continue;
    if (eq_of_root[root] < 0) eq_of_root[root] = next++;
    c->node_eq[a] = eq_of_root[root];
  }
  c->neq = next;
  free(parent);
  free(fixed_root);
  free(eq_of_root);
  {
    const int refused = check_handler(m, c->node_rank, c->node_eq);
    if (refused) {
      fxo_compiled_free(c);
      return refused;
    }
  }
  /* --- coordinates, SoA by internal node (batches.cpp:66-70) ----------------------------- */
  c->coords = (double*)malloc(((size_t)3 * N + 1) * sizeof(double));
  for (int32_t i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k) c->coords[(int64_t)k * N + i] = m->xyz[3 * (int64_t)c->node_order[i] + k];
  /* --- batch maps (one stdBrick batch), batches.cpp:153-203 ------------------------------ */
  int w = simd_width < 1 ? 1 : simd_width;
  const int32_t P = (E + w - 1) / w * w;
  c->padded = P;
  c->node = (int32_t*)malloc(((size_t)8 * P + 1) * sizeof(int32_t));
  c->eq = (int32_t*)malloc(((size_t)24 * P + 1) * sizeof(int32_t));
  c->param = (int32_t*)calloc((size_t)P + 1, sizeof(int32_t));
  for (int64_t i = 0; i < (int64_t)8 * P; ++i) c->node[i] = -1;
  for (int64_t i = 0; i < (int64_t)24 * P; ++i) c->eq[i] = -1;
  /* param sets keyed by (group, region) in first-appearance order */
  ivec pg = {0}, pr = {0};
  for (int32_t e = 0; e < E; ++e) {
    const int32_t ge = c->elem_order[e];
    const int32_t grp = m->group ? m->group[ge] : 0;
    const int32_t reg = m->region ? m->region[ge] : -1;
    int32_t pi = -1;
    for (int64_t k = 0; k < pg.n; ++k)
      if (pg.a[k] == grp && pr.a[k] == reg) {
        pi = (int32_t)k;
        break;
      }
    if (pi < 0) {
      pi = (int32_t)pg.n;
      iv_push(&pg, grp);
      iv_push(&pr, reg);
    }
    c->param[e] = pi;
    for (int a = 0; a < 8; ++a) {
      const int32_t node = c->node_rank[m->conn[8 * (int64_t)ge + a]];
      c->node[(int64_t)a * P + e] = node;
      for (int d = 0; d < 3; ++d) c->eq[(int64_t)(3 * a + d) * P + e] = c->node_eq[3 * node + d];
    }
  }
  c->nparam = (int32_t)pg.n;
  c->param_group = pg.a ? pg.a : (int32_t*)calloc(1, sizeof(int32_t));
  c->param_region = pr.a ? pr.a : (int32_t*)calloc(1, sizeof(int32_t));
  /* --- greedy colouring in batch order (batches.cpp:23-50) ------------------------------- */
  if (build_coloring) {
    int32_t* ecol = (int32_t*)malloc(((size_t)E + 1) * sizeof(int32_t));
    ivec* used = (ivec*)calloc((size_t)c->neq + 1, sizeof(ivec));
    int32_t ncol = 0;
    char* taken = (char*)calloc(1024, 1);
    int32_t taken_cap = 1024;
    for (int32_t e = 0; e < E; ++e) {
      if (ncol + 1 > taken_cap) {
        taken = (char*)realloc(taken, (size_t)taken_cap * 2);
        taken_cap *= 2;
      }
      memset(taken, 0, (size_t)ncol + 1);
      for (int k = 0; k < 24; ++k) {
        int32_t qq = c->eq[(int64_t)k * P + e];
        if (qq < 0) continue;
        for (int64_t t = 0; t < used[qq].n; ++t) taken[used[qq].a[t]] = 1;
      }
      int32_t col = 0;
      while (taken[col]) ++col;
      ecol[e] = col;
      if (col + 1 > ncol) ncol = col + 1;
      for (int k = 0; k < 24; ++k) {
        int32_t qq = c->eq[(int64_t)k * P + e];
        if (qq < 0) continue;
        int found = 0;
        for (int64_t t = 0; t < used[qq].n; ++t)
          if (used[qq].a[t] == col) {
            found = 1;
            break;
          }
        if (!found) iv_push(&used[qq], col);
      }
    }
    c->ncolor = ncol;
    c->color_start = (int32_t*)calloc((size_t)ncol + 2, sizeof(int32_t));
    for (int32_t e = 0; e < E; ++e) ++c->color_start[ecol[e] + 1];
    for (int32_t k = 0; k < ncol; ++k) c->color_start[k + 1] += c->color_start[k];
    c->color_elements = (int32_t*)malloc(((size_t)E + 1) * sizeof(int32_t));
    int32_t* fillc = (int32_t*)malloc(((size_t)ncol + 1) * sizeof(int32_t));
    for (int32_t k = 0; k < ncol; ++k) fillc[k] = c->color_start[k];
    for (int32_t e = 0; e < E; ++e) c->color_elements[fillc[ecol[e]]++] = e;
    for (int32_t qq = 0; qq < c->neq; ++qq) free(used[qq].a);
    free(used);
    free(taken);
    free(fillc);
    free(ecol);
  }
  /* --- gather pairs: each equation's slots in (element, DOF) order ------------------------ */
  c->gather_start = (int64_t*)calloc((size_t)c->neq + 2, sizeof(int64_t));
  for (int32_t e = 0; e < E; ++e)
    for (int k = 0; k < 24; ++k) {
      const int32_t qq = c->eq[(int64_t)k * P + e];
      if (qq >= 0) ++c->gather_start[qq + 1];
    }
  for (int32_t qq = 0; qq < c->neq; ++qq) c->gather_start[qq + 1] += c->gather_start[qq];
  c->gather_list = (int64_t*)malloc(((size_t)c->gather_start[c->neq] + 1) * sizeof(int64_t));
  int64_t* gfill = (int64_t*)malloc(((size_t)c->neq + 1) * sizeof(int64_t));
  for (int32_t qq = 0; qq < c->neq; ++qq) gfill[qq] = c->gather_start[qq];
  for (int32_t e = 0; e < E; ++e)
    for (int k = 0; k < 24; ++k) {
      const int32_t qq = c->eq[(int64_t)k * P + e];
      if (qq >= 0) c->gather_list[gfill[qq]++] = (int64_t)k * P + e;
    }
  free(gfill);
  c->slots = (double*)calloc((size_t)24 * P + 1, sizeof(double));
  return 0;
}

void fxo_compiled_free(fxo_compiled* c) {
  free(c->node_order);
  free(c->node_rank);
  free(c->elem_order);
  free(c->node_eq);
  free(c->coords);
  free(c->node);
  free(c->eq);
  free(c->param);
  free(c->param_group);
  free(c->param_region);
  free(c->color_start);
  free(c->color_elements);
  free(c->gather_start);
  free(c->gather_list);
  free(c->slots);
  memset(c, 0, sizeof(*c));
}

int32_t fxo_node_bandwidth(const fxo_model* m, const int32_t* rank) {
  int32_t bw = 0;
  for (int32_t e = 0; e < m->nelem; ++e)
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        int32_t d = rank[m->conn[8 * (int64_t)e + a]] - rank[m->conn[8 * (int64_t)e + b]];
        if (d < 0) d = -d;
        if (d > bw) bw = d;
      }
  return bw;
}
